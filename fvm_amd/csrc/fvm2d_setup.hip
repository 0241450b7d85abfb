// Host-side geometry of one 2D case in the reference's operations and order (cylinder/
// cylinder_v3.0.3_AF_original 4step_face_nl/: GRID grid.f90:47-115, METRIC metric.f90:18-168,
// the angles and wall areas of INIT init.f90:71-129, the wall-cell geometry of CALCCOEF
// calccoef.f90:46-79) and the line factors of the two AF stages (iter_u.f90:12-66, 68-119 with
// the folds of bcu.f90:10-38).  Compiled with -ffp-contract=off, so every product and sum
// rounds as the reference's does.
#include <math.h>

#include <vector>

#include "fvm2d.h"
#include "fvm2d_host.h"

namespace {
inline long ix(int sj, int i, int j) { return (long)i + (long)sj * j; }
}

int geo2d_build(int ni, int nj, const double* xg_in, const double* yg_in, Geo2D& g) {
  if (ni < 4 || nj < 4) return FVM_ERR_ARG;
  g.ni = ni;
  g.nj = nj;
  const int sj = ni + 1;
  const long n = (long)(ni + 1) * (nj + 1);
  const double pi = 4.0 * atan(1.0);
  std::vector<double> XG(n, 0.0), YG(n, 0.0);
  for (int j = 1; j <= nj; ++j)
    for (int i = 1; i <= ni; ++i) {
      XG[ix(sj, i, j)] = xg_in[(i - 1) + (long)ni * (j - 1)];
      YG[ix(sj, i, j)] = yg_in[(i - 1) + (long)ni * (j - 1)];
    }
  auto& X = g.x;
  auto& Y = g.y;
  X.assign(n, 0.0);
  Y.assign(n, 0.0);
  // cell centres (grid.f90:58-64)
  for (int j = 1; j <= nj - 1; ++j)
    for (int i = 1; i <= ni - 1; ++i) {
      const long c = ix(sj, i, j);
      X[c] = 0.25 * (((XG[ix(sj, i + 1, j + 1)] + XG[ix(sj, i + 1, j)]) + XG[ix(sj, i, j + 1)]) +
                     XG[c]);
      Y[c] = 0.25 * (((YG[ix(sj, i + 1, j + 1)] + YG[ix(sj, i + 1, j)]) + YG[ix(sj, i, j + 1)]) +
                     YG[c]);
    }
  // branch-cut ghost centres (grid.f90:66-73)
  for (int i = 1; i <= ni - 1; ++i) {
    X[ix(sj, i, 0)] = X[ix(sj, i, nj - 1)];
    Y[ix(sj, i, 0)] = Y[ix(sj, i, nj - 1)];
    X[ix(sj, i, nj)] = X[ix(sj, i, 1)];
    Y[ix(sj, i, nj)] = Y[ix(sj, i, 1)];
  }
  // mirrored wall and far-field ghost centres (grid.f90:75-91)
  for (int j = 1; j <= nj - 1; ++j) {
    long c0 = ix(sj, 0, j);
    X[c0] = (XG[ix(sj, 1, j + 1)] + XG[ix(sj, 1, j)]) * 0.5;
    Y[c0] = (YG[ix(sj, 1, j + 1)] + YG[ix(sj, 1, j)]) * 0.5;
    X[c0] = X[c0] - (X[ix(sj, 1, j)] - X[c0]);
    Y[c0] = Y[c0] - (Y[ix(sj, 1, j)] - Y[c0]);
    long cn = ix(sj, ni, j);
    X[cn] = (XG[ix(sj, ni, j + 1)] + XG[ix(sj, ni, j)]) * 0.5;
    Y[cn] = (YG[ix(sj, ni, j + 1)] + YG[ix(sj, ni, j)]) * 0.5;
    X[cn] = X[cn] + (X[cn] - X[ix(sj, ni - 1, j)]);
    Y[cn] = Y[cn] + (Y[cn] - Y[ix(sj, ni - 1, j)]);
  }
  // corners (grid.f90:94-113)
  for (int i : {0, ni}) {
    X[ix(sj, i, 0)] = X[ix(sj, i, nj - 1)];
    Y[ix(sj, i, 0)] = Y[ix(sj, i, nj - 1)];
    X[ix(sj, i, nj)] = X[ix(sj, i, 1)];
    Y[ix(sj, i, nj)] = Y[ix(sj, i, 1)];
  }
  // coordinate derivatives (metric.f90:18-124)
  std::vector<double> xxi(n, 0.0), yxi(n, 0.0), xet(n, 0.0), yet(n, 0.0);
  for (int j = 1; j <= nj - 1; ++j)
    for (int i = 1; i <= ni - 1; ++i) {
      const long c = ix(sj, i, j);
      xxi[c] = (X[c + 1] - X[c - 1]) / 2.0;
      xet[c] = (X[c + sj] - X[c - sj]) / 2.0;
      yxi[c] = (Y[c + 1] - Y[c - 1]) / 2.0;
      yet[c] = (Y[c + sj] - Y[c - sj]) / 2.0;
    }
  for (int i = 1; i <= ni - 1; ++i) {  // south (J = 0) and north (J = NJ) ghost rows
    long c = ix(sj, i, 0);
    xxi[c] = (X[c + 1] - X[c - 1]) / 2.0;
    xet[c] = (X[ix(sj, i, nj)] - X[ix(sj, i, nj - 2)]) / 2.0;
    yxi[c] = (Y[c + 1] - Y[c - 1]) / 2.0;
    yet[c] = (Y[ix(sj, i, nj)] - Y[ix(sj, i, nj - 2)]) / 2.0;
    c = ix(sj, i, nj);
    xxi[c] = (X[c + 1] - X[c - 1]) / 2.0;
    xet[c] = (X[ix(sj, i, 2)] - X[ix(sj, i, 0)]) / 2.0;
    yxi[c] = (Y[c + 1] - Y[c - 1]) / 2.0;
    yet[c] = (Y[ix(sj, i, 2)] - Y[ix(sj, i, 0)]) / 2.0;
  }
  for (int j = 1; j <= nj - 1; ++j) {  // west (I = 0) and east (I = NI) ghost columns
    long c = ix(sj, 0, j);
    xxi[c] = ((-3.0 * X[c] + 4.0 * X[c + 1]) - X[c + 2]) / 2.0;
    xet[c] = (X[c + sj] - X[c - sj]) / 2.0;
    yxi[c] = ((-3.0 * Y[c] + 4.0 * Y[c + 1]) - Y[c + 2]) / 2.0;
    yet[c] = (Y[c + sj] - Y[c - sj]) / 2.0;
    c = ix(sj, ni, j);
    xxi[c] = ((3.0 * X[c] - 4.0 * X[c - 1]) + X[c - 2]) / 2.0;
    xet[c] = (X[c + sj] - X[c - sj]) / 2.0;
    yxi[c] = ((3.0 * Y[c] - 4.0 * Y[c - 1]) + Y[c - 2]) / 2.0;
    yet[c] = (Y[c + sj] - Y[c - sj]) / 2.0;
  }
  for (int i : {0, ni})
    for (int j : {0, nj}) {  // corner lines (metric.f90:76-115)
      const long c = ix(sj, i, j);
      if (i == 0) {
        xxi[c] = ((-3.0 * X[c] + 4.0 * X[c + 1]) - X[c + 2]) / 2.0;
        yxi[c] = ((-3.0 * Y[c] + 4.0 * Y[c + 1]) - Y[c + 2]) / 2.0;
      } else {
        xxi[c] = ((3.0 * X[c] - 4.0 * X[c - 1]) + X[c - 2]) / 2.0;
        yxi[c] = ((3.0 * Y[c] - 4.0 * Y[c - 1]) + Y[c - 2]) / 2.0;
      }
      if (j == 0) {
        xet[c] = (X[ix(sj, i, nj)] - X[ix(sj, i, nj - 2)]) / 2.0;
        yet[c] = (Y[ix(sj, i, nj)] - Y[ix(sj, i, nj - 2)]) / 2.0;
      } else {
        xet[c] = (X[ix(sj, i, 2)] - X[ix(sj, i, 0)]) / 2.0;
        yet[c] = (Y[ix(sj, i, 2)] - Y[ix(sj, i, 0)]) / 2.0;
      }
    }
  g.djac.assign(n, 0.0);
  g.xix.assign(n, 0.0);
  g.xiy.assign(n, 0.0);
  g.etx.assign(n, 0.0);
  g.ety.assign(n, 0.0);
  std::vector<double> q11(n), q12(n), q22(n);
  for (long c = 0; c < n; ++c) {  // metric.f90:117-145
    g.djac[c] = xxi[c] * yet[c] - yxi[c] * xet[c];
    g.xix[c] = yet[c] / g.djac[c];
    g.xiy[c] = -xet[c] / g.djac[c];
    g.etx[c] = -yxi[c] / g.djac[c];
    g.ety[c] = xxi[c] / g.djac[c];
    q11[c] = g.xix[c] * g.xix[c] + g.xiy[c] * g.xiy[c];
    q12[c] = g.xix[c] * g.etx[c] + g.xiy[c] * g.ety[c];
    q22[c] = g.etx[c] * g.etx[c] + g.ety[c] * g.ety[c];
  }
  // DIF face coefficients (metric.f90:156-165), extended by one face so DW(i) = DE(i-1) and
  // DS(j) = DN(j-1) cover the first interior cell (the same sums, operands swapped)
  g.de.assign(n, 0.0);
  g.dn.assign(n, 0.0);
  g.jq12.assign(n, 0.0);
  for (int j = 0; j <= nj; ++j)
    for (int i = 0; i <= ni; ++i) {
      const long c = ix(sj, i, j);
      g.jq12[c] = g.djac[c] * q12[c];
      if (i < ni) g.de[c] = 0.5 * (g.djac[c] * q11[c] + g.djac[c + 1] * q11[c + 1]);
      if (j < nj) g.dn[c] = 0.5 * (g.djac[c] * q22[c] + g.djac[c + sj] * q22[c + sj]);
    }
  // angles (init.f90:71-99)
  g.theta.assign(n, 0.0);
  for (int j = 0; j <= nj; ++j)
    for (int i = 0; i <= ni; ++i) {
      const long c = ix(sj, i, j);
      const double rad = pow(X[c] * X[c] + Y[c] * Y[c], 0.5);
      double th = atan(Y[c] / X[c]);
      const double as = asin(Y[c] / rad);
      if (th >= 0.0 && as > 0.0) {
      } else if (th < 0.0 && as > 0.0) th = pi + th;
      else if (th >= 0.0 && as <= 0.0) th = pi + th;
      else if (th < 0.0 && as < 0.0) th = 2.0 * pi + th;
      g.theta[c] = th;
    }
  // wall-cell geometry of CALCCOEF (calccoef.f90:46-79) and the wall areas of INIT
  // (init.f90:122-129)
  g.wall.assign(6L * (nj + 1), 0.0);
  for (int j = 1; j <= nj - 1; ++j) {
    const long cc = ix(sj, 1, j);
    const double Xc = X[cc], Yc = Y[cc];
    const double xa = XG[ix(sj, 1, j)], ya = YG[ix(sj, 1, j)];
    const double xb = XG[ix(sj, 1, j + 1)], yb = YG[ix(sj, 1, j + 1)];
    const double l1 = pow(pow(Xc - xa, 2.0) + pow(Yc - ya, 2.0), 0.5);
    const double l2 = pow(pow(Xc - xb, 2.0) + pow(Yc - yb, 2.0), 0.5);
    const double l3 = pow(pow(xb - xa, 2.0) + pow(yb - ya, 2.0), 0.5);
    const double th1 = acos((l1 * l1 + l2 * l2 - l3 * l3) / (2.0 * l1 * l2));
    const double dn1 = l1 * l2 / l3 * sin(th1);
    const double la = (yb - ya) / (xb - xa);
    const double lb = ya - la * xa;
    const double la_ = -1.0 / la;
    const double lb_ = Yc - la_ * Xc;
    const double x1 = (lb_ - lb) / (la - la_);
    const double y1 = la * (lb_ - lb) / (la - la_) + lb;
    const double dn2 = pow(pow(Yc - y1, 2.0) + pow(Xc - x1, 2.0), 0.5);
    double th2 = atan((Yc - y1) / (Xc - x1));
    const double as = asin((Yc - y1) / dn2);
    if (th2 >= 0.0 && as > 0.0) {
    } else if (th2 < 0.0 && as > 0.0) th2 = pi + th2;
    else if (th2 >= 0.0 && as <= 0.0) th2 = pi + th2;
    else if (th2 < 0.0 && as < 0.0) th2 = 2.0 * pi + th2;
    double* o = g.wall.data() + 6L * j;
    o[0] = dn1;
    o[1] = sin(th2);
    o[2] = cos(th2);
    o[3] = l3;  // AREA(1, J)
    o[4] = cos(g.theta[cc]);
    o[5] = sin(g.theta[cc]);
  }
  return FVM_OK;
}

// Coefficients and line factors of the two AF stages for DTIME dt (iter_u.f90:12-66, 68-119;
// identical for U and V, iter_v.f90): stage s uses (XIX, ETX) for s = 0 and (XIY, ETY) for
// s = 1; the wall fold AP += AW, AW = 0 at I = 1 and the far-field fold AP += AE, AE = 0 at
// I = NI-1 (bcu.f90:14-17, 33-35).  Stage 0 sweeps I lines: A = -AW, D = AP - A B(i-1),
// B = -AE / D; stage 1 sweeps J lines: A = -AS, D = AP - A B(j-1), B = -AN / D with B(0) = 0.
// Stored: 1/D, -A/D, B; the lagged couplings (stage 0: AN; stage 1: AE, AW after the folds);
// the pre-fold AW(1), AE(NI-1) per J for the right-hand-side folds; 2 Re DJAC / dt; the cross
// weights (DJAC XIX) ETX, (DJAC XIY) ETY.
void geo2d_af(const Geo2D& g, double re, double dt, AF2D& a) {
  const int ni = g.ni, nj = g.nj, sj = ni + 1;
  const long n = (long)(ni + 1) * (nj + 1);
  a.ap0.assign(n, 0.0);
  a.an1.assign(n, 0.0);
  a.ae2.assign(n, 0.0);
  a.aw2.assign(n, 0.0);
  a.ga2.assign(n, 0.0);
  a.gc2.assign(n, 0.0);
  for (int s = 0; s < 2; ++s) {
    a.nat[s].assign(n, 0.0);
    a.bt[s].assign(n, 0.0);
    a.invd[s].assign(n, 0.0);
    a.na[s].assign(n, 0.0);
    a.b[s].assign(n, 0.0);
    a.jxe[s].assign(n, 0.0);
    a.awwall[s].assign(nj + 1, 0.0);
    a.aefar[s].assign(nj + 1, 0.0);
  }
  a.ae1.assign(n, 0.0);
  std::vector<double> tx(n), te(n);
  for (int s = 0; s < 2; ++s) {
    const std::vector<double>& mx = s == 0 ? g.xix : g.xiy;
    const std::vector<double>& me = s == 0 ? g.etx : g.ety;
    for (long c = 0; c < n; ++c) {
      tx[c] = g.djac[c] * (mx[c] * mx[c]);
      te[c] = g.djac[c] * (me[c] * me[c]);
      a.jxe[s][c] = g.djac[c] * mx[c] * me[c];
    }
    std::vector<double> AE(n, 0.0), AW(n, 0.0), AN(n, 0.0), AS(n, 0.0), AP(n, 0.0);
    for (int j = 1; j <= nj - 1; ++j)
      for (int i = 1; i <= ni - 1; ++i) {
        const long c = ix(sj, i, j);
        AE[c] = 0.5 * (tx[c] + tx[c + 1]);
        AW[c] = 0.5 * (tx[c] + tx[c - 1]);
        AN[c] = 0.5 * (te[c] + te[c + sj]);
        AS[c] = 0.5 * (te[c] + te[c - sj]);
        a.ap0[c] = 2.0 * re * g.djac[c] / dt;
        AP[c] = (((a.ap0[c] + AE[c]) + AW[c]) + AN[c]) + AS[c];
      }
    for (int j = 1; j <= nj - 1; ++j) {
      const long cw = ix(sj, 1, j), cf = ix(sj, ni - 1, j);
      a.awwall[s][j] = AW[cw];
      AP[cw] = AP[cw] + AW[cw];
      AW[cw] = 0.0;
      a.aefar[s][j] = AE[cf];
      AP[cf] = AP[cf] + AE[cf];
      AE[cf] = 0.0;
    }
    if (s == 0)
      for (int j = 1; j <= nj - 1; ++j)
        for (int i = 1; i <= ni - 1; ++i) a.ae1[ix(sj, i, j)] = AE[ix(sj, i, j)];
    // transposed factors along the sweep lines (a = -AW or -AS, c = -AE or -AN)
    {
      const int nl = s == 0 ? nj - 1 : ni - 1, nr = s == 0 ? ni - 1 : nj - 1;
      for (int L = 1; L <= nl; ++L) {
        std::vector<double> av(nr + 2, 0.0), cv(nr + 2, 0.0), dv(nr + 2, 0.0);
        for (int r = 1; r <= nr; ++r) {
          const long c = s == 0 ? ix(sj, r, L) : ix(sj, L, r);
          av[r] = s == 0 ? -AW[c] : -AS[c];
          cv[r] = s == 0 ? -AE[c] : -AN[c];
        }
        double bprev = 0.0;
        for (int r = 1; r <= nr; ++r) {
          const long c = s == 0 ? ix(sj, r, L) : ix(sj, L, r);
          dv[r] = AP[c] - av[r] * bprev;
          bprev = cv[r] / dv[r];
        }
        for (int r = 1; r <= nr; ++r) {
          const long c = s == 0 ? ix(sj, r, L) : ix(sj, L, r);
          a.nat[s][c] = r == 1 ? 0.0 : -cv[r - 1] / dv[r];
          a.bt[s][c] = r == nr ? 0.0 : av[r + 1] / dv[r];
          if (s == 1) {
            a.ga2[c] = -av[r];
            a.gc2[c] = -cv[r];
          }
        }
      }
    }
    if (s == 0) {
      // AN at J = 0 too: the first line's lagged south coupling AS(i, 1) = AN(i, 0)
      for (int i = 1; i <= ni - 1; ++i) {
        const long c = ix(sj, i, 0);
        a.an1[c] = 0.5 * (te[c] + te[c + sj]);
      }
      for (int j = 1; j <= nj - 1; ++j) {
        double bprev = 0.0;
        for (int i = 1; i <= ni - 1; ++i) {
          const long c = ix(sj, i, j);
          const double A = -AW[c];
          const double D = AP[c] - A * bprev;
          const double B = -AE[c] / D;
          a.invd[0][c] = 1.0 / D;
          a.na[0][c] = -A / D;
          a.b[0][c] = B;
          a.an1[c] = AN[c];
          bprev = B;
        }
      }
    } else {
      for (int i = 1; i <= ni - 1; ++i) {
        double bprev = 0.0;
        for (int j = 1; j <= nj - 1; ++j) {
          const long c = ix(sj, i, j);
          const double A = -AS[c];
          const double D = AP[c] - A * bprev;
          const double B = -AN[c] / D;
          a.invd[1][c] = 1.0 / D;
          a.na[1][c] = -A / D;
          a.b[1][c] = B;
          a.ae2[c] = AE[c];
          a.aw2[c] = AW[c];
          bprev = B;
        }
      }
    }
  }
}

// Host-only export (no HIP call): the set-up arrays of a 2D case, for checks against the
// reference's own set-up.  out: 9 planes of (NI+1)(NJ+1): X, Y, DJAC, XIX, XIY, ETX, ETY,
// THETA, DJAC Q12.
extern "C" int fvm_geometry2d(int ni, int nj, const double* xg, const double* yg, double* out) {
  if (!xg || !yg || !out) return FVM_ERR_ARG;
  Geo2D g;
  const int r = geo2d_build(ni, nj, xg, yg, g);
  if (r) return r;
  const long n = (long)(ni + 1) * (nj + 1);
  const std::vector<double>* src[9] = {&g.x, &g.y, &g.djac, &g.xix, &g.xiy,
                                       &g.etx, &g.ety, &g.theta, &g.jq12};
  for (int q = 0; q < 9; ++q)
    for (long c = 0; c < n; ++c) out[q * n + c] = (*src[q])[c];
  return FVM_OK;
}
