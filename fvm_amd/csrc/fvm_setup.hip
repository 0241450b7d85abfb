// Geometry set-up on the device: cell centres and ghost centres, metric terms, the constant
// coefficient arrays of the time step, wall geometry.  Runs once per context.
//
// Follows LES/v2.3_LES-Reynolds stresses/grid.f90:69-346 (centres, mirrored boundary-face
// ghosts, periodic copies, span re-extrapolation of z), metric.f90:32-518 (central / one-sided
// coordinate derivatives, inverse Jacobian, contravariant metrics, Q_ij, face diffusion
// coefficients), init.f90:210-259 (theta, wall triangle areas) and calccoef.f90:88-162 (wall
// normals).  The reference overwrites its edge and corner one-sided derivatives with periodic
// copies (metric.f90:400-410), so only the I-face one-sided formulas are evaluated here.
#include "fvm_ctx.h"

#include <algorithm>

namespace {

struct Box {
  int i0, j0, k0, ni, nj, nk;
  __host__ __device__ long count() const { return (long)ni * nj * nk; }
};

__device__ inline void box_ijk(const Box& b, long t, int& i, int& j, int& k) {
  i = b.i0 + (int)(t % b.ni);
  long r = t / b.ni;
  j = b.j0 + (int)(r % b.nj);
  k = b.k0 + (int)(r / b.nj);
}

#define BOX_LOOP(b, t)                                                                   \
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < (b).count();           \
       t += (long)gridDim.x * blockDim.x)

__global__ void k_centre_interior(Dims d, const double* __restrict__ g, double* __restrict__ c) {
  Box b{1, 1, 1, d.ni - 1, d.nj - 1, d.nk - 1};
  BOX_LOOP(b, t) {
    int i, j, k;
    box_ijk(b, t, i, j, k);
#define G(a, bb, cc) g[cidx(d, a, bb, cc)]
    c[cidx(d, i, j, k)] = 0.125 * (G(i + 1, j + 1, k) + G(i + 1, j, k) + G(i, j + 1, k) +
                                   G(i, j, k) + G(i + 1, j + 1, k + 1) + G(i + 1, j, k + 1) +
                                   G(i, j + 1, k + 1) + G(i, j, k + 1));
  }
}

// mirrored centres behind the I, J and K boundary faces (grid.f90:80-151)
__global__ void k_centre_faces(Dims d, const double* __restrict__ g, double* __restrict__ c,
                               int which) {
  if (which == 0) {
    Box b{0, 1, 1, 2, d.nj - 1, d.nk - 1};
    BOX_LOOP(b, t) {
      int s, j, k;
      box_ijk(b, t, s, j, k);
      if (s == 0) {
        int i = 0;
        double v = 0.25 * (G(i + 1, j + 1, k) + G(i + 1, j, k) + G(i + 1, j + 1, k + 1) +
                           G(i + 1, j, k + 1));
        c[cidx(d, i, j, k)] = v - (c[cidx(d, i + 1, j, k)] - v);
      } else {
        int i = d.ni;
        double v = 0.25 * (G(i, j + 1, k) + G(i, j, k) + G(i, j + 1, k + 1) + G(i, j, k + 1));
        c[cidx(d, i, j, k)] = v + (v - c[cidx(d, i - 1, j, k)]);
      }
    }
  } else if (which == 1) {
    Box b{1, 0, 1, d.ni - 1, 2, d.nk - 1};
    BOX_LOOP(b, t) {
      int i, s, k;
      box_ijk(b, t, i, s, k);
      if (s == 0) {
        int j = 0;
        double v = 0.25 * (G(i + 1, j + 1, k + 1) + G(i + 1, j + 1, k) + G(i, j + 1, k + 1) +
                           G(i, j + 1, k));
        c[cidx(d, i, j, k)] = v - (c[cidx(d, i, j + 1, k)] - v);
      } else {
        int j = d.nj;
        double v = 0.25 * (G(i + 1, j, k + 1) + G(i + 1, j, k) + G(i, j, k + 1) + G(i, j, k));
        c[cidx(d, i, j, k)] = v - (c[cidx(d, i, j - 1, k)] - v);
      }
    }
  } else {
    Box b{1, 1, 0, d.ni - 1, d.nj - 1, 2};
    BOX_LOOP(b, t) {
      int i, j, s;
      box_ijk(b, t, i, j, s);
      if (s == 0) {
        int k = 0;
        double v = 0.25 * (G(i, j, k + 1) + G(i + 1, j, k + 1) + G(i + 1, j + 1, k + 1) +
                           G(i, j + 1, k + 1));
        c[cidx(d, i, j, k)] = v - (c[cidx(d, i, j, k + 1)] - v);
      } else {
        int k = d.nk;
        double v = 0.25 * (G(i, j, k) + G(i + 1, j, k) + G(i + 1, j + 1, k) + G(i, j + 1, k));
        c[cidx(d, i, j, k)] = v + (v - c[cidx(d, i, j, k - 1)]);
      }
    }
  }
}

// span ghost planes after the periodic copy: x, y repeat the neighbour plane, z is
// re-extrapolated from the end node planes (grid.f90:275-346)
__global__ void k_zfix_xy(Dims d, double* x, double* y) {
  Box b{0, 0, 0, d.ni + 1, d.nj + 1, 2};
  BOX_LOOP(b, t) {
    int i, j, s;
    box_ijk(b, t, i, j, s);
    int k = s == 0 ? 0 : d.nk, kn = s == 0 ? 1 : d.nk - 1;
    x[cidx(d, i, j, k)] = x[cidx(d, i, j, kn)];
    y[cidx(d, i, j, k)] = y[cidx(d, i, j, kn)];
  }
}

__global__ void k_zfix_z(Dims d, const double* __restrict__ g, double* z) {
  Box b{1, 1, 0, d.ni - 1, d.nj - 1, 2};
  BOX_LOOP(b, t) {
    int i, j, s;
    box_ijk(b, t, i, j, s);
    if (s == 0) {
      int k = 0;
      double v = 0.25 * (G(i, j, k + 1) + G(i + 1, j, k + 1) + G(i + 1, j + 1, k + 1) +
                         G(i, j + 1, k + 1));
      z[cidx(d, i, j, k)] = v - (z[cidx(d, i, j, k + 1)] - v);
    } else {
      int k = d.nk;
      double v = 0.25 * (G(i, j, k) + G(i + 1, j, k) + G(i + 1, j + 1, k) + G(i, j + 1, k));
      z[cidx(d, i, j, k)] = v + (v - z[cidx(d, i, j, k - 1)]);
    }
  }
#undef G
}

__global__ void k_zfix_z_edges(Dims d, double* z) {
  // edges read only interior (i, j) values of the same plane
  long nt = 2L * ((d.nj - 1) * 2 + (d.ni - 1) * 2 + 4);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nt;
       t += (long)gridDim.x * blockDim.x) {
    int s = (int)(t % 2);
    long r = t / 2;
    int k = s == 0 ? 0 : d.nk;
    long ne = (long)(d.nj - 1) * 2;
    if (r < ne) {
      int j = 1 + (int)(r / 2);
      if (r % 2 == 0) z[cidx(d, 0, j, k)] = z[cidx(d, 1, j, k)];
      else z[cidx(d, d.ni, j, k)] = z[cidx(d, d.ni - 1, j, k)];
      continue;
    }
    r -= ne;
    long ni2 = (long)(d.ni - 1) * 2;
    if (r < ni2) {
      int i = 1 + (int)(r / 2);
      if (r % 2 == 0) z[cidx(d, i, 0, k)] = z[cidx(d, i, 1, k)];
      else z[cidx(d, i, d.nj, k)] = z[cidx(d, i, d.nj - 1, k)];
      continue;
    }
    r -= ni2;
    if (r == 0) z[cidx(d, 0, 0, k)] = z[cidx(d, 1, 1, k)];
    if (r == 1) z[cidx(d, d.ni, 0, k)] = z[cidx(d, d.ni - 1, 1, k)];
    if (r == 2) z[cidx(d, d.ni, d.nj, k)] = z[cidx(d, d.ni - 1, d.nj - 1, k)];
    if (r == 3) z[cidx(d, 0, d.nj, k)] = z[cidx(d, 1, d.nj - 1, k)];
  }
}

// coordinate derivatives (d/dxi, d/deta, d/dzeta) of one coordinate on i = 0..NI, interior
// j, k: central inside, second-order one-sided on the I boundary planes (metric.f90:33-79)
__global__ void k_deriv(Dims d, const double* __restrict__ c, double* dxi, double* det,
                        double* dzt) {
  Box b{0, 1, 1, d.ni + 1, d.nj - 1, d.nk - 1};
  BOX_LOOP(b, t) {
    int i, j, k;
    box_ijk(b, t, i, j, k);
    long q = cidx(d, i, j, k);
    if (i == 0) {
      dxi[q] = (-3.0 * c[q] + 4.0 * c[q + 1] - c[q + 2]) / 2.0;
      det[q] = (c[q + d.sj] - c[q - d.sj]) / 2.0;
      dzt[q] = (c[q + d.sk] - c[q - d.sk]) / 2.0;
    } else if (i == d.ni) {
      dxi[q] = (3.0 * c[q] - 4.0 * c[q - 1] + c[q - 2]) / 2.0;
      det[q] = (c[q + d.sj] - c[q - d.sj]) / 2.0;
      dzt[q] = (c[q + d.sk] - c[q - d.sk]) / 2.0;
    } else {
      dxi[q] = 0.5 * (c[q + 1] - c[q - 1]);
      det[q] = 0.5 * (c[q + d.sj] - c[q - d.sj]);
      dzt[q] = 0.5 * (c[q + d.sk] - c[q - d.sk]);
    }
  }
}

// inverse Jacobian, contravariant metrics and Q_ij at every cell (metric.f90:412-497)
__global__ void k_metric(Dims d, const double* const* __restrict__ D, double* djac,
                         double* const* __restrict__ m, double* jq11, double* jq22,
                         double* jq33, double* jq12, double* jq13, double* jq23) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x) {
    double xxi = D[0][q], xet = D[1][q], xzt = D[2][q];
    double yxi = D[3][q], yet = D[4][q], yzt = D[5][q];
    double zxi = D[6][q], zet = D[7][q], zzt = D[8][q];
    double dj = xxi * (yet * zzt - yzt * zet) - yxi * (xet * zzt - zet * xzt) +
                zxi * (xet * yzt - xzt * yet);
    double xix = (yet * zzt - yzt * zet) / dj;
    double xiy = (zet * xzt - zzt * xet) / dj;
    double xiz = (xet * yzt - xzt * yet) / dj;
    double etx = (yzt * zxi - yxi * zzt) / dj;
    double ety = (zzt * xxi - zxi * xzt) / dj;
    double etz = (xzt * yxi - xxi * yzt) / dj;
    double ztx = (yxi * zet - yet * zxi) / dj;
    double zty = (zxi * xet - zet * xxi) / dj;
    double ztz = (xxi * yet - xet * yxi) / dj;
    djac[q] = dj;
    m[0][q] = xix; m[1][q] = etx; m[2][q] = ztx;
    m[3][q] = xiy; m[4][q] = ety; m[5][q] = zty;
    m[6][q] = xiz; m[7][q] = etz; m[8][q] = ztz;
    double q11 = xix * xix + xiy * xiy + xiz * xiz;
    double q12 = xix * etx + xiy * ety + xiz * etz;
    double q13 = ztx * xix + zty * xiy + ztz * xiz;
    double q22 = etx * etx + ety * ety + etz * etz;
    double q23 = etx * ztx + ety * zty + etz * ztz;
    double q33 = ztx * ztx + zty * zty + ztz * ztz;
    jq11[q] = dj * q11; jq22[q] = dj * q22; jq33[q] = dj * q33;
    jq12[q] = dj * q12; jq13[q] = dj * q13; jq23[q] = dj * q23;
  }
}

// 2 Re J^-1/dt alone, after a change of the time step
__global__ void k_set_ap0(long n, const double* __restrict__ djac, double re, double dtime,
                          double* ap0) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
       q += (long)gridDim.x * blockDim.x)
    ap0[q] = 2.0 * re * djac[q] / dtime;
}

// constant per-cell coefficients: 2 Re J^-1/dt, squared test-filter width (J^-1)^(2/3),
// circumferential angle of the cell centre (init.f90:210-237)
__global__ void k_cell_consts(Dims d, const double* __restrict__ djac, const double* x,
                              const double* y, double re, double dtime, double pi, double* ap0,
                              double* filt2, double* theta) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x) {
    ap0[q] = 2.0 * re * djac[q] / dtime;
    double fw = pow(djac[q], 1.0 / 3.0);
    filt2[q] = fw * fw;
    double X = x[q], Y = y[q];
    double rad = sqrt(X * X + Y * Y);
    double th = atan(Y / X);
    double as = asin(Y / rad);
    if (th >= 0.0 && as > 0.0) {
      // first quadrant: angle as returned
    } else if (th < 0.0 && as > 0.0) th = pi + th;
    else if (th >= 0.0 && as <= 0.0) th = pi + th;
    else if (th < 0.0 && as < 0.0) th = 2.0 * pi + th;
    theta[q] = th;
  }
}

// face coefficients: f(c) = 0.5 (g(c) + g(c + off)) on every cell whose neighbour exists.
// AF stages (iter_u.f90:23-39): g = J^-1 xi_s^2 etc.; Poisson/DIF (metric.f90:500-511):
// g = J^-1 Q_ii.  The west/south/bottom coefficient of a cell is the face value of its
// lower neighbour.
__global__ void k_face(Dims d, const double* __restrict__ djac, const double* __restrict__ mm,
                       const double* __restrict__ jq, long off, int dir, double* __restrict__ f) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x) {
    long r = q / d.sj;
    int i = (int)(q - r * d.sj);
    int k = (int)(r / (d.nj + 1));
    int j = (int)(r - (long)k * (d.nj + 1));
    bool ok = dir == 0 ? i < d.ni : (dir == 1 ? j < d.nj : k < d.nk);
    if (!ok) { f[q] = 0.0; continue; }
    double a, b;
    if (mm) {
      a = djac[q] * (mm[q] * mm[q]);
      b = djac[q + off] * (mm[q + off] * mm[q + off]);
    } else {
      a = jq[q];
      b = jq[q + off];
    }
    f[q] = 0.5 * (a + b);
  }
}

// wall triangles of every wall cell (init.f90:239-259, calccoef.f90:88-162): plane
// coefficients A, B, C, D-offset distance d_n, |n| and area, two triangles per (j, k)
__global__ void k_wall(Dims d, const double* __restrict__ xg, const double* __restrict__ yg,
                       const double* __restrict__ zg, const double* __restrict__ x,
                       const double* __restrict__ y, const double* __restrict__ z, double* w) {
  long nw = (long)(d.nj - 1) * (d.nk - 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 2 * nw;
       t += (long)gridDim.x * blockDim.x) {
    int tri = (int)(t / nw);
    long r = t - tri * nw;
    int j = 1 + (int)(r % (d.nj - 1));
    int k = 1 + (int)(r / (d.nj - 1));
    int i = 1;
#define XG(a, b, c) xg[cidx(d, a, b, c)]
#define YG(a, b, c) yg[cidx(d, a, b, c)]
#define ZG(a, b, c) zg[cidx(d, a, b, c)]
#define SQ(a) ((a) * (a))
    double A, B, C, D, area;
    double l3 = sqrt(SQ(XG(i, j + 1, k) - XG(i, j, k + 1)) + SQ(YG(i, j + 1, k) - YG(i, j, k + 1)) +
                     SQ(ZG(i, j + 1, k) - ZG(i, j, k + 1)));
    if (tri == 0) {
      double l1 = sqrt(SQ(XG(i, j + 1, k) - XG(i, j, k)) + SQ(YG(i, j + 1, k) - YG(i, j, k)) +
                       SQ(ZG(i, j + 1, k) - ZG(i, j, k)));
      double l2 = sqrt(SQ(XG(i, j, k + 1) - XG(i, j, k)) + SQ(YG(i, j, k + 1) - YG(i, j, k)) +
                       SQ(ZG(i, j, k + 1) - ZG(i, j, k)));
      double th1 = acos((l3 * l3 - l2 * l2 - l1 * l1) / (-2.0 * l1 * l2));
      area = 0.5 * l1 * l2 * sin(th1);
      A = YG(i, j, k) * (ZG(i, j + 1, k) - ZG(i, j, k + 1)) +
          YG(i, j + 1, k) * (ZG(i, j, k + 1) - ZG(i, j, k)) +
          YG(i, j, k + 1) * (ZG(i, j, k) - ZG(i, j + 1, k));
      B = ZG(i, j, k) * (XG(i, j + 1, k) - XG(i, j, k + 1)) +
          ZG(i, j + 1, k) * (XG(i, j, k + 1) - XG(i, j, k)) +
          ZG(i, j, k + 1) * (XG(i, j, k) - XG(i, j + 1, k));
      C = XG(i, j, k) * (YG(i, j + 1, k) - YG(i, j, k + 1)) +
          XG(i, j + 1, k) * (YG(i, j, k + 1) - YG(i, j, k)) +
          XG(i, j, k + 1) * (YG(i, j, k) - YG(i, j + 1, k));
      D = XG(i, j, k) * (YG(i, j + 1, k) * ZG(i, j, k + 1) - YG(i, j, k + 1) * ZG(i, j + 1, k)) +
          XG(i, j + 1, k) * (YG(i, j, k + 1) * ZG(i, j, k) - YG(i, j, k) * ZG(i, j, k + 1)) +
          XG(i, j, k + 1) * (YG(i, j, k) * ZG(i, j + 1, k) - YG(i, j + 1, k) * ZG(i, j, k));
    } else {
      double l1 = sqrt(SQ(XG(i, j + 1, k + 1) - XG(i, j + 1, k)) +
                       SQ(YG(i, j + 1, k + 1) - YG(i, j + 1, k)) +
                       SQ(ZG(i, j + 1, k + 1) - ZG(i, j + 1, k)));
      double l2 = sqrt(SQ(XG(i, j + 1, k + 1) - XG(i, j, k + 1)) +
                       SQ(YG(i, j + 1, k + 1) - YG(i, j, k + 1)) +
                       SQ(ZG(i, j + 1, k + 1) - ZG(i, j, k + 1)));
      double th1 = acos((l3 * l3 - l2 * l2 - l1 * l1) / (-2.0 * l1 * l2));
      area = 0.5 * l1 * l2 * sin(th1);
      A = YG(i, j + 1, k + 1) * (ZG(i, j, k + 1) - ZG(i, j, k)) +
          YG(i, j, k + 1) * (ZG(i, j, k) - ZG(i, j + 1, k + 1)) +
          YG(i, j, k) * (ZG(i, j + 1, k + 1) - ZG(i, j, k + 1));
      B = ZG(i, j + 1, k + 1) * (XG(i, j, k + 1) - XG(i, j, k)) +
          ZG(i, j, k + 1) * (XG(i, j, k) - XG(i, j + 1, k + 1)) +
          ZG(i, j, k) * (XG(i, j + 1, k + 1) - XG(i, j, k + 1));
      C = XG(i, j + 1, k + 1) * (YG(i, j, k + 1) - YG(i, j, k)) +
          XG(i, j, k + 1) * (YG(i, j, k) - YG(i, j + 1, k + 1)) +
          XG(i, j, k) * (YG(i, j + 1, k + 1) - YG(i, j, k + 1));
      D = XG(i, j + 1, k + 1) * (YG(i, j, k + 1) * ZG(i, j, k) - YG(i, j, k) * ZG(i, j, k + 1)) +
          XG(i, j, k + 1) * (YG(i, j, k) * ZG(i, j + 1, k + 1) - YG(i, j + 1, k + 1) * ZG(i, j, k)) +
          XG(i, j, k) * (YG(i, j + 1, k + 1) * ZG(i, j, k + 1) - YG(i, j, k + 1) * ZG(i, j + 1, k + 1));
    }
#undef XG
#undef YG
#undef ZG
#undef SQ
    D = -1.0 * D;
    double dum = sqrt(A * A + B * B + C * C);
    long c = cidx(d, i, j, k);
    double dn = fabs(A * x[c] + B * y[c] + C * z[c] + D) / dum;
    double* o = w + 6 * t;
    o[0] = A; o[1] = B; o[2] = C; o[3] = dum; o[4] = dn; o[5] = area;
  }
}

}  // namespace

int fvm_ghost_fill(FvmCtx* c, double* a) {
  hipLaunchKernelGGL(k_ghost_fill_t<double>, grid1(ghost_count(c->d), 256), dim3(256), 0,
                     c->stream, c->d, a);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_ghost_fill_n(FvmCtx* c, double** a, int cnt) {
  for (int q0 = 0; q0 < cnt; q0 += 8) {
    PtrSet8 set{};
    const int m = std::min(8, cnt - q0);
    for (int q = 0; q < m; ++q) set.p[q] = a[q0 + q];
    hipLaunchKernelGGL(k_ghost_fill_multi, dim3(grid1(ghost_count(c->d), 256).x, m), dim3(256),
                       0, c->stream, c->d, set);
    HIP_CHECK(hipGetLastError());
  }
  return FVM_OK;
}

int fvm_geom_setup(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  dim3 B(256);
  dim3 G = grid1(d.n, 256);
  double* co[3] = {c->x, c->y, c->z};
  double* g[3] = {c->xg, c->yg, c->zg};
  for (int a = 0; a < 3; ++a) {
    hipLaunchKernelGGL(k_centre_interior, G, B, 0, st, d, g[a], co[a]);
    for (int w = 0; w < 3; ++w) hipLaunchKernelGGL(k_centre_faces, G, B, 0, st, d, g[a], co[a], w);
    int r = fvm_ghost_fill(c, co[a]);
    if (r) return r;
  }
  hipLaunchKernelGGL(k_zfix_xy, G, B, 0, st, d, c->x, c->y);
  hipLaunchKernelGGL(k_zfix_z, G, B, 0, st, d, c->zg, c->z);
  hipLaunchKernelGGL(k_zfix_z_edges, G, B, 0, st, d, c->z);
  HIP_CHECK(hipGetLastError());

  // nine derivative arrays (temporaries)
  double* D[9];
  for (int a = 0; a < 9; ++a) HIP_CHECK(hipMalloc((void**)&D[a], d.n * sizeof(double)));
  for (int a = 0; a < 9; ++a) HIP_CHECK(hipMemsetAsync(D[a], 0, d.n * sizeof(double), st));
  for (int a = 0; a < 3; ++a)
    hipLaunchKernelGGL(k_deriv, G, B, 0, st, d, co[a], D[3 * a], D[3 * a + 1], D[3 * a + 2]);
  int r = fvm_ghost_fill_n(c, D, 9);
  if (r) return r;
  double** dD;
  double** dM;
  HIP_CHECK(hipMalloc((void**)&dD, 9 * sizeof(double*)));
  HIP_CHECK(hipMalloc((void**)&dM, 9 * sizeof(double*)));
  HIP_CHECK(hipMemcpyAsync(dD, D, 9 * sizeof(double*), hipMemcpyHostToDevice, st));
  HIP_CHECK(hipMemcpyAsync(dM, c->m, 9 * sizeof(double*), hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_metric, G, B, 0, st, d, (const double* const*)dD, c->djac,
                     (double* const*)dM, c->jq11, c->jq22, c->jq33, c->jq12, c->jq13, c->jq23);
  HIP_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_cell_consts, G, B, 0, st, d, c->djac, c->x, c->y, c->prm.re,
                     c->prm.dtime, c->pi, c->ap0, c->filt2, c->theta);
  long off[3] = {1, d.sj, d.sk};
  for (int s = 0; s < 3; ++s)
    for (int dir = 0; dir < 3; ++dir)
      hipLaunchKernelGGL(k_face, G, B, 0, st, d, c->djac, c->m[3 * s + dir], nullptr, off[dir],
                         dir, c->fa[s][dir]);
  hipLaunchKernelGGL(k_face, G, B, 0, st, d, c->djac, nullptr, c->jq11, off[0], 0, c->pdI);
  hipLaunchKernelGGL(k_face, G, B, 0, st, d, c->djac, nullptr, c->jq22, off[1], 1, c->pdJ);
  hipLaunchKernelGGL(k_face, G, B, 0, st, d, c->djac, nullptr, c->jq33, off[2], 2, c->pdK);
  hipLaunchKernelGGL(k_wall, grid1(2L * (d.nj - 1) * (d.nk - 1), 256), B, 0, st, d, c->xg, c->yg,
                     c->zg, c->x, c->y, c->z, c->wall);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(st));
  for (int a = 0; a < 9; ++a) HIP_CHECK(hipFree(D[a]));
  HIP_CHECK(hipFree(dD));
  HIP_CHECK(hipFree(dM));
  {  // xi_z and eta_z identically zero (z independent of xi and eta: an extruded grid)
    std::vector<double> h(d.n);
    bool z = true;
    for (int a = 6; a < 8 && z; ++a) {
      HIP_CHECK(hipMemcpy(h.data(), c->m[a], d.n * sizeof(double), hipMemcpyDeviceToHost));
      for (long q = 0; q < d.n && z; ++q) z = (h[q] == 0.0);
    }
    c->xz_zero = z;
  }
  return FVM_OK;
}

int fvm_k_set_dtime(FvmCtx* c, double dtime) {
  c->prm.dtime = dtime;
  hipLaunchKernelGGL(k_set_ap0, grid1(c->d.n, 256), dim3(256), 0, c->stream, c->d.n, c->djac,
                     c->prm.re, dtime, c->ap0);
  if (c->thermal)  // 2 PE DJAC/DTIME of ITER_T (iter_t.f90:31)
    hipLaunchKernelGGL(k_set_ap0, grid1(c->d.n, 256), dim3(256), 0, c->stream, c->d.n, c->djac,
                       c->pe, dtime, c->ap0t);
  HIP_CHECK(hipGetLastError());
  return fvm_af_setup(c);
}
