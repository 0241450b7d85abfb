// The batched 2D time step: kernels (one launch per phase over all cases), the fixed-shape step
// captured as a hipGraph, and the C ABI of include/fvm_capi.h (fvm_batch2d_*).
//
// Step (Main.f90:48-118): U_N copy, CONTRA_VEL + NL + DIF + P_SOURCE + the stage-1 right-hand
// side of ITER_U/ITER_V (one fused stencil pass), the stage-1 BCU and I-line sweep, the
// stage-2 right-hand side, BCU and J-line sweep, the predictor correction with BCU, the
// Poisson right-hand side, ITER_P as FP64 defect correction with the separable inverse (fixed
// pass count, converged cases skip), BCP, P_SOURCE + projection + BCU + CFL, CALCCOEF.
#include <math.h>
#include <string.h>

#include <algorithm>
#include <type_traits>
#include <string>
#include <vector>

#include "../../include/fvm_capi.h"
#include "fvm2d_host.h"
#include "fvm2d_kernels.h"
#include "fvm_wave.h"

namespace {

// complex values of a global-address-space array, moved as their two components
template <typename T>
__device__ __forceinline__ Cx<T> cx_ld(gptr<const T> p, long o) {
  return Cx<T>{p[2 * o], p[2 * o + 1]};
}
template <typename T>
__device__ __forceinline__ void cx_st(gptr<T> p, long o, Cx<T> v) {
  p[2 * o] = v.x;
  p[2 * o + 1] = v.y;
}

// ---------------------------------------------------------------------------------------------
// Phase A: U_N = U (whole plane), CONTRA_VEL, NL(U_N), NL(V_N), DIF(U_N), DIF(V_N), P_SOURCE,
// the stage-1 right-hand sides SU (iter_u.f90:43-46 / iter_v.f90) with the old NL (NL_NM1, held
// in a register between the read of NL and its update, not stored) and the cross
// terms DD of the current field, the far-field U_F of BCU/BCV (bcu.f90:22-33) from U_N and the
// stage-1 right-hand-side folds (bcu.f90:15, 34).  AF32: instead of SU, the FP64 residual of the
// stage input against the stage-1 line operator, in FP32, for the change form of the sweep.
// ---------------------------------------------------------------------------------------------
template <bool AF32>
__global__ void __launch_bounds__(512) k2_explicit(const Case2D* __restrict__ tab,
                                                   const int* __restrict__ bcase,
                                                   const int* __restrict__ btile) {
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  if (i > ni - 1 || j > nj - 1) return;
  const long c = (long)i + (long)sj * j;
  const auto U = glb(cs.f[F_U]);
  const auto V = glb(cs.f[F_V]);
  const auto P = glb(cs.f[F_P]);
  const auto UN = glb(cs.f[F_UN]);
  const auto VN = glb(cs.f[F_VN]);
  // U_N = U over the whole plane: every interior thread copies its cell and the ghosts next to
  // it (corners from the corner cells)
  {
    UN[c] = U[c];
    VN[c] = V[c];
    const int gi = i == 1 ? -1 : (i == ni - 1 ? 1 : 0);
    const int gj = j == 1 ? -1 : (j == nj - 1 ? 1 : 0);
    if (gi) { UN[c + gi] = U[c + gi]; VN[c + gi] = V[c + gi]; }
    if (gj) { UN[c + gj * sj] = U[c + gj * sj]; VN[c + gj * sj] = V[c + gj * sj]; }
    if (gi && gj) {
      const long q = c + gi + (long)gj * sj;
      UN[q] = U[q];
      VN[q] = V[q];
    }
  }
  const GeomG g = geom2(cs);
  const double j1 = ju1(g, U, V, c), j2 = ju2(g, U, V, c);
  const double ue = 0.5 * (j1 + ju1(g, U, V, c + 1));
  const double uw = i == 1 ? 0.0 : 0.5 * (j1 + ju1(g, U, V, c - 1));  // UW(1) = 0
  const double vn = 0.5 * (j2 + ju2(g, U, V, c + sj));
  const double vs = 0.5 * (j2 + ju2(g, U, V, c - sj));
  const auto NLX = glb(cs.f[F_NLX]);
  const auto NLY = glb(cs.f[F_NLY]);
  const double nlx1 = NLX[c], nly1 = NLY[c];
  const double nlx = nl2(U, c, sj, ue, uw, vn, vs);
  const double nly = nl2(V, c, sj, ue, uw, vn, vs);
  const auto de = glb(cs.f[G_DE]);
  const auto dn = glb(cs.f[G_DN]);
  const auto jq = glb(cs.f[G_JQ12]);
  const double difx = dif2(de, dn, jq, U, c, sj);
  const double dify = dif2(de, dn, jq, V, c, sj);
  const double su = psrc(g.dj, g.xix, g.etx, P, c, sj);
  const double sv = psrc(g.dj, g.xiy, g.ety, P, c, sj);
  const double re = cs.re;
  // the change form leaves out the 2 Re DJAC / dt U term: it cancels against the diagonal
  const double ap0 = AF32 ? 0.0 : glb(cs.f[G_AP0])[c];
  double r1u = ap0 * U[c] - 3.0 * re * nlx + 1.0 * re * nlx1 + difx + 2.0 * re * su +
               cross2(glb(cs.f[G_JXE1]), U, c, sj);
  double r1v = ap0 * V[c] - 3.0 * re * nly + 1.0 * re * nly1 + dify + 2.0 * re * sv +
               cross2(glb(cs.f[G_JXE1]), V, c, sj);
  if (i == 1) {  // wall fold SU(1) += 2 AW(1) U_F (change form: 2 AW(1) (U_F - U))
    const double uw = AF32 ? glb(cs.b[B_UWALL])[j] - U[c] : glb(cs.b[B_UWALL])[j];
    const double vw = AF32 ? glb(cs.b[B_VWALL])[j] - V[c] : glb(cs.b[B_VWALL])[j];
    r1u = r1u + 2.0 * glb(cs.b[B_AW1W])[j] * uw;
    r1v = r1v + 2.0 * glb(cs.b[B_AW1W])[j] * vw;
  }
  if (i == ni - 1) {
    // far-field U_F: inflow Dirichlet (UIN = 1, VIN = 0) where THETA(NI, J) is in [pi/2, 3pi/2],
    // else the convective outflow value from U_N with convection speed UIN (bcu.f90:26-33)
    const double pi = 3.14159265358979323846;
    const long f = c + 1;  // (NI, J)
    const double th = glb(cs.f[G_THETA])[f];
    double ufu, ufv;
    if (th >= 0.5 * pi && th <= 1.5 * pi) {
      ufu = 1.0;
      ufv = 0.0;
    } else {
      const double ax = 0.5 * (g.xix[f] + g.xix[c]);
      const double ae = 0.5 * (g.etx[f] + g.etx[c]);
      const gptr<double> F2[2] = {U, V};
      double uf[2];
      for (int q = 0; q < 2; ++q) {
        const auto W = F2[q];
        const double s = W[f + sj] + W[c + sj] - W[f - sj] - W[c - sj];
        uf[q] = 0.5 * (W[f] + W[c]) - cs.dt * 1.0 * (ax * (W[f] - W[c]) + ae * 0.25 * s);
      }
      ufu = uf[0];
      ufv = uf[1];
    }
    glb(cs.b[B_UFAR])[j] = ufu;
    glb(cs.b[B_VFAR])[j] = ufv;
    r1u = r1u + 2.0 * glb(cs.b[B_AE1F])[j] * (AF32 ? ufu - U[c] : ufu);
    r1v = r1v + 2.0 * glb(cs.b[B_AE1F])[j] * (AF32 ? ufv - V[c] : ufv);
  }
  NLX[c] = nlx;
  NLY[c] = nly;
  glb(cs.f[F_DIFX])[c] = difx;
  glb(cs.f[F_DIFY])[c] = dify;
  glb(cs.f[F_SU])[c] = su;
  glb(cs.f[F_SV])[c] = sv;
  // the old next line's coupling AN U(J+1), taken into the right-hand side before the sweep; the
  // branch-cut ghost J = NJ is BCU's periodic image J = 1
  const long cn = j == nj - 1 ? (long)i + sj : c + sj;
  const double an = glb(cs.f[G_AN1])[c];
  if (AF32) {
    // the residual r = SU + AS x0(J-1) + AN x0(J+1) + AW x0(I-1) + AE x0(I+1) - AP x0 of the
    // stage input x0 = U, with AP = 2 Re DJAC / dt + AE + AW + AN + AS (+ the folds above), as
    // differences: AN (x0(J+1) - x0) + AS (x0(J-1) - x0) + AE (..) + AW (..) + the rest of SU
    // (folded AW(1) = AE(NI-1) = 0; the ghost line J = 0 is BCU's periodic image J = NJ-1, as
    // the sweep sees it)
    const double as = glb(cs.f[G_AN1])[c - sj];  // AS(J) = AN(J-1)
    const double ae = glb(cs.f[G_AE1])[c];
    const double aw = glb(cs.f[G_AE1])[c - 1];  // AW(I) = AE(I-1)
    const long cs0 = j == 1 ? (long)i + (long)sj * (nj - 1) : c - sj;
    const double ru = r1u + an * (U[cn] - U[c]) + as * (U[cs0] - U[c]) +
                      aw * (U[c - 1] - U[c]) + ae * (U[c + 1] - U[c]);
    const double rv = r1v + an * (V[cn] - V[c]) + as * (V[cs0] - V[c]) +
                      aw * (V[c - 1] - V[c]) + ae * (V[c + 1] - V[c]);
    glb(cs.h[H_RU])[c] = (float)ru;
    glb(cs.h[H_RV])[c] = (float)rv;
  } else {
    r1u = r1u + an * U[cn];
    r1v = r1v + an * V[cn];
    glb(cs.f[F_SU1])[c] = r1u;
    glb(cs.f[F_SV1])[c] = r1v;
  }
}

// BCU / BCV ghosts (bcu.f90:10-51) of field F from its interior: wall row I = 0 and far-field
// row I = NI for J = 0..NJ, then the branch-cut columns J = 0, NJ for I = 0..NI.  One
// workgroup of 64 lanes.
template <class P, class B>
__device__ inline void bcu_ghosts_wave(P F, int ni, int nj, int sj, B uwall, B ufar) {
  const int l = threadIdx.x;
  for (int jj = l; jj <= nj; jj += 64) {
    const bool in = jj >= 1 && jj <= nj - 1;
    const long c = (long)sj * jj;
    F[c] = -F[c + 1] + 2.0 * (in ? uwall[jj] : 0.0);
    F[c + ni] = -F[c + ni - 1] + 2.0 * (in ? ufar[jj] : 0.0);
  }
  __syncthreads();
  for (int ii = l; ii <= ni; ii += 64) {
    F[ii] = F[ii + (long)sj * (nj - 1)];
    F[ii + (long)sj * nj] = F[ii + sj];
  }
  __syncthreads();
}

// ---------------------------------------------------------------------------------------------
// The AF line Gauss-Seidel sweep of one stage (iter_u.f90:50-66 I lines, 103-119 J lines): lines
// in the reference's order, each solved with the already updated previous line and the old next
// line.  The old next line's coupling is part of the right-hand side Q formed before the sweep
// (k2_explicit, k2_rhs2), so a line needs only the previous line.  Per line the Thomas recursion
// runs as two affine-map scans over the 64 lanes (lane l holds rows l R .. l R + R - 1,
// R = ceil(rows / 64) of the case).
//
// Chunked form: the lines of a case are cut into chunks of SW_C consecutive lines and each
// (case, component, chunk) is one wavefront; a chunk takes the line before it from the
// preceding chunk's last line of the previous pass (block Jacobi across chunks, the stage input
// in the first pass).  The passes are nilpotent (exact after as many passes as chunks) and each
// contracts the chunk-boundary error by rho^SW_C, rho < 1 the ratio of the lagged coupling to
// the diagonal dominance of the line operator; a pass after the first re-solves a chunk's lines
// only until a line stops changing (max change <= tol max |x|), the rest keep their values.
// The last pass books the largest relative change of a chunk's last line in the case's
// ST_AFCH.  The operands of line L + 1 are loaded while line L is solved.
// DIR 0: I lines (rows along I, lines J = 1..NJ-1, lagged coupling AS = AN(J-1));
// DIR 1: J lines (rows along J, lines I = 1..NI-1, lagged coupling AW); the line ends take the
// ghost values (BCU, applied by k2_sweep_seed) as Dirichlet data.
// ---------------------------------------------------------------------------------------------
template <int R, typename T>
struct LineOps {
  T cdn[R], q[R], iv[R], na[R], bf[R], xo[R];
  T g0, g1;                                        // ghosts before / after the line
};

// R consecutive floats from a 16-byte (R % 4 == 0) or 8-byte (R % 2 == 0) aligned address
typedef float sw_f4 __attribute__((ext_vector_type(4)));
typedef float sw_f2 __attribute__((ext_vector_type(2)));
template <int R>
__device__ __forceinline__ void load_rows(float (&d)[R], gptr<const float> p) {
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int k = 0; k < R / 4; ++k) {
      const sw_f4 v = ((gptr<const sw_f4>)p)[k];
      d[4 * k] = v.x; d[4 * k + 1] = v.y; d[4 * k + 2] = v.z; d[4 * k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const sw_f2 v = ((gptr<const sw_f2>)p)[k];
      d[2 * k] = v.x; d[2 * k + 1] = v.y;
    }
  }
}

// roff: the offset of row 0 in a line (the padded FP32 J-line-major layout)
template <int R, int DIR, bool EARLY, typename T>
__device__ inline void sweep_load(LineOps<R, T>& o, gptr<const T> F, gptr<const T> Q,
                                  gptr<const T> invd, gptr<const T> na, gptr<const T> bf,
                                  gptr<const T> cdn, long base, long rs, long roff, long dnoff,
                                  int n, int r0, int l) {
  constexpr bool DELTA = std::is_same<T, float>::value;
  if constexpr (DELTA && DIR == 1 && R % 2 == 0) {
    // a lane's rows as vectors (row r at base + roff + 1 + r; dnoff = 0 in this direction).
    // The last lane's rows past n lie inside the padded plane and are never used.
    const long c0 = base + roff + 1 + r0;
    load_rows<R>(o.cdn, cdn + c0);
    load_rows<R>(o.q, Q + c0);
    load_rows<R>(o.iv, invd + c0);
    load_rows<R>(o.na, na + c0);
    load_rows<R>(o.bf, bf + c0);
    if (EARLY) load_rows<R>(o.xo, F + c0);
    else {
#pragma unroll
      for (int q = 0; q < R; ++q) o.xo[q] = (T)0;
    }
  } else {
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = r0 + q;
      if (r < n) {
        const long c = base + roff + rs * (r + 1);
        o.cdn[q] = cdn[c + dnoff];
        o.q[q] = Q[c];
        o.iv[q] = invd[c];
        o.na[q] = na[c];
        o.bf[q] = bf[c];
        o.xo[q] = EARLY ? F[c] : (T)0;  // the previous pass's value of the row
      }
    }
  }
  // the change form's ghosts do not change
  o.g0 = !DELTA && l == 0 ? F[base] : (T)0;
  o.g1 = !DELTA && l == 63 ? F[base + rs * (n + 1)] : (T)0;
}

// T = double: the sweep of the stage on the field itself (DIR 1 on the J-line-major copies: the
// iterate F_TX*, right-hand side F_TQ*, factors and AW G_T*, made by k2_tr_in and returned by
// k2_tr_out).  T = float: the change form.  The sweep solves T_L d_L = r_L + AS d_(L-1) for the
// change d = x - x0 of the stage input x0, with r the FP64 residual of x0 (k2_explicit,
// k2_rhs2; the J-line stage's J-line-major copy by k2_tr_in32), zero ghosts and a zero ghost
// line; the iterate is H_DU / H_DV (J-line-major for DIR 1).  The I-line stage's last pass adds
// the change to the FP64 field, the J-line stage's k2_tr_out32 does.
template <int R, int DIR, bool EARLY, typename T>
__device__ void sweep_chunk(const Case2D& cs, int comp, int ca, int slot, double tol, int final) {
  constexpr bool DELTA = std::is_same<T, float>::value;
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  T *F0, *Q0, *iv0, *na0, *bf0, *cd0;
  if constexpr (DELTA) {
    F0 = cs.h[comp ? H_DV : H_DU];
    Q0 = cs.h[DIR == 0 ? (comp ? H_RV : H_RU) : (comp ? H_TRV : H_TRU)];
    iv0 = cs.h[DIR == 0 ? H_INVD1 : H_TINVD2];
    na0 = cs.h[DIR == 0 ? H_NA1 : H_TNA2];
    bf0 = cs.h[DIR == 0 ? H_B1 : H_TB2];
    cd0 = cs.h[DIR == 0 ? H_AN1 : H_TAW2];
  } else {
    F0 = DIR == 0 ? cs.f[comp ? F_V : F_U] : cs.f[comp ? F_TXV : F_TXU];
    Q0 = DIR == 0 ? cs.f[comp ? F_SV1 : F_SU1] : cs.f[comp ? F_TQV : F_TQU];
    iv0 = cs.f[DIR == 0 ? G_INVD1 : G_TINVD2];
    na0 = cs.f[DIR == 0 ? G_NA1 : G_TNA2];
    bf0 = cs.f[DIR == 0 ? G_B1 : G_TB2];
    cd0 = cs.f[DIR == 0 ? G_AN1 : G_TAW2];
  }
  // global-address-space views: global (not flat) loads, which do not wait on LDS and scalar
  // traffic, so the next line's loads stay in flight while this line is solved
  const gptr<T> F = glb(F0);
  const gptr<const T> Q = glb((const T*)Q0), invd = glb((const T*)iv0), na = glb((const T*)na0);
  const gptr<const T> bf = glb((const T*)bf0), cdn = glb((const T*)cd0);
  // the change form's I-line stage adds its last pass's change to the field
  const bool add = DELTA && DIR == 0 && final;
  const gptr<double> F64 = glb(add ? cs.f[comp ? F_V : F_U] : nullptr);
  const int n = DIR == 0 ? ni - 1 : nj - 1;       // rows per line
  const int nl = DIR == 0 ? nj - 1 : ni - 1;      // lines
  const long rs = 1;                              // row stride
  const long ls = DIR == 0 ? sj : (DELTA ? tls32(nj) : nj + 1);  // line stride
  const long roff = DIR == 1 && DELTA ? J32_OFF : 0;  // row 0's offset in a line
  const long dnoff = DIR == 0 ? -(long)sj : 0;    // lower coupling of row c at cdn[c + dnoff]
  const int nch = sw_nchunk(nl);
  const int a0 = 1 + ca * SW_C, a1 = min(a0 + SW_C - 1, nl);
  const gptr<double> bnd = glb(cs.swb);
  const long bin = ((long)(slot * 2 + comp) * nch) * n, bout = ((long)((slot ^ 1) * 2 + comp) * nch) * n;
  const int l = threadIdx.x;
  const int r0 = l * R;
  T prev[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = r0 + q;
    // the line before the chunk: the ghost line 0 (no change in the change form), or the
    // preceding chunk's last line
    T v = (T)0;
    if (r < n) {
      if (ca > 0) v = (T)bnd[bin + (long)(ca - 1) * n + r];
      else if (!DELTA) v = F[rs * (r + 1)];
    }
    prev[q] = v;
  }
  LineOps<R, T> cur, nxt;
  sweep_load<R, DIR, EARLY>(cur, F, Q, invd, na, bf, cdn, ls * a0, rs, roff, dnoff, n, r0, l);
  bool stopped = false;
  T xrun = (T)0;
  int L = a0;
  for (; L <= a1; ++L) {
    const long base = ls * L;
    if (L < a1)
      sweep_load<R, DIR, EARLY>(nxt, F, Q, invd, na, bf, cdn, base + ls, rs, roff, dnoff, n, r0,
                                l);
    T al[R], be[R], bb[R];
    T At = (T)0, Bt = (T)1;
#pragma unroll
    for (int q = 0; q < R; ++q) {
      if (r0 + q < n) {
        const T C = cur.cdn[q] * prev[q] + cur.q[q];
        al[q] = C * cur.iv[q];
        be[q] = cur.na[q];
        bb[q] = cur.bf[q];
      } else {
        al[q] = (T)0; be[q] = (T)1; bb[q] = (T)0;
      }
      At = al[q] + be[q] * At;
      Bt = be[q] * Bt;
    }
    // lane l starts from the exclusive prefix applied to the ghost before the first row
    T yin = wave::prefix_apply(At, Bt, wave::lane_value(cur.g0, 0));
    T y[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      y[q] = al[q] + be[q] * yin;
      yin = y[q];
    }
    T Ct = (T)1, Dt = (T)0;
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
      if (r0 + q < n) {
        Dt = y[q] - bb[q] * Dt;
        Ct = -bb[q] * Ct;
      }
    }
    // and from the exclusive suffix applied to the ghost after the last row
    T xin = wave::suffix_apply(Dt, Ct, wave::lane_value(cur.g1, 63));
    T dl = (T)0;
#pragma unroll
    for (int q = R - 1; q >= 0; --q) {
      const int r = r0 + q;
      if (r < n) {
        const T x = y[q] - bb[q] * xin;
        xin = x;
        prev[q] = x;
        const long c = base + roff + rs * (r + 1);
        if (add)
          F64[c] = F64[c] + (double)x;
        else
          F[c] = x;
        if (EARLY) {
          dl = fmax(dl, fabs(x - cur.xo[q]));
          xrun = fmax(xrun, fabs(x));
        }
      }
    }
    if (EARLY && L < a1) {
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) {
        dl = fmax(dl, __shfl_xor(dl, o, 64));
        xrun = fmax(xrun, __shfl_xor(xrun, o, 64));
      }
      if (dl <= (T)tol * xrun) { stopped = true; break; }  // uniform over the wavefront
    }
    cur = nxt;
  }
  if (add && stopped) {
    // the lines after the stop keep the previous pass's change
    for (int M = L + 1; M <= a1; ++M) {
#pragma unroll
      for (int q = 0; q < R; ++q) {
        const int r = r0 + q;
        if (r < n) {
          const long c = ls * M + rs * (r + 1);
          F64[c] = F64[c] + (double)F[c];
        }
      }
    }
  }
  double dm = 0.0, xm = 0.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = r0 + q;
    if (r < n) {
      const double old = bnd[bin + (long)ca * n + r];
      const double x = stopped ? old : (double)prev[q];  // a stopped chunk's last line keeps its value
      dm = fmax(dm, fabs(x - old));
      xm = fmax(xm, fabs(x));
      bnd[bout + (long)ca * n + r] = x;
    }
  }
  if (final) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      dm = fmax(dm, __shfl_xor(dm, o, 64));
      xm = fmax(xm, __shfl_xor(xm, o, 64));
    }
    if (l == 0 && ca + 1 < nch) {  // the last chunk's boundary line feeds no other chunk
      const double rel = dm / fmax(xm, 1e-300);
      atomicMax(reinterpret_cast<unsigned long long*>(cs.stat + ST_AFCH),
                (unsigned long long)__double_as_longlong(rel));
    }
  }
}

// J-line-major copies for the J-line sweep: the right-hand sides and the stage input (ghosts
// included, after BCU) of both components, 32 x 32 tiles through LDS; block table (case, tile).
__global__ void __launch_bounds__(256) k2_tr_in(const Case2D* __restrict__ tab,
                                                const int* __restrict__ bcase,
                                                const int* __restrict__ btile) {
  __shared__ double tile[4][32][33];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int tx = (ni + 1 + 31) / 32, t = btile[blockIdx.x];
  const int i0 = (t % tx) * 32, j0 = (t / tx) * 32;
  const gptr<double> src[4] = {glb(cs.f[F_SU1]), glb(cs.f[F_SV1]), glb(cs.f[F_U]), glb(cs.f[F_V])};
  const gptr<double> dst[4] = {glb(cs.f[F_TQU]), glb(cs.f[F_TQV]), glb(cs.f[F_TXU]), glb(cs.f[F_TXV])};
  const int l = threadIdx.x;
  for (int r = 0; r < 4; ++r) {
    const int li = l % 32, lj = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 4; ++a) tile[a][lj][li] = src[a][i + (long)sj * j];
    }
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int lj = l % 32, li = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 4; ++a) dst[a][j + (long)(nj + 1) * i] = tile[a][lj][li];
    }
  }
}

// The J-line sweep's result back into the field interior.
__global__ void __launch_bounds__(256) k2_tr_out(const Case2D* __restrict__ tab,
                                                 const int* __restrict__ bcase,
                                                 const int* __restrict__ btile) {
  __shared__ double tile[2][32][33];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int tx = (ni + 1 + 31) / 32, t = btile[blockIdx.x];
  const int i0 = (t % tx) * 32, j0 = (t / tx) * 32;
  const gptr<double> src[2] = {glb(cs.f[F_TXU]), glb(cs.f[F_TXV])};
  const gptr<double> dst[2] = {glb(cs.f[F_U]), glb(cs.f[F_V])};
  const int l = threadIdx.x;
  for (int r = 0; r < 4; ++r) {
    const int lj = l % 32, li = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 2; ++a) tile[a][lj][li] = src[a][j + (long)(nj + 1) * i];
    }
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int li = l % 32, lj = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i >= 1 && i <= ni - 1 && j >= 1 && j <= nj - 1) {
#pragma unroll
      for (int a = 0; a < 2; ++a) dst[a][i + (long)sj * j] = tile[a][lj][li];
    }
  }
}

// The change form's J-line-major copy of the stage-2 residual (FP32), as k2_tr_in.
__global__ void __launch_bounds__(256) k2_tr_in32(const Case2D* __restrict__ tab,
                                                  const int* __restrict__ bcase,
                                                  const int* __restrict__ btile) {
  __shared__ float tile[2][32][33];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int tx = (ni + 1 + 31) / 32, t = btile[blockIdx.x];
  const int i0 = (t % tx) * 32, j0 = (t / tx) * 32;
  const gptr<float> src[2] = {glb(cs.h[H_RU]), glb(cs.h[H_RV])};
  const gptr<float> dst[2] = {glb(cs.h[H_TRU]), glb(cs.h[H_TRV])};
  const int l = threadIdx.x;
  for (int r = 0; r < 4; ++r) {
    const int li = l % 32, lj = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 2; ++a) tile[a][lj][li] = src[a][i + (long)sj * j];
    }
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int lj = l % 32, li = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 2; ++a) dst[a][tidx32(nj, i, j)] = tile[a][lj][li];
    }
  }
}

// The change form's J-line sweep result: the J-line-major change added to the field interior.
__global__ void __launch_bounds__(256) k2_tr_out32(const Case2D* __restrict__ tab,
                                                   const int* __restrict__ bcase,
                                                   const int* __restrict__ btile) {
  __shared__ float tile[2][32][33];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int tx = (ni + 1 + 31) / 32, t = btile[blockIdx.x];
  const int i0 = (t % tx) * 32, j0 = (t / tx) * 32;
  const gptr<float> src[2] = {glb(cs.h[H_DU]), glb(cs.h[H_DV])};
  const gptr<double> dst[2] = {glb(cs.f[F_U]), glb(cs.f[F_V])};
  const int l = threadIdx.x;
  for (int r = 0; r < 4; ++r) {
    const int lj = l % 32, li = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= ni && j <= nj) {
#pragma unroll
      for (int a = 0; a < 2; ++a) tile[a][lj][li] = src[a][tidx32(nj, i, j)];
    }
  }
  __syncthreads();
  for (int r = 0; r < 4; ++r) {
    const int li = l % 32, lj = l / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i >= 1 && i <= ni - 1 && j >= 1 && j <= nj - 1) {
#pragma unroll
      for (int a = 0; a < 2; ++a) {
        const long c = i + (long)sj * j;
        dst[a][c] = dst[a][c] + (double)tile[a][lj][li];
      }
    }
  }
}

// The first pass's chunk boundary lines (the stage input's last line of every chunk) and BCU's
// ghosts of the stage input (by each (case, component)'s chunk-0 wavefront; the ghosts and the
// interior lines copied here do not overlap); block table as the passes'.
template <int DIR, bool DELTA>
__global__ void __launch_bounds__(64) k2_sweep_seed(const Case2D* __restrict__ tab,
                                                    const int* __restrict__ bcase,
                                                    const int* __restrict__ btile) {
  const int cc = bcase[blockIdx.x];
  if (cc < 0) return;
  const Case2D& cs = tab[cc];
  const int t = btile[blockIdx.x];
  const int comp = t & 1, ca = t >> 1;
  const auto F = glb(cs.f[comp ? F_V : F_U]);
  if (ca == 0)
    bcu_ghosts_wave(F, cs.ni, cs.nj, cs.sj, glb(cs.b[comp ? B_VWALL : B_UWALL]),
                    glb(cs.b[comp ? B_VFAR : B_UFAR]));
  const int n = DIR == 0 ? cs.ni - 1 : cs.nj - 1;
  const int nl = DIR == 0 ? cs.nj - 1 : cs.ni - 1;
  const long rs = DIR == 0 ? 1 : cs.sj, ls = DIR == 0 ? cs.sj : 1;
  const int nch = sw_nchunk(nl);
  const int a1 = min((ca + 1) * SW_C, nl);
  const auto dst = glb(cs.swb) + ((long)comp * nch + ca) * n;  // slot 0
  // the change form starts from no change
  for (int r = threadIdx.x; r < n; r += 64) dst[r] = DELTA ? 0.0 : F[ls * a1 + rs * (r + 1)];
  if (DIR == 0 && comp == 0 && ca == 0 && threadIdx.x == 0) cs.stat[ST_AFCH] = 0.0;
}

// One pass over the chunks of every case (block table: case, 2 chunk + component; entries with
// case -1 pad the table to whole XCD rounds).
template <int DIR, bool EARLY, typename T>
__global__ void __launch_bounds__(64) k2_sweep_pass(const Case2D* __restrict__ tab,
                                                    const int* __restrict__ bcase,
                                                    const int* __restrict__ btile, int slot,
                                                    double tol, int final) {
  const int cc = bcase[blockIdx.x];
  if (cc < 0) return;
  const Case2D& cs = tab[cc];
  const int t = btile[blockIdx.x];
  const int comp = t & 1, ca = t >> 1;
  const int n = DIR == 0 ? cs.ni - 1 : cs.nj - 1;
  switch ((n + 63) >> 6) {  // the case's own rows per lane
    case 1: sweep_chunk<1, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 2: sweep_chunk<2, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 3: sweep_chunk<3, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 4: sweep_chunk<4, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 5: sweep_chunk<5, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 6: sweep_chunk<6, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    case 7: sweep_chunk<7, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
    default: sweep_chunk<8, DIR, EARLY, T>(cs, comp, ca, slot, tol, final); break;
  }
}

// Stage-2 right-hand side SU = 2 Re DJAC / dt U + DD(U) with the stage-2 cross weights
// (iter_u.f90:84-98) and BCU's folds with the stage-2 AW(1), AE(NI-1) (bcu.f90:15, 34).  AF32:
// instead of SU, the FP64 residual of the stage input against the stage-2 line operator, in FP32.
template <bool AF32>
__global__ void __launch_bounds__(512) k2_rhs2(const Case2D* __restrict__ tab,
                                               const int* __restrict__ bcase,
                                               const int* __restrict__ btile) {
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  if (i > ni - 1 || j > nj - 1) return;
  const long c = (long)i + (long)sj * j;
  const auto U = glb(cs.f[F_U]);
  const auto V = glb(cs.f[F_V]);
  // the change form leaves out the 2 Re DJAC / dt U term: it cancels against the diagonal
  const double ap0 = AF32 ? 0.0 : glb(cs.f[G_AP0])[c];
  double ru = ap0 * U[c] + cross2(glb(cs.f[G_JXE2]), U, c, sj);
  double rv = ap0 * V[c] + cross2(glb(cs.f[G_JXE2]), V, c, sj);
  if (i == 1) {  // (change form: 2 AW(1) (U_F - U))
    ru = ru + 2.0 * glb(cs.b[B_AW2W])[j] * (AF32 ? glb(cs.b[B_UWALL])[j] - U[c] : glb(cs.b[B_UWALL])[j]);
    rv = rv + 2.0 * glb(cs.b[B_AW2W])[j] * (AF32 ? glb(cs.b[B_VWALL])[j] - V[c] : glb(cs.b[B_VWALL])[j]);
  }
  if (i == ni - 1) {
    ru = ru + 2.0 * glb(cs.b[B_AE2F])[j] * (AF32 ? glb(cs.b[B_UFAR])[j] - U[c] : glb(cs.b[B_UFAR])[j]);
    rv = rv + 2.0 * glb(cs.b[B_AE2F])[j] * (AF32 ? glb(cs.b[B_VFAR])[j] - V[c] : glb(cs.b[B_VFAR])[j]);
  }
  const double ae = glb(cs.f[G_AE2])[c];
  if (AF32) {
    // the residual r = SU + AW x0(I-1) + AE x0(I+1) + AS x0(J-1) + AN x0(J+1) - AP x0 of the
    // stage input x0 = U as differences (AP = 2 Re DJAC / dt + AE + AW + AN + AS + the folds
    // above; folded AW(1) = AE(NI-1) = 0, AW(I) = AE(I-1), AS(J) = AN(J-1) below J = 1); the
    // line ends' ghosts J = 0, NJ are BCU's periodic images J = NJ-1, 1, as the sweep sees them
    const double aw = i == 1 ? 0.0 : glb(cs.f[G_AE2])[c - 1];
    const double an = glb(cs.f[G_GC2])[c];
    const double as = j == 1 ? glb(cs.f[G_GA2])[c] : glb(cs.f[G_GC2])[c - sj];
    const long cs0 = j == 1 ? (long)i + (long)sj * (nj - 1) : c - sj;
    const long cn0 = j == nj - 1 ? (long)i + sj : c + sj;
    const double eu = ru + ae * (U[c + 1] - U[c]) + aw * (U[c - 1] - U[c]) +
                      as * (U[cs0] - U[c]) + an * (U[cn0] - U[c]);
    const double ev = rv + ae * (V[c + 1] - V[c]) + aw * (V[c - 1] - V[c]) +
                      as * (V[cs0] - V[c]) + an * (V[cn0] - V[c]);
    glb(cs.h[H_RU])[c] = (float)eu;
    glb(cs.h[H_RV])[c] = (float)ev;
  } else {
    // the old next line's coupling AE U(I+1) (the far-field fold leaves AE(NI-1) = 0)
    ru = ru + ae * U[c + 1];
    rv = rv + ae * V[c + 1];
    glb(cs.f[F_SU1])[c] = ru;
    glb(cs.f[F_SV1])[c] = rv;
  }
}

// BCU ghosts written by the interior thread that owns their value after a full BCU: the wall
// and far-field rows, the branch-cut columns and the corners (bcu.f90:10-51)
template <class P, class B>
__device__ inline void own_ghosts(P F, const Case2D& cs, int i, int j, long c, double v, B uwall,
                                  B ufar) {
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  double gw = 0.0, gf = 0.0;
  if (i == 1) { gw = -v + 2.0 * uwall[j]; F[c - 1] = gw; }
  if (i == ni - 1) { gf = -v + 2.0 * ufar[j]; F[c + 1] = gf; }
  if (j == nj - 1) {  // J = 0 <- NJ-1
    const long g = c - (long)sj * (nj - 1);
    F[g] = v;
    if (i == 1) F[g - 1] = gw;
    if (i == ni - 1) F[g + 1] = gf;
  }
  if (j == 1) {  // J = NJ <- 1
    const long g = c + (long)sj * (nj - 1);
    F[g] = v;
    if (i == 1) F[g - 1] = gw;
    if (i == ni - 1) F[g + 1] = gf;
  }
}

// Predictor correction U = U - S_U DTIME / DJAC (Main.f90:80-84) and the BCU ghosts
// (Main.f90:86-87).
__global__ void __launch_bounds__(512) k2_correct(const Case2D* __restrict__ tab,
                                                  const int* __restrict__ bcase,
                                                  const int* __restrict__ btile) {
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  if (i > ni - 1 || j > nj - 1) return;
  const long c = (long)i + (long)sj * j;
  const double dj = glb(cs.f[G_DJAC])[c];
  const auto U = glb(cs.f[F_U]);
  const auto V = glb(cs.f[F_V]);
  const double u = U[c] - glb(cs.f[F_SU])[c] * cs.dt / dj;
  const double v = V[c] - glb(cs.f[F_SV])[c] * cs.dt / dj;
  U[c] = u;
  V[c] = v;
  own_ghosts(U, cs, i, j, c, u, glb(cs.b[B_UWALL]), glb(cs.b[B_UFAR]));
  own_ghosts(V, cs, i, j, c, v, glb(cs.b[B_VWALL]), glb(cs.b[B_VFAR]));
}

// CONTRA_VEL of the corrected velocity and the source of ITER_P (iter_p.f90:29-32):
// RHS = 1 / DTIME (UE - UW + VN - VS)
__global__ void __launch_bounds__(512) k2_prhs(const Case2D* __restrict__ tab,
                                               const int* __restrict__ bcase,
                                               const int* __restrict__ btile) {
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  if (i > ni - 1 || j > nj - 1) return;
  const long c = (long)i + (long)sj * j;
  const auto U = glb(cs.f[F_U]);
  const auto V = glb(cs.f[F_V]);
  const GeomG g = geom2(cs);
  const double j1 = ju1(g, U, V, c), j2 = ju2(g, U, V, c);
  const double ue = 0.5 * (j1 + ju1(g, U, V, c + 1));
  const double uw = i == 1 ? 0.0 : 0.5 * (j1 + ju1(g, U, V, c - 1));
  const double vn = 0.5 * (j2 + ju2(g, U, V, c + sj));
  const double vs = 0.5 * (j2 + ju2(g, U, V, c - sj));
  glb(cs.f[F_RHS])[c] = 1.0 / cs.dt * (ue - uw + vn - vs);
}

// BCP (bcp.f90:12-43) of p: wall Neumann p(0) = p(1), outflow cubic extrapolation
// p(NI) = 3 p(NI-1) - 3 p(NI-2) + p(NI-3), branch cut.  One workgroup per case; skipped for a
// converged case unless force.
__global__ void __launch_bounds__(256) k2_bcp(const Case2D* __restrict__ tab, int force) {
  const Case2D& cs = tab[blockIdx.x];
  if (!force && C2_DONE(cs)) return;
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const auto P = glb(cs.f[F_P]);
  for (int j = 1 + threadIdx.x; j <= nj - 1; j += blockDim.x) {
    const long c = (long)sj * j;
    P[c] = P[c + 1];
    P[c + ni] = 3.0 * P[c + ni - 1] - 3.0 * P[c + ni - 2] + P[c + ni - 3];
  }
  __syncthreads();
  for (int i = threadIdx.x; i <= ni; i += blockDim.x) {
    P[i] = P[i + (long)sj * (nj - 1)];
    P[i + (long)sj * nj] = P[i + sj];
  }
}

// Residual of ITER_P's operator (iter_p.f90:12-51 with BCP's folds AP(1) -= AW(1), AW(1) = 0,
// AE(NI-1) = 0): r = AE p_E + AW p_W + AN p_N + AS p_S + SU - AP p, SU = RHS - DD(p), into the
// compact array R [(i-1) + (NI-1)(j-1)], with the block's sums of |r| and |AP p|.
template <typename T>
__global__ void __launch_bounds__(512) k2_residual(const Case2D* __restrict__ tab,
                                                   const int* __restrict__ bcase,
                                                   const int* __restrict__ btile) {
  __shared__ double sh[512];  // up to 64 x 8 threads
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  if (C2_DONE(cs)) return;  // uniform over the workgroup
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  double ra = 0.0, rb = 0.0;
  if (i <= ni - 1 && j <= nj - 1) {
    const long c = (long)i + (long)sj * j;
    const auto P = glb(cs.f[F_P]);
    const auto de = glb(cs.f[G_DE]);
    const auto dn = glb(cs.f[G_DN]);
    const double dE = de[c], dW = de[c - 1], dN = dn[c], dS = dn[c - sj];
    double ap = -dE - dW - dN - dS;
    double ae = -dE, aw = -dW;
    const double an = -dN, as = -dS;
    if (i == 1) { ap = ap - aw; aw = 0.0; }
    if (i == ni - 1) ae = 0.0;
    const double su = glb(cs.f[F_RHS])[c] - cross2(glb(cs.f[G_JQ12]), P, c, sj);
    const double r = ae * P[c + 1] + aw * P[c - 1] + an * P[c + sj] + as * P[c - sj] + su -
                     ap * P[c];
    ra = fabs(r);
    rb = fabs(ap * P[c]);
    glb((T*)(sizeof(T) == 8 ? cs.R : cs.R32))[(i - 1) + (long)(ni - 1) * (j - 1)] = (T)r;
  }
  ra = wg_sum(ra, sh);
  rb = wg_sum(rb, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    cs.part[2 * t.tile] = ra;
    cs.part[2 * t.tile + 1] = rb;
  }
}

// Per case: RESIP = sum|r| / sum|AP p| from the tile sums (fixed order); pass q >= 1 ends the
// case's correction when RESIP <= RESIMAX (the reference tests after a sweep,
// iter_p.f90:36-51) or when q is the last pass; q = 0 only records the starting residual.
__global__ void __launch_bounds__(256) k2_res_final(const Case2D* __restrict__ tab, int q,
                                                    int last) {
  __shared__ double sh[256];
  const Case2D& cs = tab[blockIdx.x];
  if (C2_DONE(cs)) return;
  double a = 0.0, b = 0.0;
  for (int k = threadIdx.x; k < cs.ntile; k += blockDim.x) {
    a = a + cs.part[2 * k];
    b = b + cs.part[2 * k + 1];
  }
  a = wg_sum(a, sh);
  b = wg_sum(b, sh);
  if (threadIdx.x != 0) return;
  const double res = scaled_residual(a, b);
  double* st = cs.stat;
  st[ST_S0] = a;
  st[ST_S1] = b;
  st[ST_RESIP] = res;
  st[ST_NPASS] = (double)q;
  if (q == 0) st[ST_RES0] = res;
  if (!(res == res)) st[ST_NAN] = 1.0;
  if (q >= 1 && (res <= cs.resimax || q >= last || !(res == res))) st[ST_DONE] = 1.0;
}

__global__ void k2_begin_p(const Case2D* __restrict__ tab) {
  const Case2D& cs = tab[blockIdx.x];
  if (threadIdx.x == 0) {
    cs.stat[ST_DONE] = 0.0;
    cs.stat[ST_NAN] = 0.0;
  }
}

// Forward real J transform of the residual (two radial rows per complex sequence) into the
// half spectrum S[i + n1 m], one workgroup per (case, tile of 2 ncol rows); see
// fvm_poisson_spec.hip for the 3D form.
template <typename T>
__global__ void __launch_bounds__(256) k2_fft_fwd(const Case2D* __restrict__ tab,
                                                  const int* __restrict__ bcase,
                                                  const int* __restrict__ btile, int ncol) {
  extern __shared__ __align__(16) unsigned char smem[];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  if (C2_DONE(cs)) return;
  const int N = cs.pj.n, MJ = cs.mj, n1 = cs.ni - 1;
  const auto R = glb((const T*)(sizeof(T) == 8 ? cs.R : cs.R32));
  const auto S = glb((T*)(sizeof(T) == 8 ? cs.S : cs.S32));          // complex, interleaved
  const auto twg = glb((const T*)(sizeof(T) == 8 ? cs.twj : cs.twj32));
  Cx<T>* a = (Cx<T>*)smem;
  Cx<T>* b = a + (size_t)ncol * N;
  Cx<T>* tw = b + (size_t)ncol * N;
  const int i0 = btile[blockIdx.x] * 2 * ncol;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = cx_ld<T>(twg, k);
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, j = q / ncol;
    const int i = i0 + 2 * col;
    Cx<T> z{(T)0, (T)0};
    if (i < n1) {
      const long o = (long)i + (long)n1 * j;
      z.x = R[o];
      z.y = i + 1 < n1 ? R[o + 1] : (T)0;
    }
    a[j * ncol + col] = z;
  }
  __syncthreads();
  const Cx<T>* Z = fft_lds<T, -1>(a, b, ncol, cs.pj, tw);
  for (int q = threadIdx.x; q < ncol * MJ; q += blockDim.x) {
    const int col = q % ncol, m = q / ncol;
    const int i = i0 + 2 * col;
    if (i >= n1) continue;
    const Cx<T> zm = Z[m * ncol + col], zc = Z[((N - m) % N) * ncol + col];
    const long o = (long)i + (long)n1 * m;
    cx_st<T>(S, o, Cx<T>{(T)0.5 * (zm.x + zc.x), (T)0.5 * (zm.y - zc.y)});
    if (i + 1 < n1) cx_st<T>(S, o + 1, Cx<T>{(T)0.5 * (zm.y + zc.y), (T)0.5 * (zc.x - zm.x)});
  }
}

// Radial mode lines of the half spectrum, one wavefront per (case, mode) from the block table.
template <typename T, int R>
__global__ void __launch_bounds__(64) k2_modes(const Case2D* __restrict__ tab,
                                               const int* __restrict__ bcase,
                                               const int* __restrict__ btile) {
  const Case2D& cs = tab[bcase[blockIdx.x]];
  if (C2_DONE(cs)) return;
  const int n = cs.ni - 1;
  const long base = (long)n * btile[blockIdx.x];
  const auto Z = glb((T*)(sizeof(T) == 8 ? cs.S : cs.S32));  // complex, interleaved
  const auto dW = glb((const double*)cs.pdw);
  const auto invd = glb((const double*)cs.invd);
  const auto bf = glb((const double*)cs.bf);
  const int l = threadIdx.x;
  const int r0 = l * R;
  double alx[R], aly[R], be[R], bb[R];
  double Ax = 0.0, Ay = 0.0, Bt = 1.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = r0 + q;
    if (r < n) {
      const double iv = invd[base + r];
      const Cx<T> z = cx_ld<T>(Z, base + r);
      alx[q] = (double)z.x * iv;
      aly[q] = (double)z.y * iv;
      be[q] = r == 0 ? 0.0 : -dW[r] * iv;
      bb[q] = bf[base + r];
    } else {
      alx[q] = 0.0; aly[q] = 0.0; be[q] = 1.0; bb[q] = 0.0;
    }
    Ax = alx[q] + be[q] * Ax;
    Ay = aly[q] + be[q] * Ay;
    Bt = be[q] * Bt;
  }
  double yx_in, yy_in;
  {
    double A2[2] = {Ax, Ay}, z2[2] = {0.0, 0.0}, o2[2];
    wave::prefix_apply<2>(A2, Bt, z2, o2);  // y(-1) = 0
    yx_in = o2[0];
    yy_in = o2[1];
  }
  double yx[R], yy[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    yx[q] = alx[q] + be[q] * yx_in;
    yy[q] = aly[q] + be[q] * yy_in;
    yx_in = yx[q];
    yy_in = yy[q];
  }
  double Ct = 1.0, Dx = 0.0, Dy = 0.0;
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    if (r0 + q < n) {
      Dx = yx[q] - bb[q] * Dx;
      Dy = yy[q] - bb[q] * Dy;
      Ct = -bb[q] * Ct;
    }
  }
  double xx_in, xy_in;
  {
    double D2[2] = {Dx, Dy}, z2[2] = {0.0, 0.0}, o2[2];
    wave::suffix_apply<2>(D2, Ct, z2, o2);  // x(n) = 0
    xx_in = o2[0];
    xy_in = o2[1];
  }
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    const int r = r0 + q;
    if (r < n) {
      const double xx = yx[q] - bb[q] * xx_in, xy = yy[q] - bb[q] * xy_in;
      xx_in = xx;
      xy_in = xy;
      cx_st<T>(Z, base + r, Cx<T>{(T)xx, (T)xy});
    }
  }
}

// Inverse J transform of the corrected spectrum, added to p (scale 1 / (NJ-1)).
template <typename T>
__global__ void __launch_bounds__(256) k2_fft_inv_add(const Case2D* __restrict__ tab,
                                                      const int* __restrict__ bcase,
                                                      const int* __restrict__ btile, int ncol,
                                                      int to_mu) {
  extern __shared__ __align__(16) unsigned char smem[];
  const Case2D& cs = tab[bcase[blockIdx.x]];
  if (C2_DONE(cs)) return;
  const int N = cs.pj.n, MJ = cs.mj, n1 = cs.ni - 1, sj = cs.sj;
  const auto S = glb((const T*)(sizeof(T) == 8 ? cs.S : cs.S32));  // complex, interleaved
  const auto twg = glb((const T*)(sizeof(T) == 8 ? cs.twj : cs.twj32));
  const auto P = glb(to_mu ? cs.a[A_MU] : cs.f[F_P]);  // the adjoint corrects its own solution
  Cx<T>* a = (Cx<T>*)smem;
  Cx<T>* b = a + (size_t)ncol * N;
  Cx<T>* tw = b + (size_t)ncol * N;
  const int i0 = btile[blockIdx.x] * 2 * ncol;
  for (int k = threadIdx.x; k < N; k += blockDim.x) tw[k] = cx_ld<T>(twg, k);
  for (int q = threadIdx.x; q < ncol * MJ; q += blockDim.x) {
    const int col = q % ncol, m = q / ncol;
    const int i = i0 + 2 * col;
    Cx<T> A{(T)0, (T)0}, B{(T)0, (T)0};
    if (i < n1) {
      const long o = (long)i + (long)n1 * m;
      A = cx_ld<T>(S, o);
      if (i + 1 < n1) B = cx_ld<T>(S, o + 1);
    }
    if (m == 0 || 2 * m == N) { A.y = (T)0; B.y = (T)0; }
    a[m * ncol + col] = Cx<T>{A.x - B.y, A.y + B.x};
    if (m > 0 && 2 * m < N) a[(N - m) * ncol + col] = Cx<T>{A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  const Cx<T>* Z = fft_lds<T, +1>(a, b, ncol, cs.pj, tw);
  const double scale = 1.0 / (double)N;
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, j = q / ncol;
    const int i = i0 + 2 * col;
    if (i >= n1) continue;
    const Cx<T> z = Z[j * ncol + col];
    const long c = (long)(i + 1) + (long)sj * (j + 1);
    P[c] = P[c] + scale * (double)z.x;
    if (i + 1 < n1) P[c + 1] = P[c + 1] + scale * (double)z.y;
  }
}

// P_SOURCE(p^{n+1}), the projection U = U + S_U DTIME / DJAC (Main.f90:93-103), BCU's ghosts
// and the block maxima of the local CFL number (calccoef.f90:9-17)
__global__ void __launch_bounds__(512) k2_project(const Case2D* __restrict__ tab,
                                                  const int* __restrict__ bcase,
                                                  const int* __restrict__ btile) {
  __shared__ double sh[512];  // up to 64 x 8 threads
  const TileIdx t = tile_index(bcase, btile, tab);
  if (t.c < 0) return;
  const Case2D& cs = tab[t.c];
  const int ni = cs.ni, nj = cs.nj, sj = cs.sj;
  const int i = t.i, j = t.j;
  double cfl = 0.0;
  if (i <= ni - 1 && j <= nj - 1) {
    const long c = (long)i + (long)sj * j;
    const GeomG g = geom2(cs);
    const auto P = glb(cs.f[F_P]);
    const double su = psrc(g.dj, g.xix, g.etx, P, c, sj);
    const double sv = psrc(g.dj, g.xiy, g.ety, P, c, sj);
    const auto U = glb(cs.f[F_U]);
    const auto V = glb(cs.f[F_V]);
    const double dj = g.dj[c];
    const double u = U[c] + su * cs.dt / dj;
    const double v = V[c] + sv * cs.dt / dj;
    U[c] = u;
    V[c] = v;
    glb(cs.f[F_SU])[c] = su;
    glb(cs.f[F_SV])[c] = sv;
    own_ghosts(U, cs, i, j, c, u, glb(cs.b[B_UWALL]), glb(cs.b[B_UFAR]));
    own_ghosts(V, cs, i, j, c, v, glb(cs.b[B_VWALL]), glb(cs.b[B_VFAR]));
    cfl = cs.dt * (fabs(g.xix[c] * u + g.xiy[c] * v) + fabs(g.etx[c] * u + g.ety[c] * v));
  }
  cfl = wg_max(cfl, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) cs.part[2 * t.tile] = cfl;
}

// CALCCOEF (calccoef.f90:9-119): CFL maximum, C_pb from the corner average PG(1,1) (the far-field
// PG(NI,NJ/2) is not assigned by CALCCOEF: zero), and the pressure and friction drag and lift
// summed over the wall cells in J order, each sum carried by one lane; one workgroup per case.
// The step's scalars go to stat and to the history ring at slot *ctr.
__global__ void __launch_bounds__(256) k2_calccoef(const Case2D* __restrict__ tab, int ncase,
                                                   const long* __restrict__ ctr, int hcap,
                                                   double* __restrict__ hist) {
  __shared__ double sh[256];
  __shared__ double tm[4][256];
  const Case2D& cs = tab[blockIdx.x];
  const int nj = cs.nj, sj = cs.sj;
  const int t = threadIdx.x;
  double cm = 0.0;
  for (int k = t; k < cs.ntile; k += blockDim.x) cm = fmax(cm, cs.part[2 * k]);
  cm = wg_max(cm, sh);
  const double* P = cs.f[F_P];
  const double* U = cs.f[F_U];
  const double* V = cs.f[F_V];
  const double* w = cs.b[B_WALL];
  double acc = 0.0;
  for (int j0 = 1; j0 <= nj - 1; j0 += 256) {
    const int j = j0 + t;
    if (j <= nj - 1) {
      const double* o = w + 6 * j;
      const double dn1 = o[0], s2 = o[1], c2 = o[2], area = o[3], ct = o[4], st = o[5];
      const long c = 1 + (long)sj * j;
      const double dudn = (U[c] * s2 - V[c] * c2 - 0.0) / dn1;  // VEL_ROT = 0
      tm[0][t] = 2.0 * P[c] * area * ct;
      tm[1][t] = 2.0 * P[c] * area * st;
      tm[2][t] = (2.0 / cs.re) * dudn * area * st;
      tm[3][t] = (2.0 / cs.re) * dudn * area * ct;
    }
    __syncthreads();
    if (t < 4) {
      const int m = min(256, nj - j0);
      if (t == 2)
        for (int q = 0; q < m; ++q) acc = acc + tm[2][q];
      else
        for (int q = 0; q < m; ++q) acc = acc - tm[t][q];
    }
    __syncthreads();
  }
  if (t < 4) sh[t] = acc;
  __syncthreads();
  if (t != 0) return;
  const double pd = sh[0], pl = sh[1], td = sh[2], tl = sh[3];
  const double dum_b = 0.25 * (P[1 + sj] + P[1] + P[sj] + P[0]);
  const double cpb = 2.0 * (0.0 - dum_b);
  double* st = cs.stat;
  st[ST_CFL] = cm;
  st[ST_CPB] = cpb;
  st[ST_CD] = pd + td;
  st[ST_CL] = pl + tl;
  st[ST_CDP] = pd;
  st[ST_CDV] = td;
  st[ST_CLP] = pl;
  st[ST_CLV] = tl;
  double* h = hist + ((long)(*ctr % hcap) * ncase + blockIdx.x) * 8;
  h[0] = st[ST_CD];
  h[1] = st[ST_CL];
  h[2] = st[ST_CDP];
  h[3] = st[ST_CDV];
  h[4] = st[ST_CLP];
  h[5] = st[ST_CLV];
  h[6] = st[ST_RESIP];
  h[7] = st[ST_CFL];
}

__global__ void k2_tick(long* ctr) {
  if (threadIdx.x == 0) *ctr = *ctr + 1;
}

}  // namespace

#include "fvm2d_batch.inc"
#include "fvm2d_adjoint.inc"
