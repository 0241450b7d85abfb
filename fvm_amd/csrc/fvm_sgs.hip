// Dynamic Smagorinsky subgrid-scale model on the device
// (LES/v2.3_LES-Reynolds stresses/stress_tensor_dy.f90:24-286 velocity part, test_filter.f90).
//
// Passes: (1) strain tensor S_ij and |S| on i = 0..NI (one-sided at the I boundaries), periodic
// ghosts; (2) test-filtered velocity (1,2,1)^3/64 with the velocity boundary ghosts;
// (3) per cell the Germano identity L_ij, the Lilly tensor M_ij and the contractions LM, MM,
// with every filtered product recomputed from its 27 neighbours instead of stored, summed per
// workgroup over the cells with LM/MM >= 0; (4) the reference's global conditional average
// C = <LM>/<2MM> from those partial sums (a deterministic two-level FP64 reduction, no per-cell
// LM, MM arrays); (5) eddy viscosity and the divergence of the SGS
// stress added to the convective term, NL += J^-1 SGS (RLES/main.f90:61-66).
#include "fvm_ctx.h"

namespace {

const dim3 kBlk(64, 4, 1);

// Strain S_ij and |S| of one cell (stress_tensor_dy.f90:27-76 central, :96-197 one-sided at
// I = 0 / NI); the differences are passed in, so the LDS-staged and the global-memory callers
// share one expression order.
struct Diff3 {
  double i, j, k;
};

template <bool XZ0>
__device__ inline void strain_store(const double* const* __restrict__ m, long c, bool onesided,
                                    Diff3 u, Diff3 v, Diff3 w, double* const* __restrict__ L,
                                    double* mag) {
  const double xix = m[0][c], etx = m[1][c], ztx = m[2][c];
  const double xiy = m[3][c], ety = m[4][c], zty = m[5][c];
  const double xiz = XZ0 ? 0.0 : m[6][c], etz = XZ0 ? 0.0 : m[7][c], ztz = m[8][c];
  double l11, l12, l13, l22, l23, l33;
  if (!onesided) {
    double t;
    t = xix * u.i + etx * u.j + ztx * u.k;
    l11 = 0.5 * t;
    t = xiy * u.i + ety * u.j + zty * u.k + xix * v.i + etx * v.j + ztx * v.k;
    l12 = 0.25 * t;
    t = xiz * u.i + etz * u.j + ztz * u.k + xix * w.i + etx * w.j + ztx * w.k;
    l13 = 0.25 * t;
    t = xiy * v.i + ety * v.j + zty * v.k;
    l22 = 0.5 * t;
    t = xiz * v.i + etz * v.j + ztz * v.k + xiy * w.i + ety * w.j + zty * w.k;
    l23 = 0.25 * t;
    t = xiz * w.i + etz * w.j + ztz * w.k;
    l33 = 0.5 * t;
  } else {
    l11 = 0.5 * (2.0 * xix * u.i + etx * u.j + ztx * u.k);
    l12 = 0.25 * (2.0 * xiy * u.i + ety * u.j + zty * u.k + 2.0 * xix * v.i + etx * v.j +
                  ztx * v.k);
    l13 = 0.25 * (2.0 * xiz * u.i + etz * u.j + ztz * u.k + 2.0 * xix * w.i + etx * w.j +
                  ztx * w.k);
    l22 = 0.5 * (2.0 * xiy * v.i + ety * v.j + zty * v.k);
    l23 = 0.25 * (2.0 * xiz * v.i + etz * v.j + ztz * v.k + 2.0 * xiy * w.i + ety * w.j +
                  zty * w.k);
    l33 = 0.5 * (2.0 * xiz * w.i + etz * w.j + ztz * w.k);
  }
  L[0][c] = l11; L[1][c] = l12; L[2][c] = l13; L[3][c] = l22; L[4][c] = l23; L[5][c] = l33;
  double s = l11 * l11 + l12 * l12 + l13 * l13 + l12 * l12 + l22 * l22 + l23 * l23 +
             l13 * l13 + l23 * l23 + l33 * l33;
  mag[c] = sqrt(2.0 * s);
}

// Strain on the I ghost planes i = 0 and i = NI (one-sided radial difference), interior j, k.
__global__ void k_strain_iplanes(Dims d, const double* const* __restrict__ m,
                                 const double* __restrict__ U, const double* __restrict__ V,
                                 const double* __restrict__ W, double* const* __restrict__ L,
                                 double* mag, const double* __restrict__ T, double* dtx,
                                 double* dty, double* dtz) {
  const long nl = (long)(d.nj - 1) * (d.nk - 1);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= 2 * nl) return;
  const bool lo = t < nl;
  const long r = lo ? t : t - nl;
  const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
  const long c = cidx(d, lo ? 0 : d.ni, j, k);
  const long sj = d.sj, sk = d.sk;
  auto diff = [&](const double* F) {
    Diff3 q;
    q.i = lo ? (F[c + 1] - F[c]) : (F[c] - F[c - 1]);
    q.j = F[c + sj] - F[c - sj];
    q.k = F[c + sk] - F[c - sk];
    return q;
  };
  strain_store<false>(m, c, true, diff(U), diff(V), diff(W), L, mag);
  if (T) {  // one-sided temperature gradient (stress_tensor_dy.f90:138-146, 188-196)
    const Diff3 g = diff(T);
    dtx[c] = 0.5 * (2.0 * gp(m[0])[c] * g.i + gp(m[1])[c] * g.j + gp(m[2])[c] * g.k);
    dty[c] = 0.5 * (2.0 * gp(m[3])[c] * g.i + gp(m[4])[c] * g.j + gp(m[5])[c] * g.k);
    dtz[c] = 0.5 * (2.0 * gp(m[6])[c] * g.i + gp(m[7])[c] * g.j + gp(m[8])[c] * g.k);
  }
}

// Strain on the interior cells fused with the (1,2,1)^3/64 test filter of u, v, w
// (test_filter.f90:13-30).  A workgroup owns a 64 x 4 (I x J) column of cells over SF_KC
// consecutive K planes and marches up K with the last three planes of u, v, w (one-cell halo)
// in LDS, so every velocity value is fetched from HBM once per column.  The filter sums its 27
// terms in the reference's order: K-1 plane, K plane, K+1 plane; rows J-1, J, J+1; columns
// I-1, I, I+1.
// With NF = 4 the temperature rides along: its gradient dT/dx_i = 0.5 (xi_i dT/dxi +
// eta_i dT/deta + zeta_i dT/dzeta) (stress_tensor_dy.f90:78-94) and its test filter T^.
constexpr int SF_KC = 16;

struct SfOut {
  const double* in[4];  // u, v, w, T
  double* hat[4];       // u^, v^, w^, T^
  double* dt[3];        // dT/dx, dT/dy, dT/dz
};

template <int NF, bool XZ0>
__global__ void __launch_bounds__(256) k_strain_filter(Dims d, const double* const* __restrict__ m,
                                                       SfOut io, double* const* __restrict__ L,
                                                       double* mag) {
  const uint3 bb = xcd_block();
  __shared__ double sh[NF][3][SF_P];  // [field][plane slot][tile cell]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * SF_KC;
  const int k1 = min(k0 + SF_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int q0 = (ty + 1) * SF_W + tx + 1;
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) % 3;
    __syncthreads();  // plane kp - 3, held in this slot, is no longer read
    for (int q = tid; q < SF_P; q += 256) {
      const int gi = min(bb.x * 64 + q % SF_W, d.ni);
      const int gj = min(bb.y * 4 + q / SF_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
#pragma unroll
      for (int f = 0; f < NF; ++f) sh[f][slot][q] = io.in[f][c];
    }
    __syncthreads();
    if (kp < k0 + 1 || !act) continue;
    // cell (i, j, kp - 1): planes k-1, k, k+1 sit in slots slot+1, slot+2, slot (mod 3)
    const int k = kp - 1;
    const int sm = (slot + 1) % 3, s0 = (slot + 2) % 3, sp = slot;
    const long c = cidx(d, i, j, k);
    Diff3 g[3];
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      g[f].i = sh[f][s0][q0 + 1] - sh[f][s0][q0 - 1];
      g[f].j = sh[f][s0][q0 + SF_W] - sh[f][s0][q0 - SF_W];
      g[f].k = sh[f][sp][q0] - sh[f][sm][q0];
    }
    strain_store<XZ0>(m, c, false, g[0], g[1], g[2], L, mag);
    if (NF == 4) {
      const double ti = sh[3][s0][q0 + 1] - sh[3][s0][q0 - 1];
      const double tj = sh[3][s0][q0 + SF_W] - sh[3][s0][q0 - SF_W];
      const double tk = sh[3][sp][q0] - sh[3][sm][q0];
      double t = gp(m[0])[c] * ti + gp(m[1])[c] * tj + gp(m[2])[c] * tk;
      io.dt[0][c] = 0.5 * t;
      t = gp(m[3])[c] * ti + gp(m[4])[c] * tj + gp(m[5])[c] * tk;
      io.dt[1][c] = 0.5 * t;
      t = mget<XZ0>(m, 6, c) * ti + mget<XZ0>(m, 7, c) * tj + gp(m[8])[c] * tk;
      io.dt[2][c] = 0.5 * t;
    }
    const int pl[3] = {sm, s0, sp};
#pragma unroll
    for (int f = 0; f < NF; ++f) {
      double t = 0.0;
      bool first = true;
#pragma unroll
      for (int dk = -1; dk <= 1; ++dk)
#pragma unroll
        for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
          for (int di = -1; di <= 1; ++di) {
            const int wgt = (di == 0 ? 2 : 1) * (dj == 0 ? 2 : 1) * (dk == 0 ? 2 : 1);
            const double v = sh[f][pl[dk + 1]][q0 + dj * SF_W + di];
            const double term = wgt == 1 ? v : (double)wgt * v;
            t = first ? term : t + term;
            first = false;
          }
      io.hat[f][c] = t / 64.0;
    }
  }
}

// Germano/Lilly contractions LM and MM per interior cell (stress_tensor_dy.f90:236-483).
// The twelve test-filtered products (u_i u_j and |S| S_ij) are never stored: a workgroup owns a
// 64 x 4 (I x J) column of cells over LM_KC consecutive K planes, stages each plane of the ten
// factor fields (with a one-cell halo) in LDS (|S| formed from the six S_ij as it is staged, so
// only nine fields are read), forms the in-plane (1,2,1)x(1,2,1) sums of the
// products once per plane and combines three consecutive planes with (1,2,1) weights.  The sum
// is the reference's 27-point filter with its terms grouped by plane (rounding-level
// differences only).
// Latency: the loads of plane k+1's tile and of the finishing cell's own operands (metrics,
// test-filtered velocity stencil, Delta^2) are issued before plane k's in-plane sums are formed,
// so they arrive while the LDS work runs; the three-plane combination is carried as two running
// sums, X = P(k-1) + 2 P(k) for the cell finished next and Y = P(k) for the one after.
constexpr int LM_KC = 16;
constexpr int SD_KC = 16;  // planes per workgroup of the K-marching SGS divergences
constexpr int LM_W = 66, LM_H = 6;  // tile with halo
constexpr int LM_NQ = (LM_W * LM_H + 255) / 256;  // tile cells staged per thread

template <bool XZ0>
__global__ void __launch_bounds__(256, 2) k_lm(Dims d, const double* const* __restrict__ m,
                                            const double* __restrict__ U,
                                            const double* __restrict__ V,
                                            const double* __restrict__ W,
                                            const double* __restrict__ UH,
                                            const double* __restrict__ VH,
                                            const double* __restrict__ WH,
                                            const double* const* __restrict__ L,
                                            const double* __restrict__ filt2, double* part,
                                            double* magh_out, int kc) {
  const uint3 bb = xcd_block();
  __shared__ double sh[10][LM_W * LM_H];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * kc;
  const int k1 = min(k0 + kc - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const GP L0 = gp(L[0]), L1 = gp(L[1]), L2 = gp(L[2]);
  const GP L3 = gp(L[3]), L4 = gp(L[4]), L5 = gp(L[5]);
  const long sj = d.sj, sk = d.sk;
  // staged tile values of the next plane: u v w S11 S12 S13 S22 S23 S33 per tile cell
  double tv[LM_NQ][9];
  auto load_tile = [&](int kp) {
#pragma unroll
    for (int h = 0; h < LM_NQ; ++h) {
      const int q = min(tid + 256 * h, LM_W * LM_H - 1);
      const int gi = min(bb.x * 64 + q % LM_W, d.ni);
      const int gj = min(bb.y * 4 + q / LM_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
      tv[h][0] = U[c]; tv[h][1] = V[c]; tv[h][2] = W[c];
      tv[h][3] = L0[c]; tv[h][4] = L1[c]; tv[h][5] = L2[c];
      tv[h][6] = L3[c]; tv[h][7] = L4[c]; tv[h][8] = L5[c];
    }
  };
  // running plane sums in product order uu vv ww uv uw vw, |S|S11 |S|S22 |S|S33 |S|S12 |S|S13
  // |S|S23
  double X[12], Y[12];
#pragma unroll
  for (int q = 0; q < 12; ++q) X[q] = Y[q] = 0.0;
  // this thread's sums of LM, MM and the cell count over the cells with 0.5 LM/MM >= 0
  double sa = 0.0, sb = 0.0, sn = 0.0;
  load_tile(k0 - 1);
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    __syncthreads();
#pragma unroll
    for (int h = 0; h < LM_NQ; ++h) {
      const int q = tid + 256 * h;
      if (q < LM_W * LM_H) {
        const double l11 = tv[h][3], l12 = tv[h][4], l13 = tv[h][5], l22 = tv[h][6],
                     l23 = tv[h][7], l33 = tv[h][8];
        sh[0][q] = tv[h][0]; sh[1][q] = tv[h][1]; sh[2][q] = tv[h][2];
        sh[4][q] = l11; sh[5][q] = l12; sh[6][q] = l13;
        sh[7][q] = l22; sh[8][q] = l23; sh[9][q] = l33;
        // |S| from the stored strain rates in strain_store's order: the value it wrote to |S|
        const double s2 = l11 * l11 + l12 * l12 + l13 * l13 + l12 * l12 + l22 * l22 +
                          l23 * l23 + l13 * l13 + l23 * l23 + l33 * l33;
        sh[3][q] = sqrt(2.0 * s2);
      }
    }
    __syncthreads();
    if (kp + 1 <= k1 + 1) load_tile(kp + 1);  // in flight during this plane's sums
    // operands of the cell finished in this iteration, (i, j, kp - 1)
    const bool fin = kp >= k0 + 1 && act;
    const long c = fin ? cidx(d, i, j, kp - 1) : cidx(d, 1, 1, 1);
    const double xix = gp(m[0])[c], etx = gp(m[1])[c], ztx = gp(m[2])[c];
    const double xiy = gp(m[3])[c], ety = gp(m[4])[c], zty = gp(m[5])[c];
    const double xiz = mget<XZ0>(m, 6, c), etz = mget<XZ0>(m, 7, c), ztz = gp(m[8])[c];
    double hs[3][6];  // u^, v^, w^ at i+1, i-1, j+1, j-1, k+1, k-1
    const double* hf[3] = {UH, VH, WH};
#pragma unroll
    for (int f = 0; f < 3; ++f) {
      const GP h = gp(hf[f]);
      hs[f][0] = h[c + 1]; hs[f][1] = h[c - 1];
      hs[f][2] = h[c + sj]; hs[f][3] = h[c - sj];
      hs[f][4] = h[c + sk]; hs[f][5] = h[c - sk];
    }
    const double uh = gp(UH)[c], vh = gp(VH)[c], wh = gp(WH)[c];
    const double fw2 = gp(filt2)[c];
    double Pn[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) Pn[q] = 0.0;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        const double wgt = (double)((di == 0 ? 2 : 1) * (dj == 0 ? 2 : 1));
        const int q = (ty + 1 + dj) * LM_W + tx + 1 + di;
        const double u = sh[0][q], v = sh[1][q], w = sh[2][q], mg = sh[3][q];
        Pn[0] += wgt * (u * u);
        Pn[1] += wgt * (v * v);
        Pn[2] += wgt * (w * w);
        Pn[3] += wgt * (u * v);
        Pn[4] += wgt * (u * w);
        Pn[5] += wgt * (v * w);
        Pn[6] += wgt * (mg * sh[4][q]);
        Pn[7] += wgt * (mg * sh[7][q]);
        Pn[8] += wgt * (mg * sh[9][q]);
        Pn[9] += wgt * (mg * sh[5][q]);
        Pn[10] += wgt * (mg * sh[6][q]);
        Pn[11] += wgt * (mg * sh[8][q]);
      }
    double F[12];
#pragma unroll
    for (int q = 0; q < 12; ++q) {
      F[q] = (X[q] + Pn[q]) / 64.0;  // (P(k-1) + 2 P(k)) + P(k+1)
      X[q] = Y[q] + 2.0 * Pn[q];
      Y[q] = Pn[q];
    }
    if (!fin) continue;
    // raw (unscaled) test-filtered strain sums; S11 = 0.5 r11, S12 = 0.25 r12, ...
    const double dui = hs[0][0] - hs[0][1], duj = hs[0][2] - hs[0][3], duk = hs[0][4] - hs[0][5];
    const double dvi = hs[1][0] - hs[1][1], dvj = hs[1][2] - hs[1][3], dvk = hs[1][4] - hs[1][5];
    const double dwi = hs[2][0] - hs[2][1], dwj = hs[2][2] - hs[2][3], dwk = hs[2][4] - hs[2][5];
    const double r11 = xix * dui + etx * duj + ztx * duk;
    const double r12 = xiy * dui + ety * duj + zty * duk + xix * dvi + etx * dvj + ztx * dvk;
    const double r13 = xiz * dui + etz * duj + ztz * duk + xix * dwi + etx * dwj + ztx * dwk;
    const double r22 = xiy * dvi + ety * dvj + zty * dvk;
    const double r23 = xiz * dvi + etz * dvj + ztz * dvk + xiy * dwi + ety * dwj + zty * dwk;
    const double r33 = xiz * dwi + etz * dwj + ztz * dwk;
    const double s11 = 0.5 * r11, s12 = 0.25 * r12, s13 = 0.25 * r13, s22 = 0.5 * r22,
                 s23 = 0.25 * r23, s33 = 0.5 * r33;
    const double dum = 2.0 * (s11 * s11 + s12 * s12 + s13 * s13 + s12 * s12 + s22 * s22 +
                              s23 * s23 + s13 * s13 + s23 * s23 + s33 * s33);
    const double mh = sqrt(dum);
    if (magh_out) magh_out[c] = mh;
    // reference component order: 11, 22, 33, 12, 13, 23
    const double hA[6] = {uh, vh, wh, uh, uh, vh};
    const double hB[6] = {uh, vh, wh, vh, wh, wh};
    const double raw[6] = {r11, r22, r33, r12, r13, r23};
    double lm = 0.0, mm = 0.0;
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      const double Lij = -F[q] + hA[q] * hB[q];
      const double mt = q < 3 ? 4.0 * mh * 0.5 * raw[q] : 4.0 * mh * 0.25 * raw[q];
      const double M = fw2 * (mt - F[6 + q]);
      if (q == 0) {
        lm = Lij * M;
        mm = M * M;
      } else if (q < 3) {
        lm = lm + Lij * M;
        mm = mm + M * M;
      } else {
        lm = lm + 2.0 * Lij * M;
        mm = mm + 2.0 * M * M;
      }
    }
    if (0.5 * lm / mm >= 0.0) {  // stress_tensor_dy.f90:486-508
      sa = sa + lm;
      sb = sb + mm;
      sn = sn + 1.0;
    }
  }
  // workgroup partial sums of the conditional average, reduced by k_lm_final
  __syncthreads();
  double* red = &sh[0][0];
  red[tid] = sa; red[256 + tid] = sb; red[512 + tid] = sn;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) {
      red[tid] += red[tid + st];
      red[256 + tid] += red[256 + tid + st];
      red[512 + tid] += red[512 + tid + st];
    }
    __syncthreads();
  }
  if (tid == 0) {
    const long bid = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    part[3 * bid] = red[0]; part[3 * bid + 1] = red[256]; part[3 * bid + 2] = red[512];
  }
}

// Germano identity of the subgrid heat flux per interior cell (stress_tensor_dy.f90:589-690):
// P_i = -(u_i T)^ + u_i^ T^, R_i = Delta^2 (4 |S^| dT^/dx_i - (|S| dT/dx_i)^), with
// dT^/dx_i = 0.5 (xi_i (T^(I+1) - T^(I-1)) + eta_i (..J..) + zeta_i (..K..)), and the sums
// PR = P_1 R_1 + P_2 R_2 + P_3 R_3, RR = R_1^2 + R_2^2 + R_3^2.  Same structure as k_lm: the six
// filtered products are never stored; each plane of u, v, w, T, |S|, dT/dx_i is staged in LDS
// (one-cell halo), its in-plane (1,2,1)x(1,2,1) sums are formed once, and three consecutive
// planes are combined with (1,2,1) weights.
__global__ void __launch_bounds__(256) k_lm_t(Dims d, const double* const* __restrict__ m,
                                              const double* __restrict__ U,
                                              const double* __restrict__ V,
                                              const double* __restrict__ W,
                                              const double* __restrict__ T,
                                              const double* __restrict__ UH,
                                              const double* __restrict__ VH,
                                              const double* __restrict__ WH,
                                              const double* __restrict__ TH,
                                              const double* __restrict__ mag,
                                              const double* __restrict__ DX,
                                              const double* __restrict__ DY,
                                              const double* __restrict__ DZ,
                                              const double* __restrict__ magh,
                                              const double* __restrict__ filt2, double* lm_out,
                                              double* mm_out) {
  const uint3 bb = xcd_block();
  __shared__ double sh[8][LM_W * LM_H];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * LM_KC;
  const int k1 = min(k0 + LM_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  // in-plane sums of uT vT wT |S|dT/dx |S|dT/dy |S|dT/dz for planes k-1, k, k+1
  double Pa[6], Pb[6], Pc[6];
#pragma unroll
  for (int q = 0; q < 6; ++q) Pa[q] = Pb[q] = Pc[q] = 0.0;
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    __syncthreads();
    for (int q = tid; q < LM_W * LM_H; q += 256) {
      const int gi = min(bb.x * 64 + q % LM_W, d.ni);
      const int gj = min(bb.y * 4 + q / LM_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
      sh[0][q] = U[c]; sh[1][q] = V[c]; sh[2][q] = W[c]; sh[3][q] = T[c];
      sh[4][q] = mag[c]; sh[5][q] = DX[c]; sh[6][q] = DY[c]; sh[7][q] = DZ[c];
    }
    __syncthreads();
    double Pn[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) Pn[q] = 0.0;
#pragma unroll
    for (int dj = -1; dj <= 1; ++dj)
#pragma unroll
      for (int di = -1; di <= 1; ++di) {
        const double wgt = (double)((di == 0 ? 2 : 1) * (dj == 0 ? 2 : 1));
        const int q = (ty + 1 + dj) * LM_W + tx + 1 + di;
        const double t = sh[3][q], mg = sh[4][q];
        Pn[0] += wgt * (sh[0][q] * t);
        Pn[1] += wgt * (sh[1][q] * t);
        Pn[2] += wgt * (sh[2][q] * t);
        Pn[3] += wgt * (mg * sh[5][q]);
        Pn[4] += wgt * (mg * sh[6][q]);
        Pn[5] += wgt * (mg * sh[7][q]);
      }
#pragma unroll
    for (int q = 0; q < 6; ++q) { Pa[q] = Pb[q]; Pb[q] = Pc[q]; Pc[q] = Pn[q]; }
    if (kp < k0 + 1 || !act) continue;
    const int k = kp - 1;
    const long c = cidx(d, i, j, k);
    const long sj = d.sj, sk = d.sk;
    const double gi_ = TH[c + 1] - TH[c - 1], gj_ = TH[c + sj] - TH[c - sj],
                 gk_ = TH[c + sk] - TH[c - sk];
    const double th = TH[c], mh = magh[c], fw2 = filt2[c];
    const double hat[3] = {UH[c], VH[c], WH[c]};
    double lm = 0.0, mm = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double Fp = (Pa[q] + 2.0 * Pb[q] + Pc[q]) / 64.0;
      const double Fr = (Pa[3 + q] + 2.0 * Pb[3 + q] + Pc[3 + q]) / 64.0;
      const double Lq = -Fp + hat[q] * th;
      const double g =
          gp(m[3 * q])[c] * gi_ + gp(m[3 * q + 1])[c] * gj_ + gp(m[3 * q + 2])[c] * gk_;
      const double M = fw2 * (4.0 * mh * 0.5 * g - Fr);
      if (q == 0) {
        lm = Lq * M;
        mm = M * M;
      } else {
        lm = lm + Lq * M;
        mm = mm + M * M;
      }
    }
    lm_out[c] = lm;
    mm_out[c] = mm;
  }
}

// divergence of the subgrid heat flux -2 kappa dT/dx_i (stress_tensor_dy.f90:719-736), added to
// ADV with the inverse Jacobian (RLES/main.f90:67-68).  K-marching like k_sgs_div with the three
// products (2 kappa) dT/dx_i.
__global__ void __launch_bounds__(256) k_sgs_t_div(Dims d, const double* const* __restrict__ m,
                                                   const double* __restrict__ kappa,
                                                   const double* __restrict__ DX,
                                                   const double* __restrict__ DY,
                                                   const double* __restrict__ DZ,
                                                   const double* __restrict__ djac, double* adv,
                                                   double* sgst) {
  const uint3 bb = xcd_block();
  __shared__ double sh[2][3][SF_P];  // [plane slot][component][tile cell]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * SD_KC;
  const int k1 = min(k0 + SD_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int e0 = (ty + 1) * SF_W + tx + 1;
  double Pm[3] = {0.0, 0.0, 0.0};  // own-column products of plane k-1
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) & 1;
    __syncthreads();  // plane kp - 2, held in this slot, is no longer read
    for (int e = tid; e < SF_P; e += 256) {
      const int gi = min(bb.x * 64 + e % SF_W, d.ni);
      const int gj = min(bb.y * 4 + e / SF_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
      const double tk = 2.0 * kappa[c];
      sh[slot][0][e] = tk * DX[c];
      sh[slot][1][e] = tk * DY[c];
      sh[slot][2][e] = tk * DZ[c];
    }
    __syncthreads();
    if (kp == k0 - 1) {
#pragma unroll
      for (int q = 0; q < 3; ++q) Pm[q] = sh[slot][q][e0];
      continue;
    }
    if (kp == k0 || !act) continue;
    const int sc = slot ^ 1;
    const int k = kp - 1;
    const long c = cidx(d, i, j, k);
    double t = 0.0;
#pragma unroll
    for (int q = 0; q < 3; ++q) {
#pragma unroll
      for (int dir = 0; dir < 3; ++dir) {
        double pp, pn;
        if (dir == 0) { pp = sh[sc][q][e0 + 1]; pn = sh[sc][q][e0 - 1]; }
        else if (dir == 1) { pp = sh[sc][q][e0 + SF_W]; pn = sh[sc][q][e0 - SF_W]; }
        else { pp = sh[slot][q][e0]; pn = Pm[q]; }
        const double inner = -pp + pn;
        const double term = gp(m[3 * q + dir])[c] * inner;
        t = (q == 0 && dir == 0) ? term : t + term;
      }
    }
    const double sg = 0.5 * t;
    sgst[c] = sg;
    adv[c] = adv[c] + sg * djac[c];
#pragma unroll
    for (int q = 0; q < 3; ++q) Pm[q] = sh[sc][q][e0];
  }
}

// conditional sums over cells with 0.5 LM/MM >= 0 (stress_tensor_dy.f90:486-507):
// per-block partial (sum LM, sum MM, count)
__global__ void k_lm_reduce(Dims d, const double* __restrict__ lm, const double* __restrict__ mm,
                            double* part) {
  __shared__ double s0[256], s1[256], s2[256];
  const long nint = (long)(d.ni - 1) * (d.nj - 1) * (d.nk - 1);
  double a = 0.0, b = 0.0, n = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint;
       t += (long)gridDim.x * blockDim.x) {
    int i = 1 + (int)(t % (d.ni - 1));
    long r = t / (d.ni - 1);
    int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    long c = cidx(d, i, j, k);
    double cs = 0.5 * lm[c] / mm[c];
    if (cs >= 0.0) {
      a = a + lm[c];
      b = b + mm[c];
      n = n + 1.0;
    }
  }
  const int t = threadIdx.x;
  s0[t] = a; s1[t] = b; s2[t] = n;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) { s0[t] += s0[t + s]; s1[t] += s1[t + s]; s2[t] += s2[t + s]; }
    __syncthreads();
  }
  if (t == 0) { part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0]; }
}

// C = 0.5 <LM>/<MM>, zero when <MM> < 1e-6 (stress_tensor_dy.f90:509-518)
__global__ void k_lm_final(const double* __restrict__ part, int nb, double* out) {
  __shared__ double s0[256], s1[256], s2[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0, n = 0.0;
  for (int q = t; q < nb; q += blockDim.x) { a += part[3 * q]; b += part[3 * q + 1]; n += part[3 * q + 2]; }
  s0[t] = a; s1[t] = b; s2[t] = n;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) { s0[t] += s0[t + s]; s1[t] += s1[t + s]; s2[t] += s2[t + s]; }
    __syncthreads();
  }
  if (t == 0) {
    double tlm = s0[0] / s2[0], tmm = s1[0] / s2[0];
    double cavg = 0.5 * tlm / tmm;
    if (tmm < 1e-6) cavg = 0.0;
    out[0] = cavg;
  }
}

// eddy viscosity nu_t = C Delta^2 |S| on every cell (stress_tensor_dy.f90:527)
__global__ void k_nut(Dims d, const double* __restrict__ cavg, const double* __restrict__ filt2,
                      const double* __restrict__ mag, double* nut) {
  const double C = cavg[0];
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x)
    nut[q] = C * filt2[q] * mag[q];
}

// divergence of the SGS stress -2 nu_t S_ij (stress_tensor_dy.f90:541-583), added to NL
// with the inverse Jacobian (RLES/main.f90:61-66).  A workgroup owns a 64 x 4 (I x J) column
// of cells over SD_KC consecutive K planes and marches up K: the six products
// P_q = (2 nu_t) S_q of each plane are formed once, with a one-cell halo, in one of two LDS
// slots (the I and J differences of plane k and the own-column value of plane k+1), and the
// own-column products of plane k-1 are kept in registers.  Differences, products and sums are
// the reference's, in its order.
template <bool XZ0>
__global__ void __launch_bounds__(256) k_sgs_div(Dims d, const double* const* __restrict__ m,
                                                 const double* __restrict__ filt2,
                                                 const double* __restrict__ cavg,
                                                 const double* const* __restrict__ L,
                                                 const double* __restrict__ djac, double* nlx,
                                                 double* nly, double* nlz, double* sx, double* sy,
                                                 double* sz, int mode) {
  const uint3 bb = xcd_block();
  __shared__ double sh[2][6][SF_P];  // [plane slot][product][tile cell]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * SD_KC;
  const int k1 = min(k0 + SD_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int e0 = (ty + 1) * SF_W + tx + 1;
  const GP S0 = gp(L[0]), S1 = gp(L[1]), S2 = gp(L[2]);
  const GP S3 = gp(L[3]), S4 = gp(L[4]), S5 = gp(L[5]);
  const double C = cavg[0];
  double* nl[3] = {nlx, nly, nlz};
  double* so[3] = {sx, sy, sz};
  // symmetric strain components stored as 11,12,13,22,23,33: row/col -> product index
  const int idx[3][3] = {{0, 1, 2}, {1, 3, 4}, {2, 4, 5}};
  double Pm[6];  // own-column products of plane k-1
#pragma unroll
  for (int q = 0; q < 6; ++q) Pm[q] = 0.0;
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) & 1;
    __syncthreads();  // plane kp - 2, held in this slot, is no longer read
    for (int e = tid; e < SF_P; e += 256) {
      const int gi = min(bb.x * 64 + e % SF_W, d.ni);
      const int gj = min(bb.y * 4 + e / SF_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
      const double l11 = S0[c], l12 = S1[c], l13 = S2[c], l22 = S3[c], l23 = S4[c], l33 = S5[c];
      // nu_t = C Delta^2 |S| as k_nut forms it, with |S| from the strain rates in
      // strain_store's order (the value it stored)
      const double s2 = l11 * l11 + l12 * l12 + l13 * l13 + l12 * l12 + l22 * l22 + l23 * l23 +
                        l13 * l13 + l23 * l23 + l33 * l33;
      const double tn = 2.0 * (C * filt2[c] * sqrt(2.0 * s2));
      sh[slot][0][e] = tn * l11;
      sh[slot][1][e] = tn * l12;
      sh[slot][2][e] = tn * l13;
      sh[slot][3][e] = tn * l22;
      sh[slot][4][e] = tn * l23;
      sh[slot][5][e] = tn * l33;
    }
    __syncthreads();
    if (kp == k0 - 1) {
#pragma unroll
      for (int q = 0; q < 6; ++q) Pm[q] = sh[slot][q][e0];
      continue;
    }
    if (kp == k0 || !act) continue;
    // cell (i, j, kp - 1): plane k in slot ^ 1, plane k + 1 in slot
    const int sc = slot ^ 1;
    const int k = kp - 1;
    const long c = cidx(d, i, j, k);
    const double J = (mode & 1) ? djac[c] : 0.0;
    // the metrics and NL of the cell are loaded before its stores (which might alias them)
    double mv[9], nv[3];
#pragma unroll
    for (int a = 0; a < 9; ++a) mv[a] = mget<XZ0>(m, a, c);
#pragma unroll
    for (int row = 0; row < 3; ++row) nv[row] = (mode & 1) ? nl[row][c] : 0.0;
#pragma unroll
    for (int row = 0; row < 3; ++row) {
      double t = 0.0;
#pragma unroll
      for (int col = 0; col < 3; ++col) {
        const int q = idx[row][col];
#pragma unroll
        for (int dir = 0; dir < 3; ++dir) {
          double pp, pn;
          if (dir == 0) { pp = sh[sc][q][e0 + 1]; pn = sh[sc][q][e0 - 1]; }
          else if (dir == 1) { pp = sh[sc][q][e0 + SF_W]; pn = sh[sc][q][e0 - SF_W]; }
          else { pp = sh[slot][q][e0]; pn = Pm[q]; }
          const double inner = -pp + pn;
          const double term = mv[3 * col + dir] * inner;
          t = (col == 0 && dir == 0) ? term : t + term;
        }
      }
      const double sg = 0.5 * t;
      if (mode & 2) so[row][c] = sg;
      if (mode & 1) nl[row][c] = nv[row] + sg * J;
    }
#pragma unroll
    for (int q = 0; q < 6; ++q) Pm[q] = sh[sc][q][e0];
  }
}

}  // namespace

// SGS_x, SGS_y, SGS_z from the step's nu_t and S_ij, which stay in place until the next SGS
// pass: a step's NL update does not write them, the first read of the fields does
int fvm_k_nut_out(FvmCtx* c) {
  if (!c->nut_stale) return FVM_OK;
  hipLaunchKernelGGL(k_nut, grid1(c->d.n, 256), dim3(256), 0, c->stream, c->d, c->scal + 0,
                     c->filt2, c->mag, c->nut);
  HIP_CHECK(hipGetLastError());
  c->nut_stale = false;
  return FVM_OK;
}

int fvm_k_sgs_out(FvmCtx* c) {
  if (!c->sgs_stale) return FVM_OK;
  const Dims d = c->d;
  const dim3 gs((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + SD_KC - 1) / SD_KC);
  auto kern = c->xz_zero ? k_sgs_div<true> : k_sgs_div<false>;
  hipLaunchKernelGGL(kern, gs, kBlk, 0, c->stream, d, (const double* const*)c->dm, c->filt2,
                     c->scal + 0, (const double* const*)c->dlss, c->djac, c->nl[0], c->nl[1],
                     c->nl[2], c->sgs[0], c->sgs[1], c->sgs[2], 2);
  HIP_CHECK(hipGetLastError());
  c->sgs_stale = false;
  return FVM_OK;
}

int fvm_k_sgs(FvmCtx* c, bool sgs_out, bool nl_update) {
  const Dims d = c->d;
  double** m = c->dm;
  ProfScope ps(c, "sgs");
  hipStream_t st = c->stream;
  const bool th = c->thermal;
  {
    ProfScope pk(c, "k_strain_filter");
    const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + SF_KC - 1) / SF_KC);
    const SfOut io = {{c->u, c->v, c->w, c->t}, {c->uh[0], c->uh[1], c->uh[2], c->that},
                      {c->dtd[0], c->dtd[1], c->dtd[2]}};
    auto kern = th ? k_strain_filter<4, false>
                   : (c->xz_zero ? k_strain_filter<3, true> : k_strain_filter<3, false>);
    hipLaunchKernelGGL(kern, g, kBlk, 0, st, d, (const double* const*)m, io,
                       (double* const*)c->dlss, c->mag);
    const long nb = 2L * (d.nj - 1) * (d.nk - 1);
    hipLaunchKernelGGL(k_strain_iplanes, grid1(nb, 256), dim3(256), 0, st, d,
                       (const double* const*)m, c->u, c->v, c->w, (double* const*)c->dlss,
                       c->mag, th ? c->t : nullptr, c->dtd[0], c->dtd[1], c->dtd[2]);
    HIP_CHECK(hipGetLastError());
  }
  double* gl[10] = {c->lss[0], c->lss[1], c->lss[2], c->lss[3], c->lss[4], c->lss[5], c->mag,
                    c->dtd[0], c->dtd[1], c->dtd[2]};
  int r = fvm_ghost_fill_n(c, gl, th ? 10 : 7);
  if (r) return r;
  if (th) {  // BCT(T_HAT, T_N) (stress_tensor_dy.f90:233)
    r = fvm_k_bct_ghosts(c, c->that);
    if (r) return r;
  }
  r = fvm_k_bcu_ghosts(c, c->uh, 3);
  if (r) return r;
  {
    ProfScope pk(c, "k_lm");
    const int kc = LM_KC;
    const dim3 gl((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + kc - 1) / kc);
    auto kern = c->xz_zero ? k_lm<true> : k_lm<false>;
    hipLaunchKernelGGL(kern, gl, kBlk, 0, st, d, (const double* const*)m, c->u, c->v, c->w,
                       c->uh[0], c->uh[1], c->uh[2], (const double* const*)c->dlss, c->filt2,
                       c->red, th ? c->magh : nullptr, kc);
    HIP_CHECK(hipGetLastError());
    hipLaunchKernelGGL(k_lm_final, dim3(1), dim3(256), 0, st, c->red, (int)(gl.x * gl.y * gl.z),
                       c->scal + 0);
  }
  const int nb = 512;
  // nu_t itself is formed here only for a standalone SGS call; in the step k_sgs_div forms it
  // from C, Delta^2 and S_ij, and FVM_NUT is filled when read
  if (sgs_out)
    hipLaunchKernelGGL(k_nut, grid1(d.n, 256), dim3(256), 0, st, d, c->scal + 0, c->filt2,
                       c->mag, c->nut);
  c->nut_stale = !sgs_out;
  {
    ProfScope pk(c, "k_sgs_div");
    const dim3 gs((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + SD_KC - 1) / SD_KC);
    auto kern = c->xz_zero ? k_sgs_div<true> : k_sgs_div<false>;
    hipLaunchKernelGGL(kern, gs, kBlk, 0, st, d, (const double* const*)m, c->filt2, c->scal + 0,
                       (const double* const*)c->dlss, c->djac, c->nl[0], c->nl[1], c->nl[2],
                       c->sgs[0], c->sgs[1], c->sgs[2],
                       (nl_update ? 1 : 0) | (sgs_out ? 2 : 0));
    HIP_CHECK(hipGetLastError());
    c->sgs_stale = !sgs_out;
  }
  if (th) {  // the temperature's dynamic model (stress_tensor_dy.f90:589-738), main.f90:67-68
    ProfScope pk(c, "sgs_t");
    const dim3 gl((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + LM_KC - 1) / LM_KC);
    hipLaunchKernelGGL(k_lm_t, gl, kBlk, 0, st, d, (const double* const*)m, c->u, c->v, c->w,
                       c->t, c->uh[0], c->uh[1], c->uh[2], c->that, c->mag, c->dtd[0],
                       c->dtd[1], c->dtd[2], c->magh, c->filt2, c->su[0], c->su[1]);
    hipLaunchKernelGGL(k_lm_reduce, dim3(nb), dim3(256), 0, st, d, c->su[0], c->su[1], c->red);
    hipLaunchKernelGGL(k_lm_final, dim3(1), dim3(256), 0, st, c->red, nb, c->scal + 1);
    hipLaunchKernelGGL(k_nut, grid1(d.n, 256), dim3(256), 0, st, d, c->scal + 1, c->filt2,
                       c->mag, c->kappa);
    const dim3 gs((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + SD_KC - 1) / SD_KC);
    hipLaunchKernelGGL(k_sgs_t_div, gs, kBlk, 0, st, d, (const double* const*)m, c->kappa,
                       c->dtd[0], c->dtd[1], c->dtd[2], c->djac, c->adv, c->sgst);
    HIP_CHECK(hipGetLastError());
  }
  return FVM_OK;
}
