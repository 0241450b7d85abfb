// Pressure Poisson equation ITER_P + LISOLV + BCP
// (LES/v2.3_LES-Reynolds stresses/iter_p.f90:12-63, lisolv.f90:15-30, bcp.f90:15-36).
//
// Operator (iter_p.f90:12-21): AP = -(DE+DW+DN+DS+DT+DB), A_nb = -D_nb; wall Neumann fold
// AP(1) -= AW(1), AW(1) = 0, outflow AE(NI-1) = 0 (bcp.f90:22-29); explicit cross-derivative
// terms of p on the right-hand side, SU = div/dt - DD(p).  The reference relaxes it with
// lexicographic radial-line Gauss-Seidel; here the radial lines are visited in zebra (red-black
// in J+K parity) order, so every line of one colour is independent and each half-sweep is one
// launch with one 64-lane wavefront per line (coefficients and right-hand side staged in LDS,
// parallel cyclic reduction).  Both orders relax the same discrete equation; the iteration
// stops on the reference's scaled-L1 residual sum|r| / sum|AP p| (iter_p.f90:45-59), evaluated
// in FP64 for the current iterate with fresh ghost values.
//
// The mixed-precision form keeps p, the residual, the metric coefficients and the norms in
// FP64 and runs the line sweeps in FP32 on the correction e of a defect-correction loop:
// r = b - A p (FP64), e ~ A^-1 r by FP32 zebra sweeps, p += e (FP64).
#include "fvm_ctx.h"

namespace {

const dim3 kBlk(64, 4, 1);

// div / dt from the face volume fluxes (iter_p.f90:31-34)
__global__ void k_p_rhs(Dims d, const double* __restrict__ ju0, const double* __restrict__ ju1,
                        const double* __restrict__ ju2, double dtime, double* rhs) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  double ue = 0.5 * (ju0[c] + ju0[c + 1]);
  double uw = i == 1 ? 0.0 : 0.5 * (ju0[c] + ju0[c - 1]);
  double vn = 0.5 * (ju1[c] + ju1[c + d.sj]);
  double vs = 0.5 * (ju1[c] + ju1[c - d.sj]);
  double wt = 0.5 * (ju2[c] + ju2[c + d.sk]);
  double wb = 0.5 * (ju2[c] + ju2[c - d.sk]);
  double dv = ue - uw + vn - vs + wt - wb;
  rhs[c] = 1.0 / dtime * dv;
}

// The discrete divergence of the face volume fluxes (the source ITER_P drives to zero for the
// corrected velocity, iter_p.f90:31-34) for the interior cells of one 64 x 4 x 1 block: per block
// max |div|, sum |div| and the flux scale sum (|UE| + |VN| + |WT|), fixed-order trees.
__global__ void __launch_bounds__(256) k_div_part(Dims d, const double* __restrict__ ju0,
                                                  const double* __restrict__ ju1,
                                                  const double* __restrict__ ju2, double* part) {
  __shared__ double sh[3][256];
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + blockIdx.z;
  const int t = threadIdx.x + 64 * threadIdx.y;
  double m = 0.0, a = 0.0, f = 0.0;
  if (i <= d.ni - 1 && j <= d.nj - 1) {
    const long c = cidx(d, i, j, k);
    const double ue = 0.5 * (ju0[c] + ju0[c + 1]);
    const double uw = i == 1 ? 0.0 : 0.5 * (ju0[c] + ju0[c - 1]);
    const double vn = 0.5 * (ju1[c] + ju1[c + d.sj]);
    const double vs = 0.5 * (ju1[c] + ju1[c - d.sj]);
    const double wt = 0.5 * (ju2[c] + ju2[c + d.sk]);
    const double wb = 0.5 * (ju2[c] + ju2[c - d.sk]);
    const double dv = ue - uw + vn - vs + wt - wb;
    m = fabs(dv);
    a = fabs(dv);
    f = fabs(ue) + fabs(vn) + fabs(wt);
  }
  sh[0][t] = m; sh[1][t] = a; sh[2][t] = f;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      sh[0][t] = fmax(sh[0][t], sh[0][t + s]);
      sh[1][t] = sh[1][t] + sh[1][t + s];
      sh[2][t] = sh[2][t] + sh[2][t + s];
    }
    __syncthreads();
  }
  if (t == 0) {
    const long b = blockIdx.x + (long)gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z);
    part[3 * b] = sh[0][0]; part[3 * b + 1] = sh[1][0]; part[3 * b + 2] = sh[2][0];
  }
}

__global__ void __launch_bounds__(256) k_div_final(const double* __restrict__ part, long nb,
                                                   double* out) {
  __shared__ double sh[3][256];
  const int t = threadIdx.x;
  double m = 0.0, a = 0.0, f = 0.0;
  for (long q = t; q < nb; q += blockDim.x) {
    m = fmax(m, part[3 * q]);
    a = a + part[3 * q + 1];
    f = f + part[3 * q + 2];
  }
  sh[0][t] = m; sh[1][t] = a; sh[2][t] = f;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) {
      sh[0][t] = fmax(sh[0][t], sh[0][t + s]);
      sh[1][t] = sh[1][t] + sh[1][t + s];
      sh[2][t] = sh[2][t] + sh[2][t + s];
    }
    __syncthreads();
  }
  if (t == 0) { out[0] = sh[0][0]; out[1] = sh[1][0]; out[2] = sh[2][0]; }
}

template <typename T>
__device__ inline double cross_p(const Dims& d, long c, const T* __restrict__ f,
                                 const double* __restrict__ jq12,
                                 const double* __restrict__ jq13,
                                 const double* __restrict__ jq23) {
  const long sj = d.sj, sk = d.sk;
  const long e = c + 1, w = c - 1, n = c + sj, s = c - sj, t = c + sk, b = c - sk;
  double r = jq12[e] * ((double)f[e + sj] - (double)f[e - sj]);
  r = r - jq12[w] * ((double)f[w + sj] - (double)f[w - sj]);
  r = r + jq13[e] * ((double)f[e + sk] - (double)f[e - sk]);
  r = r - jq13[w] * ((double)f[w + sk] - (double)f[w - sk]);
  r = r + jq12[n] * ((double)f[n + 1] - (double)f[n - 1]);
  r = r - jq12[s] * ((double)f[s + 1] - (double)f[s - 1]);
  r = r + jq23[n] * ((double)f[n + sk] - (double)f[n - sk]);
  r = r - jq23[s] * ((double)f[s + sk] - (double)f[s - sk]);
  r = r + jq13[t] * ((double)f[t + 1] - (double)f[t - 1]);
  r = r - jq13[b] * ((double)f[b + 1] - (double)f[b - 1]);
  r = r + jq23[t] * ((double)f[t + sj] - (double)f[t - sj]);
  r = r - jq23[b] * ((double)f[b + sj] - (double)f[b - sj]);
  return 0.25 * r;
}

struct PCoef {
  double ap, ae, aw, an, as, at, ab;
};

// operator coefficients of cell (i, c) with the BCP folds
__device__ inline PCoef pcoef(const Dims& d, int i, long c, const double* __restrict__ pdI,
                              const double* __restrict__ pdJ, const double* __restrict__ pdK) {
  const double de = pdI[c], dw = pdI[c - 1], dn = pdJ[c], ds = pdJ[c - d.sj], dt = pdK[c],
               db = pdK[c - d.sk];
  PCoef q;
  q.ap = -de - dw - dn - ds - dt - db;
  q.ae = -de; q.aw = -dw; q.an = -dn; q.as = -ds; q.at = -dt; q.ab = -db;
  if (i == 1) { q.ap = q.ap - q.aw; q.aw = 0.0; }
  if (i == d.ni - 1) q.ae = 0.0;
  return q;
}

// SU = rhs - DD(p) and, when asked, the per-block sums of |r| and |AP p| of the current
// iterate: r = sum A_nb p_nb + SU - AP p (iter_p.f90:45-58)
__global__ void k_p_su(Dims d, const double* __restrict__ p, const double* __restrict__ rhs,
                       const double* __restrict__ pdI, const double* __restrict__ pdJ,
                       const double* __restrict__ pdK, const double* __restrict__ jq12,
                       const double* __restrict__ jq13, const double* __restrict__ jq23,
                       double* psu, int want_res, double* part) {
  const uint3 bb = xcd_block();
  __shared__ double s0[256], s1[256];
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  const int t = threadIdx.x + 64 * threadIdx.y;
  double ra = 0.0, rb = 0.0;
  if (i <= d.ni - 1 && j <= d.nj - 1) {
    const long c = cidx(d, i, j, k);
    double su = rhs[c] - cross_p(d, c, p, jq12, jq13, jq23);
    psu[c] = su;
    if (want_res) {
      PCoef q = pcoef(d, i, c, pdI, pdJ, pdK);
      double r = q.ae * p[c + 1] + q.aw * p[c - 1] + q.an * p[c + d.sj] + q.as * p[c - d.sj] +
                 q.at * p[c + d.sk] + q.ab * p[c - d.sk] + su - q.ap * p[c];
      ra = fabs(r);
      rb = fabs(q.ap * p[c]);
    }
  }
  if (!want_res) return;
  s0[t] = ra; s1[t] = rb;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { s0[t] += s0[t + s]; s1[t] += s1[t + s]; }
    __syncthreads();
  }
  if (t == 0) {
    long bid = bb.x + (long)gridDim.x * (bb.y + (long)gridDim.y * bb.z);
    part[2 * bid] = s0[0];
    part[2 * bid + 1] = s1[0];
  }
}

__global__ void k_p_res_final(const double* __restrict__ part, long nb, double* out) {
  __shared__ double s0[256], s1[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (long q = t; q < nb; q += blockDim.x) { a += part[2 * q]; b += part[2 * q + 1]; }
  s0[t] = a; s1[t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { s0[t] += s0[t + s]; s1[t] += s1[t + s]; }
    __syncthreads();
  }
  if (t == 0) { out[0] = s0[0]; out[1] = s1[0]; out[2] = scaled_residual(s0[0], s1[0]); }
}

// BCP ghosts: wall p(0) = p(1), outflow cubic extrapolation (bcp.f90:22-32); the periodic
// ghosts follow (fvm_ghost_fill)
template <typename T>
__global__ void k_bcp(Dims d, T* p) {
  const long nf = (long)(d.nj - 1) * (d.nk - 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nf;
       t += (long)gridDim.x * blockDim.x) {
    int j = 1 + (int)(t % (d.nj - 1)), k = 1 + (int)(t / (d.nj - 1));
    long c0 = cidx(d, 0, j, k), cn = cidx(d, d.ni, j, k);
    p[c0] = p[c0 + 1];
    p[cn] = (T)3.0 * p[cn - 1] - (T)3.0 * p[cn - 2] + p[cn - 3];
  }
}

// One zebra half-sweep: every radial line (j,k) with (j+k) % 2 == colour solved for its new
// values with the current neighbours (lisolv.f90:15-30 line equation).
template <typename T>
__global__ void __launch_bounds__(64) k_p_sweep(Dims d, T* p, const T* __restrict__ su,
                                                const double* __restrict__ pdI,
                                                const double* __restrict__ pdJ,
                                                const double* __restrict__ pdK, int colour) {
  __shared__ T la[FVM_MAX_LINE], lb[FVM_MAX_LINE], lc[FVM_MAX_LINE], ld[FVM_MAX_LINE];
  // line index: lines of one colour, (NJ-1) is split in halves per K row
  const int nj1 = d.nj - 1;
  const int k = 1 + blockIdx.y;
  const int jh = blockIdx.x;                       // 0 .. ceil(nj1/2)-1
  const int jstart = 1 + ((colour + 1 + k) & 1);   // first j with (j+k)%2 == colour
  const int j = jstart + 2 * jh;
  if (j > nj1) return;
  const int n = d.ni - 1;
  const long base = cidx(d, 0, j, k);
  for (int r = threadIdx.x; r < n; r += 64) {
    const int i = r + 1;
    const long c = base + i;
    PCoef q = pcoef(d, i, c, pdI, pdJ, pdK);
    double rhs = q.an * (double)p[c + d.sj] + q.as * (double)p[c - d.sj] +
                 q.at * (double)p[c + d.sk] + q.ab * (double)p[c - d.sk] + (double)su[c];
    la[r] = (T)(-q.aw); lb[r] = (T)q.ap; lc[r] = (T)(-q.ae); ld[r] = (T)rhs;
  }
  __syncthreads();
  pcr_solve_any(la, lb, lc, ld, n);
  for (int r = threadIdx.x; r < n; r += 64) p[base + r + 1] = ld[r];
}

// Mixed precision: FP32 defect r32 = (b - A p) with b = SU from k_p_su; the correction
// equation A e = r keeps the operator's cross terms on e as well.
__global__ void k_p_defect32(Dims d, const double* __restrict__ p, const double* __restrict__ psu,
                             const double* __restrict__ pdI, const double* __restrict__ pdJ,
                             const double* __restrict__ pdK, float* r32, float* e32) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  PCoef q = pcoef(d, i, c, pdI, pdJ, pdK);
  double r = q.ae * p[c + 1] + q.aw * p[c - 1] + q.an * p[c + d.sj] + q.as * p[c - d.sj] +
             q.at * p[c + d.sk] + q.ab * p[c - d.sk] + psu[c] - q.ap * p[c];
  // correction equation L(e) = r with L(p) = AP p - sum A_nb p_nb + DD(p)
  r32[c] = (float)r;
  e32[c] = 0.0f;
}

// FP32 source of the correction equation: s = defect - DD(e)
__global__ void k_e_su(Dims d, const float* __restrict__ e, const float* __restrict__ r32,
                       const double* __restrict__ jq12, const double* __restrict__ jq13,
                       const double* __restrict__ jq23, float* esu) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  esu[c] = (float)((double)r32[c] - cross_p(d, c, e, jq12, jq13, jq23));
}

__global__ void k_p_add(Dims d, double* p, const float* __restrict__ e) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  p[c] = p[c] + (double)e[c];
}

}  // namespace

int fvm_k_bcp(FvmCtx* c) {
  long nf = (long)(c->d.nj - 1) * (c->d.nk - 1);
  hipLaunchKernelGGL(k_bcp<double>, grid1(nf, 256), dim3(256), 0, c->stream, c->d, c->p);
  HIP_CHECK(hipGetLastError());
  return fvm_ghost_fill(c, c->p);
}

static int bcp32(FvmCtx* c, float* e) {
  const Dims d = c->d;
  long nf = (long)(d.nj - 1) * (d.nk - 1);
  hipLaunchKernelGGL(k_bcp<float>, grid1(nf, 256), dim3(256), 0, c->stream, d, e);
  hipLaunchKernelGGL(k_ghost_fill_t<float>, grid1(ghost_count(d), 256), dim3(256), 0, c->stream, d, e);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// residual sums of the current iterate (after fresh BCP ghosts), SU for the next sweep
static int su_and_residual(FvmCtx* c, bool want, double* res) {
  const Dims d = c->d;
  dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  long nb = (long)g.x * g.y * g.z;
  {
    ProfScope ps(c, "p_su");
    hipLaunchKernelGGL(k_p_su, g, kBlk, 0, c->stream, d, c->p, c->rhs, c->pdI, c->pdJ, c->pdK,
                       c->jq12, c->jq13, c->jq23, c->psu, want ? 1 : 0, c->red);
    HIP_CHECK(hipGetLastError());
  }
  if (want) {
    hipLaunchKernelGGL(k_p_res_final, dim3(1), dim3(256), 0, c->stream, c->red, nb, c->scal + 4);
    HIP_CHECK(hipGetLastError());
    const int rf = fvm_fetch(c, c->scal + 4, 3, c->hscal + 4);
    if (rf) return rf;
    *res = c->hscal[6];
  }
  return FVM_OK;
}

template <typename T>
static int zebra(FvmCtx* c, T* p, const T* su) {
  const Dims d = c->d;
  dim3 g((d.nj - 1 + 1) / 2, d.nk - 1);
  ProfScope ps(c, sizeof(T) == 8 ? "p_sweep64" : "p_sweep32");
  for (int colour = 0; colour < 2; ++colour)
    hipLaunchKernelGGL(k_p_sweep<T>, g, dim3(64), 0, c->stream, d, p, su, c->pdI, c->pdJ,
                       c->pdK, colour);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// out: max |div|, mean |div| over the interior cells, sum |div| / sum (|UE| + |VN| + |WT|), of
// the face fluxes of the current velocity (CONTRA_VEL recomputed; the next step recomputes it too)
int fvm_k_div_norm(FvmCtx* c, double* out) {
  const Dims d = c->d;
  int r = fvm_k_contra_vel(c);
  if (r) return r;
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  const long nb = (long)g.x * g.y * g.z;
  if (3 * nb + 3 > (long)c->nred) return FVM_ERR_SIZE;
  hipLaunchKernelGGL(k_div_part, g, kBlk, 0, c->stream, d, c->ju[0], c->ju[1], c->ju[2], c->red);
  hipLaunchKernelGGL(k_div_final, dim3(1), dim3(256), 0, c->stream, c->red, nb, c->red + 3 * nb);
  HIP_CHECK(hipGetLastError());
  double h[3];
  HIP_CHECK(hipMemcpyAsync(h, c->red + 3 * nb, 3 * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  const double ncell = (double)(d.ni - 1) * (d.nj - 1) * (d.nk - 1);
  out[0] = h[0];
  out[1] = h[1] / ncell;
  out[2] = h[2] > 0.0 ? h[1] / h[2] : 0.0;
  return FVM_OK;
}

int fvm_k_iter_p(FvmCtx* c, double* resip, int* niter) {
  const Dims d = c->d;
  if (d.ni - 1 > FVM_MAX_LINE) return FVM_ERR_SIZE;
  dim3 gi((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  {
    ProfScope ps(c, "p_rhs");
    hipLaunchKernelGGL(k_p_rhs, gi, kBlk, 0, c->stream, d, c->ju[0], c->ju[1], c->ju[2],
                       c->prm.dtime, c->rhs);
    HIP_CHECK(hipGetLastError());
  }
  if (c->prm.poisson_mode >= 2) {
    int rs = fvm_k_iter_p_sep(c, c->prm.poisson_mode == 3, resip, niter);
    if (rs && rs != FVM_ERR_NOCONV) return rs;
    return rs;
  }
  const int every = c->prm.check_every > 0 ? c->prm.check_every : 1;
  const int cap = c->prm.nswpp;
  const double tol = c->prm.resimax;
  int sweeps = 0, outer = 0;
  double res = INFINITY;
  int r;
  if (c->prm.poisson_mode == 0) {
    // FP64 zebra line Gauss-Seidel
    for (;;) {
      r = fvm_k_bcp(c);
      if (r) return r;
      bool want = (sweeps % every == 0) || sweeps >= cap;
      r = su_and_residual(c, want, &res);
      if (r) return r;
      if (want) {
        ++outer;
        if (!(res == res)) return FVM_ERR_NAN;
        if (res <= tol || sweeps >= cap) break;
      }
      r = zebra<double>(c, c->p, c->psu);
      if (r) return r;
      ++sweeps;
    }
  } else {
    // FP32 zebra sweeps on the correction inside FP64 defect correction
    const int inner = every;
    float* e = c->p32;
    float* esu = c->su32;
    float* r32 = (float*)c->su[2];  // defect, FP32 view of a scratch array
    for (;;) {
      r = fvm_k_bcp(c);
      if (r) return r;
      r = su_and_residual(c, true, &res);
      if (r) return r;
      ++outer;
      if (!(res == res)) return FVM_ERR_NAN;
      if (res <= tol || sweeps >= cap) break;
      {
        ProfScope ps(c, "p_defect");
        hipLaunchKernelGGL(k_p_defect32, gi, kBlk, 0, c->stream, d, c->p, c->psu, c->pdI,
                           c->pdJ, c->pdK, r32, e);
        HIP_CHECK(hipGetLastError());
      }
      for (int it = 0; it < inner; ++it) {
        r = bcp32(c, e);
        if (r) return r;
        {
          ProfScope ps(c, "e_su");
          hipLaunchKernelGGL(k_e_su, gi, kBlk, 0, c->stream, d, e, r32, c->jq12, c->jq13,
                             c->jq23, esu);
          HIP_CHECK(hipGetLastError());
        }
        r = zebra<float>(c, e, esu);
        if (r) return r;
        ++sweeps;
      }
      {
        ProfScope ps(c, "p_add");
        hipLaunchKernelGGL(k_p_add, gi, kBlk, 0, c->stream, d, c->p, e);
        HIP_CHECK(hipGetLastError());
      }
    }
  }
  r = fvm_k_bcp(c);
  if (r) return r;
  *resip = res;
  *niter = sweeps;
  c->last.nouter = outer;
  if (sweeps >= cap && res > tol) return FVM_ERR_NOCONV;
  return FVM_OK;
}
