// Time averages of POST_PROCESS on the device (LES/v2.3_LES-Reynolds stresses/
// post_process.f90:1-283) and their accumulation at the end of CALCCOEF (calccoef.f90:286-307).
//
// Seventeen running sums live on every cell, ghosts included: u, v, w, p, T; the products uu,
// uv, uw, vv, vw, ww; and the SGS stresses -2 nu_t S_ij (11, 12, 13, 22, 23, 33).  One
// elementwise pass per step adds the state the step leaves; nu_t = C Delta^2 |S| is formed in
// that pass from the step's C, Delta^2 and |S| exactly as the step's own nu_t is.  The
// finalization divides by the number of steps, forms the rms and the Reynolds stresses in the
// sum arrays (as the reference does in U_AVG..WW_AVG), averages them to the nodes and evaluates
// the wall coefficients from the means with the wall triangle table of CALCCOEF.
#include "fvm_ctx.h"

namespace {

constexpr int kSums = 17;

struct PostSums {
  double* s[kSums];
};

struct PostIn {
  const double* f[5];    // u v w p T (T may be null: zero)
  const double* lss[6];  // S_11 S_12 S_13 S_22 S_23 S_33
};

// calccoef.f90:286-307 on every cell: all of a cell's operands are loaded before its stores
__global__ void __launch_bounds__(256) k_post_accum(long n, PostIn in,
                                                    const double* __restrict__ filt2,
                                                    const double* __restrict__ mag,
                                                    const double* __restrict__ cavg, PostSums a,
                                                    int les) {
  const double C = les ? cavg[0] : 0.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
       q += (long)gridDim.x * blockDim.x) {
    double s[kSums];
#pragma unroll
    for (int r = 0; r < kSums; ++r) s[r] = a.s[r][q];
    const double u = in.f[0][q], v = in.f[1][q], w = in.f[2][q], p = in.f[3][q];
    const double t = in.f[4] ? in.f[4][q] : 0.0;
    s[0] = s[0] + u;
    s[1] = s[1] + v;
    s[2] = s[2] + w;
    s[3] = s[3] + p;
    s[4] = s[4] + t;
    s[5] = s[5] + u * u;
    s[6] = s[6] + u * v;
    s[7] = s[7] + u * w;
    s[8] = s[8] + v * v;
    s[9] = s[9] + v * w;
    s[10] = s[10] + w * w;
    if (les) {
      const double nut = C * filt2[q] * mag[q];  // stress_tensor_dy.f90:527
#pragma unroll
      for (int r = 0; r < 6; ++r) s[11 + r] = s[11 + r] - 2.0 * nut * in.lss[r][q];
    }
#pragma unroll
    for (int r = 0; r < kSums; ++r) a.s[r][q] = s[r];
  }
}

// post_process.f90:63-67, 160-174, 212-217 on every cell: the means in place, the rms
// sqrt(<u_i u_i>) into rms[], the Reynolds stresses <u_i u_j> - <u_i><u_j> + <tau_ij> in place
// of the product sums, the SGS means in place
__global__ void k_post_cells(long n, PostSums a, double np, double* rmsu, double* rmsv,
                             double* rmsw) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n;
       q += (long)gridDim.x * blockDim.x) {
    double s[kSums];
#pragma unroll
    for (int r = 0; r < kSums; ++r) s[r] = a.s[r][q];
#pragma unroll
    for (int r = 0; r < kSums; ++r) s[r] = s[r] / np;
    rmsu[q] = sqrt(s[5]);
    rmsv[q] = sqrt(s[8]);
    rmsw[q] = sqrt(s[10]);
    s[5] = s[5] - s[0] * s[0] + s[11];
    s[6] = s[6] - s[0] * s[1] + s[12];
    s[7] = s[7] - s[0] * s[2] + s[13];
    s[8] = s[8] - s[1] * s[1] + s[14];
    s[9] = s[9] - s[1] * s[2] + s[15];
    s[10] = s[10] - s[2] * s[2] + s[16];
#pragma unroll
    for (int r = 0; r < kSums; ++r) a.s[r][q] = s[r];
  }
}

// 0.125 x the eight cells around each node I = 1..NI, J = 1..NJ, K = 1..NK in
// post_process.f90's order (:125-134), stored at the node's cell index; blockIdx.y selects
// the field
struct NodalSet14 {
  const double* src[14];
  double* dst[14];
};

__global__ void k_post_nodal(Dims d, NodalSet14 ns) {
  const double* __restrict__ P = ns.src[blockIdx.y];
  double* __restrict__ out = ns.dst[blockIdx.y];
  const long nn = (long)d.ni * d.nj * d.nk;
  const long sj = d.sj, sk = d.sk;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nn;
       t += (long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(t % d.ni);
    const long r = t / d.ni;
    const int j = 1 + (int)(r % d.nj), k = 1 + (int)(r / d.nj);
    const long c = cidx(d, i, j, k);
    out[c] = 0.125 * (P[c] + P[c - sj] + P[c - 1] + P[c - 1 - sj] + P[c - sk] + P[c - sj - sk] +
                      P[c - 1 - sk] + P[c - 1 - sj - sk]);
  }
}

// post_process.f90:69-123, 252-256 on the wall cells (J, K): per triangle the shear rate
// |u x n|/d_n and the temperature gradient (T - T_WALL)/d_n of the means, weighted by the
// triangle's share of the cell's wall area (the first assigns, the second adds), the 8-cell
// mean pressure at I - 1/2, J - 1/2, K + 1/2; then Cf = 2/sqrt(Re) x, Cp = 2 x, Nu = -x.  The
// wall table holds A, B, C, |n|, d_n, area per triangle, the second triangles after the first.
__global__ void k_post_surface(Dims d, const double* __restrict__ wall,
                               const double* __restrict__ um, const double* __restrict__ vm,
                               const double* __restrict__ wm, const double* __restrict__ pm,
                               const double* __restrict__ tm, double re, double* cf, double* cp,
                               double* nus) {
  const long nw = (long)(d.nj - 1) * (d.nk - 1);
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= nw) return;
  const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
  const long c = cidx(d, 1, j, k), sj = d.sj, sk = d.sk;
  const double U = um[c], V = vm[c], W = wm[c], T = tm[c];
  const double a1 = wall[6 * r + 5], a2 = wall[6 * (nw + r) + 5];
  double sh = 0.0, nu = 0.0;
  for (int tri = 0; tri < 2; ++tri) {
    const double* o = wall + 6 * (tri * nw + r);
    const double A = o[0], B = o[1], C = o[2], dum = o[3], dn = o[4], ar = o[5];
    const double t1 = V * C / dum - W * B / dum, t2 = U * C / dum - W * A / dum,
                 t3 = U * B / dum - V * A / dum;
    double dudn = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    dudn = dudn / dn;
    const double dtdn = (T - 1.0) / dn;  // T_WALL = 1 (comdat_param.f90:40)
    if (tri == 0) {
      sh = dudn * ar / (a1 + a2);
      nu = dtdn * ar / (a1 + a2);
    } else {
      sh = sh + dudn * ar / (a1 + a2);
      nu = nu + dtdn * ar / (a1 + a2);
    }
  }
  const double P = 0.125 * (pm[c] + pm[c - sj] + pm[c - 1] + pm[c - 1 - sj] + pm[c + sk] +
                            pm[c - sj + sk] + pm[c - 1 + sk] + pm[c - 1 - sj + sk]);
  cf[c] = 2.0 / sqrt(re) * sh;
  cp[c] = 2.0 * P;
  nus[c] = -nu;
}

int post_alloc(FvmCtx* c, double** p) {
  if (*p) return FVM_OK;
  if (hipMalloc((void**)p, c->d.n * sizeof(double)) != hipSuccess) return FVM_ERR_HIP;
  c->allocs.push_back(*p);
  return FVM_OK;
}

}  // namespace

// POST_PROCESS's first call (post_process.f90:9-61): the arrays (once), the sums zeroed
int fvm_k_post_begin(FvmCtx* c, const double* t_const) {
  if (c->two_d) return FVM_ERR_ARG;
  const long n = c->d.n;
  int r = FVM_OK;
  for (double*& p : c->post_sum) r |= post_alloc(c, &p);
  for (double*& p : c->post_rms) r |= post_alloc(c, &p);
  for (double*& p : c->post_node) r |= post_alloc(c, &p);
  for (double*& p : c->post_surf) r |= post_alloc(c, &p);
  if (r) return FVM_ERR_HIP;
  for (double* p : c->post_sum) HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), c->stream));
  for (double* p : c->post_node) HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), c->stream));
  for (double* p : c->post_surf) HIP_CHECK(hipMemsetAsync(p, 0, n * sizeof(double), c->stream));
  if (t_const && !c->thermal) {
    if (post_alloc(c, &c->post_t)) return FVM_ERR_HIP;
    HIP_CHECK(hipMemcpyAsync(c->post_t, t_const, n * sizeof(double), hipMemcpyHostToDevice,
                             c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  } else if (!c->thermal && c->post_t) {
    HIP_CHECK(hipMemsetAsync(c->post_t, 0, n * sizeof(double), c->stream));
  }
  c->niter_pos = 0;
  c->post_on = true;
  return FVM_OK;
}

// the accumulation of CALCCOEF while IPOST = 1 (calccoef.f90:286-307)
int fvm_k_post_accum(FvmCtx* c) {
  if (!c->post_on) return FVM_OK;
  ProfScope ps(c, "post_accum");
  PostIn in{};
  in.f[0] = c->u; in.f[1] = c->v; in.f[2] = c->w; in.f[3] = c->p;
  in.f[4] = c->thermal ? c->t : c->post_t;
  for (int q = 0; q < 6; ++q) in.lss[q] = c->lss[q];
  PostSums a{};
  for (int q = 0; q < kSums; ++q) a.s[q] = c->post_sum[q];
  const long n = c->d.n;
  hipLaunchKernelGGL(k_post_accum, grid1(n, 256), dim3(256), 0, c->stream, n, in, c->filt2,
                     c->mag, c->scal + 0, a, c->prm.use_les ? 1 : 0);
  HIP_CHECK(hipGetLastError());
  c->niter_pos = c->niter_pos + 1;
  return FVM_OK;
}

// POST_PROCESS's second call (post_process.f90:62-276)
int fvm_k_post_finalize(FvmCtx* c) {
  if (!c->post_on || c->niter_pos <= 0) return FVM_ERR_ARG;
  const Dims d = c->d;
  hipStream_t st = c->stream;
  PostSums a{};
  for (int q = 0; q < kSums; ++q) a.s[q] = c->post_sum[q];
  hipLaunchKernelGGL(k_post_cells, grid1(d.n, 256), dim3(256), 0, st, d.n, a,
                     (double)c->niter_pos, c->post_rms[0], c->post_rms[1], c->post_rms[2]);
  const long nw = (long)(d.nj - 1) * (d.nk - 1);
  double** s = c->post_sum;
  hipLaunchKernelGGL(k_post_surface, grid1(nw, 256), dim3(256), 0, st, d, c->wall, s[0], s[1],
                     s[2], s[3], s[4], c->prm.re, c->post_surf[0], c->post_surf[1],
                     c->post_surf[2]);
  // Avrg: p u v w T; Urms: u v w; Strs: uu uv uw vv vw ww
  NodalSet14 ns = {{s[3], s[0], s[1], s[2], s[4], c->post_rms[0], c->post_rms[1],
                    c->post_rms[2], s[5], s[6], s[7], s[8], s[9], s[10]},
                   {}};
  for (int q = 0; q < 14; ++q) ns.dst[q] = c->post_node[q];
  const long nn = (long)d.ni * d.nj * d.nk;
  hipLaunchKernelGGL(k_post_nodal, dim3(grid1(nn, 256).x, 14), dim3(256), 0, st, d, ns);
  HIP_CHECK(hipGetLastError());
  c->post_on = false;
  return FVM_OK;
}
