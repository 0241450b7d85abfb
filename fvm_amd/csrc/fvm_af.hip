// Approximate-factorisation momentum step ITER_U (LES/v2.3_LES-Reynolds stresses/iter_u.f90).
//
// Each of the three stages (s = x, y, z) assembles the explicit right-hand side (predictor
// terms in stage 1, the stage input otherwise, plus the explicit cross-derivative terms DD of
// the stage input) and then performs ONE in-place line Gauss-Seidel sweep: Thomas solves along
// I (stage 1), J (stage 2) or K (stage 3) with the two other directions' neighbours taken in
// place, lines visited in the reference's lexicographic order (iter_u.f90:86-101, 163-178,
// 240-255).  A line depends on its predecessor in the inner line index and in the outer one
// only, so all lines on an anti-diagonal h = (a-1)+(b-1) of the line plane are independent and
// the reference order is reproduced exactly by sweeping the anti-diagonals in order.  Every
// line is solved by one 64-lane wavefront: its coefficients and right-hand side are staged in
// LDS and the tridiagonal system is reduced by parallel cyclic reduction.  The wall and
// far-field boundary folds of BCU (bcu.f90:15-44) are applied while the coefficients are
// formed; J/K line ends take the periodic ghost values, refreshed before the sweep, as
// Dirichlet data (the reference's lines are not cyclic).  u, v and w are independent and share
// each launch.
#include "fvm_ctx.h"

namespace {

const dim3 kBlk(64, 4, 1);

struct AFArgs {
  double* f[3];           // stage input / output field per component
  const double* fn[3];    // step-start field U_N per component
  const double* nl[3];
  const double* nl1[3];
  const double* dif[3];
  const double* s[3];
  double* su[3];
  const double* ufar;     // far-field boundary values, (NJ+1)(NK+1) per component
};

// right-hand side of stage s (iter_u.f90:49-81, 129-158, 206-235)
template <int S>
__global__ void k_af_rhs(Dims d, AFArgs A, const double* const* __restrict__ m,
                         const double* __restrict__ djac, const double* __restrict__ ap0,
                         double re) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int k = 1 + (int)(blockIdx.z % (d.nk - 1));
  const int comp = (int)(blockIdx.z / (d.nk - 1));
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  const long sj = d.sj, sk = d.sk;
  const double* X = m[3 * S];
  const double* E = m[3 * S + 1];
  const double* Z = m[3 * S + 2];
  const double* J = djac;
  const double* f = A.f[comp];
  const long e = c + 1, w = c - 1, nn = c + sj, ss = c - sj, t = c + sk, b = c - sk;
  double r = J[e] * X[e] * E[e] * (f[e + sj] - f[e - sj]);
  r = r - J[w] * X[w] * E[w] * (f[w + sj] - f[w - sj]);
  r = r + J[e] * X[e] * Z[e] * (f[e + sk] - f[e - sk]);
  r = r - J[w] * X[w] * Z[w] * (f[w + sk] - f[w - sk]);
  r = r + J[nn] * E[nn] * X[nn] * (f[nn + 1] - f[nn - 1]);
  r = r - J[ss] * E[ss] * X[ss] * (f[ss + 1] - f[ss - 1]);
  r = r + J[nn] * E[nn] * Z[nn] * (f[nn + sk] - f[nn - sk]);
  r = r - J[ss] * E[ss] * Z[ss] * (f[ss + sk] - f[ss - sk]);
  r = r + J[t] * Z[t] * X[t] * (f[t + 1] - f[t - 1]);
  r = r - J[b] * Z[b] * X[b] * (f[b + 1] - f[b - 1]);
  r = r + J[t] * Z[t] * E[t] * (f[t + sj] - f[t - sj]);
  r = r - J[b] * Z[b] * E[b] * (f[b + sj] - f[b - sj]);
  const double dd = 0.25 * r;
  double su;
  if (S == 0) {
    su = ap0[c] * A.fn[comp][c];
    su = su - 3.0 * re * A.nl[comp][c];
    su = su + 1.0 * re * A.nl1[comp][c];
    su = su + A.dif[comp][c];
    su = su + 2.0 * re * A.s[comp][c];
    su = su + dd;
  } else {
    su = ap0[c] * f[c] + dd;
  }
  A.su[comp][c] = su;
}

// Stage-s line sweep over the lines of anti-diagonal h of the line plane.
// Line plane: stage 0 lines (j,k), inner a=j, outer b=k; stage 1 lines (i,k), inner a=i,
// outer b=k; stage 2 lines (i,j), inner a=i, outer b=j.
template <int S>
__global__ void __launch_bounds__(64) k_af_sweep(Dims d, AFArgs A, const double* const* fa,
                                                 const double* __restrict__ ap0, int h) {
  __shared__ double la[FVM_MAX_LINE], lb[FVM_MAX_LINE], lc[FVM_MAX_LINE], ld[FVM_MAX_LINE];
  const int comp = blockIdx.y;
  const int na = S == 0 ? d.nj - 1 : d.ni - 1;
  const int nb = S == 2 ? d.nj - 1 : d.nk - 1;
  const int a0 = max(0, h - (nb - 1));
  const int a = a0 + blockIdx.x + 1;  // inner line index (1-based)
  const int b = h - (a - 1) + 1;      // outer line index (1-based)
  if (a > na || b < 1 || b > nb) return;
  int n;
  long base, st;
  if (S == 0) { n = d.ni - 1; st = 1; base = cidx(d, 0, a, b); }
  else if (S == 1) { n = d.nj - 1; st = d.sj; base = cidx(d, a, 0, b); }
  else { n = d.nk - 1; st = d.sk; base = cidx(d, a, b, 0); }
  const double* fE = fa[0];
  const double* fN = fa[1];
  const double* fT = fa[2];
  double* f = A.f[comp];
  const double* su = A.su[comp];
  const long sj = d.sj, sk = d.sk;
  const long nf = (long)(d.nj + 1) * (d.nk + 1);
  for (int r = threadIdx.x; r < n; r += 64) {
    const long c = base + (long)(r + 1) * st;
    const int i = S == 0 ? r + 1 : a;
    double ae = fE[c], aw = fE[c - 1], an = fN[c], as = fN[c - sj], at = fT[c], ab = fT[c - sk];
    double ap = ap0[c] + ae + aw + an + as + at + ab;
    double s = su[c];
    if (i == 1) {  // wall: no-slip, U_F = 0 (bcu.f90:15-20)
      ap = ap + aw;
      s = s + 2.0 * aw * 0.0;
      aw = 0.0;
    }
    if (i == d.ni - 1) {  // far field (bcu.f90:38-40)
      int jj = S == 0 ? a : (S == 1 ? r + 1 : b);
      int kk = S == 0 ? b : (S == 1 ? b : r + 1);
      const double uf = A.ufar[comp * nf + jj + (long)(d.nj + 1) * kk];
      ap = ap + ae;
      s = s + 2.0 * ae * uf;
      ae = 0.0;
    }
    double lo, up, rhs;
    if (S == 0) {
      lo = -aw; up = -ae;
      rhs = an * f[c + sj] + as * f[c - sj] + at * f[c + sk] + ab * f[c - sk] + s;
    } else if (S == 1) {
      lo = -as; up = -an;
      rhs = ae * f[c + 1] + aw * f[c - 1] + at * f[c + sk] + ab * f[c - sk] + s;
    } else {
      lo = -ab; up = -at;
      rhs = ae * f[c + 1] + aw * f[c - 1] + an * f[c + sj] + as * f[c - sj] + s;
    }
    if (r == 0) { rhs = rhs - lo * f[c - st]; lo = 0.0; }
    if (r == n - 1) { rhs = rhs - up * f[c + st]; up = 0.0; }
    la[r] = lo; lb[r] = ap; lc[r] = up; ld[r] = rhs;
  }
  __syncthreads();
  pcr_solve_any(la, lb, lc, ld, n);
  for (int r = threadIdx.x; r < n; r += 64) f[base + (long)(r + 1) * st] = ld[r];
}

}  // namespace

int fvm_k_iter_u(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  AFArgs A;
  double* fld[3] = {c->u, c->v, c->w};
  const double* fn[3] = {c->un, c->vn, c->wn};
  for (int q = 0; q < 3; ++q) {
    A.f[q] = fld[q]; A.fn[q] = fn[q]; A.nl[q] = c->nl[q]; A.nl1[q] = c->nl1[q];
    A.dif[q] = c->dif[q]; A.s[q] = c->s[q]; A.su[q] = c->su[q];
  }
  A.ufar = c->ufar;
  if (d.ni - 1 > FVM_MAX_LINE || d.nj - 1 > FVM_MAX_LINE || d.nk - 1 > FVM_MAX_LINE)
    return FVM_ERR_SIZE;
  const dim3 gi((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, 3 * (d.nk - 1));
  for (int s = 0; s < 3; ++s) {
    {
      ProfScope ps(c, "af_rhs");
      if (s == 0) hipLaunchKernelGGL(k_af_rhs<0>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm, c->djac, c->ap0, c->prm.re);
      if (s == 1) hipLaunchKernelGGL(k_af_rhs<1>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm, c->djac, c->ap0, c->prm.re);
      if (s == 2) hipLaunchKernelGGL(k_af_rhs<2>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm, c->djac, c->ap0, c->prm.re);
      HIP_CHECK(hipGetLastError());
    }
    // boundary ghosts of the stage input (BCU, iter_u.f90:83,160,237)
    int r = fvm_k_bcu_ghosts(c, fld, 3);
    if (r) return r;
    const int na = s == 0 ? d.nj - 1 : d.ni - 1;
    const int nb = s == 2 ? d.nj - 1 : d.nk - 1;
    const double* const* fa = (const double* const*)c->dfa[s];
    ProfScope ps(c, s == 0 ? "af_sweep_i" : (s == 1 ? "af_sweep_j" : "af_sweep_k"));
    for (int h = 0; h <= na + nb - 2; ++h) {
      int lo = h - (nb - 1) > 0 ? h - (nb - 1) : 0;
      int hi = h < na - 1 ? h : na - 1;
      dim3 g(hi - lo + 1, 3);
      if (s == 0) hipLaunchKernelGGL(k_af_sweep<0>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
      if (s == 1) hipLaunchKernelGGL(k_af_sweep<1>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
      if (s == 2) hipLaunchKernelGGL(k_af_sweep<2>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
    }
    HIP_CHECK(hipGetLastError());
  }
  return FVM_OK;
}
