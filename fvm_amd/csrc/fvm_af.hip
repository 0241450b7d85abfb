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

#include <algorithm>
#include <vector>

namespace {

const dim3 kBlk(64, 4, 1);

// right-hand side of stage s (iter_u.f90:49-81, 129-158, 206-235); one thread per cell and all
// three components, so the metric products of the cross terms are loaded once
// With qform set (line-scan form of stages 0 and 1) the kernel stores instead
// Q = SU' + c_(a+1) x_in(a+1) + AT x_in(K+1) + AB x_in(K-1): SU' carries BCU's far-field
// addition 2 AE U_F at I = NI-1 (bcu.f90:38-40; the wall addition multiplies U_F = 0), c_(a+1)
// is AN (stage 0) or the folded AE (stage 1).
// With ddzero set (stage 2 on a grid whose xi_z and eta_z vanish identically around every cell)
// each of the twelve cross terms is an exact zero and SU = 2 Re J^-1/dt phi exactly.
template <int S>
__global__ void k_af_rhs(Dims d, AFArgs A, const double* const* __restrict__ m,
                         const double* __restrict__ djac, const double* __restrict__ ap0,
                         double re, const double* const* __restrict__ fa, int qform,
                         int ddzero) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  if (S == 2 && ddzero) {
    const double a0 = ap0[c];
#pragma unroll
    for (int comp = 0; comp < A.nc; ++comp) gpw(A.su[comp])[c] = a0 * gp(A.f[comp])[c];
    return;
  }
  const long sj = d.sj, sk = d.sk;
  const GP X = gp(m[3 * S]);
  const GP E = gp(m[3 * S + 1]);
  const GP Z = gp(m[3 * S + 2]);
  const double* J = djac;
  const long e = c + 1, w = c - 1, nn = c + sj, ss = c - sj, t = c + sk, b = c - sk;
  // metric products in the reference's association (J X) E etc.
  const double pe_xe = J[e] * X[e] * E[e], pw_xe = J[w] * X[w] * E[w];
  const double pe_xz = J[e] * X[e] * Z[e], pw_xz = J[w] * X[w] * Z[w];
  const double pn_ex = J[nn] * E[nn] * X[nn], ps_ex = J[ss] * E[ss] * X[ss];
  const double pn_ez = J[nn] * E[nn] * Z[nn], ps_ez = J[ss] * E[ss] * Z[ss];
  const double pt_zx = J[t] * Z[t] * X[t], pb_zx = J[b] * Z[b] * X[b];
  const double pt_ze = J[t] * Z[t] * E[t], pb_ze = J[b] * Z[b] * E[b];
  const double a0 = ap0[c];
  for (int comp = 0; comp < A.nc; ++comp) {
    const GP f = gp(A.f[comp]);
    double r = pe_xe * (f[e + sj] - f[e - sj]);
    r = r - pw_xe * (f[w + sj] - f[w - sj]);
    r = r + pe_xz * (f[e + sk] - f[e - sk]);
    r = r - pw_xz * (f[w + sk] - f[w - sk]);
    r = r + pn_ex * (f[nn + 1] - f[nn - 1]);
    r = r - ps_ex * (f[ss + 1] - f[ss - 1]);
    r = r + pn_ez * (f[nn + sk] - f[nn - sk]);
    r = r - ps_ez * (f[ss + sk] - f[ss - sk]);
    r = r + pt_zx * (f[t + 1] - f[t - 1]);
    r = r - pb_zx * (f[b + 1] - f[b - 1]);
    r = r + pt_ze * (f[t + sj] - f[t - sj]);
    r = r - pb_ze * (f[b + sj] - f[b - sj]);
    const double dd = 0.25 * r;
    double su;
    if (S == 0) {
      su = a0 * gp(A.fn[comp])[c];
      su = su - 3.0 * re * gp(A.nl[comp])[c];
      su = su + 1.0 * re * gp(A.nl1[comp])[c];
      su = su + gp(A.dif[comp])[c];
      su = su + 2.0 * re * gp(A.s[comp])[c];
      su = su + dd;
    } else {
      su = a0 * f[c] + dd;
    }
    if (S < 2 && qform) {
      double ae = gp(fa[0])[c];
      if (i == 1) su = su + 2.0 * gp(fa[0])[c - 1] * 0.0;
      if (i == d.ni - 1) {
        const long nf = (long)(d.nj + 1) * (d.nk + 1);
        su = su + 2.0 * ae * A.ufar[comp * nf + j + (long)(d.nj + 1) * k];
        ae = 0.0;
      }
      su = S == 0 ? su + gp(fa[1])[c] * f[nn] : su + ae * f[e];
      su = su + gp(fa[2])[c] * f[t];
      su = su + gp(fa[2])[b] * f[b];
    }
    gpw(A.su[comp])[c] = su;
  }
}

// Metric products of the cross terms of stage s, once per cell including ghosts, in the
// association of k_af_rhs: (J X) E, (J X) Z, (J E) X, (J E) Z, (J Z) X, (J Z) E with X, E, Z the
// stage's xi_s, eta_s, zeta_s.
__global__ void k_af_products(Dims d, const double* __restrict__ J, const double* __restrict__ X,
                              const double* __restrict__ E, const double* __restrict__ Z,
                              double* const* __restrict__ P) {
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x) {
    const double j = J[q], x = X[q], e = E[q], z = Z[q];
    P[0][q] = j * x * e;
    P[1][q] = j * x * z;
    P[2][q] = j * e * x;
    P[3][q] = j * e * z;
    P[4][q] = j * z * x;
    P[5][q] = j * z * e;
  }
}

// k_af_rhs for stages 0 and 1, K-marching: a workgroup owns a 64 x 4 (I x J) column of cells
// over RK_KC consecutive K planes and keeps the last three planes of the stage inputs (one-cell
// halo) in LDS, so the twelve edge neighbours of the cross terms come from LDS and every input
// value is fetched from HBM once per column.  The metric products come from the per-stage
// arrays of k_af_products (equal, bit for bit, to the products k_af_rhs forms); those at K+-1
// and AT at K-1 ride along in registers.  Expressions and their order are k_af_rhs's.
// TH selects ITER_T (RLES/iter_t.f90:49-55): re is PE, the stage-1 source is
// 2 PE J^-1/dt T - 3 PE ADV + PE ADV_NM1 + DIFT + DD (A.nl = ADV, A.nl1 = ADV_NM1,
// A.dif = DIFT) and BCT's wall fold adds 2 AW T_WALL (bct.f90:16-19).
constexpr int RK_KC = 16;

// With DR set (stages 0 and 1 of the FP32 correction form, qform on) the kernel stores, as FP32
// in SU's memory, the residual of the stage's line system at the stage input instead of Q:
// r = Q + c_(a-1) x_in(a-1) - (AP x_in - lo x_in(p-1) - up x_in(p+1)), with k_scan_factor's AP,
// folds and couplings; the branch-cut ghosts enter as the periodic images of the interior, the
// values BCU gives them before the sweep.  The line sweep then solves for the stage's change
// delta = x - x_in in FP32 and adds it to the FP64 field.
template <int S, bool TH, int NCV = 3, bool DR = false>
__global__ void __launch_bounds__(256) k_af_rhs_km(Dims d, AFArgs A,
                                                   const double* const* __restrict__ P,
                                                   const double* __restrict__ ap0, double re,
                                                   const double* const* __restrict__ fa,
                                                   int qform, int fn_alias) {
  constexpr double kTWall = 1.0;  // T_WALL (comdat_param.f90:40)
  constexpr int NC = TH ? 1 : NCV;  // components: T, or u, v, w (NCV = 1: one of them)
  const uint3 bb = xcd_block();
  __shared__ double sh[NC][3][SF_P];  // [component][plane slot][tile cell]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * RK_KC;
  const int k1 = min(k0 + RK_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int q0 = (ty + 1) * SF_W + tx + 1;
  const long sj = d.sj, sk = d.sk;
  const GP PXE = gp(P[0]), PXZ = gp(P[1]), PEX = gp(P[2]), PEZ = gp(P[3]), PZX = gp(P[4]),
           PZE = gp(P[5]);
  const GP fE = gp(fa[0]), fN = gp(fa[1]), fT = gp(fa[2]);
  // own-column values of planes kp-2, kp-1, kp: (J Z) X, (J Z) E and AT
  double zx[3] = {0.0, 0.0, 0.0}, ze[3] = {0.0, 0.0, 0.0}, at[3] = {0.0, 0.0, 0.0};
  // the next plane's tile values and own-column values, loaded one plane ahead (in flight
  // while the current plane's cells are formed) in stage 1; stage 0's cells hold many more
  // operands, and there the registers are worth more than the overlap
  constexpr bool PF = S == 1;
  constexpr int NQ = (SF_P + 255) / 256;
  double tv[NQ][NC], nzx = 0.0, nze = 0.0, nat = 0.0;
  auto load = [&](int kp) {
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const int q = min(tid + 256 * h, SF_P - 1);
      const int gi = min(bb.x * 64 + q % SF_W, d.ni);
      const int gj = min(bb.y * 4 + q / SF_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) tv[h][comp] = gp(A.f[comp])[c];
    }
    const long cp = cidx(d, act ? i : 1, act ? j : 1, kp);
    nzx = PZX[cp];
    nze = PZE[cp];
    nat = (S < 2 && qform) ? fT[cp] : 0.0;
  };
  if (PF) load(k0 - 1);
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) % 3;
    __syncthreads();  // plane kp - 3, held in this slot, is no longer read
    if constexpr (PF) {
#pragma unroll
      for (int h = 0; h < NQ; ++h) {
        const int q = tid + 256 * h;
        if (q < SF_P) {
#pragma unroll
          for (int comp = 0; comp < NC; ++comp) sh[comp][slot][q] = tv[h][comp];
        }
      }
      zx[0] = zx[1]; zx[1] = zx[2]; zx[2] = nzx;
      ze[0] = ze[1]; ze[1] = ze[2]; ze[2] = nze;
      at[0] = at[1]; at[1] = at[2]; at[2] = nat;
    } else {
      for (int q = tid; q < SF_P; q += 256) {
        const int gi = min(bb.x * 64 + q % SF_W, d.ni);
        const int gj = min(bb.y * 4 + q / SF_W, d.nj);
        const long c = cidx(d, gi, gj, kp);
#pragma unroll
        for (int comp = 0; comp < NC; ++comp) sh[comp][slot][q] = gp(A.f[comp])[c];
      }
      zx[0] = zx[1]; zx[1] = zx[2];
      ze[0] = ze[1]; ze[1] = ze[2];
      at[0] = at[1]; at[1] = at[2];
      if (act) {
        const long cp = cidx(d, i, j, kp);
        zx[2] = PZX[cp];
        ze[2] = PZE[cp];
        at[2] = (S < 2 && qform) ? fT[cp] : 0.0;
      }
    }
    __syncthreads();
    if (PF && kp + 1 <= k1 + 1) load(kp + 1);
    if (kp < k0 + 1 || !act) continue;
    // cell (i, j, kp - 1): planes k-1, k, k+1 sit in slots slot+1, slot+2, slot (mod 3)
    const int k = kp - 1;
    const int pl[3] = {(slot + 1) % 3, (slot + 2) % 3, slot};
    const long c = cidx(d, i, j, k);
    const double pe_xe = PXE[c + 1], pw_xe = PXE[c - 1];
    const double pe_xz = PXZ[c + 1], pw_xz = PXZ[c - 1];
    const double pn_ex = PEX[c + sj], ps_ex = PEX[c - sj];
    const double pn_ez = PEZ[c + sj], ps_ez = PEZ[c - sj];
    const double pt_zx = zx[2], pb_zx = zx[0];
    const double pt_ze = ze[2], pb_ze = ze[0];
    const double a0 = ap0[c];
    // every global input of the cell is loaded before the first SU store: the stores could alias
    // the inputs as far as the compiler knows, and loads after them would wait for them
    double vnl[NC], vnl1[NC], vdif[NC], vs[NC], vfn[NC], vuf[NC];
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
      vnl[comp] = vnl1[comp] = vdif[comp] = vs[comp] = vfn[comp] = vuf[comp] = 0.0;
      if (S == 0) {
        vnl[comp] = gp(A.nl[comp])[c];
        vnl1[comp] = gp(A.nl1[comp])[c];
        vdif[comp] = gp(A.dif[comp])[c];
        if (!TH) {
          vs[comp] = gp(A.s[comp])[c];
          if (!fn_alias) vfn[comp] = gp(A.fn[comp])[c];
        }
      }
    }
    double qE = 0.0, qW = 0.0, qN = 0.0, rN = 0.0, rS = 0.0;
    const long nf = (long)(d.nj + 1) * (d.nk + 1);
    if (DR) {
      rN = fN[c];
      rS = fN[c - sj];
    }
    if (qform) {
      qE = fE[c];
      qW = fE[c - 1];
      if (S == 0) qN = fN[c];
      if (i == d.ni - 1) {
#pragma unroll
        for (int comp = 0; comp < NC; ++comp)
          vuf[comp] = A.ufar[comp * nf + j + (long)(d.nj + 1) * k];
      }
    }
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
#define F(dk, dj, di) sh[comp][pl[(dk) + 1]][q0 + (dj) * SF_W + (di)]
      double r = pe_xe * (F(0, 1, 1) - F(0, -1, 1));
      r = r - pw_xe * (F(0, 1, -1) - F(0, -1, -1));
      r = r + pe_xz * (F(1, 0, 1) - F(-1, 0, 1));
      r = r - pw_xz * (F(1, 0, -1) - F(-1, 0, -1));
      r = r + pn_ex * (F(0, 1, 1) - F(0, 1, -1));
      r = r - ps_ex * (F(0, -1, 1) - F(0, -1, -1));
      r = r + pn_ez * (F(1, 1, 0) - F(-1, 1, 0));
      r = r - ps_ez * (F(1, -1, 0) - F(-1, -1, 0));
      r = r + pt_zx * (F(1, 0, 1) - F(1, 0, -1));
      r = r - pb_zx * (F(-1, 0, 1) - F(-1, 0, -1));
      r = r + pt_ze * (F(1, 1, 0) - F(1, -1, 0));
      r = r - pb_ze * (F(-1, 1, 0) - F(-1, -1, 0));
      const double dd = 0.25 * r;
      double su;
      if (S == 0 && TH) {
        su = a0 * F(0, 0, 0);
        su = su - 3.0 * re * vnl[comp];
        su = su + 1.0 * re * vnl1[comp];
        su = su + vdif[comp];
        su = su + dd;
      } else if (S == 0) {
        // U_N: equal to the staged stage input when nothing has written u since the step start
        su = a0 * (fn_alias ? F(0, 0, 0) : vfn[comp]);
        su = su - 3.0 * re * vnl[comp];
        su = su + 1.0 * re * vnl1[comp];
        su = su + vdif[comp];
        su = su + 2.0 * re * vs[comp];
        su = su + dd;
      } else {
        su = a0 * F(0, 0, 0) + dd;
      }
      if (qform) {
        double ae = qE;
        if (i == 1) su = su + 2.0 * qW * (TH ? kTWall : 0.0);
        if (i == d.ni - 1) {
          su = su + 2.0 * ae * vuf[comp];
          ae = 0.0;
        }
        su = S == 0 ? su + qN * F(0, 1, 0) : su + ae * F(0, 0, 1);
        su = su + at[1] * F(1, 0, 0);
        su = su + at[0] * F(-1, 0, 0);
      }
      if (DR) {
        // the periodic images of the branch-cut ghosts j = 0 and j = NJ
        const GP fc = gp(A.f[comp]);
        const double xjm = j == 1 ? fc[cidx(d, i, d.nj - 1, k)] : F(0, -1, 0);
        const double xjp = j == d.nj - 1 ? fc[cidx(d, i, 1, k)] : F(0, 1, 0);
        double fe = qE, fw = qW, ap = a0 + fe + fw + rN + rS + at[1] + at[0];
        if (i == 1) { ap = ap + fw; fw = 0.0; }
        if (i == d.ni - 1) { ap = ap + fe; fe = 0.0; }
        double r;
        if (S == 0) {  // rows along I, lines along J: lo = AW, up = AE, c_(a-1) = AS
          r = su + rS * xjm;
          r = r - (ap * F(0, 0, 0) - fw * F(0, 0, -1) - fe * F(0, 0, 1));
        } else {       // rows along J, lines along I: lo = AS, up = AN, c_(a-1) = AW
          r = su + fw * F(0, 0, -1);
          r = r - (ap * F(0, 0, 0) - rS * xjm - rN * xjp);
        }
        reinterpret_cast<float*>(A.su[comp])[c] = (float)r;
        continue;
      }
#undef F
      gpw(A.su[comp])[c] = su;
    }
  }
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


// Chunked iteration of the stage-s line Gauss-Seidel sweep.  The sweep is the block lower
// triangular system (D - L) x = U x_old + b over lines: D = line tridiagonals, L = couplings to
// the two "earlier" neighbours (a-1, b-1), U = couplings to the "later" ones (a+1, b+1).  A
// workgroup owns C consecutive lines along a at one b and sweeps them in the reference order
// (the a-1 value of each line is the line just solved, kept in registers); the first line's
// a-1 value and every b-1 value come from the previous iterate.  The iteration is nilpotent
// (exact after at most #chunks + nb passes) and contractive (|| D^-1 L || < 1 by the diagonal
// dominance AP = 2 Re J^-1/dt + sum A), so it converges to the exact lexicographic sweep; the
// host stops it once the largest change is below the tolerance.  Threads own rows p of the line
// (blockDim >= n) and load their row of all C lines up front: for the strided J and K lines the
// C values of one row are contiguous (the chunk runs along I).
template <int S, int C>
__global__ void __launch_bounds__(512) k_af_chunk(Dims d, AFArgs A, const double* const* fa,
                                                  const double* __restrict__ ap0,
                                                  const double* const* prevs,
                                                  double* const* nexts, int nchunk, long nwork,
                                                  double* part) {
  __shared__ double la[FVM_MAX_LINE], lb[FVM_MAX_LINE], lc[FVM_MAX_LINE], ld[FVM_MAX_LINE];
  __shared__ double r0[FVM_MAX_LINE], r1[FVM_MAX_LINE];
  // XCD-aware mapping: block ids are dealt round-robin over the 8 XCDs; consecutive work items
  // (neighbouring chunks sharing cache lines) are placed on the same XCD.
  const long G = (long)gridDim.x;
  const long lin = blockIdx.x;
  const long w = (lin % 8) * (G / 8) + lin / 8;
  const int t = threadIdx.x;
  double dmax = 0.0, xmax = 0.0;
  if (w < nwork) {
    const int na = S == 0 ? d.nj - 1 : d.ni - 1;
    const int nb = S == 2 ? d.nj - 1 : d.nk - 1;
    const int n = S == 0 ? d.ni - 1 : (S == 1 ? d.nj - 1 : d.nk - 1);
    const int ca = (int)(w % nchunk);
    const long rest = w / nchunk;
    const int b = 1 + (int)(rest % nb);
    const int comp = (int)(rest / nb);
    const long sa = S == 0 ? d.sj : 1;
    const long sb = S == 2 ? d.sj : d.sk;
    const long sp = S == 0 ? 1 : (S == 1 ? d.sj : d.sk);
    const double* fo = A.f[comp];  // stage input, frozen during the sweep
    const double* pv = prevs[comp];
    double* nx = nexts[comp];
    const double* su = A.su[comp];
    const long nf = (long)(d.nj + 1) * (d.nk + 1);
    const int p = t + 1;
    const bool row = p <= n;
    const int a0 = 1 + ca * C;
    double AP[C], LO[C], UP[C], T1[C], AM[C], T3[C], T4[C], SU[C], GE[C];
#pragma unroll
    for (int q = 0; q < C; ++q) {
      AP[q] = 1.0; LO[q] = UP[q] = T1[q] = AM[q] = T3[q] = T4[q] = SU[q] = GE[q] = 0.0;
      const int a = a0 + q;
      if (a > na || !row) continue;
      const long c = sa * a + sb * b + sp * p;
      // coefficients in the reference's order AE, AW, AN, AS, AT, AB (iter_u.f90:23-47)
      double ae = gp(fa[0])[c], aw = gp(fa[0])[c - 1], an = gp(fa[1])[c], as = gp(fa[1])[c - d.sj],
             at = gp(fa[2])[c], ab = gp(fa[2])[c - d.sk];
      double ap = ap0[c] + ae + aw + an + as + at + ab;
      double s = su[c];
      const int i = S == 0 ? p : a;
      if (i == 1) { ap = ap + aw; s = s + 2.0 * aw * 0.0; aw = 0.0; }  // bcu.f90:18-20
      if (i == d.ni - 1) {                                            // bcu.f90:38-40
        const int jj = S == 0 ? a : (S == 1 ? p : b);
        const int kk = S == 0 ? b : (S == 1 ? b : p);
        const double uf = A.ufar[comp * nf + jj + (long)(d.nj + 1) * kk];
        ap = ap + ae;
        s = s + 2.0 * ae * uf;
        ae = 0.0;
      }
      double cap, cam, cbp, cbm, lo, up;
      if (S == 0) { cap = an; cam = as; cbp = at; cbm = ab; lo = -aw; up = -ae; }
      else if (S == 1) { cap = ae; cam = aw; cbp = at; cbm = ab; lo = -as; up = -an; }
      else { cap = ae; cam = aw; cbp = an; cbm = as; lo = -ab; up = -at; }
      T1[q] = cap * fo[c + sa];
      AM[q] = cam;
      T3[q] = cbp * fo[c + sb];
      T4[q] = cbm * (b > 1 ? pv[c - sb] : fo[c - sb]);
      SU[q] = s;
      if (p == 1) { GE[q] = lo * fo[c - sp]; lo = 0.0; }  // line-end Dirichlet data
      if (p == n) { GE[q] = up * fo[c + sp]; up = 0.0; }
      AP[q] = ap; LO[q] = lo; UP[q] = up;
    }
    double xe = 0.0;
    if (row) {
      const long c0 = sa * a0 + sb * b + sp * p;
      xe = a0 > 1 ? pv[c0 - sa] : fo[c0 - sa];
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int a = a0 + q;
      if (a > na) break;  // uniform over the block
      if (row) {
        // C = A_a+ x(a+1) + A_a- x(a-1) + A_b+ x(b+1) + A_b- x(b-1) + SU, minus the end term
        double rhs = T1[q] + AM[q] * xe;
        rhs = rhs + T3[q];
        rhs = rhs + T4[q];
        rhs = rhs + SU[q];
        rhs = rhs - GE[q];
        la[t] = LO[q]; lb[t] = AP[q]; lc[t] = UP[q]; ld[t] = rhs;
      }
      __syncthreads();
      pcr_solve_rows(la, lb, lc, ld, n);
      if (row) {
        const long c = sa * a + sb * b + sp * p;
        const double x = ld[t];
        dmax = fmax(dmax, fabs(x - pv[c]));
        xmax = fmax(xmax, fabs(x));
        nx[c] = x;
        xe = x;
      }
      __syncthreads();
    }
  }
  r0[t] = dmax;
  r1[t] = xmax;
  __syncthreads();
  for (int st = 256; st > 0; st >>= 1) {
    if (t < st && t + st < (int)blockDim.x) {
      r0[t] = fmax(r0[t], r0[t + st]);
      r1[t] = fmax(r1[t], r1[t + st]);
    }
    __syncthreads();
  }
  if (t == 0) { part[2 * lin] = r0[0]; part[2 * lin + 1] = r1[0]; }
}

__global__ void k_max_final(const double* __restrict__ part, long nb, double* out) {
  __shared__ double s0[256], s1[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (long q = t; q < nb; q += blockDim.x) { a = fmax(a, part[2 * q]); b = fmax(b, part[2 * q + 1]); }
  s0[t] = a; s1[t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { s0[t] = fmax(s0[t], s0[t + s]); s1[t] = fmax(s1[t], s1[t + s]); }
    __syncthreads();
  }
  if (t == 0) { out[0] = s0[0]; out[1] = s1[0]; }
}

// copy every ghost cell (I, J and K boundary planes) of src into dst
__global__ void k_copy_ghosts(Dims d, const double* __restrict__ src, double* dst) {
  const long nI = 2L * (d.nj + 1) * (d.nk + 1);
  const long nJ = 2L * (d.ni + 1) * (d.nk + 1);
  const long nK = 2L * (d.ni + 1) * (d.nj + 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nI + nJ + nK;
       t += (long)gridDim.x * blockDim.x) {
    long c;
    if (t < nI) {
      long r = t / 2;
      int i = (t % 2) ? d.ni : 0;
      c = cidx(d, i, (int)(r % (d.nj + 1)), (int)(r / (d.nj + 1)));
    } else if (t < nI + nJ) {
      long q = t - nI, r = q / 2;
      int j = (q % 2) ? d.nj : 0;
      c = cidx(d, (int)(r % (d.ni + 1)), j, (int)(r / (d.ni + 1)));
    } else {
      long q = t - nI - nJ, r = q / 2;
      int k = (q % 2) ? d.nk : 0;
      c = cidx(d, (int)(r % (d.ni + 1)), (int)(r / (d.ni + 1)), k);
    }
    dst[c] = src[c];
  }
}

// The same sweep with the forward results held in registers (lines of at most NMAX rows): one
// read of SU and the factors, one write of the result.  The components of a line sit in the
// wavefronts of one workgroup (threadIdx.y), which load the shared factors together: one HBM
// read of them serves all three.
template <int NMAX>
__global__ void __launch_bounds__(192) k_af_thomas_k_reg(Dims d, AFArgs A,
                                                         const double* __restrict__ ta,
                                                         const double* __restrict__ td,
                                                         const double* __restrict__ tb) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int comp = threadIdx.y;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const GPW f = gpw(A.f[comp]);
  const GP su = gp(A.su[comp]);
  const int n = d.nk - 1;
  const long c0 = cidx(d, i, j, 0);
  // Rows are handled in blocks of KB: the operands of block b+1 are loaded before block b is
  // solved (the scheduling barrier keeps them there), so one memory latency covers KB rows
  // instead of each row waiting for its own loads.  Rows past n repeat row n and are never kept,
  // so no load sits behind a branch on n.
  constexpr int KB = 8, NB = NMAX / KB;
  static_assert(NMAX % KB == 0, "row blocks");
  double xs[NMAX];
  double x = f[c0];  // K ghost of the stage input
  double bs[KB], ba[KB], bd[KB];
#pragma unroll
  for (int r = 0; r < KB; ++r) {
    const long c = c0 + (long)min(r + 1, n) * d.sk;
    bs[r] = su[c]; ba[r] = ta[c]; bd[r] = td[c];
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    double ns[KB], na[KB], nd[KB];
    if (b + 1 < NB) {
#pragma unroll
      for (int r = 0; r < KB; ++r) {
        const long c = c0 + (long)min((b + 1) * KB + r + 1, n) * d.sk;
        ns[r] = su[c]; na[r] = ta[c]; nd[r] = td[c];
      }
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < KB; ++r) {
      const int k = b * KB + r + 1;
      const double y = (bs[r] - ba[r] * x) / bd[r];
      x = k <= n ? y : x;
      xs[k - 1] = x;
    }
    if (b + 1 < NB) {
#pragma unroll
      for (int r = 0; r < KB; ++r) { bs[r] = ns[r]; ba[r] = na[r]; bd[r] = nd[r]; }
    }
  }
  double xn = f[c0 + (long)(n + 1) * d.sk];
  double bt[KB];
#pragma unroll
  for (int r = 0; r < KB; ++r) bt[r] = tb[c0 + (long)min(NMAX - r, n) * d.sk];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    double nt[KB];
    if (b + 1 < NB) {
#pragma unroll
      for (int r = 0; r < KB; ++r) nt[r] = tb[c0 + (long)min(NMAX - (b + 1) * KB - r, n) * d.sk];
    }
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int r = 0; r < KB; ++r) {
      const int k = NMAX - b * KB - r;
      const double y = xs[k - 1] - bt[r] * xn;
      if (k <= n) {
        f[c0 + (long)k * d.sk] = y;
        xn = y;
      }
    }
    if (b + 1 < NB) {
#pragma unroll
      for (int r = 0; r < KB; ++r) bt[r] = nt[r];
    }
  }
}

// Largest |coupling| of stage S over the interior: part[2*block] along the sweep direction a,
// part[2*block+1] across it (b).  Exact zeros (extruded grid: xi_z = eta_z = zeta_x = zeta_y = 0)
// make the lines of a stage independent, and the sweep is then exact in one pass.
template <int S>
__global__ void k_af_couplings(Dims d, const double* const* __restrict__ fa,
                               const double* __restrict__ ap0, double* part) {
  __shared__ double s0[256], s1[256], s2[256];
  const long nint = (long)(d.ni - 1) * (d.nj - 1) * (d.nk - 1);
  double ma = 0.0, mb = 0.0, rb = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint;
       t += (long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(t % (d.ni - 1));
    const long r = t / (d.ni - 1);
    const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    const long c = cidx(d, i, j, k);
    const double ae = gp(fa[0])[c], aw = gp(fa[0])[c - 1], an = gp(fa[1])[c], as = gp(fa[1])[c - d.sj],
                 at = gp(fa[2])[c], ab = gp(fa[2])[c - d.sk];
    const double ca = S == 0 ? fmax(fabs(an), fabs(as)) : fmax(fabs(ae), fabs(aw));
    const double cb = S == 2 ? fmax(fabs(an), fabs(as)) : fmax(fabs(at), fabs(ab));
    const double sb = S == 2 ? fabs(an) + fabs(as) : fabs(at) + fabs(ab);
    ma = fmax(ma, ca);
    mb = fmax(mb, cb);
    rb = fmax(rb, sb / ap0[c]);
  }
  const int t = threadIdx.x;
  s0[t] = ma; s1[t] = mb; s2[t] = rb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) {
      s0[t] = fmax(s0[t], s0[t + st]);
      s1[t] = fmax(s1[t], s1[t + st]);
      s2[t] = fmax(s2[t], s2[t + st]);
    }
    __syncthreads();
  }
  if (t == 0) {
    part[3 * blockIdx.x] = s0[0]; part[3 * blockIdx.x + 1] = s1[0]; part[3 * blockIdx.x + 2] = s2[0];
  }
}

// Largest |xi_z|, |eta_z| over every interior cell and its six neighbours (part[block])
__global__ void k_zmetric_max(Dims d, const double* __restrict__ xiz,
                              const double* __restrict__ etz, double* part) {
  __shared__ double s0[256];
  const long nint = (long)(d.ni - 1) * (d.nj - 1) * (d.nk - 1);
  double mx = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint;
       t += (long)gridDim.x * blockDim.x) {
    const int i = 1 + (int)(t % (d.ni - 1));
    const long r = t / (d.ni - 1);
    const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    const long c = cidx(d, i, j, k);
    const long nb[7] = {c, c + 1, c - 1, c + d.sj, c - d.sj, c + d.sk, c - d.sk};
    for (int q = 0; q < 7; ++q) mx = fmax(mx, fmax(fabs(xiz[nb[q]]), fabs(etz[nb[q]])));
  }
  const int t = threadIdx.x;
  s0[t] = mx;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) s0[t] = fmax(s0[t], s0[t + st]);
    __syncthreads();
  }
  if (t == 0) part[blockIdx.x] = s0[0];
}

// Factors of the reference's Thomas recursion along a K-line (iter_u.f90:240-255):
// A = -AB, D = AP - A B(K-1), B = -AT/D with B(0) = 0 (the SAVE'd scratch is never written at
// index 0).  AP carries the wall / far-field folds of BCU.  One thread per (I, J) line.
__global__ void k_af_factor_k(Dims d, const double* const* __restrict__ fa,
                              const double* __restrict__ ap0, double* ta, double* td,
                              double* tb) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  double bprev = 0.0;
  for (int k = 1; k <= d.nk - 1; ++k) {
    const long c = cidx(d, i, j, k);
    const double ae = gp(fa[0])[c], aw = gp(fa[0])[c - 1], an = gp(fa[1])[c], as = gp(fa[1])[c - d.sj],
                 at = gp(fa[2])[c], ab = gp(fa[2])[c - d.sk];
    double ap = ap0[c] + ae + aw + an + as + at + ab;
    if (i == 1) ap = ap + aw;          // bcu.f90:18
    if (i == d.ni - 1) ap = ap + ae;   // bcu.f90:38
    const double A = -ab;
    const double D = ap - A * bprev;
    const double B = -at / D;
    ta[c] = A; td[c] = D; tb[c] = B;
    bprev = B;
  }
}

// The K-line sweep when its lines are independent: one thread per line and component, lanes
// along I (coalesced), the reference's forward elimination and back substitution in its
// operation order, in place.  With the I and J couplings identically zero the lagged-neighbour
// sum C of the reference is exactly SU (the BCU additions to SU multiply the same zeros).
__global__ void __launch_bounds__(192) k_af_thomas_k(Dims d, AFArgs A,
                                                     const double* __restrict__ ta,
                                                     const double* __restrict__ td,
                                                     const double* __restrict__ tb) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y;
  const int comp = threadIdx.y;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  double* f = A.f[comp];
  const double* su = A.su[comp];
  const int n = d.nk - 1;
  const long c0 = cidx(d, i, j, 0);
  double x = f[c0];  // K ghost of the stage input
  for (int k = 1; k <= n; ++k) {
    const long c = c0 + (long)k * d.sk;
    x = (su[c] - ta[c] * x) / td[c];
    f[c] = x;
  }
  double xn = f[c0 + (long)(n + 1) * d.sk];
  for (int k = n; k >= 1; --k) {
    const long c = c0 + (long)k * d.sk;
    x = f[c] - tb[c] * xn;
    f[c] = x;
    xn = x;
  }
}

}  // namespace

template <int S, int C>
static int chunk_stage(FvmCtx* c, AFArgs& A, int* iters) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  const int na = S == 0 ? d.nj - 1 : d.ni - 1;
  const int nb = S == 2 ? d.nj - 1 : d.nk - 1;
  const int n = S == 0 ? d.ni - 1 : (S == 1 ? d.nj - 1 : d.nk - 1);
  const int nchunk = (na + C - 1) / C;
  const long nwork = (long)nchunk * nb * 3;
  const long G = (nwork + 7) / 8 * 8;
  if (2 * G > c->nred) return FVM_ERR_SIZE;
  const int threads = (n + 63) / 64 * 64;
  double** fld[3] = {&c->u, &c->v, &c->w};
  double* cur[3] = {c->u, c->v, c->w};
  double* prev[3] = {cur[0], cur[1], cur[2]};
  double* next[3];
  int slot = 0;
  int it = 0;
  const int cap = nchunk + nb + 1;
  const double tol = c->prm.af_tol > 0 ? c->prm.af_tol : 1e-14;
  for (;;) {
    for (int q = 0; q < 3; ++q) next[q] = c->afb[q][slot];
    HIP_CHECK(hipMemcpyAsync(c->afptr, prev, 3 * sizeof(double*), hipMemcpyHostToDevice, st));
    HIP_CHECK(hipMemcpyAsync(c->afptr + 3, next, 3 * sizeof(double*), hipMemcpyHostToDevice, st));
    {
      ProfScope pk(c, S == 0 ? "k_af_chunk_i" : (S == 1 ? "k_af_chunk_j" : "k_af_chunk_k"));
      hipLaunchKernelGGL((k_af_chunk<S, C>), dim3((unsigned)G), dim3(threads), 0, st, d, A,
                         (const double* const*)c->dfa[S], c->ap0, (const double* const*)c->afptr,
                         (double* const*)(c->afptr + 3), nchunk, nwork, c->red);
    }
    hipLaunchKernelGGL(k_max_final, dim3(1), dim3(256), 0, st, c->red, G, c->scal + 20);
    HIP_CHECK(hipGetLastError());
    const int rf = fvm_fetch(c, c->scal + 20, 2, c->hscal + 20);
    if (rf) return rf;
    ++it;
    const double dm = c->hscal[20], xm = c->hscal[21];
    if (!(dm == dm) || !(xm == xm)) {
      fprintf(stderr, "fvm_amd: non-finite AF stage %d iterate (pass %d)\n", S, it);
      return FVM_ERR_NAN;
    }
    for (int q = 0; q < 3; ++q) prev[q] = next[q];
    slot ^= 1;
    if ((it >= 2 && dm <= tol * xm) || it >= cap) break;
  }
  // the sweep leaves the ghost values of the stage input untouched: carry them over, then
  // the converged iterate becomes the field
  for (int q = 0; q < 3; ++q) {
    long ng = 2L * ((d.nj + 1) * (d.nk + 1) + (d.ni + 1) * (d.nk + 1) + (d.ni + 1) * (d.nj + 1));
    hipLaunchKernelGGL(k_copy_ghosts, grid1(ng, 256), dim3(256), 0, st, d, cur[q], prev[q]);
    HIP_CHECK(hipGetLastError());
    for (int sl = 0; sl < 2; ++sl)
      if (c->afb[q][sl] == prev[q]) c->afb[q][sl] = cur[q];
    *fld[q] = prev[q];
  }
  *iters = it;
  return FVM_OK;
}

int fvm_k_iter_u(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  const int fn_alias = c->u_is_un ? 1 : 0;
  c->u_is_un = false;  // the stages write u, v, w
  if (d.ni - 1 > FVM_MAX_LINE || d.nj - 1 > FVM_MAX_LINE || d.nk - 1 > FVM_MAX_LINE)
    return FVM_ERR_SIZE;
  const dim3 gi((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  // the 2D scheme has two sweeps per component, I then J (2D iter_u.f90 / iter_v.f90): the
  // one-cell span of the 2D slab gets no K stage
  const int nstage = c->two_d ? 2 : 3;
  if (c->two_d) c->af_iters[2] = 0;
  // af_mode 0 and 2: chunked line scans; 2: the I and J stages in the FP32 change form
  const bool chunked = c->prm.af_mode == 0 || c->prm.af_mode == 2;
  const bool f32 = c->prm.af_mode == 2 && c->lf32[0][0] && c->lf32[1][0];
  for (int s = 0; s < nstage; ++s) {
    AFArgs A;
    double* fld[3] = {c->u, c->v, c->w};
    const double* fn[3] = {c->un, c->vn, c->wn};
    for (int q = 0; q < 3; ++q) {
      A.f[q] = fld[q]; A.fn[q] = fn[q]; A.nl[q] = c->nl[q]; A.nl1[q] = c->nl1[q];
      A.dif[q] = c->dif[q]; A.s[q] = c->s[q]; A.su[q] = c->su[q];
    }
    A.ufar = c->ufar;
    {
      ProfScope ps(c, "af_rhs");
      const int qf = (s < 2 && chunked && c->af_scan[s]) ? 1 : 0;
      const double* const* fas = (const double* const*)c->dfa[s];
      const int dz = c->dd2_zero ? 1 : 0;
      const dim3 gk((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + RK_KC - 1) / RK_KC);
      if (s == 0) {
        ProfScope pk(c, "k_af_rhs_i");
        auto kern = (qf && f32) ? k_af_rhs_km<0, false, 3, true> : k_af_rhs_km<0, false, 3, false>;
        hipLaunchKernelGGL(kern, gk, kBlk, 0, st, d, A, (const double* const*)c->dafp[0], c->ap0, c->prm.re, fas, qf, fn_alias);
      }
      if (s == 1) {
        ProfScope pk(c, "k_af_rhs_j");
        auto kern = (qf && f32) ? k_af_rhs_km<1, false, 3, true> : k_af_rhs_km<1, false, 3, false>;
        hipLaunchKernelGGL(kern, gk, kBlk, 0, st, d, A, (const double* const*)c->dafp[1], c->ap0, c->prm.re, fas, qf, 0);
      }
      if (s == 2) hipLaunchKernelGGL(k_af_rhs<2>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm, c->djac, c->ap0, c->prm.re, fas, qf, dz);
      HIP_CHECK(hipGetLastError());
    }
    // boundary ghosts of the stage input (BCU, iter_u.f90:83,160,237)
    int r = fvm_k_bcu_ghosts(c, fld, 3);
    if (r) return r;
    ProfScope ps(c, s == 0 ? "af_sweep_i" : (s == 1 ? "af_sweep_j" : "af_sweep_k"));
    if (s == 2 && chunked && c->tk_a && c->af_zero_a[2] && c->af_zero_b[2]) {
      // independent K-lines: the exact sweep in one pass
      ProfScope pk(c, "k_af_thomas_k");
      const dim3 g((d.ni - 1 + 63) / 64, d.nj - 1, 1);
      if (d.nk - 1 <= 64)
        hipLaunchKernelGGL(k_af_thomas_k_reg<64>, g, dim3(64, A.nc, 1), 0, st, d, A, c->tk_a,
                           c->tk_d, c->tk_b);
      else
        hipLaunchKernelGGL(k_af_thomas_k, g, dim3(64, A.nc, 1), 0, st, d, A, c->tk_a, c->tk_d,
                           c->tk_b);
      HIP_CHECK(hipGetLastError());
      c->af_iters[s] = 1;
    } else if (s < 2 && chunked && c->af_scan[s]) {
      int it = 0;
      r = fvm_af_scan_stage(c, s, A, &it, false, f32);
      if (r) return r;
      c->af_iters[s] = it;
    } else if (c->prm.af_mode == 1) {
      // anti-diagonal wavefront: one launch per anti-diagonal of the line plane
      const int na = s == 0 ? d.nj - 1 : d.ni - 1;
      const int nb = s == 2 ? d.nj - 1 : d.nk - 1;
      const double* const* fa = (const double* const*)c->dfa[s];
      for (int h = 0; h <= na + nb - 2; ++h) {
        int lo = h - (nb - 1) > 0 ? h - (nb - 1) : 0;
        int hi = h < na - 1 ? h : na - 1;
        dim3 g(hi - lo + 1, 3);
        if (s == 0) hipLaunchKernelGGL(k_af_sweep<0>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
        if (s == 1) hipLaunchKernelGGL(k_af_sweep<1>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
        if (s == 2) hipLaunchKernelGGL(k_af_sweep<2>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
      }
      HIP_CHECK(hipGetLastError());
      c->af_iters[s] = na + nb - 1;
    } else {
      int it = 0;
      if (s == 0) r = chunk_stage<0, 8>(c, A, &it);
      if (s == 1) r = chunk_stage<1, 8>(c, A, &it);
      if (s == 2) r = chunk_stage<2, 8>(c, A, &it);
      if (r) return r;
      c->af_iters[s] = it;
    }
  }
  return FVM_OK;
}

// ITER_U(IN, INM, INA, INB, INC, IND) for one velocity component (RLES/iter_u.f90:1-262 as
// MAIN calls it, main.f90:82-88): f = IN (the field, updated), fn = INM (its step-start copy),
// nl / nl1 / dif / s = INA..IND; comp (0 u, 1 v, 2 w) picks BCU's far-field values, i.e. the UIN
// MAIN sets before the call.  The same stages as fvm_k_iter_u with one component per launch;
// grids whose stage couplings need the chunked fallback are refused (FVM_ERR_ARG).
int fvm_k_iter_u_one(FvmCtx* c, int comp, double* f, const double* fn, const double* nl,
                     const double* nl1, const double* dif, const double* s_) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  if (comp < 0 || comp > 2) return FVM_ERR_ARG;
  if (d.ni - 1 > FVM_MAX_LINE || d.nj - 1 > FVM_MAX_LINE || d.nk - 1 > FVM_MAX_LINE)
    return FVM_ERR_SIZE;
  c->u_is_un = false;
  const dim3 gi((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  const int nstage = c->two_d ? 2 : 3;
  const long nf = (long)(d.nj + 1) * (d.nk + 1);
  for (int s = 0; s < nstage; ++s) {
    const bool chunked = c->prm.af_mode == 0 || c->prm.af_mode == 2;  // FP64 scans here
    const bool exact_k = s == 2 && chunked && c->tk_a && c->af_zero_a[2] && c->af_zero_b[2];
    const bool scan = s < 2 && chunked && c->af_scan[s];
    if (!exact_k && !scan && c->prm.af_mode != 1) {
      fprintf(stderr, "fvm_amd: ITER_U of one component needs line-separable stage couplings "
                      "(stage %d is not); use fvm_iter_u\n", s);
      return FVM_ERR_ARG;
    }
    AFArgs A;
    A.nc = 1;
    for (int q = 0; q < 3; ++q) {
      A.f[q] = f; A.fn[q] = fn; A.nl[q] = nl; A.nl1[q] = nl1; A.dif[q] = dif; A.s[q] = s_;
      A.su[q] = c->su[comp];
    }
    A.ufar = c->ufar + comp * nf;
    {
      ProfScope ps(c, "af_rhs");
      const int qf = scan ? 1 : 0;
      const double* const* fas = (const double* const*)c->dfa[s];
      const int dz = c->dd2_zero ? 1 : 0;
      const dim3 gk((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + RK_KC - 1) / RK_KC);
      if (s == 0)
        hipLaunchKernelGGL((k_af_rhs_km<0, false, 1>), gk, kBlk, 0, st, d, A,
                           (const double* const*)c->dafp[0], c->ap0, c->prm.re, fas, qf, 0);
      if (s == 1)
        hipLaunchKernelGGL((k_af_rhs_km<1, false, 1>), gk, kBlk, 0, st, d, A,
                           (const double* const*)c->dafp[1], c->ap0, c->prm.re, fas, qf, 0);
      if (s == 2)
        hipLaunchKernelGGL(k_af_rhs<2>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm,
                           c->djac, c->ap0, c->prm.re, fas, qf, dz);
      HIP_CHECK(hipGetLastError());
    }
    int r = fvm_k_bcu_ghosts_one(c, f, comp);  // BCU(IN, INM), iter_u.f90:83,160,237
    if (r) return r;
    ProfScope ps(c, s == 0 ? "af_sweep_i" : (s == 1 ? "af_sweep_j" : "af_sweep_k"));
    if (exact_k) {
      const dim3 g((d.ni - 1 + 63) / 64, d.nj - 1, 1);
      if (d.nk - 1 <= 64)
        hipLaunchKernelGGL(k_af_thomas_k_reg<64>, g, dim3(64, 1, 1), 0, st, d, A, c->tk_a,
                           c->tk_d, c->tk_b);
      else
        hipLaunchKernelGGL(k_af_thomas_k, g, dim3(64, 1, 1), 0, st, d, A, c->tk_a, c->tk_d,
                           c->tk_b);
      HIP_CHECK(hipGetLastError());
    } else if (scan) {
      int it = 0;
      r = fvm_af_scan_stage(c, s, A, &it);
      if (r) return r;
    } else {  // af_mode 1: anti-diagonal wavefront launches
      const int na = s == 0 ? d.nj - 1 : d.ni - 1;
      const int nb = s == 2 ? d.nj - 1 : d.nk - 1;
      const double* const* fa = (const double* const*)c->dfa[s];
      for (int h = 0; h <= na + nb - 2; ++h) {
        int lo = h - (nb - 1) > 0 ? h - (nb - 1) : 0;
        int hi = h < na - 1 ? h : na - 1;
        dim3 g(hi - lo + 1, 1);
        if (s == 0) hipLaunchKernelGGL(k_af_sweep<0>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
        if (s == 1) hipLaunchKernelGGL(k_af_sweep<1>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
        if (s == 2) hipLaunchKernelGGL(k_af_sweep<2>, g, dim3(64), 0, st, d, A, fa, c->ap0, h);
      }
      HIP_CHECK(hipGetLastError());
    }
  }
  return FVM_OK;
}

// Which stage couplings vanish identically, and the K-line Thomas factors when that stage's
// lines are independent.  Called after the geometry set-up and after every change of dt.
int fvm_af_setup(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  const int nb = 256;
  if (!c->afp[0][0]) {  // metric products of the stage 0 and 1 cross terms (geometry only)
    for (int s = 0; s < 2; ++s) {
      for (int q = 0; q < 6; ++q) {
        HIP_CHECK(hipMalloc((void**)&c->afp[s][q], d.n * sizeof(double)));
        c->allocs.push_back(c->afp[s][q]);
      }
      HIP_CHECK(hipMalloc((void**)&c->dafp[s], 6 * sizeof(double*)));
      c->allocs.push_back(c->dafp[s]);
      HIP_CHECK(hipMemcpy(c->dafp[s], c->afp[s], 6 * sizeof(double*), hipMemcpyHostToDevice));
      hipLaunchKernelGGL(k_af_products, grid1(d.n, 256), dim3(256), 0, st, d, c->djac,
                         c->m[3 * s], c->m[3 * s + 1], c->m[3 * s + 2],
                         (double* const*)c->dafp[s]);
      HIP_CHECK(hipGetLastError());
    }
  }
  std::vector<double> h(3 * nb);
  for (int s = 0; s < 3; ++s) {
    if (s == 0) hipLaunchKernelGGL(k_af_couplings<0>, dim3(nb), dim3(256), 0, st, d, (const double* const*)c->dfa[0], c->ap0, c->red);
    if (s == 1) hipLaunchKernelGGL(k_af_couplings<1>, dim3(nb), dim3(256), 0, st, d, (const double* const*)c->dfa[1], c->ap0, c->red);
    if (s == 2) hipLaunchKernelGGL(k_af_couplings<2>, dim3(nb), dim3(256), 0, st, d, (const double* const*)c->dfa[2], c->ap0, c->red);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(h.data(), c->red, 3 * nb * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double ma = 0.0, mb = 0.0, rb = 0.0;
    for (int q = 0; q < nb; ++q) {
      ma = std::max(ma, h[3 * q]); mb = std::max(mb, h[3 * q + 1]); rb = std::max(rb, h[3 * q + 2]);
    }
    c->af_zero_a[s] = (ma == 0.0);
    c->af_zero_b[s] = (mb == 0.0);
    c->af_b_ratio[s] = rb;
  }
  {
    hipLaunchKernelGGL(k_zmetric_max, dim3(nb), dim3(256), 0, st, d, c->m[6], c->m[7], c->red);
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipMemcpyAsync(h.data(), c->red, nb * sizeof(double), hipMemcpyDeviceToHost, st));
    HIP_CHECK(hipStreamSynchronize(st));
    double mz = 0.0;
    for (int q = 0; q < nb; ++q) mz = std::max(mz, h[q]);
    c->dd2_zero = (mz == 0.0);
  }
  if (c->af_zero_a[2] && c->af_zero_b[2]) {
    if (!c->tk_a) {
      double** arr[3] = {&c->tk_a, &c->tk_d, &c->tk_b};
      for (double** q : arr) {
        HIP_CHECK(hipMalloc((void**)q, d.n * sizeof(double)));
        c->allocs.push_back(*q);
      }
    }
    const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, 1);
    hipLaunchKernelGGL(k_af_factor_k, g, dim3(64, 4, 1), 0, st, d,
                       (const double* const*)c->dfa[2], c->ap0, c->tk_a, c->tk_d, c->tk_b);
    if (c->thermal) {  // the temperature's K-line factors (2 Pe J^-1/dt on the diagonal)
      if (!c->tk_at) {
        double** arr[3] = {&c->tk_at, &c->tk_dt, &c->tk_bt};
        for (double** q : arr) {
          HIP_CHECK(hipMalloc((void**)q, d.n * sizeof(double)));
          c->allocs.push_back(*q);
        }
      }
      hipLaunchKernelGGL(k_af_factor_k, g, dim3(64, 4, 1), 0, st, d,
                         (const double* const*)c->dfa[2], c->ap0t, c->tk_at, c->tk_dt,
                         c->tk_bt);
    }
    HIP_CHECK(hipGetLastError());
    HIP_CHECK(hipStreamSynchronize(st));
  }
  return fvm_af_scan_setup(c);
}

// ITER_T (RLES/iter_t.f90:1-247) followed by BCT(T, T_N) (main.f90:145-147): the three AF stages
// of the momentum step on the temperature alone, with 2 Pe J^-1/dt for 2 Re J^-1/dt and BCT's
// folds.  Needs the exact K-line stage and the line-scan form of the I and J stages (checked
// by fvm_thermal_setup).
int fvm_k_iter_t(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  if (!c->thermal) return FVM_ERR_ARG;
  AFArgs A;
  A.nc = 1;
  for (int q = 0; q < 3; ++q) {
    A.f[q] = c->t; A.fn[q] = c->tn; A.nl[q] = c->adv; A.nl1[q] = c->adv1;
    A.dif[q] = c->dift; A.s[q] = nullptr; A.su[q] = c->su[0];
  }
  A.ufar = c->tfar;
  const dim3 gi((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  const dim3 gk((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + RK_KC - 1) / RK_KC);
  for (int s = 0; s < 3; ++s) {
    {
      ProfScope ps(c, "t_rhs");
      const double* const* fas = (const double* const*)c->dfa[s];
      if (s == 0) hipLaunchKernelGGL((k_af_rhs_km<0, true>), gk, kBlk, 0, st, d, A, (const double* const*)c->dafp[0], c->ap0t, c->pe, fas, 1, 0);
      if (s == 1) hipLaunchKernelGGL((k_af_rhs_km<1, true>), gk, kBlk, 0, st, d, A, (const double* const*)c->dafp[1], c->ap0t, c->pe, fas, 1, 0);
      if (s == 2) hipLaunchKernelGGL(k_af_rhs<2>, gi, kBlk, 0, st, d, A, (const double* const*)c->dm, c->djac, c->ap0t, c->pe, fas, 0, c->dd2_zero ? 1 : 0);
      HIP_CHECK(hipGetLastError());
    }
    int r = fvm_k_bct_ghosts(c, c->t);  // iter_t.f90:84,150,226
    if (r) return r;
    ProfScope ps(c, "t_sweep");
    if (s < 2) {
      int it = 0;
      r = fvm_af_scan_stage(c, s, A, &it, true);
      if (r) return r;
    } else {
      const dim3 g((d.ni - 1 + 63) / 64, d.nj - 1, 1);
      if (d.nk - 1 <= 64)
        hipLaunchKernelGGL(k_af_thomas_k_reg<64>, g, dim3(64, 1, 1), 0, st, d, A, c->tk_at,
                           c->tk_dt, c->tk_bt);
      else
        hipLaunchKernelGGL(k_af_thomas_k, g, dim3(64, 1, 1), 0, st, d, A, c->tk_at, c->tk_dt,
                           c->tk_bt);
      HIP_CHECK(hipGetLastError());
    }
  }
  return fvm_k_bct_ghosts(c, c->t);  // BCT(T, T_N), main.f90:147
}
