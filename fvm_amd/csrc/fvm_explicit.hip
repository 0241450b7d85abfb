// Explicit stencils of the time step on interior cells, boundary conditions, projection,
// pressure reference and force coefficients.
//
// Every kernel evaluates the reference's expressions in the reference's association order
// (FP contraction is off for this library), so its FP64 results equal the reference's bit for
// bit where the operation sequence is the same.
#include "fvm_ctx.h"

#include <math.h>
#include <vector>

namespace {

// interior launch: x -> i (64 wide), y -> j (4), z -> k
inline dim3 interior_grid(const Dims& d, int nz = 1) {
  return dim3((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1) * nz);
}
const dim3 kBlk(64, 4, 1);

#define INTERIOR_IJK(d)                                   \
  const uint3 bb = xcd_block();                           \
  const int i = 1 + bb.x * 64 + threadIdx.x;              \
  const int j = 1 + bb.y * 4 + threadIdx.y;               \
  const int k = 1 + (int)(bb.z % (d.nk - 1));             \
  if (i > d.ni - 1 || j > d.nj - 1) return;               \
  const long c = cidx(d, i, j, k);

// Contravariant velocity times J^-1 on i = 0..NI, interior j,k (RLES/contra_vel.f90:13-23);
// the face volume fluxes are 0.5 (JU(c) + JU(c+1)) (contra_vel.f90:37-53).
template <bool XZ0>
__global__ void k_contra(Dims d, const double* const* __restrict__ m, const double* __restrict__ djac,
                         const double* __restrict__ u, const double* __restrict__ v,
                         const double* __restrict__ w, double* ju0, double* ju1, double* ju2) {
  const uint3 bb = xcd_block();
  const int i = bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  double U = u[c], V = v[c], W = w[c], J = djac[c];
  double u1 = gp(m[0])[c] * U + gp(m[3])[c] * V + mget<XZ0>(m, 6, c) * W;
  double u2 = gp(m[1])[c] * U + gp(m[4])[c] * V + mget<XZ0>(m, 7, c) * W;
  double u3 = gp(m[2])[c] * U + gp(m[5])[c] * V + gp(m[8])[c] * W;
  ju0[c] = J * u1;
  ju1[c] = J * u2;
  ju2[c] = J * u3;
}

struct Faces {
  double ue, uw, vn, vs, wt, wb;
};

__device__ inline Faces faces(const Dims& d, int i, long c, const double* __restrict__ ju0,
                              const double* __restrict__ ju1, const double* __restrict__ ju2) {
  Faces f;
  f.ue = 0.5 * (ju0[c] + ju0[c + 1]);
  f.uw = i == 1 ? 0.0 : 0.5 * (ju0[c] + ju0[c - 1]);  // UW(1) = 0 (contra_vel.f90:56)
  f.vn = 0.5 * (ju1[c] + ju1[c + d.sj]);
  f.vs = 0.5 * (ju1[c] + ju1[c - d.sj]);
  f.wt = 0.5 * (ju2[c] + ju2[c + d.sk]);
  f.wb = 0.5 * (ju2[c] + ju2[c - d.sk]);
  return f;
}

// fields in and out of the per-component stencils: u_n, v_n, w_n (and T_N when the passive
// temperature is on)
struct CompSet {
  const double* in[4];
  double* out[4];
};

// NL(IN,OUT) for the velocity components and, with nc = 4, the temperature (NL(T_N,ADV),
// RLES/main.f90:58) on the same face fluxes (RLES/nl.f90:23-31)
__global__ void k_nl(Dims d, const double* __restrict__ ju0, const double* __restrict__ ju1,
                     const double* __restrict__ ju2, CompSet cs, int nc) {
  INTERIOR_IJK(d);
  Faces F = faces(d, i, c, ju0, ju1, ju2);
  for (int q = 0; q < nc; ++q) {
    const GP f = gp(cs.in[q]);
    double fc = f[c];
    double t = F.ue * (f[c + 1] + fc);
    t = t - F.uw * (f[c - 1] + fc);
    t = t + F.vn * (f[c + d.sj] + fc);
    t = t - F.vs * (f[c - d.sj] + fc);
    t = t + F.wt * (f[c + d.sk] + fc);
    t = t - F.wb * (f[c - d.sk] + fc);
    gpw(cs.out[q])[c] = 0.5 * t;
  }
}

// DIF(IN,OUT) for the three components (RLES/dif.f90:13-23): the orthogonal part
// sum D_f (phi_nb - phi_P) plus the cross-derivative terms DIF_DIAGONAL
// (RLES/dif_diagonal.f90:13-32, jqAB = J^-1 Q_AB), in the reference's order.  A workgroup owns a
// 64 x 4 (I x J) column of cells over DF_KC consecutive planes and marches up K with the last
// three planes of u_n, v_n, w_n (one-cell halo) in LDS, so the seven-point and twelve-edge
// stencils are read from HBM once per column; J^-1 Q13, J^-1 Q23 and D_T of the planes K+-1
// ride along in registers.
constexpr int DF_KC = 4;

// With WNL set the same pass also forms NL(IN, NLOUT) of k_nl (RLES/nl.f90:23-31, its operation
// order) from the staged planes and the face fluxes of ju (ju2 of the planes K+-1 in registers):
// the whole step runs NL and DIF in one pass over u_n, v_n, w_n (DIF does not depend on NL or on
// the SGS term added to NL in between).
template <int NC, bool WNL = false>
__global__ void __launch_bounds__(256) k_dif_km(Dims d, const double* __restrict__ pdI,
                                                const double* __restrict__ pdJ,
                                                const double* __restrict__ pdK,
                                                const double* __restrict__ jq12,
                                                const double* __restrict__ jq13,
                                                const double* __restrict__ jq23, CompSet cs,
                                                const double* __restrict__ ju0 = nullptr,
                                                const double* __restrict__ ju1 = nullptr,
                                                const double* __restrict__ ju2 = nullptr,
                                                CompSet nlo = CompSet{}) {
  const uint3 bb = xcd_block();
  __shared__ double sh[NC][3][SF_P];  // [component][plane slot][tile cell]
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * DF_KC;
  const int k1 = min(k0 + DF_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int q0 = (ty + 1) * SF_W + tx + 1;
  const long sj = d.sj;
  double j13[3] = {0.0, 0.0, 0.0}, j23[3] = {0.0, 0.0, 0.0}, pk[3] = {0.0, 0.0, 0.0};
  double z2[3] = {0.0, 0.0, 0.0};  // ju2 of the own column (NL's K faces)
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) % 3;
    __syncthreads();  // plane kp - 3, held in this slot, is no longer read
    for (int q = tid; q < SF_P; q += 256) {
      const int gi = min(bb.x * 64 + q % SF_W, d.ni);
      const int gj = min(bb.y * 4 + q / SF_W, d.nj);
      const long c = cidx(d, gi, gj, kp);
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) sh[comp][slot][q] = gp(cs.in[comp])[c];
    }
    j13[0] = j13[1]; j13[1] = j13[2];
    j23[0] = j23[1]; j23[1] = j23[2];
    pk[0] = pk[1]; pk[1] = pk[2];
    z2[0] = z2[1]; z2[1] = z2[2];
    if (act) {
      const long cp = cidx(d, i, j, kp);
      j13[2] = jq13[cp];
      j23[2] = jq23[cp];
      pk[2] = pdK[cp];
      if (WNL) z2[2] = ju2[cp];
    }
    __syncthreads();
    if (kp < k0 + 1 || !act) continue;
    const int k = kp - 1;
    const int pl[3] = {(slot + 1) % 3, (slot + 2) % 3, slot};
    const long c = cidx(d, i, j, k);
    if (WNL) {  // NL(IN, NLOUT): the face fluxes of k_nl's faces(), in its order
      const double ue = 0.5 * (ju0[c] + ju0[c + 1]);
      const double uw = i == 1 ? 0.0 : 0.5 * (ju0[c] + ju0[c - 1]);  // UW(1) = 0
      const double vn = 0.5 * (ju1[c] + ju1[c + sj]);
      const double vs = 0.5 * (ju1[c] + ju1[c - sj]);
      const double wt = 0.5 * (z2[1] + z2[2]);
      const double wb = 0.5 * (z2[1] + z2[0]);
#pragma unroll
      for (int comp = 0; comp < NC; ++comp) {
#define F(dk, dj, di) sh[comp][pl[(dk) + 1]][q0 + (dj) * SF_W + (di)]
        const double fc = F(0, 0, 0);
        double t = ue * (F(0, 0, 1) + fc);
        t = t - uw * (F(0, 0, -1) + fc);
        t = t + vn * (F(0, 1, 0) + fc);
        t = t - vs * (F(0, -1, 0) + fc);
        t = t + wt * (F(1, 0, 0) + fc);
        t = t - wb * (F(-1, 0, 0) + fc);
#undef F
        gpw(nlo.out[comp])[c] = 0.5 * t;
      }
    }
    const double de = pdI[c], dw = pdI[c - 1], dn = pdJ[c], ds = pdJ[c - sj], dt = pk[1],
                 db = pk[0];
    const double e12 = jq12[c + 1], w12 = jq12[c - 1], e13 = jq13[c + 1], w13 = jq13[c - 1];
    const double n12 = jq12[c + sj], s12 = jq12[c - sj], n23 = jq23[c + sj], s23 = jq23[c - sj];
#pragma unroll
    for (int comp = 0; comp < NC; ++comp) {
#define F(dk, dj, di) sh[comp][pl[(dk) + 1]][q0 + (dj) * SF_W + (di)]
      const double fc = F(0, 0, 0);
      double t = de * (F(0, 0, 1) - fc);
      t = t + dw * (F(0, 0, -1) - fc);
      t = t + dn * (F(0, 1, 0) - fc);
      t = t + ds * (F(0, -1, 0) - fc);
      t = t + dt * (F(1, 0, 0) - fc);
      t = t + db * (F(-1, 0, 0) - fc);
      double r = e12 * (F(0, 1, 1) - F(0, -1, 1));
      r = r - w12 * (F(0, 1, -1) - F(0, -1, -1));
      r = r + e13 * (F(1, 0, 1) - F(-1, 0, 1));
      r = r - w13 * (F(1, 0, -1) - F(-1, 0, -1));
      r = r + n12 * (F(0, 1, 1) - F(0, 1, -1));
      r = r - s12 * (F(0, -1, 1) - F(0, -1, -1));
      r = r + n23 * (F(1, 1, 0) - F(-1, 1, 0));
      r = r - s23 * (F(1, -1, 0) - F(-1, -1, 0));
      r = r + j13[2] * (F(1, 0, 1) - F(1, 0, -1));
      r = r - j13[0] * (F(-1, 0, 1) - F(-1, 0, -1));
      r = r + j23[2] * (F(1, 1, 0) - F(1, -1, 0));
      r = r - j23[0] * (F(-1, 1, 0) - F(-1, -1, 0));
#undef F
      t = t + 0.25 * r;
      gpw(cs.out[comp])[c] = 1.0 * t;
    }
  }
}

// P_SOURCE(IN): S = -0.5 sum_dir [(J xi_s p)_+ - (J xi_s p)_-] (RLES/p_source.f90:14-42).
// With CORR set, followed on the same cell by U = U + sign DTIME S / J^-1 (RLES/main.f90:108-111;
// 90-92 with sign -1), S kept for the next step's predictor, and, when cflpart is given, the
// local CFL number of the corrected velocity (CALCCOEF, k_cfl's formula) reduced to one maximum
// per workgroup.  A workgroup owns a 64 x 4 (I x J) column of cells over PS_KC
// consecutive planes and marches up K; the products (J zeta_s p) of the planes K+-1 ride along
// in registers, so only the I and J neighbours are fetched per cell.
constexpr int PS_KC = 16;

template <int CORR, bool XZ0>
__global__ void __launch_bounds__(256) k_psource_km(Dims d, const double* const* __restrict__ m,
                                                    const double* __restrict__ djac,
                                                    const double* __restrict__ p, double sign,
                                                    double dtime, double* s0, double* s1,
                                                    double* s2, double* u, double* v, double* w,
                                                    double* cflpart) {
  __shared__ double sh[256];
  const uint3 bb = xcd_block();
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * PS_KC;
  const int k1 = min(k0 + PS_KC - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const long sj = d.sj;
  const GPW out[3] = {gpw(s0), gpw(s1), gpw(s2)};
  const GPW f[3] = {gpw(u), gpw(v), gpw(w)};
  // (J zeta_s p) of the own column, planes kp-2, kp-1, kp, per component s
  double zp[3][3] = {{0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}, {0.0, 0.0, 0.0}};
  double cmax = 0.0;  // this thread's largest local CFL number
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
#pragma unroll
    for (int q = 0; q < 3; ++q) { zp[q][0] = zp[q][1]; zp[q][1] = zp[q][2]; }
    if (act) {
      const long cp = cidx(d, i, j, kp);
      const double jc = djac[cp], pc = p[cp];
#pragma unroll
      for (int q = 0; q < 3; ++q) zp[q][2] = jc * gp(m[3 * q + 2])[cp] * pc;
    }
    if (kp < k0 + 1) continue;  // uniform over the workgroup
    const int k = kp - 1;
    if (act) {
      const long c = cidx(d, i, j, k);
      double nv[3], sv[3], fv[3];
      const double J = djac[c];
      // all loads of the cell precede its stores (which might alias them for the compiler)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double t = djac[c + 1] * mget<XZ0>(m, 3 * q, c + 1) * p[c + 1];
        t = t - djac[c - 1] * mget<XZ0>(m, 3 * q, c - 1) * p[c - 1];
        t = t + djac[c + sj] * mget<XZ0>(m, 3 * q + 1, c + sj) * p[c + sj];
        t = t - djac[c - sj] * mget<XZ0>(m, 3 * q + 1, c - sj) * p[c - sj];
        t = t + zp[q][2];
        t = t - zp[q][0];
        sv[q] = -(0.5 * t);
        fv[q] = CORR ? f[q][c] : 0.0;
      }
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        out[q][c] = sv[q];
        if (CORR) {
          nv[q] = sign < 0 ? fv[q] - dtime * sv[q] / J : fv[q] + dtime * sv[q] / J;
          f[q][c] = nv[q];
        }
      }
      if (CORR && cflpart) {
        const double U = nv[0], V = nv[1], W = nv[2];
        const double ca = fabs(gp(m[0])[c] * U + gp(m[3])[c] * V + mget<XZ0>(m, 6, c) * W);
        const double cb = fabs(gp(m[1])[c] * U + gp(m[4])[c] * V + mget<XZ0>(m, 7, c) * W);
        const double ce = fabs(gp(m[2])[c] * U + gp(m[5])[c] * V + gp(m[8])[c] * W);
        cmax = fmax(cmax, dtime * (ca + cb + ce));
      }
    }
  }
  if (!(CORR && cflpart)) return;  // uniform over the workgroup
  sh[tid] = cmax;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (tid < s) sh[tid] = fmax(sh[tid], sh[tid + s]);
    __syncthreads();
  }
  if (tid == 0) cflpart[blockIdx.x + gridDim.x * (blockIdx.y + (long)gridDim.y * blockIdx.z)] = sh[0];
}

// predictor correction U = U - DTIME S / J^-1 fused with CONTRA_VEL on the interior cells
// (RLES/main.f90:90-92, contra_vel.f90:13-23); the I ghost planes follow after BCU
template <bool XZ0>
__global__ void k_correct_contra(Dims d, const double* const* __restrict__ m,
                                 const double* __restrict__ djac, double dtime,
                                 const double* __restrict__ s0, const double* __restrict__ s1,
                                 const double* __restrict__ s2, double* u, double* v, double* w,
                                 double* ju0, double* ju1, double* ju2) {
  INTERIOR_IJK(d);
  const double J = djac[c];
  const double U = u[c] - dtime * s0[c] / J;
  const double V = v[c] - dtime * s1[c] / J;
  const double W = w[c] - dtime * s2[c] / J;
  u[c] = U; v[c] = V; w[c] = W;
  const double u1 = gp(m[0])[c] * U + gp(m[3])[c] * V + mget<XZ0>(m, 6, c) * W;
  const double u2 = gp(m[1])[c] * U + gp(m[4])[c] * V + mget<XZ0>(m, 7, c) * W;
  const double u3 = gp(m[2])[c] * U + gp(m[5])[c] * V + gp(m[8])[c] * W;
  ju0[c] = J * u1;
  ju1[c] = J * u2;
  ju2[c] = J * u3;
}

// CONTRA_VEL on the I ghost planes i = 0 and i = NI only
__global__ void k_contra_iplanes(Dims d, const double* const* __restrict__ m,
                                 const double* __restrict__ djac, const double* __restrict__ u,
                                 const double* __restrict__ v, const double* __restrict__ w,
                                 double* ju0, double* ju1, double* ju2) {
  const long nf = (long)(d.nj - 1) * (d.nk - 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < 2 * nf;
       t += (long)gridDim.x * blockDim.x) {
    const int i = t < nf ? 0 : d.ni;
    const long r = t % nf;
    const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    const long c = cidx(d, i, j, k);
    const double U = u[c], V = v[c], W = w[c], J = djac[c];
    const double u1 = gp(m[0])[c] * U + gp(m[3])[c] * V + gp(m[6])[c] * W;
    const double u2 = gp(m[1])[c] * U + gp(m[4])[c] * V + gp(m[7])[c] * W;
    const double u3 = gp(m[2])[c] * U + gp(m[5])[c] * V + gp(m[8])[c] * W;
    ju0[c] = J * u1;
    ju1[c] = J * u2;
    ju2[c] = J * u3;
  }
}

// Far-field boundary value U_F per (j,k) and component: inflow Dirichlet UIN where the ghost
// centre's angle lies in [pi/2, 3pi/2], else the convective outflow condition advanced with
// the step-start velocity B = U_N (RLES/bcu.f90:27-37).  Constant within a step.
// (nc components of b with inflow values uin[]: u, v, w with UIN = 1, 0, 0, or the
// temperature with TIN = 0, bct.f90:31-37)
struct FarSet {
  const double* b[3];
  double uin[3];
};

__global__ void k_ufar(Dims d, const double* const* __restrict__ m, const double* __restrict__ theta,
                       FarSet fs, int nc, double dtime, double pi, double* ufar) {
  const long nf = (long)(d.nj + 1) * (d.nk + 1);
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nc * nf;
       t += (long)gridDim.x * blockDim.x) {
    int comp = (int)(t / nf);
    long r = t - comp * nf;
    int j = (int)(r % (d.nj + 1)), k = (int)(r / (d.nj + 1));
    if (j < 1 || j > d.nj - 1 || k < 1 || k > d.nk - 1) { ufar[t] = 0.0; continue; }
    const GP b = gp(fs.b[comp]);
    const double uin = fs.uin[comp];
    long c = cidx(d, d.ni, j, k), mm = c - 1;
    double th = theta[c], uf;
    if (th >= 0.5 * pi && th <= 1.5 * pi) {
      uf = uin;
    } else {
      const double *xix = m[0], *etx = m[1], *ztx = m[2];
      const long sj = d.sj, sk = d.sk;
      double t1 = 0.5 * (xix[c] + xix[mm]) * (b[c] - b[mm]);
      double t2 = 0.5 * (etx[c] + etx[mm]) * 0.25 * (b[c + sj] + b[mm + sj] - b[c - sj] - b[mm - sj]);
      double t3 = 0.5 * (ztx[c] + ztx[mm]) * 0.25 * (b[c + sk] + b[mm + sk] - b[c - sk] - b[mm - sk]);
      uf = 0.5 * (b[c] + b[mm]) - dtime * 1.0 * (t1 + t2 + t3);
    }
    ufar[t] = uf;
  }
}

// dst[q] = src[q] for whole cell arrays, ghosts included (U_N = U, ..., T_N = T), one launch;
// two doubles per lane per step of the grid-stride loop
__global__ void k_copy_multi(long n, PtrSet8 dst, PtrSet8 src) {
  const double2* __restrict__ s = reinterpret_cast<const double2*>(src.p[blockIdx.y]);
  double2* __restrict__ t = reinterpret_cast<double2*>(dst.p[blockIdx.y]);
  const long n2 = n / 2;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < n2;
       q += (long)gridDim.x * blockDim.x)
    t[q] = s[q];
  if (blockIdx.x == 0 && threadIdx.x == 0 && (n & 1))
    dst.p[blockIdx.y][n - 1] = src.p[blockIdx.y][n - 1];
}

// BCU / BCT ghosts and the periodic MPI_X copy in one pass over the ghost surface.  Every ghost
// is a function of one interior cell: with (js, ks) the periodic partner of (j, k) (the cell
// itself when j, k are interior), the wall ghost is -a(1, js, ks) + 2 wall, the far-field ghost
// -a(NI-1, js, ks) + 2 U_F(js, ks) and any other ghost a(i, js, ks) -- exactly the values the
// reference's BCU (wall and far-field ghosts of the interior rows) followed by MPI_X leaves.
__global__ void k_bc_ghosts(Dims d, PtrSet8 set, const double* __restrict__ ufar, double wall) {
  double* __restrict__ a = set.p[blockIdx.y];
  const double* __restrict__ uf = ufar + blockIdx.y * (long)(d.nj + 1) * (d.nk + 1);
  const long nI = 2L * (d.nj - 1) * (d.nk - 1);          // I ghost planes, interior j, k
  const long nJ = (long)(d.ni + 1) * 2 * (d.nk - 1);      // J ghost rows, interior k
  const long nK = (long)(d.ni + 1) * (d.nj + 1) * 2;      // K ghost planes
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nI + nJ + nK;
       t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    if (t < nI) {
      const long q = t % (nI / 2);
      i = t < nI / 2 ? 0 : d.ni;
      j = 1 + (int)(q % (d.nj - 1));
      k = 1 + (int)(q / (d.nj - 1));
    } else if (t < nI + nJ) {
      const long q = t - nI;
      i = (int)(q % (d.ni + 1));
      const long r = q / (d.ni + 1);
      j = (r % 2) == 0 ? 0 : d.nj;
      k = 1 + (int)(r / 2);
    } else {
      const long q = t - nI - nJ;
      i = (int)(q % (d.ni + 1));
      const long r = q / (d.ni + 1);
      j = (int)(r % (d.nj + 1));
      k = (r / (d.nj + 1)) == 0 ? 0 : d.nk;
    }
    const int js = j == 0 ? d.nj - 1 : (j == d.nj ? 1 : j);
    const int ks = k == 0 ? d.nk - 1 : (k == d.nk ? 1 : k);
    double v;
    if (i == 0)
      v = -a[cidx(d, 1, js, ks)] + 2.0 * wall;
    else if (i == d.ni)
      v = -a[cidx(d, d.ni - 1, js, ks)] + 2.0 * uf[js + (long)(d.nj + 1) * ks];
    else
      v = a[cidx(d, i, js, ks)];
    a[cidx(d, i, j, k)] = v;
  }
}

// U = U + sign DTIME S / J^-1 (RLES/main.f90:90-92 with sign -1, 109-111 with +1)
__global__ void k_correct(Dims d, double sign, double dtime, const double* __restrict__ djac,
                          double* u, double* v, double* w, const double* __restrict__ s0,
                          const double* __restrict__ s1, const double* __restrict__ s2) {
  INTERIOR_IJK(d);
  double J = djac[c];
  if (sign < 0) {
    u[c] = u[c] - dtime * s0[c] / J;
    v[c] = v[c] - dtime * s1[c] / J;
    w[c] = w[c] - dtime * s2[c] / J;
  } else {
    u[c] = u[c] + dtime * s0[c] / J;
    v[c] = v[c] + dtime * s1[c] / J;
    w[c] = w[c] + dtime * s2[c] / J;
  }
}

// one component of the predictor correction / projection: f = f + sign DTIME s / J
// (MAIN's array statements, RLES/main.f90:90-92, 108-111)
__global__ void k_correct1(Dims d, double sign, double dtime, const double* __restrict__ djac,
                           double* f, const double* __restrict__ s) {
  INTERIOR_IJK(d);
  const double J = djac[c];
  if (sign < 0)
    f[c] = f[c] - dtime * s[c] / J;
  else
    f[c] = f[c] + dtime * s[c] / J;
}

// NL = NL + SGS J (RLES/main.f90:61-66) for one component
__global__ void k_add_sgs(Dims d, const double* __restrict__ djac, double* nl,
                          const double* __restrict__ sgs) {
  INTERIOR_IJK(d);
  nl[c] = nl[c] + sgs[c] * djac[c];
}

// span mean of p at (IE, NJ/4, K), summed in the reference's order by one lane
// (RLES/main.f90:125-131)
__global__ void k_pref(Dims d, const double* __restrict__ p, double* out) {
  // the span values are fetched by the whole workgroup at once, then summed in K order by one
  // lane (the reference's order)
  __shared__ double sh[256];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int k0 = 1; k0 <= d.nk - 1; k0 += 256) {
    const int k = k0 + t;
    if (k <= d.nk - 1) sh[t] = p[cidx(d, d.ni, d.nj / 4, k)];
    __syncthreads();
    if (t == 0) {
      const int cnt = min(256, d.nk - k0);
      for (int q = 0; q < cnt; ++q) s = s + sh[q];
    }
    __syncthreads();
  }
  if (t == 0) out[0] = s / (double)(d.nk - 1);
}

__global__ void k_sub_pref(Dims d, double* p, const double* __restrict__ pref) {
  const double r = pref[0] / 1.0;
  for (long q = blockIdx.x * (long)blockDim.x + threadIdx.x; q < d.n;
       q += (long)gridDim.x * blockDim.x)
    p[q] = p[q] - r;
}

// ---------------------------------------------------------------------------------------
// CALCCOEF (RLES/calccoef.f90:11-201)
// ---------------------------------------------------------------------------------------
__device__ inline double block_reduce_sum(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = sh[t] + sh[t + s];
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

__device__ inline double block_reduce_max(double v, double* sh) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if (t < s) sh[t] = fmax(sh[t], sh[t + s]);
    __syncthreads();
  }
  double r = sh[0];
  __syncthreads();
  return r;
}

// local CFL number max over the interior, per block partial maxima
__global__ void k_cfl(Dims d, const double* const* __restrict__ m, const double* __restrict__ u,
                      const double* __restrict__ v, const double* __restrict__ w, double dtime,
                      double* part) {
  __shared__ double sh[256];
  const long nint = (long)(d.ni - 1) * (d.nj - 1) * (d.nk - 1);
  double mx = 0.0;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nint;
       t += (long)gridDim.x * blockDim.x) {
    int i = 1 + (int)(t % (d.ni - 1));
    long r = t / (d.ni - 1);
    int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    long c = cidx(d, i, j, k);
    double U = u[c], V = v[c], W = w[c];
    double a = fabs(gp(m[0])[c] * U + gp(m[3])[c] * V + gp(m[6])[c] * W);
    double b = fabs(gp(m[1])[c] * U + gp(m[4])[c] * V + gp(m[7])[c] * W);
    double e = fabs(gp(m[2])[c] * U + gp(m[5])[c] * V + gp(m[8])[c] * W);
    double l = dtime * (a + b + e);
    mx = mx > l ? mx : l;
  }
  double r = block_reduce_max(mx, sh);
  if (threadIdx.x == 0) part[blockIdx.x] = r;
}

// surface-force partial sums over the wall triangles (calccoef.f90:76-183): one element per
// lane, a fixed-order tree per block; part[5 b + q] for pressure drag / lift, viscous drag /
// lift and the wall heat flux sum (T - T_WALL)/D_N AREA of the Nusselt number (calccoef.f90:117-119,
// 155-157; zero without the temperature, T = nullptr)
__global__ void k_forces_part(Dims d, const double* __restrict__ wall, const double* __restrict__ p,
                              const double* __restrict__ u, const double* __restrict__ v,
                              const double* __restrict__ w, const double* __restrict__ T,
                              double re, double* part) {
  __shared__ double sh[256];
  const long nw = (long)(d.nj - 1) * (d.nk - 1);
  double pd = 0.0, ply = 0.0, td = 0.0, tly = 0.0, snu = 0.0;
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t < 2 * nw) {
    int tri = (int)(t / nw);
    long r = t - tri * nw;
    int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
    long c = cidx(d, 1, j, k);
    const double* o = wall + 6 * t;
    double A = o[0], B = o[1], C = o[2], dum = o[3], dn = o[4], ar = o[5];
    double U = u[c], V = v[c], W = w[c];
    double t1 = V * C / dum - W * B / dum, t2 = U * C / dum - W * A / dum,
           t3 = U * B / dum - V * A / dum;
    double dudn = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    dudn = dudn / dn;
    double um = sqrt(U * U + V * V + W * W);
    pd = -2.0 * p[c] * ar * A / dum;
    ply = -2.0 * p[c] * ar * B / dum;
    td = (2.0 / re) * dudn * ar * U / um;
    tly = (2.0 / re) * dudn * ar * V / um;
    if (T) {
      const double dtdn = (T[c] - 1.0) / dn;  // T_WALL = 1
      snu = dtdn * ar;
    }
  }
  pd = block_reduce_sum(pd, sh);
  ply = block_reduce_sum(ply, sh);
  td = block_reduce_sum(td, sh);
  tly = block_reduce_sum(tly, sh);
  snu = block_reduce_sum(snu, sh);
  if (threadIdx.x == 0) {
    part[5 * blockIdx.x] = pd; part[5 * blockIdx.x + 1] = ply;
    part[5 * blockIdx.x + 2] = td; part[5 * blockIdx.x + 3] = tly;
    part[5 * blockIdx.x + 4] = snu;
  }
}

// force coefficients, base pressure, CFL maximum and Nusselt number NU = -sum / SUM_AREA
// (calccoef.f90:11-184), one block; out[0..7] as fvm_step_info, out[8] = NU
__global__ void k_forces_final(Dims d, const double* __restrict__ p,
                               const double* __restrict__ fpart, int nfb, double ref_a,
                               double sum_area, const double* __restrict__ cflpart, int ncfl,
                               double* out, Mailbox mb) {
  __shared__ double sh[256];
  double acc[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
  for (int b = threadIdx.x; b < nfb; b += blockDim.x)
    for (int q = 0; q < 5; ++q) acc[q] = acc[q] + fpart[5 * b + q];
  const double pd = block_reduce_sum(acc[0], sh);
  const double ply = block_reduce_sum(acc[1], sh);
  const double td = block_reduce_sum(acc[2], sh);
  const double tly = block_reduce_sum(acc[3], sh);
  const double snu = block_reduce_sum(acc[4], sh);
  double cm[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};  // eight loads in flight per lane
  int q = threadIdx.x;
  for (; q + 7 * (int)blockDim.x < ncfl; q += 8 * blockDim.x) {
#pragma unroll
    for (int u = 0; u < 8; ++u) cm[u] = fmax(cm[u], cflpart[q + u * blockDim.x]);
  }
  for (; q < ncfl; q += blockDim.x) cm[0] = fmax(cm[0], cflpart[q]);
  double cmax = cm[0];
#pragma unroll
  for (int u = 1; u < 8; ++u) cmax = fmax(cmax, cm[u]);
  cmax = block_reduce_max(cmax, sh);
  // base pressure: 8-cell average at the first wall node line, span mean (calccoef.f90:39-74);
  // the node averages are formed by one lane per K, the span sum in K order by lane 0
  double dumb = 0.0;
  const long sj = d.sj, sk = d.sk;
  for (int k0 = 1; k0 <= d.nk - 1; k0 += blockDim.x) {
    const int k = k0 + (int)threadIdx.x;
    if (k <= d.nk - 1) {
      const long c = cidx(d, 1, 1, k);
      sh[threadIdx.x] = 0.125 * (p[c] + p[c - sj] + p[c - 1] + p[c - 1 - sj] + p[c + sk] +
                                 p[c - sj + sk] + p[c - 1 + sk] + p[c - 1 - sj + sk]);
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      const int cnt = min((int)blockDim.x, d.nk - k0);
      for (int q = 0; q < cnt; ++q) dumb = dumb + sh[q];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    dumb = dumb / (double)(d.nk - 1);
    double cdp = pd / ref_a, clp = ply / ref_a, cdv = td / ref_a, clv = tly / ref_a;
    out[0] = cmax;
    out[1] = 2.0 * (0.0 - dumb);
    out[2] = cdp + cdv;
    out[3] = clp + clv;
    out[4] = cdp;
    out[5] = cdv;
    out[6] = clp;
    out[7] = clv;
    out[8] = sum_area > 0.0 ? -snu / sum_area : 0.0;
    if (mb.dst) {  // scal[0..16]: out is scal + 8; the rest was written by earlier kernels
      const double* sc = out - 8;
      double v[17];
      for (int q = 0; q < 17; ++q) v[q] = sc[q];
      publish(mb, v, 17);
    }
  }
}

// CALCCOEF of the 2D solver (cylinder/cylinder_v3.0.3_AF_original 4step_face_nl/
// calccoef.f90:11-144) on the slab's plane K = 1, one lane, the reference's summation order.
// wall2d per J: segment length AREA, centre-to-wall distance DN1, sin/cos TH2 (direction of the
// wall normal through the centre), cos/sin THETA of the wall cell.
// One workgroup of 256: the CFL maximum over the partial maxima (order-free) and the wall terms
// of every wall cell formed in parallel, then added by one lane in the reference's J order
// (calccoef.f90:46-79), 256 wall cells at a time through LDS, so the sums round as the
// reference's loop does.
__global__ void __launch_bounds__(256) k_calccoef2d(Dims d, const double* __restrict__ wall2d,
                                                    const double* __restrict__ p,
                                                    const double* __restrict__ u,
                                                    const double* __restrict__ v, double re,
                                                    double vel_rot,
                                                    const double* __restrict__ cflpart, int ncfl,
                                                    double* scal, Mailbox mb) {
  double* out = scal + 8;
  __shared__ double tm[4][256];
  __shared__ double cm[256];
  const int t = threadIdx.x;
  const int k = 1;
  double cmax = 0.0;
  for (int q = t; q < ncfl; q += 256) cmax = fmax(cmax, cflpart[q]);
  cm[t] = cmax;
  // lanes 0..3 each carry one of the four sums (pressure drag, pressure lift, viscous drag,
  // viscous lift) through every wall cell in J order
  double acc = 0.0;
  for (int j0 = 1; j0 <= d.nj - 1; j0 += 256) {
    const int j = j0 + t;
    if (j <= d.nj - 1) {
      const double* o = wall2d + 6 * (j - 1);
      const double area = o[0], dn1 = o[1], s2 = o[2], c2 = o[3], ct = o[4], st = o[5];
      const long c = cidx(d, 1, j, k);
      const double dudn = (u[c] * s2 - v[c] * c2 - vel_rot) / dn1;
      tm[0][t] = 2.0 * p[c] * area * ct;
      tm[1][t] = 2.0 * p[c] * area * st;
      tm[2][t] = (2.0 / re) * dudn * area * st;
      tm[3][t] = (2.0 / re) * dudn * area * ct;
    }
    __syncthreads();
    if (t < 4) {
      const int m = min(256, d.nj - j0);
      if (t == 2)
        for (int q = 0; q < m; ++q) acc = acc + tm[2][q];
      else
        for (int q = 0; q < m; ++q) acc = acc - tm[t][q];
    }
    __syncthreads();
  }
  if (t < 4) tm[t][0] = acc;
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) cm[t] = fmax(cm[t], cm[t + s]);
    __syncthreads();
  }
  if (t != 0) return;
  const double pdrag = tm[0][0], plift = tm[1][0], tdrag = tm[2][0], tlift = tm[3][0];
  // base pressure: corner average PG(1,1) (calccoef.f90:30-44).  The reference's far-field
  // value PG(NI,NJ/2) is not assigned by CALCCOEF (its loop covers I = 1..3); it holds what the
  // last field dump left there (writedata.f90:20), zero before the first dump, and zero here.
  const double dum_a = 0.0;
  const double dum_b = 0.25 * (p[cidx(d, 1, 1, k)] + p[cidx(d, 1, 0, k)] + p[cidx(d, 0, 1, k)] +
                               p[cidx(d, 0, 0, k)]);
  const double cpb = 2.0 * (dum_a - dum_b);
  double sv[17];
  for (int q = 0; q < 8; ++q) sv[q] = scal[q];  // the step's earlier scalars (stream order)
  sv[8] = cm[0];
  sv[9] = cpb;
  sv[10] = pdrag + tdrag;
  sv[11] = plift + tlift;
  sv[12] = pdrag;
  sv[13] = tdrag;
  sv[14] = plift;
  sv[15] = tlift;
  sv[16] = scal[16];
  for (int q = 0; q < 8; ++q) out[q] = sv[8 + q];
  publish(mb, sv, 17);  // scal[0..16] to the host mailbox (no-op without one)
}

}  // namespace

// ----------------------------------------------------------------------------------------
// launchers
// ----------------------------------------------------------------------------------------
int fvm_k_contra_vel(FvmCtx* c) {
  const Dims d = c->d;
  double** m = c->dm;
  {
    ProfScope ps(c, "contra_vel");
    dim3 g((d.ni + 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
    if (c->xz_zero)
      hipLaunchKernelGGL(k_contra<true>, g, kBlk, 0, c->stream, d, (const double* const*)m,
                         c->djac, c->u, c->v, c->w, c->ju[0], c->ju[1], c->ju[2]);
    else
      hipLaunchKernelGGL(k_contra<false>, g, kBlk, 0, c->stream, d, (const double* const*)m,
                         c->djac, c->u, c->v, c->w, c->ju[0], c->ju[1], c->ju[2]);
    HIP_CHECK(hipGetLastError());
  }
  return fvm_ghost_fill_n(c, c->ju, 3);
}

int fvm_k_nl(FvmCtx* c) {
  for (int q = 0; q < 3; ++q) std::swap(c->nl[q], c->nl1[q]);
  if (c->thermal) std::swap(c->adv, c->adv1);
  ProfScope ps(c, "nl");
  const CompSet cs = {{c->un, c->vn, c->wn, c->tn}, {c->nl[0], c->nl[1], c->nl[2], c->adv}};
  hipLaunchKernelGGL(k_nl, interior_grid(c->d), kBlk, 0, c->stream, c->d, c->ju[0], c->ju[1],
                     c->ju[2], cs, c->thermal ? 4 : 3);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_dif(FvmCtx* c) {
  ProfScope ps(c, "dif");
  const Dims d = c->d;
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + DF_KC - 1) / DF_KC);
  const CompSet cs = {{c->un, c->vn, c->wn, c->tn}, {c->dif[0], c->dif[1], c->dif[2], c->dift}};
  if (c->thermal)  // DIF(T_N, DIFT), RLES/main.f90:75
    hipLaunchKernelGGL(k_dif_km<4>, g, kBlk, 0, c->stream, c->d, c->pdI, c->pdJ, c->pdK,
                       c->jq12, c->jq13, c->jq23, cs);
  else
    hipLaunchKernelGGL(k_dif_km<3>, g, kBlk, 0, c->stream, c->d, c->pdI, c->pdJ, c->pdK,
                       c->jq12, c->jq13, c->jq23, cs);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// NL and DIF of the step in one pass (k_dif_km with WNL): the AB2 history swap of fvm_k_nl, then
// NL(U_N..) and DIF(U_N..) of every component (and of T_N with the passive temperature)
int fvm_k_nl_dif(FvmCtx* c) {
  for (int q = 0; q < 3; ++q) std::swap(c->nl[q], c->nl1[q]);
  if (c->thermal) std::swap(c->adv, c->adv1);
  ProfScope ps(c, "nl_dif");
  const Dims d = c->d;
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + DF_KC - 1) / DF_KC);
  const CompSet cs = {{c->un, c->vn, c->wn, c->tn}, {c->dif[0], c->dif[1], c->dif[2], c->dift}};
  const CompSet nlo = {{nullptr, nullptr, nullptr, nullptr}, {c->nl[0], c->nl[1], c->nl[2], c->adv}};
  if (c->thermal)
    hipLaunchKernelGGL((k_dif_km<4, true>), g, kBlk, 0, c->stream, c->d, c->pdI, c->pdJ, c->pdK,
                       c->jq12, c->jq13, c->jq23, cs, c->ju[0], c->ju[1], c->ju[2], nlo);
  else
    hipLaunchKernelGGL((k_dif_km<3, true>), g, kBlk, 0, c->stream, c->d, c->pdI, c->pdJ, c->pdK,
                       c->jq12, c->jq13, c->jq23, cs, c->ju[0], c->ju[1], c->ju[2], nlo);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_p_source(FvmCtx* c) {
  double** m = c->dm;
  ProfScope ps(c, "p_source");
  const Dims d = c->d;
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + PS_KC - 1) / PS_KC);
  if (c->xz_zero)
    hipLaunchKernelGGL((k_psource_km<0, true>), g, kBlk, 0, c->stream, c->d,
                       (const double* const*)m, c->djac, c->p, 0.0, 0.0, c->s[0], c->s[1],
                       c->s[2], nullptr, nullptr, nullptr, nullptr);
  else
    hipLaunchKernelGGL((k_psource_km<0, false>), g, kBlk, 0, c->stream, c->d,
                       (const double* const*)m, c->djac, c->p, 0.0, 0.0, c->s[0], c->s[1],
                       c->s[2], nullptr, nullptr, nullptr, nullptr);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_ufar(FvmCtx* c) {
  double** m = c->dm;
  long nf = 3L * (c->d.nj + 1) * (c->d.nk + 1);
  const FarSet fs = {{c->un, c->vn, c->wn}, {1.0, 0.0, 0.0}};
  hipLaunchKernelGGL(k_ufar, grid1(nf, 256), dim3(256), 0, c->stream, c->d,
                     (const double* const*)m, c->theta, fs, 3, c->prm.dtime, c->pi, c->ufar);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// U_N = U, V_N = V, W_N = W (and T_N = T) at the start of the step (RLES/main.f90:41-44)
int fvm_k_step_copy(FvmCtx* c) {
  PtrSet8 dst{}, src{};
  dst.p[0] = c->un; dst.p[1] = c->vn; dst.p[2] = c->wn; dst.p[3] = c->tn;
  src.p[0] = c->u; src.p[1] = c->v; src.p[2] = c->w; src.p[3] = c->t;
  const int cnt = c->thermal ? 4 : 3;
  hipLaunchKernelGGL(k_copy_multi, dim3(2048, cnt), dim3(256), 0, c->stream, c->d.n, dst, src);
  HIP_CHECK(hipGetLastError());
  c->u_is_un = true;
  return FVM_OK;
}

// cells of the ghost surface enumerated by k_bc_ghosts
inline long bc_ghost_count(const Dims& d) {
  return 2L * (d.nj - 1) * (d.nk - 1) + (long)(d.ni + 1) * 2 * (d.nk - 1) +
         (long)(d.ni + 1) * (d.nj + 1) * 2;
}

// far-field values of T for BCT from T_N (bct.f90:28-37)
int fvm_k_tfar(FvmCtx* c) {
  long nf = (long)(c->d.nj + 1) * (c->d.nk + 1);
  const FarSet fs = {{c->tn, nullptr, nullptr}, {0.0, 0.0, 0.0}};
  hipLaunchKernelGGL(k_ufar, grid1(nf, 256), dim3(256), 0, c->stream, c->d,
                     (const double* const*)c->dm, c->theta, fs, 1, c->prm.dtime, c->pi, c->tfar);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// BCT ghosts of one temperature array (wall T_WALL = 1, far field from tfar) and MPI_X
int fvm_k_bct_ghosts(FvmCtx* c, double* f) {
  ProfScope ps(c, "bct");
  PtrSet8 set{};
  set.p[0] = f;
  hipLaunchKernelGGL(k_bc_ghosts, dim3(grid1(bc_ghost_count(c->d), 256).x, 1), dim3(256), 0,
                     c->stream, c->d, set, c->tfar, 1.0);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// f[q] takes the far-field values of component q (cnt <= 3)
int fvm_k_bcu_ghosts(FvmCtx* c, double** f, int cnt) {
  if (cnt < 1 || cnt > 3) return FVM_ERR_ARG;
  ProfScope ps(c, "bcu");
  PtrSet8 set{};
  for (int q = 0; q < cnt; ++q) set.p[q] = f[q];
  hipLaunchKernelGGL(k_bc_ghosts, dim3(grid1(bc_ghost_count(c->d), 256).x, cnt), dim3(256), 0,
                     c->stream, c->d, set, c->ufar, 0.0);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// ----------------------------------------------------------------------------------------
// one component at a time: the reference's subroutines with their own arguments
// ----------------------------------------------------------------------------------------
int fvm_k_nl_one(FvmCtx* c, const double* in, double* out) {
  ProfScope ps(c, "nl");
  CompSet cs{};
  cs.in[0] = in;
  cs.out[0] = out;
  hipLaunchKernelGGL(k_nl, interior_grid(c->d), kBlk, 0, c->stream, c->d, c->ju[0], c->ju[1],
                     c->ju[2], cs, 1);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_dif_one(FvmCtx* c, const double* in, double* out) {
  ProfScope ps(c, "dif");
  const Dims d = c->d;
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + DF_KC - 1) / DF_KC);
  CompSet cs{};
  cs.in[0] = in;
  cs.out[0] = out;
  hipLaunchKernelGGL(k_dif_km<1>, g, kBlk, 0, c->stream, c->d, c->pdI, c->pdJ, c->pdK, c->jq12,
                     c->jq13, c->jq23, cs);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_correct_one(FvmCtx* c, double* f, const double* s, double sign) {
  ProfScope ps(c, "correct");
  hipLaunchKernelGGL(k_correct1, interior_grid(c->d), kBlk, 0, c->stream, c->d, sign,
                     c->prm.dtime, c->djac, f, s);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_add_sgs(FvmCtx* c, double* nl, const double* sgs) {
  hipLaunchKernelGGL(k_add_sgs, interior_grid(c->d), kBlk, 0, c->stream, c->d, c->djac, nl, sgs);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// BCU ghosts of one array with the far-field values of component comp (0 u, 1 v, 2 w)
int fvm_k_bcu_ghosts_one(FvmCtx* c, double* f, int comp) {
  if (comp < 0 || comp > 2) return FVM_ERR_ARG;
  ProfScope ps(c, "bcu");
  PtrSet8 set{};
  set.p[0] = f;
  const long nf = (long)(c->d.nj + 1) * (c->d.nk + 1);
  hipLaunchKernelGGL(k_bc_ghosts, dim3(grid1(bc_ghost_count(c->d), 256).x, 1), dim3(256), 0,
                     c->stream, c->d, set, c->ufar + comp * nf, 0.0);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_correct(FvmCtx* c, double sign) {
  ProfScope ps(c, "correct");
  hipLaunchKernelGGL(k_correct, interior_grid(c->d), kBlk, 0, c->stream, c->d, sign,
                     c->prm.dtime, c->djac, c->u, c->v, c->w, c->s[0], c->s[1], c->s[2]);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

int fvm_k_pref(FvmCtx* c, double* pref_host) {
  ProfScope ps(c, "pref");
  hipLaunchKernelGGL(k_pref, dim3(1), dim3(256), 0, c->stream, c->d, c->p, c->scal + 2);
  hipLaunchKernelGGL(k_sub_pref, grid1(c->d.n, 256), dim3(256), 0, c->stream, c->d, c->p,
                     c->scal + 2);
  HIP_CHECK(hipGetLastError());
  if (pref_host) {
    HIP_CHECK(hipMemcpyAsync(pref_host, c->scal + 2, sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return FVM_OK;
}

// cfl_ready: the CFL partial maxima were left in c->cflpart by the projection of this step
long fvm_calccoef_scratch(const FvmCtx* c, bool cfl_ready) {
  const long ncfl = cfl_ready ? 0 : 512;
  if (c->two_d) return ncfl;
  const long nw2 = 2L * (c->d.nj - 1) * (c->d.nk - 1);
  return ncfl + 5L * ((nw2 + 255) / 256);
}

int fvm_k_calccoef(FvmCtx* c, double* out8, bool cfl_ready, const Mailbox* mb, double* part,
                   long cap) {
  double** m = c->dm;
  ProfScope ps(c, "calccoef");
  if (!part) { part = c->red; cap = c->nred; }
  if (fvm_calccoef_scratch(c, cfl_ready) > cap) return FVM_ERR_SIZE;
  const int ncfl = cfl_ready ? c->ncflused : 512;
  const double* cfl = cfl_ready ? c->cflpart : part;
  const int off = cfl_ready ? 0 : ncfl;
  if (!cfl_ready)
    hipLaunchKernelGGL(k_cfl, dim3(ncfl), dim3(256), 0, c->stream, c->d, (const double* const*)m,
                       c->u, c->v, c->w, c->prm.dtime, part);
  const long nw2 = 2L * (c->d.nj - 1) * (c->d.nk - 1);
  const int nfb = (int)((nw2 + 255) / 256);
  hipLaunchKernelGGL(k_forces_part, dim3(nfb), dim3(256), 0, c->stream, c->d, c->wall, c->p,
                     c->u, c->v, c->w, c->thermal ? c->t : nullptr, c->prm.re, part + off);
  hipLaunchKernelGGL(k_forces_final, dim3(1), dim3(256), 0, c->stream, c->d, c->p, part + off,
                     nfb, c->prm.ref_a, c->thermal ? c->sum_area : 0.0, cfl, ncfl, c->scal + 8,
                     mb ? *mb : Mailbox{nullptr, nullptr, 0});
  HIP_CHECK(hipGetLastError());
  if (out8) {
    HIP_CHECK(hipMemcpyAsync(out8, c->scal + 8, 8 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return FVM_OK;
}

int fvm_k_calccoef2d(FvmCtx* c, double* out8, bool cfl_ready, const Mailbox* mb, double* part,
                     long cap) {
  double** m = c->dm;
  ProfScope ps(c, "calccoef");
  if (!part) { part = c->red; cap = c->nred; }
  if (fvm_calccoef_scratch(c, cfl_ready) > cap) return FVM_ERR_SIZE;
  const int ncfl = cfl_ready ? c->ncflused : 512;
  const double* cfl = cfl_ready ? c->cflpart : part;
  if (!cfl_ready)
    hipLaunchKernelGGL(k_cfl, dim3(ncfl), dim3(256), 0, c->stream, c->d, (const double* const*)m,
                       c->u, c->v, c->w, c->prm.dtime, part);
  const Mailbox none{nullptr, nullptr, 0};
  hipLaunchKernelGGL(k_calccoef2d, dim3(1), dim3(256), 0, c->stream, c->d, c->wall2d, c->p, c->u,
                     c->v, c->prm.re, 0.0, cfl, ncfl, c->scal, mb ? *mb : none);
  HIP_CHECK(hipGetLastError());
  if (out8) {
    HIP_CHECK(hipMemcpyAsync(out8, c->scal + 8, 8 * sizeof(double), hipMemcpyDeviceToHost,
                             c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
  }
  return FVM_OK;
}

// Wall-cell geometry of the 2D force integrals (calccoef.f90:46-79, init.f90:122-129), from the
// wall nodes (xg, yg: NI*NJ, I fastest) and the device cell centres and angles, in the
// reference's operations.
int fvm_wall2d_setup(FvmCtx* c, const double* xg, const double* yg) {
  const Dims d = c->d;
  const int ni = d.ni, nj = d.nj;
  std::vector<double> x(d.n), y(d.n), th(d.n);
  HIP_CHECK(hipMemcpy(x.data(), c->x, d.n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(y.data(), c->y, d.n * sizeof(double), hipMemcpyDeviceToHost));
  HIP_CHECK(hipMemcpy(th.data(), c->theta, d.n * sizeof(double), hipMemcpyDeviceToHost));
  const double pi = c->pi;
  std::vector<double> w(6L * (nj - 1));
  for (int j = 1; j <= nj - 1; ++j) {
    const long cc = cidx(d, 1, j, 1);
    const double X = x[cc], Y = y[cc];
    const double xa = xg[(long)(j - 1) * ni], ya = yg[(long)(j - 1) * ni];  // node (1, j)
    const double xb = xg[(long)j * ni], yb = yg[(long)j * ni];              // node (1, j+1)
    const double l1 = pow(pow(X - xa, 2.0) + pow(Y - ya, 2.0), 0.5);
    const double l2 = pow(pow(X - xb, 2.0) + pow(Y - yb, 2.0), 0.5);
    const double l3 = pow(pow(xb - xa, 2.0) + pow(yb - ya, 2.0), 0.5);
    const double th1 = acos((l1 * l1 + l2 * l2 - l3 * l3) / (2.0 * l1 * l2));
    const double dn1 = l1 * l2 / l3 * sin(th1);
    const double la = (yb - ya) / (xb - xa);
    const double lb = ya - la * xa;
    const double la_ = -1.0 / la;
    const double lb_ = Y - la_ * X;
    const double x1 = (lb_ - lb) / (la - la_);
    const double y1 = la * (lb_ - lb) / (la - la_) + lb;
    const double dn2 = pow(pow(Y - y1, 2.0) + pow(X - x1, 2.0), 0.5);
    double th2 = atan((Y - y1) / (X - x1));
    const double as = asin((Y - y1) / dn2);
    if (th2 >= 0.0 && as > 0.0) {
      // first quadrant
    } else if (th2 < 0.0 && as > 0.0) th2 = pi + th2;
    else if (th2 >= 0.0 && as <= 0.0) th2 = pi + th2;
    else if (th2 < 0.0 && as < 0.0) th2 = 2.0 * pi + th2;
    double* o = w.data() + 6 * (j - 1);
    o[0] = l3;            // AREA(1, J) = |XG(1,J+1) - XG(1,J)|
    o[1] = dn1;
    o[2] = sin(th2);
    o[3] = cos(th2);
    o[4] = cos(th[cc]);
    o[5] = sin(th[cc]);
  }
  HIP_CHECK(hipMalloc((void**)&c->wall2d, w.size() * sizeof(double)));
  c->allocs.push_back(c->wall2d);
  HIP_CHECK(hipMemcpy(c->wall2d, w.data(), w.size() * sizeof(double), hipMemcpyHostToDevice));
  return FVM_OK;
}

// with_cfl: also leave the per-workgroup CFL maxima of the projected velocity in c->cflpart for the
// CALCCOEF that ends the step (the pressure reference and BCU that come between change neither
// the interior velocity nor the metrics)
int fvm_k_psource_correct(FvmCtx* c, double sign, bool with_cfl) {
  ProfScope ps(c, "psource_correct");
  const Dims d = c->d;
  const dim3 gt = interior_grid(d);
  const dim3 g(gt.x, gt.y, (d.nk - 1 + PS_KC - 1) / PS_KC);  // CFL maxima: one per workgroup
  if (with_cfl && (long)g.x * g.y * g.z > c->ncflpart) return FVM_ERR_SIZE;
  if (with_cfl) c->ncflused = (int)(g.x * g.y * g.z);
  if (c->xz_zero)
    hipLaunchKernelGGL((k_psource_km<1, true>), g, kBlk, 0, c->stream, c->d,
                       (const double* const*)c->dm, c->djac, c->p, sign, c->prm.dtime, c->s[0],
                       c->s[1], c->s[2], c->u, c->v, c->w, with_cfl ? c->cflpart : nullptr);
  else
    hipLaunchKernelGGL((k_psource_km<1, false>), g, kBlk, 0, c->stream, c->d,
                       (const double* const*)c->dm, c->djac, c->p, sign, c->prm.dtime, c->s[0],
                       c->s[1], c->s[2], c->u, c->v, c->w, with_cfl ? c->cflpart : nullptr);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}

// U = U - DTIME S / J^-1, BCU on u, v, w, CONTRA_VEL (RLES/main.f90:90-101)
int fvm_k_correct_bcu_contra(FvmCtx* c) {
  const Dims d = c->d;
  {
    ProfScope ps(c, "correct_contra");
    if (c->xz_zero)
      hipLaunchKernelGGL(k_correct_contra<true>, interior_grid(d), kBlk, 0, c->stream, d,
                         (const double* const*)c->dm, c->djac, c->prm.dtime, c->s[0], c->s[1],
                         c->s[2], c->u, c->v, c->w, c->ju[0], c->ju[1], c->ju[2]);
    else
      hipLaunchKernelGGL(k_correct_contra<false>, interior_grid(d), kBlk, 0, c->stream, d,
                         (const double* const*)c->dm, c->djac, c->prm.dtime, c->s[0], c->s[1],
                         c->s[2], c->u, c->v, c->w, c->ju[0], c->ju[1], c->ju[2]);
    HIP_CHECK(hipGetLastError());
  }
  double* f[3] = {c->u, c->v, c->w};
  int r = fvm_k_bcu_ghosts(c, f, 3);
  if (r) return r;
  const long nf = (long)(d.nj - 1) * (d.nk - 1);
  hipLaunchKernelGGL(k_contra_iplanes, grid1(2 * nf, 256), dim3(256), 0, c->stream, d,
                     (const double* const*)c->dm, c->djac, c->u, c->v, c->w, c->ju[0], c->ju[1],
                     c->ju[2]);
  HIP_CHECK(hipGetLastError());
  return fvm_ghost_fill_n(c, c->ju, 3);
}
