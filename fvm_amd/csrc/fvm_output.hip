// Field output of WRITEDATA on the device (LES/v2.3_LES-Reynolds stresses/writedata.f90:8-189,
// calc_vor.f90:1-141, find_root.f90:1-27): the vorticity magnitude of CALC_VOR, the lambda_2
// vortex criterion (the middle eigenvalue of S^2 + Omega^2 from the trigonometric roots of its
// characteristic cubic) and the 8-cell nodal averages of p, u, v, w, T, |omega| and lambda_2
// that the reference writes to its Tecplot files, and the wall y+ of CALCCOEF that it writes to
// its YPLUS contour.  Every expression follows the reference's association order; the
// transcendental functions are the device's.
#include "fvm_ctx.h"

namespace {

// FIND_ROOT(A,B,C,D,ROOT) (find_root.f90:10-25)
__device__ inline void find_root(double A, double B, double C, double D, double root[3]) {
  const double q = (3.0 * A * C - B * B) / (9.0 * (A * A));
  const double r = (9.0 * A * B * C - 27.0 * (A * A) * D - 2.0 * (B * B * B)) / (54.0 * (A * A * A));
  const double rho = sqrt(-(q * q * q));
  const double phi = acos(r / rho);
  const double b3 = -(B / (3.0 * A));
  const double cr = pow(rho, 1.0 / 3.0);
  const double c = cos(phi / 3.0), sn = sin(phi / 3.0), s3 = sqrt(3.0);
  root[0] = b3 + 2.0 * cr * c;
  root[1] = b3 - cr * c + s3 * cr * sn;
  root[2] = b3 - cr * c - s3 * cr * sn;
}

// |omega| (calc_vor.f90:8-102) and lambda_2 (writedata.f90:36-104) on the interior cells.
// lambda_2 keeps its previous value where no root lies strictly between the other two, as the
// reference's selection loop does (writedata.f90:101-103).
__global__ void k_vor_lambda2(Dims d, const double* const* __restrict__ m,
                              const double* __restrict__ U, const double* __restrict__ V,
                              const double* __restrict__ W, double* vorm, double* lam2) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  const long sj = d.sj, sk = d.sk;
  // metrics grouped per Cartesian component s: m[3 s + 0..2] = (xi_s, eta_s, zeta_s)
#define DI(F) (F[c + 1] - F[c - 1])
#define DJ(F) (F[c + sj] - F[c - sj])
#define DK(F) (F[c + sk] - F[c - sk])
  const double vz = m[0][c] * 0.5 * DI(V) + m[1][c] * 0.5 * DJ(V) + m[2][c] * 0.5 * DK(V) -
                    m[3][c] * 0.5 * DI(U) - m[4][c] * 0.5 * DJ(U) - m[5][c] * 0.5 * DK(U);
  const double vy = m[6][c] * 0.5 * DI(U) + m[7][c] * 0.5 * DJ(U) + m[8][c] * 0.5 * DK(U) -
                    m[0][c] * 0.5 * DI(W) - m[1][c] * 0.5 * DJ(W) - m[2][c] * 0.5 * DK(W);
  const double vx = m[3][c] * 0.5 * DI(W) + m[4][c] * 0.5 * DJ(W) + m[5][c] * 0.5 * DK(W) -
                    m[6][c] * 0.5 * DI(V) - m[7][c] * 0.5 * DJ(V) - m[8][c] * 0.5 * DK(V);
  vorm[c] = sqrt(vx * vx + vy * vy + vz * vz);
  // g[a][b] = d u_a / d x_b = 0.5 (xi_b du_a/dxi + eta_b du_a/deta + zeta_b du_a/dzeta)
  const double* F3[3] = {U, V, W};
  double g[3][3];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      const double* F = F3[a];
      g[a][b] = 0.5 * (m[3 * b][c] * DI(F) + m[3 * b + 1][c] * DJ(F) + m[3 * b + 2][c] * DK(F));
    }
#undef DI
#undef DJ
#undef DK
  double S1[3][3], O1[3][3], A2[3][3];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      S1[p][q] = 0.5 * (g[p][q] + g[q][p]);
      O1[p][q] = 0.5 * (g[p][q] - g[q][p]);
    }
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) {
      const double s2 = S1[p][0] * S1[0][q] + S1[p][1] * S1[1][q] + S1[p][2] * S1[2][q];
      const double o2 = O1[p][0] * O1[0][q] + O1[p][1] * O1[1][q] + O1[p][2] * O1[2][q];
      A2[p][q] = s2 + o2;
    }
  const double da = -1.0;
  const double db = A2[0][0] + A2[1][1] + A2[2][2];
  const double dc = -(A2[0][0] * A2[1][1]) - A2[1][1] * A2[2][2] - A2[2][2] * A2[0][0] +
                    A2[0][1] * A2[1][0] + A2[1][2] * A2[2][1] + A2[2][0] * A2[0][2];
  const double dd = A2[0][0] * A2[1][1] * A2[2][2] + A2[0][1] * A2[1][2] * A2[2][0] +
                    A2[0][2] * A2[1][0] * A2[2][1] - A2[0][0] * A2[1][2] * A2[2][1] -
                    A2[1][1] * A2[0][2] * A2[2][0] - A2[2][2] * A2[0][1] * A2[1][0];
  double lam[3];
  find_root(da, db, dc, dd, lam);
  const double l1 = fmax(fmax(lam[0], lam[1]), lam[2]);
  const double l3 = fmin(fmin(lam[0], lam[1]), lam[2]);
  double out = lam2[c];
#pragma unroll
  for (int q = 0; q < 3; ++q)
    if (lam[q] < l1 && lam[q] > l3) out = lam[q];
  lam2[c] = out;
}

// lambda_2 on the wall and far-field ghost planes by three-point extrapolation
// (writedata.f90:106-107) for the interior j, k; the J/K ghost faces of writedata.f90:108-111
// are overwritten by MPI_X (:113), which fvm_ghost_fill applies next.
__global__ void k_lambda2_iplanes(Dims d, double* L) {
  const long nf = (long)(d.nj - 1) * (d.nk - 1);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= nf) return;
  const int j = 1 + (int)(t % (d.nj - 1)), k = 1 + (int)(t / (d.nj - 1));
  const long c0 = cidx(d, 0, j, k), cn = cidx(d, d.ni, j, k);
  L[c0] = 3.0 * L[c0 + 1] - 3.0 * L[c0 + 2] + L[c0 + 3];
  L[cn] = 3.0 * L[cn - 1] - 3.0 * L[cn - 2] + L[cn - 3];
}

// 8-cell averages at the nodes I = 1..NI, J = 1..NJ, K = 1..NK (writedata.f90:115-133),
// stored at the node's cell index; blockIdx.y selects the field
struct NodalSet {
  const double* src[7];
  double* dst[7];
};

__global__ void k_nodal(Dims d, NodalSet ns) {
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
    if (!P) { out[c] = 0.0; continue; }
    out[c] = 0.125 * (P[c] + P[c - sj] + P[c - 1] + P[c - 1 - sj] + P[c - sk] + P[c - sj - sk] +
                      P[c - 1 - sk] + P[c - 1 - sj - sk]);
  }
}

// YPLUS of CALCCOEF on the wall cells (calccoef.f90:90-160): per triangle, the wall shear rate
// |u x n|/d_n of the wall cell's velocity, weighted by the triangle's share of the cell's wall
// area; the first triangle assigns, the second adds.  One thread per wall cell (J, K); the wall
// table holds A, B, C, |n|, d_n, area per triangle, the second triangles after the first.
__global__ void k_yplus(Dims d, const double* __restrict__ wall, const double* __restrict__ u,
                        const double* __restrict__ v, const double* __restrict__ w, double re,
                        double* yplus) {
  const long nw = (long)(d.nj - 1) * (d.nk - 1);
  const long r = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (r >= nw) return;
  const int j = 1 + (int)(r % (d.nj - 1)), k = 1 + (int)(r / (d.nj - 1));
  const long c = cidx(d, 1, j, k);
  const double U = u[c], V = v[c], W = w[c];
  const double a1 = wall[6 * r + 5], a2 = wall[6 * (nw + r) + 5];
  double yp = 0.0;
  for (int tri = 0; tri < 2; ++tri) {
    const double* o = wall + 6 * (tri * nw + r);
    const double A = o[0], B = o[1], C = o[2], dum = o[3], dn = o[4], ar = o[5];
    const double t1 = V * C / dum - W * B / dum, t2 = U * C / dum - W * A / dum,
                 t3 = U * B / dum - V * A / dum;
    double dudn = sqrt(t1 * t1 + t2 * t2 + t3 * t3);
    dudn = dudn / dn;
    const double y = sqrt(re * fabs(dudn)) * dn * ar / (a1 + a2);
    yp = tri == 0 ? y : yp + y;
  }
  yplus[c] = yp;
}

}  // namespace

int fvm_k_writedata_fields(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  if (d.ni < 4 || c->two_d) return FVM_ERR_ARG;  // three-point extrapolation needs I = 1..3
  if (!c->vorm) {
    double** arr[] = {&c->vorm, &c->lam2, &c->nodal[0], &c->nodal[1], &c->nodal[2],
                      &c->nodal[3], &c->nodal[4], &c->nodal[5], &c->nodal[6], &c->yplus};
    for (double** q : arr) {
      HIP_CHECK(hipMalloc((void**)q, d.n * sizeof(double)));
      c->allocs.push_back(*q);
      HIP_CHECK(hipMemsetAsync(*q, 0, d.n * sizeof(double), st));
    }
  }
  ProfScope ps(c, "writedata_fields");
  const dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
  hipLaunchKernelGGL(k_vor_lambda2, g, dim3(64, 4, 1), 0, st, d, (const double* const*)c->dm,
                     c->u, c->v, c->w, c->vorm, c->lam2);
  const long nf = (long)(d.nj - 1) * (d.nk - 1);
  hipLaunchKernelGGL(k_lambda2_iplanes, grid1(nf, 256), dim3(256), 0, st, d, c->lam2);
  HIP_CHECK(hipGetLastError());
  double* gl[2] = {c->vorm, c->lam2};
  int r = fvm_ghost_fill_n(c, gl, 2);
  if (r) return r;
  NodalSet ns = {{c->p, c->u, c->v, c->w, c->thermal ? c->t : nullptr, c->vorm, c->lam2},
                 {c->nodal[0], c->nodal[1], c->nodal[2], c->nodal[3], c->nodal[4],
                  c->nodal[5], c->nodal[6]}};
  const long nn = (long)d.ni * d.nj * d.nk;
  hipLaunchKernelGGL(k_nodal, dim3(grid1(nn, 256).x, 7), dim3(256), 0, st, d, ns);
  hipLaunchKernelGGL(k_yplus, grid1(nf, 256), dim3(256), 0, st, d, c->wall, c->u, c->v, c->w,
                     c->prm.re, c->yplus);
  HIP_CHECK(hipGetLastError());
  return FVM_OK;
}
