// Fast solve of the pressure equation by defect correction with a separable approximate inverse.
//
// The discrete operator of ITER_P (LES/v2.3_LES-Reynolds stresses/iter_p.f90:12-63) is
//   L p = sum_nb D_nb (p_nb - p) + DD(p)        (+ the BCP folds: wall Neumann, outflow AE=0)
// On an O-grid with uniform circumferential spacing extruded along a uniform span, the face
// coefficients D_nb depend on the radial index only and the cross terms DD vanish to round-off,
// so the orthogonal part L_D is diagonalised by real Fourier bases in J (periodic branch cut) and
// K (periodic span): for every mode pair (mJ, mK) what is left is ONE radial tridiagonal line
//   dW x(i-1) - (dE + dW + lamJ dN + lamK dT) x(i) + dE x(i+1) = r_hat(i),
// lam = 4 sin^2(pi m / N).  The transforms are this library's own mixed-radix FFTs in LDS
// (fvm_fft.h): a real-to-half-complex transform along J that packs two radial columns into one
// complex sequence, a complex transform along K; the radial mode lines are solved by 64-lane
// affine-map scans, and the inverse J transform adds the correction to p in the same kernel.  The outer loop keeps the reference's FP64 operator -- cross terms,
// boundary folds, periodic ghosts -- in the residual r = rhs - L p and its scaled-L1 norm
// sum|r| / sum|AP p| (iter_p.f90:45-59), so it converges to the same fixed point as the
// reference's line Gauss-Seidel; the approximate inverse runs in FP64 (plain form) or FP32
// (mixed form).  Coefficients are averaged over J and K; their largest deviation from the
// average is measured at set-up (fvm_ctx.sep_dev).
#include <type_traits>
#include <vector>

#include "fvm_ctx.h"
#include "fvm_fft.h"
#include "fvm_wave.h"

namespace {

const dim3 kBlk(64, 4, 1);

// residual r = sum A_nb p_nb + (rhs - DD(p)) - AP p of every interior cell (iter_p.f90:50-53
// with SU = rhs - DD, DD the twelve cross terms of p weighted by J^-1 Q12, Q13, Q23), written to
// the compact interior array R (I fastest), plus per-workgroup sums of |r| and |AP p|.
// K-marching: a workgroup owns a 64 x 4 (I x J) column over RS_KC planes and
// keeps three consecutive p planes (one-cell halo) in an LDS ring, so every p value is read from
// memory once per column instead of once per plane that needs it; the own-column values of the
// planes below and above (J^-1 Q13, J^-1 Q23 at K +- 1, D_T at K - 1) ride along in registers.
// One pair of partial sums per workgroup.
constexpr int RS_KC = 16;

template <typename T>
__global__ void __launch_bounds__(256) k_residual_km(Dims d, const double* __restrict__ p,
                                                     const double* __restrict__ rhs,
                                                     const double* __restrict__ pdI,
                                                     const double* __restrict__ pdJ,
                                                     const double* __restrict__ pdK,
                                                     const double* __restrict__ jq12,
                                                     const double* __restrict__ jq13,
                                                     const double* __restrict__ jq23,
                                                     T* __restrict__ R, double* part, int kc) {
  const uint3 bb = xcd_block();
  __shared__ double sh[3][SF_P];
  __shared__ double s0[256], s1[256];
  const int tx = threadIdx.x, ty = threadIdx.y, tid = ty * 64 + tx;
  const int i = 1 + bb.x * 64 + tx;
  const int j = 1 + bb.y * 4 + ty;
  const int k0 = 1 + bb.z * kc;
  const int k1 = min(k0 + kc - 1, d.nk - 1);
  const bool act = i <= d.ni - 1 && j <= d.nj - 1;
  const int q0 = (ty + 1) * SF_W + tx + 1;
  const long sj = d.sj;
  // own-column values of planes kp-2, kp-1, kp
  double z13[3] = {0.0, 0.0, 0.0}, z23[3] = {0.0, 0.0, 0.0}, zk[3] = {0.0, 0.0, 0.0};
  double ra = 0.0, rb = 0.0;
  // the next plane's tile values and own-column values, loaded one plane ahead
  constexpr int NQ = (SF_P + 255) / 256;
  double pv[NQ], n13 = 0.0, n23 = 0.0, nk = 0.0;
  auto load = [&](int kp) {
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const int q = min(tid + 256 * h, SF_P - 1);
      const int gi = min(bb.x * 64 + q % SF_W, d.ni);
      const int gj = min(bb.y * 4 + q / SF_W, d.nj);
      pv[h] = p[cidx(d, gi, gj, kp)];
    }
    const long cp = cidx(d, act ? i : 1, act ? j : 1, kp);
    n13 = jq13[cp];
    n23 = jq23[cp];
    nk = pdK[cp];
  };
  load(k0 - 1);
  for (int kp = k0 - 1; kp <= k1 + 1; ++kp) {
    const int slot = (kp - (k0 - 1)) % 3;
    __syncthreads();  // plane kp - 3, held in this slot, is no longer read
#pragma unroll
    for (int h = 0; h < NQ; ++h) {
      const int q = tid + 256 * h;
      if (q < SF_P) sh[slot][q] = pv[h];
    }
    z13[0] = z13[1]; z13[1] = z13[2]; z13[2] = n13;
    z23[0] = z23[1]; z23[1] = z23[2]; z23[2] = n23;
    zk[0] = zk[1]; zk[1] = zk[2]; zk[2] = nk;
    __syncthreads();
    if (kp + 1 <= k1 + 1) load(kp + 1);  // in flight while this plane's cell is formed
    if (kp < k0 + 1 || !act) continue;
    const int k = kp - 1;
    const int pl[3] = {(slot + 1) % 3, (slot + 2) % 3, slot};
#define P(dk, dj, di) sh[pl[(dk) + 1]][q0 + (dj) * SF_W + (di)]
    const long c = cidx(d, i, j, k);
    const double de = pdI[c], dw = pdI[c - 1], dn = pdJ[c], ds = pdJ[c - sj], dt = zk[1],
                 db = zk[0];
    double ap = -de - dw - dn - ds - dt - db;
    double ae = -de, aw = -dw, an = -dn, as = -ds, at = -dt, ab = -db;
    if (i == 1) { ap = ap - aw; aw = 0.0; }
    if (i == d.ni - 1) ae = 0.0;
    // the cross terms DD
    double x = jq12[c + 1] * (P(0, 1, 1) - P(0, -1, 1));
    x = x - jq12[c - 1] * (P(0, 1, -1) - P(0, -1, -1));
    x = x + jq13[c + 1] * (P(1, 0, 1) - P(-1, 0, 1));
    x = x - jq13[c - 1] * (P(1, 0, -1) - P(-1, 0, -1));
    x = x + jq12[c + sj] * (P(0, 1, 1) - P(0, 1, -1));
    x = x - jq12[c - sj] * (P(0, -1, 1) - P(0, -1, -1));
    x = x + jq23[c + sj] * (P(1, 1, 0) - P(-1, 1, 0));
    x = x - jq23[c - sj] * (P(1, -1, 0) - P(-1, -1, 0));
    x = x + z13[2] * (P(1, 0, 1) - P(1, 0, -1));
    x = x - z13[0] * (P(-1, 0, 1) - P(-1, 0, -1));
    x = x + z23[2] * (P(1, 1, 0) - P(1, -1, 0));
    x = x - z23[0] * (P(-1, 1, 0) - P(-1, -1, 0));
    const double su = rhs[c] - 0.25 * x;
    const double pc = P(0, 0, 0);
    const double r = ae * P(0, 0, 1) + aw * P(0, 0, -1) + an * P(0, 1, 0) + as * P(0, -1, 0) +
                     at * P(1, 0, 0) + ab * P(-1, 0, 0) + su - ap * pc;
#undef P
    ra = ra + fabs(r);
    rb = rb + fabs(ap * pc);
    const long ci = (long)(i - 1) + (long)(d.ni - 1) * ((j - 1) + (long)(d.nj - 1) * (k - 1));
    R[ci] = (T)r;
  }
  s0[tid] = ra; s1[tid] = rb;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (tid < st) { s0[tid] += s0[tid + st]; s1[tid] += s1[tid + st]; }
    __syncthreads();
  }
  if (tid == 0) {
    const long bid = bb.x + (long)gridDim.x * (bb.y + (long)gridDim.y * bb.z);
    part[2 * bid] = s0[0];
    part[2 * bid + 1] = s1[0];
  }
}

__global__ void k_res_final(const double* __restrict__ part, long nb, double* out, Mailbox mb) {
  __shared__ double s0[256], s1[256];
  const int t = threadIdx.x;
  double a = 0.0, b = 0.0;
  long q = t;
  // eight partial pairs in flight per lane, added in the same order as one at a time
  for (; q + 7L * blockDim.x < nb; q += 8L * blockDim.x) {
    double x[8], y[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      x[u] = part[2 * (q + (long)u * blockDim.x)];
      y[u] = part[2 * (q + (long)u * blockDim.x) + 1];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) { a += x[u]; b += y[u]; }
  }
  for (; q < nb; q += blockDim.x) { a += part[2 * q]; b += part[2 * q + 1]; }
  s0[t] = a; s1[t] = b;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) { s0[t] += s0[t + s]; s1[t] += s1[t + s]; }
    __syncthreads();
  }
  if (t == 0) {
    const double q = scaled_residual(s0[0], s1[0]);
    out[0] = s0[0]; out[1] = s1[0]; out[2] = q;
    const double v[3] = {s0[0], s1[0], q};
    publish(mb, v, 3);
  }
}

// Radial line operator of Fourier mode (mJ, mK): lo x(r-1) + dg x(r) + up x(r+1) with
// lo = <D_W>, up = <D_E>, dg = -(<D_E> + <D_W>) - lambda_J <D_N> - lambda_K <D_T> (the J, K
// second differences diagonalised, lambda = 4 sin^2(pi m / N)), the wall Neumann fold at r = 0
// and a zero ghost beyond the last row.  Its Thomas factors 1/D(r) and B(r) = up(r)/D(r),
// D(r) = dg(r) - lo(r) B(r-1), are stored once in the layout of the spectrum
// [r + N1 (mJ + MJ mK)] (mJ = 0..NJ1/2, mK = 0..NK1-1); one thread per line (set-up only).
__global__ void k_mode_factor(int n, int mj_n, int nj1, int nk1, const double* __restrict__ dE,
                              const double* __restrict__ dW, const double* __restrict__ dN,
                              const double* __restrict__ dT, double* invd, double* bf) {
  const int mj = blockIdx.x * 64 + threadIdx.x, mk = blockIdx.y;
  if (mj >= mj_n) return;
  const double pi = 3.14159265358979323846;
  const double sj = sin(pi * (double)mj / (double)nj1), sk = sin(pi * (double)mk / (double)nk1);
  const double lj = 4.0 * sj * sj, lk = 4.0 * sk * sk;
  const long base = (long)n * (mj + (long)mj_n * mk);
  double bprev = 0.0;
  for (int r = 0; r < n; ++r) {
    const double lo = r == 0 ? 0.0 : dW[r];
    const double up = r == n - 1 ? 0.0 : dE[r];
    double dg = -(dE[r] + dW[r]) - lj * dN[r] - lk * dT[r];
    if (r == 0) dg = dg + dW[r];  // wall Neumann fold
    const double D = dg - lo * bprev;
    const double iv = 1.0 / D;
    const double B = up * iv;
    invd[base + r] = iv;
    bf[base + r] = B;
    bprev = B;
  }
}

// Radial mode-line solves of the complex spectrum with the stored factors, in FP64 whatever the
// storage type T: one 64-lane wavefront per line, lane l owning rows l*R .. l*R+R-1; forward
// elimination y(r) = z(r)/D(r) - lo(r)/D(r) y(r-1) and back substitution
// x(r) = y(r) - B(r) x(r+1) as affine-map scans across the lanes (DPP lane moves, no LDS); the real
// and imaginary parts share the real line operator.
template <typename T, int R>
__global__ void __launch_bounds__(64) k_mode_scan_c(int n, const double* __restrict__ dW,
                                                    const double* __restrict__ invd,
                                                    const double* __restrict__ bf, Cx<T>* Z) {
  const long base = (long)n * blockIdx.x;
  const int l = threadIdx.x;
  const int r0 = l * R;
  double alx[R], aly[R], be[R], bb[R];
  double Ax = 0.0, Ay = 0.0, Bt = 1.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = r0 + q;
    if (r < n) {
      const double iv = invd[base + r];
      const Cx<T> z = Z[base + r];
      alx[q] = (double)z.x * iv;
      aly[q] = (double)z.y * iv;
      be[q] = r == 0 ? 0.0 : -dW[r] * iv;
      bb[q] = bf[base + r];
      Ax = alx[q] + be[q] * Ax;
      Ay = aly[q] + be[q] * Ay;
      Bt = be[q] * Bt;
    } else {
      alx[q] = 0.0; aly[q] = 0.0; be[q] = 0.0; bb[q] = 0.0;
    }
  }
  double yinx, yiny;
  {
    double A2[2] = {Ax, Ay}, z2[2] = {0.0, 0.0}, o2[2];
    wave::prefix_apply<2>(A2, Bt, z2, o2);  // y(-1) = 0
    yinx = o2[0];
    yiny = o2[1];
  }
  double yx[R], yy[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    yx[q] = alx[q] + be[q] * yinx;
    yy[q] = aly[q] + be[q] * yiny;
    if (r0 + q < n) { yinx = yx[q]; yiny = yy[q]; }
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
  double xinx, xiny;
  {
    double D2[2] = {Dx, Dy}, z2[2] = {0.0, 0.0}, o2[2];
    wave::suffix_apply<2>(D2, Ct, z2, o2);  // x(n) = 0
    xinx = o2[0];
    xiny = o2[1];
  }
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    const int r = r0 + q;
    if (r < n) {
      const double xx = yx[q] - bb[q] * xinx, xy = yy[q] - bb[q] * xiny;
      xinx = xx;
      xiny = xy;
      Z[base + r] = Cx<T>{(T)xx, (T)xy};
    }
  }
}

template <typename T>
void mode_lines(int n1, int nlines, const double* dW, const double* invd, const double* bf,
                Cx<T>* Z, hipStream_t st) {
  const int R = (n1 + 63) / 64;
  if (R <= 3) hipLaunchKernelGGL((k_mode_scan_c<T, 3>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else if (R <= 4) hipLaunchKernelGGL((k_mode_scan_c<T, 4>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else if (R <= 6) hipLaunchKernelGGL((k_mode_scan_c<T, 6>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else hipLaunchKernelGGL((k_mode_scan_c<T, 8>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
}

// Forward real transform along J of the compact residual R[i + N1 (j + NJ1 k)] for a tile of
// 2 ncol radial indices at one k: columns i and i+1 are packed into one complex sequence
// z_j = R(i) + i R(i+1), transformed in LDS, and split into the half spectra
// A_m = (Z_m + conj Z_{N-m}) / 2, B_m = (Z_m - conj Z_{N-m}) / 2i, m = 0..N/2, written to
// S[i + N1 (m + MJ k)].  LDS: two buffers of ncol N complex values, then the twiddles.
template <typename T>
__global__ void __launch_bounds__(256) k_fft_j_fwd(int n1, int nk1, FftPlan pl, int ncol,
                                                   const T* __restrict__ R,
                                                   const Cx<T>* __restrict__ twg, Cx<T>* S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = pl.n, MJ = N / 2 + 1;
  Cx<T>* a = (Cx<T>*)smem;
  Cx<T>* b = a + (size_t)ncol * N;
  Cx<T>* tw = b + (size_t)ncol * N;
  const int i0 = blockIdx.x * 2 * ncol, k = blockIdx.y;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, j = q / ncol;
    const int i = i0 + 2 * col;
    Cx<T> z{(T)0, (T)0};
    if (i < n1) {
      const long o = (long)i + (long)n1 * (j + (long)N * k);
      z.x = R[o];
      z.y = i + 1 < n1 ? R[o + 1] : (T)0;
    }
    a[j * ncol + col] = z;
  }
  __syncthreads();
  const Cx<T>* Z = fft_lds<T, -1>(a, b, ncol, pl, tw);
  for (int q = threadIdx.x; q < ncol * MJ; q += blockDim.x) {
    const int col = q % ncol, m = q / ncol;
    const int i = i0 + 2 * col;
    if (i >= n1) continue;
    const Cx<T> zm = Z[m * ncol + col], zc = Z[((N - m) % N) * ncol + col];
    // A = (zm + conj zc) / 2, B = (zm - conj zc) / (2 i)
    const Cx<T> A{(T)0.5 * (zm.x + zc.x), (T)0.5 * (zm.y - zc.y)};
    const Cx<T> B{(T)0.5 * (zm.y + zc.y), (T)0.5 * (zc.x - zm.x)};
    const long o = (long)i + (long)n1 * (m + (long)MJ * k);
    S[o] = A;
    if (i + 1 < n1) S[o + 1] = B;
  }
}

// Complex transform along K of S[i + N1 (mJ + MJ k)] for ncol consecutive radial indices at one
// mJ, in place.
template <typename T, int SIGN>
__global__ void __launch_bounds__(256) k_fft_k(int n1, int MJ, FftPlan pl, int ncol,
                                               const Cx<T>* __restrict__ twg, Cx<T>* S) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = pl.n;
  Cx<T>* a = (Cx<T>*)smem;
  Cx<T>* b = a + (size_t)ncol * N;
  Cx<T>* tw = b + (size_t)ncol * N;
  const int i0 = blockIdx.x * ncol, mj = blockIdx.y;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, k = q / ncol;
    const int i = i0 + col;
    Cx<T> z{(T)0, (T)0};
    if (i < n1) z = S[(long)i + (long)n1 * (mj + (long)MJ * k)];
    a[k * ncol + col] = z;
  }
  __syncthreads();
  const Cx<T>* Z = fft_lds<T, SIGN>(a, b, ncol, pl, tw);
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, k = q / ncol;
    const int i = i0 + col;
    if (i < n1) S[(long)i + (long)n1 * (mj + (long)MJ * k)] = Z[k * ncol + col];
  }
}

// Inverse of k_fft_j_fwd fused with the correction: the half spectra of columns i and i+1 at one
// k become Z_m = A_m + i B_m (m = 0..N/2, with the imaginary parts of A_0, B_0, A_{N/2},
// B_{N/2} dropped: a real sequence has real coefficients there) and Z_{N-m} = conj A_m +
// i conj B_m; after the inverse transform, p(i+1, j+1, k+1) += scale Re z_j and
// p(i+2, j+1, k+1) += scale Im z_j (scale = 1 / (NJ1 NK1)).
template <typename T>
__global__ void __launch_bounds__(256) k_fft_j_inv_add(Dims d, FftPlan pl, int ncol,
                                                       const Cx<T>* __restrict__ S,
                                                       const Cx<T>* __restrict__ twg,
                                                       double scale, double* p) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int N = pl.n, MJ = N / 2 + 1;
  const int n1 = d.ni - 1;
  Cx<T>* a = (Cx<T>*)smem;
  Cx<T>* b = a + (size_t)ncol * N;
  Cx<T>* tw = b + (size_t)ncol * N;
  const int i0 = blockIdx.x * 2 * ncol, k = blockIdx.y;
  for (int t = threadIdx.x; t < N; t += blockDim.x) tw[t] = twg[t];
  for (int q = threadIdx.x; q < ncol * MJ; q += blockDim.x) {
    const int col = q % ncol, m = q / ncol;
    const int i = i0 + 2 * col;
    Cx<T> A{(T)0, (T)0}, B{(T)0, (T)0};
    if (i < n1) {
      const long o = (long)i + (long)n1 * (m + (long)MJ * k);
      A = S[o];
      if (i + 1 < n1) B = S[o + 1];
    }
    if (m == 0 || 2 * m == N) { A.y = (T)0; B.y = (T)0; }
    a[m * ncol + col] = Cx<T>{A.x - B.y, A.y + B.x};
    if (m > 0 && 2 * m < N) a[(N - m) * ncol + col] = Cx<T>{A.x + B.y, B.x - A.y};
  }
  __syncthreads();
  const Cx<T>* Z = fft_lds<T, +1>(a, b, ncol, pl, tw);
  for (int q = threadIdx.x; q < ncol * N; q += blockDim.x) {
    const int col = q % ncol, j = q / ncol;
    const int i = i0 + 2 * col;
    if (i >= n1) continue;
    const Cx<T> z = Z[j * ncol + col];
    const long c = cidx(d, i + 1, j + 1, k + 1);
    p[c] = p[c] + scale * (double)z.x;
    if (i + 1 < n1) p[c + 1] = p[c + 1] + scale * (double)z.y;
  }
}

// J,K-averaged face coefficients per radial index and their largest relative deviation.
// Block b handles radial index i = b (0..NI-1): avg[i] = <D_E>(i+1/2); for i >= 1 also
// avg[NI + i-1] = <D_N>(i) and avg[2NI-1 + i-1] = <D_T>(i).
__global__ void k_avg_faces(Dims d, const double* __restrict__ pdI, const double* __restrict__ pdJ,
                            const double* __restrict__ pdK, double* avg, double* dev) {
  const int i = blockIdx.x;
  __shared__ double s[3][256];
  __shared__ double mx[256];
  const int t = threadIdx.x;
  const long nl = (long)(d.nj - 1) * (d.nk - 1);
  double a = 0, b = 0, c = 0;
  for (long q = t; q < nl; q += blockDim.x) {
    int j = 1 + (int)(q % (d.nj - 1)), k = 1 + (int)(q / (d.nj - 1));
    long ce = cidx(d, i, j, k);
    a += pdI[ce];
    if (i > 0) { b += pdJ[ce]; c += pdK[ce]; }
  }
  s[0][t] = a; s[1][t] = b; s[2][t] = c;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) { s[0][t] += s[0][t + st]; s[1][t] += s[1][t + st]; s[2][t] += s[2][t + st]; }
    __syncthreads();
  }
  const double ea = s[0][0] / nl, eb = s[1][0] / nl, ec = s[2][0] / nl;
  double m = 0.0;
  for (long q = t; q < nl; q += blockDim.x) {
    int j = 1 + (int)(q % (d.nj - 1)), k = 1 + (int)(q / (d.nj - 1));
    long ce = cidx(d, i, j, k);
    m = fmax(m, fabs(pdI[ce] - ea) / fabs(ea));
    if (i > 0) {
      m = fmax(m, fabs(pdJ[ce] - eb) / fabs(eb));
      m = fmax(m, fabs(pdK[ce] - ec) / fabs(ec));
    }
  }
  mx[t] = m;
  __syncthreads();
  for (int st = 128; st > 0; st >>= 1) {
    if (t < st) mx[t] = fmax(mx[t], mx[t + st]);
    __syncthreads();
  }
  if (t == 0) {
    const int n1 = d.ni - 1;
    avg[i] = ea;
    if (i > 0) {
      avg[(n1 + 1) + i - 1] = eb;
      avg[(n1 + 1) + n1 + i - 1] = ec;
    }
    dev[i] = mx[0];
  }
}

// Largest power-of-two column count whose two LDS buffers and twiddle table fit in lds_max
// bytes (at least 1).
inline int fft_cols(int n, size_t elem, size_t lds_max, int cap) {
  int c = 1;
  while (2 * c <= cap && (size_t)(2 * (2 * c) * n + n) * elem <= lds_max) c *= 2;
  return c;
}

template <typename T>
struct SepData {
  T* R = nullptr;                      // residual, compact interior array (I fastest)
  Cx<T>* S = nullptr;                  // spectrum [i + N1 (mJ + MJ mK)]
  Cx<T>*twJ = nullptr, *twK = nullptr; // exp(-2 pi i t / N) tables
  double* dW = nullptr;                // <D_W> per radial row (mode-line lower coefficient)
  double *invd = nullptr, *bf = nullptr;  // mode-line Thomas factors (FP64)
  int colsJ = 1, colsK = 1;            // columns per workgroup of the J and K transforms
};

struct SepState {
  FftPlan pJ, pK;
  int MJ = 0;
  SepData<double> d64;
  SepData<float> d32;
  bool ready = false;
};

template <typename T>
int sep_init(FvmCtx* c, SepState& st, SepData<T>& S, const std::vector<double>& avg,
             const double* dE_dev, const double* dW_dev, const double* dN_dev,
             const double* dT_dev) {
  const int n1 = c->d.ni - 1, nj1 = c->d.nj - 1, nk1 = c->d.nk - 1;
  auto dev_alloc = [&](void** p, size_t bytes) -> int {
    if (hipMalloc(p, bytes) != hipSuccess) return FVM_ERR_HIP;
    c->allocs.push_back(*p);
    return FVM_OK;
  };
  int r = 0;
  const size_t ncomp = (size_t)n1 * nj1 * nk1;
  const size_t nspec = (size_t)n1 * st.MJ * nk1;
  r |= dev_alloc((void**)&S.R, ncomp * sizeof(T));
  r |= dev_alloc((void**)&S.S, nspec * sizeof(Cx<T>));
  r |= dev_alloc((void**)&S.twJ, (size_t)nj1 * sizeof(Cx<T>));
  r |= dev_alloc((void**)&S.twK, (size_t)nk1 * sizeof(Cx<T>));
  r |= dev_alloc((void**)&S.invd, nspec * sizeof(double));
  r |= dev_alloc((void**)&S.bf, nspec * sizeof(double));
  if (r) return FVM_ERR_HIP;
  S.dW = (double*)dW_dev;
  for (int dir = 0; dir < 2; ++dir) {
    const int N = dir == 0 ? nj1 : nk1;
    std::vector<double> w = fft_twiddles(N);
    std::vector<Cx<T>> h(N);
    for (int t = 0; t < N; ++t) h[t] = Cx<T>{(T)w[2 * t], (T)w[2 * t + 1]};
    HIP_CHECK(hipMemcpy(dir == 0 ? S.twJ : S.twK, h.data(), N * sizeof(Cx<T>),
                        hipMemcpyHostToDevice));
  }
  // columns per workgroup: two buffers and the twiddles in at most 64 KB of LDS
  S.colsJ = fft_cols(nj1, sizeof(Cx<T>), 64 * 1024, 32);
  S.colsK = fft_cols(nk1, sizeof(Cx<T>), 64 * 1024, 64);
  hipLaunchKernelGGL(k_mode_factor, dim3((st.MJ + 63) / 64, nk1), dim3(64), 0, c->stream, n1,
                     st.MJ, nj1, nk1, dE_dev, dW_dev, dN_dev, dT_dev, S.invd, S.bf);
  HIP_CHECK(hipGetLastError());
  HIP_CHECK(hipStreamSynchronize(c->stream));
  return FVM_OK;
}

// p += M^-1 R: forward transforms of the residual, radial mode lines, inverse transforms with
// the correction added to p
template <typename T>
int sep_apply(FvmCtx* c, SepState& st, SepData<T>& S) {
  const Dims d = c->d;
  const int n1 = d.ni - 1, nj1 = d.nj - 1, nk1 = d.nk - 1;
  const size_t cb = sizeof(Cx<T>);
  const size_t ldsJ = (size_t)(2 * S.colsJ * nj1 + nj1) * cb;
  const size_t ldsK = (size_t)(2 * S.colsK * nk1 + nk1) * cb;
  const dim3 gJ((n1 + 2 * S.colsJ - 1) / (2 * S.colsJ), nk1);
  const dim3 gK((n1 + S.colsK - 1) / S.colsK, st.MJ);
  {
    ProfScope ps(c, "p_fft");
    hipLaunchKernelGGL(k_fft_j_fwd<T>, gJ, dim3(256), ldsJ, c->stream, n1, nk1, st.pJ, S.colsJ,
                       S.R, S.twJ, S.S);
    if (nk1 > 1)
      hipLaunchKernelGGL((k_fft_k<T, -1>), gK, dim3(256), ldsK, c->stream, n1, st.MJ, st.pK,
                         S.colsK, S.twK, S.S);
    HIP_CHECK(hipGetLastError());
  }
  {
    ProfScope ps(c, "p_mode_lines");
    mode_lines<T>(n1, st.MJ * nk1, S.dW, S.invd, S.bf, S.S, c->stream);
    HIP_CHECK(hipGetLastError());
  }
  {
    ProfScope ps(c, "p_fft");
    if (nk1 > 1)
      hipLaunchKernelGGL((k_fft_k<T, +1>), gK, dim3(256), ldsK, c->stream, n1, st.MJ, st.pK,
                         S.colsK, S.twK, S.S);
    const double scale = 1.0 / ((double)nj1 * (double)nk1);
    hipLaunchKernelGGL(k_fft_j_inv_add<T>, gJ, dim3(256), ldsJ, c->stream, d, st.pJ, S.colsJ,
                       S.S, S.twJ, scale, c->p);
    HIP_CHECK(hipGetLastError());
  }
  return FVM_OK;
}

}  // namespace

int fvm_sep_setup(FvmCtx* c) {
  const Dims d = c->d;
  const int n1 = d.ni - 1, nj1 = d.nj - 1, nk1 = d.nk - 1;
  SepState* S = new SepState();
  c->sep = S;
  // mode lines: at most 8 rows per lane of a wavefront; transforms: lengths with prime factors
  // 2, 3, 5, 7 and an even J length (two radial columns per complex sequence need none, the
  // half spectrum needs N / 2).  Other grids keep the line Gauss-Seidel forms only.
  if (n1 > 512 || !fft_plan(nj1, &S->pJ) || !fft_plan(nk1, &S->pK) || nj1 % 2 != 0) {
    fprintf(stderr, "fvm_amd: note: no separable pressure inverse for %d x %d x %d cells "
                    "(Poisson modes 0 and 1 only)\n", n1, nj1, nk1);
    return FVM_OK;
  }
  S->MJ = nj1 / 2 + 1;
  double *avg = nullptr, *dev = nullptr;
  const int navg = 3 * n1 + 1;
  HIP_CHECK(hipMalloc((void**)&avg, navg * sizeof(double)));
  c->allocs.push_back(avg);
  HIP_CHECK(hipMalloc((void**)&dev, (n1 + 1) * sizeof(double)));
  c->allocs.push_back(dev);
  hipLaunchKernelGGL(k_avg_faces, dim3(n1 + 1), dim3(256), 0, c->stream, d, c->pdI, c->pdJ,
                     c->pdK, avg, dev);
  HIP_CHECK(hipGetLastError());
  std::vector<double> hdev(n1 + 1);
  HIP_CHECK(hipMemcpyAsync(hdev.data(), dev, (n1 + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  double mx = 0.0;
  for (double v : hdev) mx = std::max(mx, v);
  c->sep_dev = mx;
  // row r (radial index i = r+1): dE = <D_E>(i+1/2) = avg[r+1], dW = <D_E>(i-1/2) = avg[r],
  // dN = avg[n1+1+r], dT = avg[2n1+1+r]
  const double* dW = avg;
  const double* dE = avg + 1;
  const double* dN = avg + n1 + 1;
  const double* dT = avg + 2 * n1 + 1;
  std::vector<double> none;
  int r = sep_init<double>(c, *S, S->d64, none, dE, dW, dN, dT);
  r |= sep_init<float>(c, *S, S->d32, none, dE, dW, dN, dT);
  if (r) return FVM_ERR_HIP;
  S->ready = true;
  return FVM_OK;
}

void fvm_sep_destroy(FvmCtx* c) {
  SepState* S = (SepState*)c->sep;
  if (!S) return;
  delete S;
  c->sep = nullptr;
}

// Defect correction p += M^-1 (rhs - L p) until sum|r| / sum|AP p| <= RESIMAX.
int fvm_k_iter_p_sep(FvmCtx* c, bool mixed, double* resip, int* niter) {
  SepState* S = (SepState*)c->sep;
  if (!S || !S->ready) {
    fprintf(stderr, "fvm_amd: Poisson modes 2 and 3 need the separable inverse, which this "
                    "grid does not have\n");
    return FVM_ERR_ARG;
  }
  const Dims d = c->d;
  const int kc = RS_KC;
  dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, (d.nk - 1 + kc - 1) / kc);
  long nb = (long)g.x * g.y * g.z;
  int outer = 0;
  double res = INFINITY, res_prev = INFINITY;
  const int cap = c->prm.nswpp;
  bool use32 = mixed;
  Mailbox mb{nullptr, nullptr, 0};
  for (;;) {
    int r = fvm_k_bcp(c);
    if (r) return r;
    {
      ProfScope ps(c, "p_residual");
      ProfScope pk(c, "k_residual");
      if (use32)
        hipLaunchKernelGGL(k_residual_km<float>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                           c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d32.R, c->red, kc);
      else
        hipLaunchKernelGGL(k_residual_km<double>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                           c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d64.R, c->red, kc);
      mb = (outer > 0 || cap < 1) ? fvm_mail(c) : Mailbox{nullptr, nullptr, 0};
      hipLaunchKernelGGL(k_res_final, dim3(1), dim3(256), 0, c->stream, c->red, nb, c->scal + 4,
                         mb);
      HIP_CHECK(hipGetLastError());
    }
    // Like ITER_P's sweep loop (iter_p.f90:36-62), at least one correction pass is made before
    // the residual is tested, so the first residual needs no round trip to the host.
    if (outer > 0 || cap < 1) {
      const int rf = fvm_mail_wait(c, mb, 3, c->hscal + 4);
      if (rf) return rf;
      res = c->hscal[6];
      if (!(res == res)) {
        fprintf(stderr, "fvm_amd: non-finite pressure residual (outer pass %d)\n", outer);
        return FVM_ERR_NAN;
      }
      if (res <= c->prm.resimax || outer >= cap) break;
    }
    if (use32 && outer > 0 && !(res < 0.5 * res_prev)) {
      // the FP32 approximate inverse stopped contracting: finish with the FP64 inverse
      use32 = false;
      hipLaunchKernelGGL(k_residual_km<double>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                         c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d64.R, c->red, kc);
      HIP_CHECK(hipGetLastError());
      ++c->sep_fallbacks;
    }
    res_prev = res;
    r = use32 ? sep_apply<float>(c, *S, S->d32) : sep_apply<double>(c, *S, S->d64);
    if (r) return r;
    ++outer;
  }
  int r = fvm_k_bcp(c);
  if (r) return r;
  *resip = res;
  *niter = outer;
  c->last.nouter = outer;
  if (outer >= cap && res > c->prm.resimax) return FVM_ERR_NOCONV;
  return FVM_OK;
}
