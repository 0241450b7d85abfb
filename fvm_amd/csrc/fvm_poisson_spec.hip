// Fast solve of the pressure equation by defect correction with a separable approximate inverse.
//
// The discrete operator of ITER_P (LES/v2.3_LES-Reynolds stresses/iter_p.f90:12-63) is
//   L p = sum_nb D_nb (p_nb - p) + DD(p)        (+ the BCP folds: wall Neumann, outflow AE=0)
// On an O-grid with uniform circumferential spacing extruded along a uniform span, the face
// coefficients D_nb depend on the radial index only and the cross terms DD vanish to round-off,
// so the orthogonal part L_D is diagonalised by real Fourier bases in J (periodic branch cut) and
// K (periodic span): for every mode pair (mJ, mK) what is left is ONE radial tridiagonal line
//   dW x(i-1) - (dE + dW + lamJ dN + lamK dT) x(i) + dE x(i+1) = r_hat(i),
// lam = 4 sin^2(pi m / N).  The transforms are dense orthonormal DFT matrices applied as GEMMs
// (rocBLAS, MFMA) on the compact interior array; the radial lines are solved by the batched
// 64-lane PCR line kernel.  The outer loop keeps the reference's FP64 operator -- cross terms,
// boundary folds, periodic ghosts -- in the residual r = rhs - L p and its scaled-L1 norm
// sum|r| / sum|AP p| (iter_p.f90:45-59), so it converges to the same fixed point as the
// reference's line Gauss-Seidel; the approximate inverse runs in FP64 (plain form) or FP32
// (mixed form).  Coefficients are averaged over J and K; their largest deviation from the
// average is measured at set-up (fvm_ctx.sep_dev).
#include <rocblas/rocblas.h>

#include <type_traits>
#include <vector>

#include "fvm_ctx.h"

namespace {

const dim3 kBlk(64, 4, 1);

template <typename T>
__device__ inline double cross_s(const Dims& d, long c, const double* __restrict__ f,
                                 const double* __restrict__ jq12, const double* __restrict__ jq13,
                                 const double* __restrict__ jq23) {
  const long sj = d.sj, sk = d.sk;
  const long e = c + 1, w = c - 1, n = c + sj, s = c - sj, t = c + sk, b = c - sk;
  double r = jq12[e] * (f[e + sj] - f[e - sj]);
  r = r - jq12[w] * (f[w + sj] - f[w - sj]);
  r = r + jq13[e] * (f[e + sk] - f[e - sk]);
  r = r - jq13[w] * (f[w + sk] - f[w - sk]);
  r = r + jq12[n] * (f[n + 1] - f[n - 1]);
  r = r - jq12[s] * (f[s + 1] - f[s - 1]);
  r = r + jq23[n] * (f[n + sk] - f[n - sk]);
  r = r - jq23[s] * (f[s + sk] - f[s - sk]);
  r = r + jq13[t] * (f[t + 1] - f[t - 1]);
  r = r - jq13[b] * (f[b + 1] - f[b - 1]);
  r = r + jq23[t] * (f[t + sj] - f[t - sj]);
  r = r - jq23[b] * (f[b + sj] - f[b - sj]);
  return 0.25 * r;
}

// residual r = sum A_nb p_nb + (rhs - DD(p)) - AP p of every interior cell (iter_p.f90:50-53
// with SU = rhs - DD), written to the compact interior array R (I fastest), plus per-block sums
// of |r| and |AP p|
template <typename T>
__global__ void k_residual_compact(Dims d, const double* __restrict__ p,
                                   const double* __restrict__ rhs,
                                   const double* __restrict__ pdI, const double* __restrict__ pdJ,
                                   const double* __restrict__ pdK, const double* __restrict__ jq12,
                                   const double* __restrict__ jq13, const double* __restrict__ jq23,
                                   T* __restrict__ R, double* part) {
  const uint3 bb = xcd_block();
  __shared__ double s0[256], s1[256];
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  const int t = threadIdx.x + 64 * threadIdx.y;
  double ra = 0.0, rb = 0.0;
  if (i <= d.ni - 1 && j <= d.nj - 1) {
    const long c = cidx(d, i, j, k);
    const double de = pdI[c], dw = pdI[c - 1], dn = pdJ[c], ds = pdJ[c - d.sj], dt = pdK[c],
                 db = pdK[c - d.sk];
    double ap = -de - dw - dn - ds - dt - db;
    double ae = -de, aw = -dw, an = -dn, as = -ds, at = -dt, ab = -db;
    if (i == 1) { ap = ap - aw; aw = 0.0; }
    if (i == d.ni - 1) ae = 0.0;
    double su = rhs[c] - cross_s<double>(d, c, p, jq12, jq13, jq23);
    double r = ae * p[c + 1] + aw * p[c - 1] + an * p[c + d.sj] + as * p[c - d.sj] +
               at * p[c + d.sk] + ab * p[c - d.sk] + su - ap * p[c];
    ra = fabs(r);
    rb = fabs(ap * p[c]);
    const long ci = (long)(i - 1) + (long)(d.ni - 1) * ((j - 1) + (long)(d.nj - 1) * (k - 1));
    R[ci] = (T)r;
  }
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
// D(r) = dg(r) - lo(r) B(r-1), are stored once in the layout of the transformed array
// [r + N1 (mJ + NJ1 mK)]; one thread per line (set-up only).
__global__ void k_mode_factor(int n, int nj1, int nk1, const double* __restrict__ dE,
                              const double* __restrict__ dW, const double* __restrict__ dN,
                              const double* __restrict__ dT, const double* __restrict__ lamJ,
                              const double* __restrict__ lamK, double* invd, double* bf) {
  const int mj = blockIdx.x * 64 + threadIdx.x, mk = blockIdx.y;
  if (mj >= nj1) return;
  const double lj = lamJ[mj], lk = lamK[mk];
  const long base = (long)n * (mj + (long)nj1 * mk);
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

// Radial mode-line solves with the stored factors, in FP64 whatever the storage type T: one
// 64-lane wavefront per line, lane l owning rows l*R .. l*R+R-1; forward elimination
// y(r) = z(r)/D(r) - lo(r)/D(r) y(r-1) and back substitution x(r) = y(r) - B(r) x(r+1) as
// affine-map scans across the lanes (shuffles, no LDS).
template <typename T, int R>
__global__ void __launch_bounds__(64) k_mode_scan(int n, const double* __restrict__ dW,
                                                  const double* __restrict__ invd,
                                                  const double* __restrict__ bf, T* Z) {
  const long base = (long)n * blockIdx.x;
  const int l = threadIdx.x;
  const int r0 = l * R;
  double al[R], be[R], bb[R];
  double At = 0.0, Bt = 1.0;
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = r0 + q;
    if (r < n) {
      const double iv = invd[base + r];
      al[q] = (double)Z[base + r] * iv;
      be[q] = r == 0 ? 0.0 : -dW[r] * iv;
      bb[q] = bf[base + r];
      At = al[q] + be[q] * At;
      Bt = be[q] * Bt;
    } else {
      al[q] = 0.0; be[q] = 0.0; bb[q] = 0.0;
    }
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double Ao = __shfl_up(At, o, 64), Bo = __shfl_up(Bt, o, 64);
    if (l >= o) { At = At + Bt * Ao; Bt = Bt * Bo; }
  }
  double yin = __shfl_up(At, 1, 64);  // y(-1) = 0: only the offset part survives
  if (l == 0) yin = 0.0;
  double y[R];
#pragma unroll
  for (int q = 0; q < R; ++q) {
    y[q] = al[q] + be[q] * yin;
    if (r0 + q < n) yin = y[q];
  }
  double Ct = 1.0, Dt = 0.0;
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    if (r0 + q < n) {
      Dt = y[q] - bb[q] * Dt;
      Ct = -bb[q] * Ct;
    }
  }
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    const double Do = __shfl_down(Dt, o, 64), Co = __shfl_down(Ct, o, 64);
    if (l + o < 64) { Dt = Dt + Ct * Do; Ct = Ct * Co; }
  }
  double xin = __shfl_down(Dt, 1, 64);  // x(n) = 0
  if (l == 63) xin = 0.0;
#pragma unroll
  for (int q = R - 1; q >= 0; --q) {
    const int r = r0 + q;
    if (r < n) {
      const double x = y[q] - bb[q] * xin;
      xin = x;
      Z[base + r] = (T)x;
    }
  }
}

template <typename T>
void mode_lines(int n1, int nlines, const double* dW, const double* invd, const double* bf, T* Z,
                hipStream_t st) {
  const int R = (n1 + 63) / 64;
  if (R <= 3) hipLaunchKernelGGL((k_mode_scan<T, 3>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else if (R <= 4) hipLaunchKernelGGL((k_mode_scan<T, 4>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else if (R <= 6) hipLaunchKernelGGL((k_mode_scan<T, 6>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
  else hipLaunchKernelGGL((k_mode_scan<T, 8>), dim3(nlines), dim3(64), 0, st, n1, dW, invd, bf, Z);
}

template <typename T>
__global__ void k_add_compact(Dims d, double* p, const T* __restrict__ X) {
  const uint3 bb = xcd_block();
  const int i = 1 + bb.x * 64 + threadIdx.x;
  const int j = 1 + bb.y * 4 + threadIdx.y;
  const int k = 1 + bb.z;
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const long c = cidx(d, i, j, k);
  const long ci = (long)(i - 1) + (long)(d.ni - 1) * ((j - 1) + (long)(d.nj - 1) * (k - 1));
  p[c] = p[c] + (double)X[ci];
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

// orthonormal real Fourier basis of length N (rows: cos 0, cos 1, sin 1, ..., [cos N/2]) and the
// eigenvalues 4 sin^2(pi m / N) of the periodic second difference
void fourier_basis(int N, std::vector<double>& F, std::vector<double>& lam) {
  F.assign((size_t)N * N, 0.0);  // column-major: F[row + N*col]
  lam.assign(N, 0.0);
  const double pi = 4.0 * atan(1.0);
  int row = 0;
  auto put = [&](int m, bool sine, double scale) {
    for (int j = 0; j < N; ++j) {
      double a = 2.0 * pi * (double)m * (double)j / (double)N;
      F[row + (size_t)N * j] = scale * (sine ? sin(a) : cos(a));
    }
    double sm = sin(pi * (double)m / (double)N);
    lam[row] = 4.0 * sm * sm;
    ++row;
  };
  put(0, false, 1.0 / sqrt((double)N));
  for (int m = 1; 2 * m < N; ++m) {
    put(m, false, sqrt(2.0 / N));
    put(m, true, sqrt(2.0 / N));
  }
  if (N % 2 == 0) put(N / 2, false, 1.0 / sqrt((double)N));
}

template <typename T>
struct Gemm;
template <>
struct Gemm<double> {
  static rocblas_status mm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                           int n, int k, const double* A, int lda, const double* B, int ldb,
                           double* C, int ldc) {
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
  }
  static rocblas_status sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                           int n, int k, const double* A, int lda, long sa, const double* B,
                           int ldb, long sb_, double* C, int ldc, long sc, int batch) {
    const double one = 1.0, zero = 0.0;
    return rocblas_dgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb_,
                                         &zero, C, ldc, sc, batch);
  }
};
template <>
struct Gemm<float> {
  static rocblas_status mm(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                           int n, int k, const float* A, int lda, const float* B, int ldb,
                           float* C, int ldc) {
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm(h, ta, tb, m, n, k, &one, A, lda, B, ldb, &zero, C, ldc);
  }
  static rocblas_status sb(rocblas_handle h, rocblas_operation ta, rocblas_operation tb, int m,
                           int n, int k, const float* A, int lda, long sa, const float* B,
                           int ldb, long sb_, float* C, int ldc, long sc, int batch) {
    const float one = 1.0f, zero = 0.0f;
    return rocblas_sgemm_strided_batched(h, ta, tb, m, n, k, &one, A, lda, sa, B, ldb, sb_,
                                         &zero, C, ldc, sc, batch);
  }
};

#define RB_CHECK(x)                                                                 \
  do {                                                                              \
    rocblas_status s_ = (x);                                                        \
    if (s_ != rocblas_status_success) {                                             \
      fprintf(stderr, "fvm_amd: rocBLAS status %d at %s:%d\n", (int)s_, __FILE__,   \
              __LINE__);                                                            \
      return FVM_ERR_HIP;                                                           \
    }                                                                               \
  } while (0)

template <typename T>
struct SepData {
  T *FJ = nullptr, *FK = nullptr;                       // transform matrices (storage type)
  double *lamJ = nullptr, *lamK = nullptr;              // line operators in FP64
  double *dE = nullptr, *dW = nullptr, *dN = nullptr, *dT = nullptr;
  T *X = nullptr, *Y = nullptr;
  double *invd = nullptr, *bf = nullptr;  // mode-line Thomas factors (FP64)
};

template <typename T>
int sep_init(FvmCtx* c, SepData<T>& S, const std::vector<double>& FJ,
             const std::vector<double>& FK, const std::vector<double>& lj,
             const std::vector<double>& lk, const std::vector<double>& avg) {
  const int n1 = c->d.ni - 1, nj1 = c->d.nj - 1, nk1 = c->d.nk - 1;
  auto up = [&](auto** dst, const std::vector<double>& src, size_t off, size_t cnt) -> int {
    using V = typename std::remove_reference<decltype(**dst)>::type;
    std::vector<V> h(cnt);
    for (size_t q = 0; q < cnt; ++q) h[q] = (V)src[off + q];
    if (hipMalloc((void**)dst, cnt * sizeof(V)) != hipSuccess) return FVM_ERR_HIP;
    c->allocs.push_back(*dst);
    if (hipMemcpy(*dst, h.data(), cnt * sizeof(V), hipMemcpyHostToDevice) != hipSuccess)
      return FVM_ERR_HIP;
    return FVM_OK;
  };
  int r = 0;
  r |= up(&S.FJ, FJ, 0, FJ.size());
  r |= up(&S.FK, FK, 0, FK.size());
  r |= up(&S.lamJ, lj, 0, lj.size());
  r |= up(&S.lamK, lk, 0, lk.size());
  // row r (radial index i = r+1): dE = <D_E>(i+1/2), dW = <D_E>(i-1/2)
  r |= up(&S.dE, avg, 1, n1);
  r |= up(&S.dW, avg, 0, n1);
  r |= up(&S.dN, avg, n1 + 1, n1);
  r |= up(&S.dT, avg, 2 * n1 + 1, n1);
  size_t ncomp = (size_t)n1 * nj1 * nk1;
  if (hipMalloc((void**)&S.X, ncomp * sizeof(T)) != hipSuccess) return FVM_ERR_HIP;
  c->allocs.push_back(S.X);
  if (hipMalloc((void**)&S.Y, ncomp * sizeof(T)) != hipSuccess) return FVM_ERR_HIP;
  c->allocs.push_back(S.Y);
  if (r) return r;
  double** fac[2] = {&S.invd, &S.bf};
  for (double** q : fac) {
    if (hipMalloc((void**)q, ncomp * sizeof(double)) != hipSuccess) return FVM_ERR_HIP;
    c->allocs.push_back(*q);
  }
  hipLaunchKernelGGL(k_mode_factor, dim3((nj1 + 63) / 64, nk1), dim3(64), 0, c->stream, n1, nj1,
                     nk1, S.dE, S.dW, S.dN, S.dT, S.lamJ, S.lamK, S.invd, S.bf);
  if (hipGetLastError() != hipSuccess) return FVM_ERR_HIP;
  if (hipStreamSynchronize(c->stream) != hipSuccess) return FVM_ERR_HIP;
  return FVM_OK;
}

// e = M^-1 r on the compact array X (in place through Y)
template <typename T>
int sep_apply(FvmCtx* c, SepData<T>& S, rocblas_handle h) {
  const int n1 = c->d.ni - 1, nj1 = c->d.nj - 1, nk1 = c->d.nk - 1;
  const int m = n1 * nj1;
  const long slab = (long)n1 * nj1;
  {
    ProfScope ps(c, "p_dft");
    // K transform: Y = X FK^T ; J transform per K slab: X = Y FJ^T
    RB_CHECK(Gemm<T>::mm(h, rocblas_operation_none, rocblas_operation_transpose, m, nk1, nk1,
                         S.X, m, S.FK, nk1, S.Y, m));
    RB_CHECK(Gemm<T>::sb(h, rocblas_operation_none, rocblas_operation_transpose, n1, nj1, nj1,
                         S.Y, n1, slab, S.FJ, nj1, 0, S.X, n1, slab, nk1));
  }
  {
    ProfScope ps(c, "p_mode_lines");
    mode_lines<T>(n1, nj1 * nk1, S.dW, S.invd, S.bf, S.X, c->stream);
    HIP_CHECK(hipGetLastError());
  }
  {
    ProfScope ps(c, "p_dft");
    RB_CHECK(Gemm<T>::sb(h, rocblas_operation_none, rocblas_operation_none, n1, nj1, nj1, S.X,
                         n1, slab, S.FJ, nj1, 0, S.Y, n1, slab, nk1));
    RB_CHECK(Gemm<T>::mm(h, rocblas_operation_none, rocblas_operation_none, m, nk1, nk1, S.Y, m,
                         S.FK, nk1, S.X, m));
  }
  return FVM_OK;
}

struct SepState {
  rocblas_handle h = nullptr;
  SepData<double> d64;
  SepData<float> d32;
  bool ready = false;
};

}  // namespace

int fvm_sep_setup(FvmCtx* c) {
  const Dims d = c->d;
  const int n1 = d.ni - 1, nj1 = d.nj - 1, nk1 = d.nk - 1;
  if (n1 > 512) return FVM_ERR_SIZE;  // mode lines: at most 8 rows per lane of a wavefront
  SepState* S = new SepState();
  c->sep = S;
  if (rocblas_create_handle(&S->h) != rocblas_status_success) return FVM_ERR_HIP;
  if (rocblas_set_stream(S->h, c->stream) != rocblas_status_success) return FVM_ERR_HIP;
  double *avg = nullptr, *dev = nullptr;
  const int navg = 3 * n1 + 1;
  HIP_CHECK(hipMalloc((void**)&avg, navg * sizeof(double)));
  HIP_CHECK(hipMalloc((void**)&dev, (n1 + 1) * sizeof(double)));
  hipLaunchKernelGGL(k_avg_faces, dim3(n1 + 1), dim3(256), 0, c->stream, d, c->pdI, c->pdJ,
                     c->pdK, avg, dev);
  HIP_CHECK(hipGetLastError());
  std::vector<double> havg(navg), hdev(n1 + 1);
  HIP_CHECK(hipMemcpyAsync(havg.data(), avg, navg * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipMemcpyAsync(hdev.data(), dev, (n1 + 1) * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
  HIP_CHECK(hipStreamSynchronize(c->stream));
  HIP_CHECK(hipFree(avg));
  HIP_CHECK(hipFree(dev));
  double mx = 0.0;
  for (double v : hdev) mx = std::max(mx, v);
  c->sep_dev = mx;
  std::vector<double> FJ, FK, lj, lk;
  fourier_basis(nj1, FJ, lj);
  fourier_basis(nk1, FK, lk);
  int r = sep_init<double>(c, S->d64, FJ, FK, lj, lk, havg);
  r |= sep_init<float>(c, S->d32, FJ, FK, lj, lk, havg);
  if (r) return FVM_ERR_HIP;
  S->ready = true;
  return FVM_OK;
}

void fvm_sep_destroy(FvmCtx* c) {
  SepState* S = (SepState*)c->sep;
  if (!S) return;
  if (S->h) rocblas_destroy_handle(S->h);
  delete S;
  c->sep = nullptr;
}

// Defect correction p += M^-1 (rhs - L p) until sum|r| / sum|AP p| <= RESIMAX.
int fvm_k_iter_p_sep(FvmCtx* c, bool mixed, double* resip, int* niter) {
  SepState* S = (SepState*)c->sep;
  if (!S || !S->ready) return FVM_ERR_ARG;
  const Dims d = c->d;
  dim3 g((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, d.nk - 1);
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
        hipLaunchKernelGGL(k_residual_compact<float>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                           c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d32.X, c->red);
      else
        hipLaunchKernelGGL(k_residual_compact<double>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                           c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d64.X, c->red);
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
      hipLaunchKernelGGL(k_residual_compact<double>, g, kBlk, 0, c->stream, d, c->p, c->rhs,
                         c->pdI, c->pdJ, c->pdK, c->jq12, c->jq13, c->jq23, S->d64.X, c->red);
      HIP_CHECK(hipGetLastError());
      ++c->sep_fallbacks;
    }
    res_prev = res;
    r = use32 ? sep_apply<float>(c, S->d32, S->h) : sep_apply<double>(c, S->d64, S->h);
    if (r) return r;
    {
      ProfScope ps(c, "p_add");
      if (use32)
        hipLaunchKernelGGL(k_add_compact<float>, g, kBlk, 0, c->stream, d, c->p, S->d32.X);
      else
        hipLaunchKernelGGL(k_add_compact<double>, g, kBlk, 0, c->stream, d, c->p, S->d64.X);
      HIP_CHECK(hipGetLastError());
    }
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
