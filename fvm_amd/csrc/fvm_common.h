// Shared definitions of the MI355X fractional-step solver: grid dimensions, HBM layout,
// launch helpers and the in-wavefront parallel-cyclic-reduction line solver.
//
// HBM layout: every cell field is one FP64 array of (NI+1)(NJ+1)(NK+1) values, I fastest,
// holding indices 0..NI, 0..NJ, 0..NK exactly as the reference's (IS:IE,JS:JE,KS:KE) arrays with
// one MPI rank (I = radial, J = circumferential, K = span; index 0 and N are ghosts).  The
// column-major order makes the reference's restart records a straight copy.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>

#define FVM_OK 0
#define FVM_ERR_ARG 1
#define FVM_ERR_HIP 2
#define FVM_ERR_NAN 3
#define FVM_ERR_NOCONV 4
#define FVM_ERR_SIZE 5

struct Dims {
  int ni, nj, nk;     // node counts NI, NJ, NK
  long sj, sk, n;     // strides of J and K, number of cells incl. ghosts
};

__host__ __device__ inline long cidx(const Dims& d, int i, int j, int k) {
  return (long)i + d.sj * (long)j + d.sk * (long)k;
}

// XCD-aware block order for the 3D stencil launches.  Workgroups are dealt round-robin, in flat
// (x fastest) order, over the 8 XCDs, each with its own L2.  This bijection of the flat block id
// hands XCD x one contiguous run of tiles in flat order (J rows, then K planes), so the J+-1 and
// K+-1 neighbours a tile reads were mostly fetched into the same L2 by the tiles next to it.
__device__ inline uint3 xcd_block() {
  const unsigned gx = gridDim.x, gy = gridDim.y;
  const unsigned nb = gx * gy * gridDim.z;
  const unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
  const unsigned x = b & 7u, slot = b >> 3, q = nb >> 3, r = nb & 7u;
  const unsigned w = (x < r ? x * (q + 1) : r * (q + 1) + (x - r) * q) + slot;
  return make_uint3(w % gx, (w / gx) % gy, w / (gx * gy));
}

// Device arrays read through pointers that come from tables in device memory (the metric, strain
// and face-coefficient tables) would be generic pointers to the compiler: their loads would be
// flat loads that may alias LDS, so each one is ordered behind the LDS stores before it.  Cast
// to the global address space they become global loads, free to be issued together.
typedef const __attribute__((address_space(1))) double* GP;
typedef __attribute__((address_space(1))) double* GPW;
__device__ inline GP gp(const double* p) { return (GP)p; }
__device__ inline GPW gpw(double* p) { return (GPW)p; }
// Line stride of the J-line-major copies of the line scans (a multiple of four elements, row 0
// at offset 3) and the element count that holds them with 128 elements of slack for the last
// line's vector loads past row NJ-1
__host__ __device__ inline long scan_jline_stride(int nj) { return ((long)nj + 1 + 3 + 3) & ~3L; }
__host__ __device__ inline long scan_jline_elems(int ni, int nj, int nk) {
  return scan_jline_stride(nj) * (ni + 1) * (nk + 1) + 128;
}

// the same for any element type: glb(p) is p in the global address space
template <typename T>
using gptr = __attribute__((address_space(1))) T*;
template <typename T>
__device__ __forceinline__ gptr<T> glb(T* p) { return (gptr<T>)p; }

// Metric a of the table (grouped per Cartesian component: xi_s, eta_s, zeta_s for s = x, y, z)
// at cell c.  With XZ0 the entries xi_z (a = 6) and eta_z (a = 7), which are exactly zero on
// every cell of a grid extruded along z (FvmCtx::xz_zero), are the constant 0.0 and are not
// loaded; a is a compile-time constant at every call, so the choice folds away.  The products
// and sums keep the reference's order, so results are unchanged.
template <bool XZ0>
__device__ inline double mget(const double* const* __restrict__ m, int a, long c) {
  return (XZ0 && (a == 6 || a == 7)) ? 0.0 : gp(m[a])[c];
}

// A 64 x 4 (I x J) tile of the K-marching stencil kernels with its one-cell halo, as staged in
// LDS: SF_W x SF_H cells, row-major.
constexpr int SF_W = 66, SF_H = 6, SF_P = SF_W * SF_H;

// Maximum line length handled by one 64-lane wavefront (8 rows per lane).
#define FVM_MAX_LINE 512

#define HIP_CHECK(x)                                                                   \
  do {                                                                                 \
    hipError_t e_ = (x);                                                               \
    if (e_ != hipSuccess) {                                                            \
      fprintf(stderr, "fvm_amd: HIP error %s at %s:%d\n", hipGetErrorString(e_),       \
              __FILE__, __LINE__);                                                     \
      return FVM_ERR_HIP;                                                              \
    }                                                                                  \
  } while (0)

// Parallel cyclic reduction of one tridiagonal system by one 64-thread workgroup.
// Rows r = 0..n-1 live in LDS: a[r] x[r-1] + b[r] x[r] + c[r] x[r+1] = d[r] with a[0] = 0 and
// c[n-1] = 0.  Lane t owns rows t, t+64, ... (R rows).  On return d[r] holds x[r].
// Each step of stride s eliminates the couplings to rows r-s and r+s; after ceil(log2 n)
// steps every row is decoupled.  Stable for the diagonally dominant systems of the AF and
// Poisson line solves.
template <int R, typename T>
__device__ inline void pcr_solve(T* a, T* b, T* c, T* d, int n) {
  const int t = threadIdx.x;
  for (int s = 1; s < n; s <<= 1) {
    T na[R], nb[R], nc[R], nd[R];
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = t + 64 * q;
      na[q] = nb[q] = nc[q] = nd[q] = T(0);
      if (r < n) {
        T ar = a[r], br = b[r], cr = c[r], dr = d[r];
        T xa = T(0), xc = T(0), xd = T(0), ya = T(0), yc = T(0), yd = T(0);
        T alpha = T(0), gamma = T(0);
        if (r >= s) {
          alpha = -ar / b[r - s];
          xa = a[r - s]; xc = c[r - s]; xd = d[r - s];
        }
        if (r + s < n) {
          gamma = -cr / b[r + s];
          ya = a[r + s]; yc = c[r + s]; yd = d[r + s];
        }
        na[q] = alpha * xa;
        nc[q] = gamma * yc;
        nb[q] = br + alpha * xc + gamma * ya;
        nd[q] = dr + alpha * xd + gamma * yd;
      }
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < R; ++q) {
      const int r = t + 64 * q;
      if (r < n) { a[r] = na[q]; b[r] = nb[q]; c[r] = nc[q]; d[r] = nd[q]; }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < R; ++q) {
    const int r = t + 64 * q;
    if (r < n) d[r] = d[r] / b[r];
  }
  __syncthreads();
}

// Dispatch on rows per lane.
template <typename T>
__device__ inline void pcr_solve_any(T* a, T* b, T* c, T* d, int n) {
  const int R = (n + 63) >> 6;
  switch (R) {
    case 1: pcr_solve<1>(a, b, c, d, n); break;
    case 2: pcr_solve<2>(a, b, c, d, n); break;
    case 3: pcr_solve<3>(a, b, c, d, n); break;
    case 4: pcr_solve<4>(a, b, c, d, n); break;
    case 5: pcr_solve<5>(a, b, c, d, n); break;
    case 6: pcr_solve<6>(a, b, c, d, n); break;
    case 7: pcr_solve<7>(a, b, c, d, n); break;
    default: pcr_solve<8>(a, b, c, d, n); break;
  }
}

// Periodic ghost copy of one cell array: J ghosts (branch cut) and K ghosts (span) take the
// value of the periodic partner for whole I rows (LES/v2.3_LES-Reynolds stresses/mpi_x.f90:26-53
// with one rank); only the ghost surface is enumerated.
template <typename T>
__global__ void k_ghost_fill_t(Dims d, T* __restrict__ a) {
  // periodic J then K ghosts of every I row: a(i,j,k) = a(i, wrap j, wrap k)
  const long nJ = (long)(d.ni + 1) * 2 * (d.nk + 1);
  const long nK = (long)(d.ni + 1) * (d.nj + 1) * 2;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nJ + nK;
       t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    if (t < nJ) {
      i = (int)(t % (d.ni + 1));
      long r = t / (d.ni + 1);
      j = (r % 2) == 0 ? 0 : d.nj;
      k = (int)(r / 2);
      if (k == 0 || k == d.nk) continue;  // corners come from the K pass
    } else {
      long q = t - nJ;
      i = (int)(q % (d.ni + 1));
      long r = q / (d.ni + 1);
      j = (int)(r % (d.nj + 1));
      k = (r / (d.nj + 1)) == 0 ? 0 : d.nk;
    }
    int js = j == 0 ? d.nj - 1 : (j == d.nj ? 1 : j);
    int ks = k == 0 ? d.nk - 1 : (k == d.nk ? 1 : k);
    a[cidx(d, i, j, k)] = a[cidx(d, i, js, ks)];
  }
}

// Host mailbox of a reduction kernel's result: fine-grained pinned memory that lane 0 of the
// kernel's last workgroup fills (values, then a system-scope fence, then the sequence number s),
// so the host reads the result without a further kernel or a stream synchronisation.
// dst == nullptr: nothing is published.
struct Mailbox {
  volatile double* dst;
  volatile unsigned long long* seq;
  unsigned long long s;
};

__device__ inline void publish(const Mailbox& mb, const double* v, int n) {
  if (!mb.dst) return;
  for (int q = 0; q < n; ++q) mb.dst[q] = v[q];
  __threadfence_system();
  *mb.seq = mb.s;
}

// up to eight arrays passed by value (no device pointer table needed)
struct PtrSet8 {
  double* p[8];
};

// the periodic ghost copy of several arrays in one launch (blockIdx.y selects the array); every
// ghost reads an interior (J, K) cell only, so the arrays and the cells are independent
static __global__ void k_ghost_fill_multi(Dims d, PtrSet8 set) {
  double* __restrict__ a = set.p[blockIdx.y];
  const long nJ = (long)(d.ni + 1) * 2 * (d.nk + 1);
  const long nK = (long)(d.ni + 1) * (d.nj + 1) * 2;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < nJ + nK;
       t += (long)gridDim.x * blockDim.x) {
    int i, j, k;
    if (t < nJ) {
      i = (int)(t % (d.ni + 1));
      long r = t / (d.ni + 1);
      j = (r % 2) == 0 ? 0 : d.nj;
      k = (int)(r / 2);
      if (k == 0 || k == d.nk) continue;
    } else {
      long q = t - nJ;
      i = (int)(q % (d.ni + 1));
      long r = q / (d.ni + 1);
      j = (int)(r % (d.nj + 1));
      k = (r / (d.nj + 1)) == 0 ? 0 : d.nk;
    }
    int js = j == 0 ? d.nj - 1 : (j == d.nj ? 1 : j);
    int ks = k == 0 ? d.nk - 1 : (k == d.nk ? 1 : k);
    a[cidx(d, i, j, k)] = a[cidx(d, i, js, ks)];
  }
}

// Scaled-L1 residual sum|r| / sum|AP p| of ITER_P (iter_p.f90:45-59).  When both sums are
// zero (p = 0 and a zero right-hand side) the reference's quotient is 0/0 = NaN, for which its
// test 'RESIP .GT. RESIMAX' is false, so its sweep loop ends: that state is converged, and 0 is
// returned for it.  Every other quotient, non-finite ones included, is returned as computed.
__host__ __device__ inline double scaled_residual(double s0, double s1) {
  return (s0 == 0.0 && s1 == 0.0) ? 0.0 : s0 / s1;
}

// 1D launch over n items.
inline dim3 grid1(long n, int block) { return dim3((unsigned)((n + block - 1) / block)); }

inline long ghost_count(const Dims& d) {
  return (long)(d.ni + 1) * 2 * (d.nk + 1) + (long)(d.ni + 1) * (d.nj + 1) * 2;
}

// Parallel cyclic reduction with one row per thread (row = threadIdx.x < n, blockDim >= n);
// same recurrence as pcr_solve.  On return d[r] holds x[r].
template <typename T>
__device__ inline void pcr_solve_rows(T* a, T* b, T* c, T* d, int n) {
  const int r = threadIdx.x;
  for (int s = 1; s < n; s <<= 1) {
    T na = T(0), nb = T(0), nc = T(0), nd = T(0);
    if (r < n) {
      T ar = a[r], br = b[r], cr = c[r], dr = d[r];
      T alpha = T(0), gamma = T(0), xa = T(0), xc = T(0), xd = T(0), ya = T(0), yc = T(0),
        yd = T(0);
      if (r >= s) { alpha = -ar / b[r - s]; xa = a[r - s]; xc = c[r - s]; xd = d[r - s]; }
      if (r + s < n) { gamma = -cr / b[r + s]; ya = a[r + s]; yc = c[r + s]; yd = d[r + s]; }
      na = alpha * xa;
      nc = gamma * yc;
      nb = br + alpha * xc + gamma * ya;
      nd = dr + alpha * xd + gamma * yd;
    }
    __syncthreads();
    if (r < n) { a[r] = na; b[r] = nb; c[r] = nc; d[r] = nd; }
    __syncthreads();
  }
  if (r < n) d[r] = d[r] / b[r];
  __syncthreads();
}
