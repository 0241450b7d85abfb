// Line-scan form of the I and J stages of ITER_U (LES/v2.3_LES-Reynolds stresses/
// iter_u.f90:86-101 for I-lines, 163-178 for J-lines) on grids whose K couplings of these
// stages are negligible: on the extruded O-grid zeta_x and zeta_y are rounding residues of the
// metric (~1e-12), so AT, AB ~ 1e-24 J^-1 against a diagonal AP >= 2 Re J^-1/dt.  The form is
// enabled when (|AT| + |AB|) / AP <= 1e-15 everywhere; those terms are then taken from the
// stage input instead of the current sweep, a change below FP64 resolution.
//
// Each K plane is then an independent line Gauss-Seidel sweep: lines a = 1..na in order, each a
// tridiagonal solve along its rows p = 1..n whose right-hand side takes the line a-1 just solved
// and the stage input at line a+1 (the reference's lexicographic order within the plane).
//
// * The tridiagonal matrices depend only on the geometry, Re and dt: the reference's Thomas
//   factors A = -(lower coefficient), D = AP - A B(p-1), B = -(upper coefficient)/D
//   (B(0) = 0) are computed once per dt and stored as beta = -A/D, 1/D, B, together with the
//   coupling to line a-1 (BCU's wall / far-field folds included, bcu.f90:15-44).
// * A line solve is the forward recurrence y(p) = alpha(p) + beta(p) y(p-1),
//   alpha = (Q + cam x_{a-1}) / D, and the back substitution x(p) = y(p) - B(p) x(p+1); both are
//   first-order affine recurrences, evaluated by one 64-lane wavefront as an affine-map scan
//   (each lane composes its block of rows, then a six-step scan of DPP lane moves, no LDS).
//   Q = SU' + (coupling to line a+1) x_in(a+1) is formed once per stage.
// * A wavefront owns SC_C consecutive lines of one plane and sweeps them in order; chunks are
//   coupled through the last line of the preceding chunk from the previous pass (block Jacobi
//   across chunks).  The passes are nilpotent (exact after nchunk passes) and contractive by the
//   diagonal dominance of AP; the host stops once the chunk boundary lines no longer change
//   (max change <= af_tol max |x|).  The result equals the reference's sweep up to rounding.
// * Rows are contiguous for every load: I-lines are contiguous in the natural layout; for
//   J-lines the line factors, Q and the iterate live in a line-major copy (J fastest), and the
//   converged iterate is transposed back into the field.
#include "fvm_ctx.h"
#include "fvm_wave.h"

#include <string.h>

#include <algorithm>
#include <type_traits>
#include <vector>

namespace {

constexpr int SC_C = 16;  // lines per chunk
constexpr int SC_R = 8;   // max rows per lane (n <= 512)

struct ScanGeo {
  int n, na, nb;          // rows per line, lines per plane, planes
  long ls_a, ls_b;        // line-major strides of a and b (row stride 1)
  long off;               // offset of row 0 in a line-major line
  long ns_p, ns_a, ns_b;  // natural strides of p, a, b
};

__host__ __device__ inline ScanGeo scan_geo(const Dims& d, int S) {
  ScanGeo g;
  if (S == 0) {  // I-lines: p = i, a = j, b = k; line-major == natural
    g.n = d.ni - 1; g.na = d.nj - 1; g.nb = d.nk - 1;
    g.ls_a = d.sj; g.ls_b = d.sk; g.off = 0;
    g.ns_p = 1; g.ns_a = d.sj; g.ns_b = d.sk;
  } else {       // J-lines: p = j, a = i, b = k; line-major index 3 + j + L (i + (NI+1) k) with
                 // L = scan_jline_stride(NJ) a multiple of four: row 1 of every line, and so a
                 // lane's first row l R + 1 for R = 4, 8, on a 32-byte boundary (vector loads)
    g.n = d.nj - 1; g.na = d.ni - 1; g.nb = d.nk - 1;
    g.ls_a = scan_jline_stride(d.nj); g.ls_b = g.ls_a * (d.ni + 1); g.off = 3;
    g.ns_p = d.sj; g.ns_a = 1; g.ns_b = d.sk;
  }
  return g;
}

// Line factors of stage S, one thread per line (a, b), rows in order (setup only).
template <int S>
__global__ void k_scan_factor(Dims d, const double* const* __restrict__ fa,
                              const double* __restrict__ ap0, double* beta, double* invd,
                              double* bb, double* cam) {
  const ScanGeo g = scan_geo(d, S);
  const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
  if (t >= (long)g.na * g.nb) return;
  const int a = 1 + (int)(t % g.na), b = 1 + (int)(t / g.na);
  double bprev = 0.0;
  for (int p = 1; p <= g.n; ++p) {
    const long c = g.ns_p * p + g.ns_a * a + g.ns_b * b;
    const long l = g.off + (long)p + g.ls_a * a + g.ls_b * b;
    const int i = S == 0 ? p : a;
    double ae = fa[0][c], aw = fa[0][c - 1], an = fa[1][c], as = fa[1][c - d.sj],
           at = fa[2][c], ab = fa[2][c - d.sk];
    double ap = ap0[c] + ae + aw + an + as + at + ab;
    if (i == 1) { ap = ap + aw; aw = 0.0; }          // bcu.f90:18-20
    if (i == d.ni - 1) { ap = ap + ae; ae = 0.0; }   // bcu.f90:38-40
    const double lo = S == 0 ? aw : as;   // reference A(p) = -lo
    const double up = S == 0 ? ae : an;   // reference B(p) = -up / D(p)
    const double A = -lo;
    const double D = ap - A * bprev;
    const double B = -up / D;
    beta[l] = -A / D;
    invd[l] = 1.0 / D;
    bb[l] = B;
    cam[l] = S == 0 ? as : aw;
    bprev = B;
  }
}

// Q of the J stage (formed by the right-hand-side kernel in the natural layout) into the
// line-major layout, 32 x 32 (I x J) tiles transposed through LDS.
template <typename T>
__global__ void __launch_bounds__(256) k_scan_q1v(Dims d, PtrSet8 qn, double* const* q) {
  __shared__ T tile[32][33];
  const int i0 = 1 + blockIdx.x * 32, j0 = 1 + blockIdx.y * 32;
  const int k = 1 + (int)(blockIdx.z % (d.nk - 1));
  const int comp = (int)(blockIdx.z / (d.nk - 1));
  const int t = threadIdx.x;
  const T* __restrict__ src = reinterpret_cast<const T*>(qn.p[comp]);
  const gptr<T> dst = glb(reinterpret_cast<T*>(q[comp]));
  for (int r = 0; r < 4; ++r) {
    const int li = t % 32, lj = t / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= d.ni - 1 && j <= d.nj - 1) tile[lj][li] = src[cidx(d, i, j, k)];
  }
  __syncthreads();
  const ScanGeo g = scan_geo(d, 1);
  for (int r = 0; r < 4; ++r) {
    const int lj = t % 32, li = t / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= d.ni - 1 && j <= d.nj - 1) dst[g.off + j + g.ls_a * i + g.ls_b * k] = tile[lj][li];
  }
}

// converged J-line iterate (line-major) back into the field interior; T = float: the iterate is
// the stage's change, added to the field.  A workgroup moves a 64 (I) x 32 (J) tile of one plane
// through LDS: 32 contiguous J rows of a line in, 64 contiguous I cells of a field row out.
template <typename T>
__global__ void __launch_bounds__(256) k_scan_unpack1(Dims d, const double* const* xo,
                                                      double* const* f) {
  __shared__ T tile[32][65];
  const int i0 = 1 + blockIdx.x * 64, j0 = 1 + blockIdx.y * 32;
  const int k = 1 + (int)(blockIdx.z % (d.nk - 1));
  const int comp = (int)(blockIdx.z / (d.nk - 1));
  const int t = threadIdx.x;
  const ScanGeo g = scan_geo(d, 1);
  const gptr<const T> src = glb(reinterpret_cast<const T*>(xo[comp]));
  const gptr<double> dst = glb(f[comp]);
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int lj = t % 32, li = t / 32 + 8 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= d.ni - 1 && j <= d.nj - 1) tile[lj][li] = src[g.off + j + g.ls_a * i + g.ls_b * k];
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < 8; ++r) {
    const int li = t % 64, lj = t / 64 + 4 * r;
    const int i = i0 + li, j = j0 + lj;
    if (i <= d.ni - 1 && j <= d.nj - 1) {
      const long c = cidx(d, i, j, k);
      dst[c] = std::is_same<T, float>::value ? dst[c] + (double)tile[lj][li]
                                             : (double)tile[lj][li];
    }
  }
}

// FP32 correction form of the I stage: the converged change (natural layout) added to the field;
// a thread adds the change of four K planes of one cell column, loads first
constexpr int ADD0_KP = 4;
__global__ void __launch_bounds__(256) k_scan_add0(Dims d, PtrSet8 dl, PtrSet8 f) {
  const int i = 1 + blockIdx.x * 64 + threadIdx.x;
  const int j = 1 + blockIdx.y * 4 + threadIdx.y;
  const int nkb = (d.nk - 1 + ADD0_KP - 1) / ADD0_KP;
  const int k0 = 1 + (int)(blockIdx.z % nkb) * ADD0_KP;
  const int comp = (int)(blockIdx.z / nkb);
  if (i > d.ni - 1 || j > d.nj - 1) return;
  const gptr<const float> x = glb(reinterpret_cast<const float*>(dl.p[comp]));
  const gptr<double> y = glb(f.p[comp]);
  double a[ADD0_KP];
  float b[ADD0_KP];
#pragma unroll
  for (int q = 0; q < ADD0_KP; ++q) {
    const long c = cidx(d, i, j, min(k0 + q, d.nk - 1));
    a[q] = y[c];
    b[q] = x[c];
  }
#pragma unroll
  for (int q = 0; q < ADD0_KP; ++q)
    if (k0 + q <= d.nk - 1) y[cidx(d, i, j, k0 + q)] = a[q] + (double)b[q];
}

__global__ void k_to_float(long n, const double* __restrict__ a, float* b) {
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < n;
       t += (long)gridDim.x * blockDim.x)
    b[t] = (float)a[t];
}

// boundary lines of the first pass: the stage input at the last line of every chunk
// (bnd[((comp nb + b-1) nchunk + ca) n + p-1])
// (T = float: the change form, whose stage input is the zero change)
template <int S, typename T>
__global__ void k_scan_seed(Dims d, AFArgs A, int nchunk, T* bnd) {
  const ScanGeo g = scan_geo(d, S);
  const long tot = (long)A.nc * g.nb * nchunk * g.n;
  for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < tot;
       t += (long)gridDim.x * blockDim.x) {
    const int p = 1 + (int)(t % g.n);
    long r = t / g.n;
    const int ca = (int)(r % nchunk);
    r /= nchunk;
    const int b = 1 + (int)(r % g.nb);
    const int comp = (int)(r / g.nb);
    const int a = min((ca + 1) * SC_C, g.na);
    bnd[t] = std::is_same<T, float>::value ? (T)0
                                           : (T)A.f[comp][g.ns_p * p + g.ns_a * a + g.ns_b * b];
  }
}

// R consecutive elements from an address aligned to 4 (R % 4 == 0) or 2 (R % 2 == 0) elements
typedef float sc_f4 __attribute__((ext_vector_type(4)));
typedef float sc_f2 __attribute__((ext_vector_type(2)));
typedef double sc_d2 __attribute__((ext_vector_type(2)));
template <int R>
__device__ __forceinline__ void load_rows(float (&d)[R], gptr<const float> p) {
  if constexpr (R % 4 == 0) {
#pragma unroll
    for (int k = 0; k < R / 4; ++k) {
      const sc_f4 v = ((gptr<const sc_f4>)p)[k];
      d[4 * k] = v.x; d[4 * k + 1] = v.y; d[4 * k + 2] = v.z; d[4 * k + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int k = 0; k < R / 2; ++k) {
      const sc_f2 v = ((gptr<const sc_f2>)p)[k];
      d[2 * k] = v.x; d[2 * k + 1] = v.y;
    }
  }
}
template <int R>
__device__ __forceinline__ void load_rows(double (&d)[R], gptr<const double> p) {
#pragma unroll
  for (int k = 0; k < R / 2; ++k) {
    const sc_d2 v = ((gptr<const sc_d2>)p)[k];
    d[2 * k] = v.x; d[2 * k + 1] = v.y;
  }
}

__device__ inline double shfl_up_d(double v, int o) { return __shfl_up(v, o, 64); }
__device__ inline double shfl_down_d(double v, int o) { return __shfl_down(v, o, 64); }

// One pass: wavefront w owns chunk ca of plane b for one component and sweeps its lines in
// order (the component is the fastest index of w, so the three wavefronts sharing the line
// factors run side by side on one XCD).  Lane l owns rows l*R+1 .. l*R+R.  The loads of line
// a+1 are issued before line a is solved.  Writes the iterate (line-major), the chunk's last line
// to bnd_out, and per wavefront the largest change of that line between passes and the largest
// |x| (part[2w], part[2w+1]).
//
// Early stop (passes after the first, early != 0): a pass only changes a chunk through its
// incoming line a0-1, and a change d of line a moves line a+1 by at most rho d, with
// rho = max_p |cam| / (D - |lower| - |upper|) < 1 by the diagonal dominance of AP (the 2 Re J^-1/dt
// term).  So once the change of a whole line from the previous pass is <= tol times the largest
// |x| the wavefront has produced (never more than the global max |x| of the host's test), the
// remaining lines of the chunk would change by less than that too and keep their values from
// the previous pass; the chunk's last line is then unchanged (bnd_out = bnd_in).  Where rho is
// small (rho ~ dt |grad a|^2 / 2 Re), a pass after the first re-solves a few lines per chunk.
//
// T = float: the FP32 correction form.  q holds the FP32 residual of the line systems at the stage
// input (k_af_rhs_km with DR), the factors are FP32 copies, the iterate is the change of the
// stage (zero at the line ends and on the ghost line before the first chunk) and the arithmetic
// is FP32; the FP64 field is updated with the converged change afterwards.
template <int S, int R, bool EARLY, typename T>
__global__ void __launch_bounds__(64) k_scan_pass(Dims d, AFArgs A, const T* __restrict__ beta,
                                                  const T* __restrict__ invd,
                                                  const T* __restrict__ bb,
                                                  const T* __restrict__ cam,
                                                  const double* const* qt, double* const* xot,
                                                  const T* __restrict__ bnd_in,
                                                  T* bnd_out, int nchunk, double* part,
                                                  double tol, double* sinkd) {
  constexpr bool DELTA = std::is_same<T, float>::value;
  const ScanGeo g = scan_geo(d, S);
  const long G = gridDim.x;
  const long lin = blockIdx.x;
  const long w = (lin % 8) * (G / 8) + lin / 8;  // XCD-aware: neighbouring items share an XCD
  const int l = threadIdx.x;
  const gptr<T> sink = glb(reinterpret_cast<T*>(sinkd));
  double dmax = 0.0, xmax = 0.0;
  if (w < (long)A.nc * nchunk * g.nb) {
    const int comp = (int)(w % A.nc);
    const long wr = w / A.nc;
    const int ca = (int)(wr % nchunk);
    const int b = 1 + (int)(wr / nchunk);
    const int a0 = 1 + ca * SC_C;
    const int a1 = min(a0 + SC_C - 1, g.na);
    const int p0 = l * R + 1;
    const gptr<const T> qc = glb(reinterpret_cast<const T*>(qt[comp]));
    const GP fc = gp(A.f[comp]);
    const gptr<T> xc = glb(reinterpret_cast<T*>(xot[comp]));
    T xp[R];
    // line a0-1: the stage input's ghost line for the first chunk, else the preceding
    // chunk's last line of the previous pass
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int p = p0 + r;
      T v = (T)0;
      if (p <= g.n) {
        if (ca == 0)
          v = DELTA ? (T)0 : (T)fc[g.ns_p * p + g.ns_a * (a0 - 1) + g.ns_b * b];
        else
          v = bnd_in[(((long)comp * g.nb + (b - 1)) * nchunk + (ca - 1)) * g.n + (p - 1)];
      }
      xp[r] = v;
    }
    // Operands of one line.  Rows past n (the last lane's tail) read row n and are never kept,
    // and the next line is always fetched (the last line again at the end of the chunk): with
    // no load or store behind a branch the compiler counts the loads in flight exactly and
    // waits only for the ones it needs, so the next line's loads overlap this line's scans.
    T be[R], iv[R], bv[R], cm[R], qv[R], xv[R], y0n, xnn;
    auto load = [&](int a) {
      const long lb = g.ls_a * a + g.ls_b * b + g.off;
      if constexpr (S == 1 && R % 2 == 0) {
        // J lines: a lane's rows p0 .. p0+R-1 as vectors (lb + p0 is a multiple of 4 for R = 4,
        // 8 and of 2 for R = 6); rows past n lie in the padded line (or the next one, or the
        // slack after the last) and are never kept
        const long o = lb + p0;
        load_rows<R>(be, glb(beta) + o);
        load_rows<R>(iv, glb(invd) + o);
        load_rows<R>(bv, glb(bb) + o);
        load_rows<R>(cm, glb(cam) + o);
        load_rows<R>(qv, qc + o);
        if (EARLY) load_rows<R>(xv, (gptr<const T>)xc + o);
        else {
#pragma unroll
          for (int r = 0; r < R; ++r) xv[r] = (T)0;
        }
      } else {
#pragma unroll
        for (int r = 0; r < R; ++r) {
          const long o = lb + min(p0 + r, g.n);
          be[r] = beta[o]; iv[r] = invd[o]; bv[r] = bb[o]; cm[r] = cam[o]; qv[r] = qc[o];
          xv[r] = EARLY ? xc[o] : (T)0;  // the previous pass's value of the line
        }
      }
      if (DELTA) {
        y0n = (T)0;
        xnn = (T)0;
      } else {
        const long nb0 = g.ns_a * a + g.ns_b * b;  // natural base of the line
        y0n = (T)fc[nb0];                          // row 0 ghost
        xnn = (T)fc[nb0 + g.ns_p * (g.n + 1)];     // row n+1 ghost
      }
    };
    load(a0);
    bool stopped = false;
    T xrun = (T)0;  // largest |x| of the lines this wavefront solved in this pass
    const T ttol = (T)tol;
    for (int a = a0; a <= a1; ++a) {
      const long lb = g.ls_a * a + g.ls_b * b + g.off;
      const T y0 = y0n, xn = xnn;
      // forward: y(p) = al(p) + be(p) y(p-1)
      T al[R], bt[R], bw[R], xo[R];
      T Bt = (T)1, At = (T)0;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        al[r] = (qv[r] + cm[r] * xp[r]) * iv[r];
        bt[r] = be[r];
        bw[r] = bv[r];
        xo[r] = xv[r];  // line a's previous value (load below refills xv)
        const bool in = p0 + r <= g.n;
        At = in ? al[r] + bt[r] * At : At;
        Bt = in ? bt[r] * Bt : Bt;
      }
      load(min(a + 1, a1));  // next line's operands in flight during the scans
      __builtin_amdgcn_sched_barrier(0);  // keep those loads ahead of the scans
      T yin = wave::prefix_apply(At, Bt, y0);  // y0, xn: the same on every lane
      T y[R];
#pragma unroll
      for (int r = 0; r < R; ++r) {
        y[r] = al[r] + bt[r] * yin;
        yin = p0 + r <= g.n ? y[r] : yin;
      }
      // backward: x(p) = y(p) - B(p) x(p+1)
      T Ct = (T)1, Dt = (T)0;
#pragma unroll
      for (int r = R - 1; r >= 0; --r) {
        const bool in = p0 + r <= g.n;
        Dt = in ? y[r] - bw[r] * Dt : Dt;
        Ct = in ? -bw[r] * Ct : Ct;
      }
      T xin = wave::suffix_apply(Dt, Ct, xn);
      T dl = (T)0;
#pragma unroll
      for (int r = R - 1; r >= 0; --r) {
        const int p = p0 + r;
        const bool in = p <= g.n;
        const T x = y[r] - bw[r] * xin;
        xin = in ? x : xin;
        dl = in ? fmax(dl, fabs(x - xo[r])) : dl;
        xrun = in ? fmax(xrun, fabs(x)) : xrun;
        // a row past n stores to this lane's slot of the sink instead of a branch
        const gptr<T> dst = in ? xc + (lb + p) : sink + l;
        *dst = x;
        xp[r] = in ? x : xp[r];
      }
      if (EARLY && a < a1) {
#pragma unroll
        for (int o = 32; o > 0; o >>= 1) {
          dl = fmax(dl, __shfl_xor(dl, o, 64));
          xrun = fmax(xrun, __shfl_xor(xrun, o, 64));
        }
        if (dl <= ttol * xrun) { stopped = true; break; }  // uniform over the wavefront
      }
    }
    if (stopped) {  // the chunk's last line keeps the previous pass's value
#pragma unroll
      for (int r = 0; r < R; ++r) {
        const int p = p0 + r;
        if (p <= g.n) xp[r] = bnd_in[(((long)comp * g.nb + (b - 1)) * nchunk + ca) * g.n + (p - 1)];
      }
    }
    // chunk boundary line: change since the previous pass
#pragma unroll
    for (int r = 0; r < R; ++r) {
      const int p = p0 + r;
      if (p <= g.n) {
        const long o = (((long)comp * g.nb + (b - 1)) * nchunk + ca) * g.n + (p - 1);
        const T x = xp[r];
        dmax = fmax(dmax, (double)fabs(x - bnd_in[o]));
        xmax = fmax(xmax, (double)fabs(x));
        bnd_out[o] = x;
      }
    }
  }
  // wavefront max
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    dmax = fmax(dmax, shfl_down_d(dmax, o));
    xmax = fmax(xmax, shfl_down_d(xmax, o));
  }
  if (l == 0) { part[2 * lin] = dmax; part[2 * lin + 1] = xmax; }
}

__global__ void k_scan_max(const double* __restrict__ part, long nb, double* out, Mailbox mb) {
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
  if (t == 0) {
    out[0] = s0[0]; out[1] = s1[0];
    const double v[2] = {s0[0], s1[0]};
    publish(mb, v, 2);
  }
}

template <int S, int R, typename T>
int scan_stage(FvmCtx* c, const AFArgs& A, int* iters, bool temp) {
  constexpr bool DELTA = std::is_same<T, float>::value;
  const Dims d = c->d;
  hipStream_t st = c->stream;
  const ScanGeo g = scan_geo(d, S);
  const int nchunk = (g.na + SC_C - 1) / SC_C;
  const int nc = A.nc;
  const long nwork = (long)nc * nchunk * g.nb;
  const long G = (nwork + 7) / 8 * 8;
  if (2 * G > c->nred) return FVM_ERR_SIZE;
  // Q (or the FP32 residual): the right-hand-side kernel left it in A.su (natural layout); the
  // J stage uses a line-major copy.  The iterate is the field itself for the FP64 I stage, else
  // a work array (line-major for J).
  double* q[3] = {};
  for (int k = 0; k < nc; ++k) q[k] = S == 0 ? A.su[k] : c->afb[k][1];
  double* xo[3] = {};
  for (int k = 0; k < nc; ++k) xo[k] = (S == 0 && !DELTA) ? A.f[k] : c->afb[k][0];
  double* fo[3] = {};
  for (int k = 0; k < nc; ++k) fo[k] = A.f[k];
  // device pointer table of this stage: q at [0..2], iterate at [3..5], field at [6..8]; the
  // buffers are the context's own, so it is uploaded once and again only if they change
  // (the temperature has a table of its own)
  double* tab[9] = {q[0], q[1], q[2], xo[0], xo[1], xo[2], fo[0], fo[1], fo[2]};
  double** dtab = temp ? c->scantab_t[S] : c->scantab[S];
  double** htab = temp ? c->scantab_th[S] : c->scantab_h[S];
  if (memcmp(tab, htab, sizeof(tab)) != 0) {
    HIP_CHECK(hipMemcpyAsync(dtab, tab, sizeof(tab), hipMemcpyHostToDevice, st));
    memcpy(htab, tab, sizeof(tab));
  }
  double* const* dq = dtab;
  double* const* dxo = dtab + 3;
  T* bnd[2] = {reinterpret_cast<T*>(c->bnd[0]), reinterpret_cast<T*>(c->bnd[1])};
  {
    ProfScope pk(c, "af_scan_q");
    if (S == 1) {
      PtrSet8 qn{};
      for (int k = 0; k < nc; ++k) qn.p[k] = A.su[k];
      hipLaunchKernelGGL(k_scan_q1v<T>, dim3((d.ni - 1 + 31) / 32, (d.nj - 1 + 31) / 32, nc * (d.nk - 1)),
                         dim3(256), 0, st, d, qn, dq);
    }
    hipLaunchKernelGGL((k_scan_seed<S, T>), grid1((long)nc * g.nb * nchunk * g.n, 256), dim3(256),
                       0, st, d, A, nchunk, bnd[0]);
    HIP_CHECK(hipGetLastError());
  }
  // the FP32 change is resolved to its own precision
  double tol = c->prm.af_tol > 0 ? c->prm.af_tol : 1e-14;
  if (DELTA) tol = std::max(tol, 6e-8);
  const int cap = nchunk + 1;
  int slot = 0, it = 0;
  const T* const* lf;
  if (DELTA)
    lf = reinterpret_cast<const T* const*>(c->lf32[S]);
  else
    lf = reinterpret_cast<const T* const*>(temp ? c->lft[S] : c->lf[S]);
  for (;;) {
    {
      ProfScope pk(c, S == 0 ? "k_scan_pass_i" : "k_scan_pass_j");
      auto kern = it > 0 ? k_scan_pass<S, R, true, T> : k_scan_pass<S, R, false, T>;
      hipLaunchKernelGGL(kern, dim3((unsigned)G), dim3(64), 0, st, d, A, lf[0], lf[1], lf[2],
                         lf[3], (const double* const*)dq, dxo, bnd[slot], bnd[slot ^ 1],
                         nchunk, c->red, tol, c->sink);
    }
    // the first pass never suffices on its own: its maxima are not read
    const Mailbox mb = it > 0 ? fvm_mail(c) : Mailbox{nullptr, nullptr, 0};
    hipLaunchKernelGGL(k_scan_max, dim3(1), dim3(256), 0, st, c->red, G, c->scal + 20, mb);
    HIP_CHECK(hipGetLastError());
    slot ^= 1;
    ++it;
    if (it < 2) continue;
    const int rf = fvm_mail_wait(c, mb, 2, c->hscal + 20);
    if (rf) return rf;
    const double dm = c->hscal[20], xm = c->hscal[21];
    if (!(dm == dm) || !(xm == xm)) {
      fprintf(stderr, "fvm_amd: non-finite AF stage %d iterate (pass %d)\n", S, it);
      return FVM_ERR_NAN;
    }
    if (dm <= tol * xm || it >= cap) break;
  }
  if (S == 1) {
    ProfScope pk(c, "af_scan_unpack");
    hipLaunchKernelGGL(k_scan_unpack1<T>, dim3((d.ni - 1 + 63) / 64, (d.nj - 1 + 31) / 32, nc * (d.nk - 1)),
                       dim3(256), 0, st, d, (const double* const*)dxo, dtab + 6);
    HIP_CHECK(hipGetLastError());
  } else if (DELTA) {
    ProfScope pk(c, "af_scan_add");
    PtrSet8 dl{}, f{};
    for (int k = 0; k < nc; ++k) { dl.p[k] = xo[k]; f.p[k] = A.f[k]; }
    const int nkb = (d.nk - 1 + ADD0_KP - 1) / ADD0_KP;
    hipLaunchKernelGGL(k_scan_add0, dim3((d.ni - 1 + 63) / 64, (d.nj - 1 + 3) / 4, nc * nkb),
                       dim3(64, 4, 1), 0, st, d, dl, f);
    HIP_CHECK(hipGetLastError());
  }
  *iters = it;
  return FVM_OK;
}

}  // namespace

// Enables the scan form of stage s (0: I-lines, 1: J-lines) when its K couplings vanish and the
// lines fit the wavefront (n <= 64 * SC_R), and computes the line factors.
int fvm_af_scan_setup(FvmCtx* c) {
  const Dims d = c->d;
  hipStream_t st = c->stream;
  long nbnd = 0;
  // the factor arrays hold the natural (I) and the padded J-line-major (J) layouts
  const long ne = std::max((long)d.n, scan_jline_elems(d.ni, d.nj, d.nk));
  for (int s = 0; s < 2; ++s) {
    const ScanGeo g = scan_geo(d, s);
    c->af_scan[s] = c->af_b_ratio[s] <= 1e-15 && g.n <= 64 * SC_R;
    if (!c->af_scan[s]) continue;
    const int nchunk = (g.na + SC_C - 1) / SC_C;
    nbnd = std::max(nbnd, 3L * g.nb * nchunk * g.n);
    const long nl = (long)g.na * g.nb;
    for (int set = 0; set < (c->thermal ? 2 : 1); ++set) {  // set 1: the temperature's factors
      double** L = set ? c->lft[s] : c->lf[s];
      const double* a0 = set ? c->ap0t : c->ap0;
      if (!L[0]) {
        for (int q = 0; q < 4; ++q) {
          HIP_CHECK(hipMalloc((void**)&L[q], ne * sizeof(double)));
          c->allocs.push_back(L[q]);
        }
      }
      if (s == 0)
        hipLaunchKernelGGL(k_scan_factor<0>, grid1(nl, 64), dim3(64), 0, st, d,
                           (const double* const*)c->dfa[0], a0, L[0], L[1], L[2], L[3]);
      else
        hipLaunchKernelGGL(k_scan_factor<1>, grid1(nl, 64), dim3(64), 0, st, d,
                           (const double* const*)c->dfa[1], a0, L[0], L[1], L[2], L[3]);
      HIP_CHECK(hipGetLastError());
    }
    if (c->prm.af_mode == 2) {  // FP32 copies of the momentum factors (correction form)
      for (int q = 0; q < 4; ++q) {
        if (!c->lf32[s][q]) {
          HIP_CHECK(hipMalloc((void**)&c->lf32[s][q], ne * sizeof(float)));
          c->allocs.push_back(c->lf32[s][q]);
        }
        hipLaunchKernelGGL(k_to_float, grid1(ne, 256), dim3(256), 0, st, ne, c->lf[s][q],
                           c->lf32[s][q]);
      }
      HIP_CHECK(hipGetLastError());
    }
  }
  if (nbnd > c->nbnd) {
    for (int q = 0; q < 2; ++q) {
      HIP_CHECK(hipMalloc((void**)&c->bnd[q], nbnd * sizeof(double)));
      c->allocs.push_back(c->bnd[q]);
    }
    c->nbnd = nbnd;
  }
  HIP_CHECK(hipStreamSynchronize(st));
  return FVM_OK;
}

template <typename T>
static int scan_stage_r(FvmCtx* c, int s, const AFArgs& A, int* iters, bool temp) {
  const ScanGeo g = scan_geo(c->d, s);
  const int R = (g.n + 63) / 64;
  if (s == 0) {
    if (R <= 3) return scan_stage<0, 3, T>(c, A, iters, temp);
    if (R <= 4) return scan_stage<0, 4, T>(c, A, iters, temp);
    if (R <= 6) return scan_stage<0, 6, T>(c, A, iters, temp);
    return scan_stage<0, SC_R, T>(c, A, iters, temp);
  }
  if (R <= 3) return scan_stage<1, 3, T>(c, A, iters, temp);
  if (R <= 4) return scan_stage<1, 4, T>(c, A, iters, temp);
  if (R <= 6) return scan_stage<1, 6, T>(c, A, iters, temp);
  return scan_stage<1, SC_R, T>(c, A, iters, temp);
}

int fvm_af_scan_stage(FvmCtx* c, int s, const AFArgs& A, int* iters, bool temp, bool delta) {
  if (A.nc < 1 || A.nc > 3 || (temp && !c->lft[s][0])) return FVM_ERR_ARG;
  if (delta) {
    if (temp || !c->lf32[s][0]) return FVM_ERR_ARG;
    return scan_stage_r<float>(c, s, A, iters, false);
  }
  return scan_stage_r<double>(c, s, A, iters, temp);
}
