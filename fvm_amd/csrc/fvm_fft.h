// Batched complex FFTs in LDS for the separable pressure inverse: one workgroup transforms a
// tile of columns held in LDS, column-interleaved (element j of column c at [j * ncol + c], so
// the lanes of a wavefront, which take consecutive columns, touch consecutive LDS words).
//
// Mixed-radix Stockham autosort: a stage of radix R combines R transforms of length Ns into
// transforms of length Ns R, reading element j + q N/R (q < R) of the current buffer and writing
// element (j / Ns) Ns R + (j mod Ns) + r Ns of the other (j < N/R, r < R), with the twiddle
// exp(sign 2 pi i q (j mod Ns) / (Ns R)) taken from a table of exp(sign 2 pi i t / N).  Radices
// 4, 2, 3, 5, 7, 11 and 13 are butterflies in registers; lengths with other prime factors are not
// planned (fft_plan returns false).  Forward transforms are unnormalised
// (X_m = sum_j x_j exp(-2 pi i j m / N)); the inverse carries sign +1 and no 1/N.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

template <typename T>
struct Cx {
  T x, y;
};

template <typename T>
__device__ __forceinline__ Cx<T> cadd(Cx<T> a, Cx<T> b) { return {a.x + b.x, a.y + b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> csub(Cx<T> a, Cx<T> b) { return {a.x - b.x, a.y - b.y}; }
template <typename T>
__device__ __forceinline__ Cx<T> cmul(Cx<T> a, Cx<T> b) {
  return {a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
// i a (sign +1) or -i a (sign -1)
template <typename T, int SIGN>
__device__ __forceinline__ Cx<T> cmuli(Cx<T> a) {
  return SIGN > 0 ? Cx<T>{-a.y, a.x} : Cx<T>{a.y, -a.x};
}

#define FFT_MAX_STAGES 12

struct FftPlan {
  int n = 0, nst = 0;
  int rad[FFT_MAX_STAGES] = {};
};

// Radix sequence for length n: fours first, then a two, then the odd primes up to 13.
inline bool fft_plan(int n, FftPlan* pl) {
  pl->n = n;
  pl->nst = 0;
  int m = n;
  auto push = [&](int r) {
    if (pl->nst >= FFT_MAX_STAGES) return false;
    pl->rad[pl->nst++] = r;
    m /= r;
    return true;
  };
  while (m % 4 == 0) if (!push(4)) return false;
  while (m % 2 == 0) if (!push(2)) return false;
  for (int r : {3, 5, 7, 11, 13})
    while (m % r == 0) if (!push(r)) return false;
  return m == 1 && n >= 1;
}

// exp(-2 pi i t / n), t = 0..n-1, in FP64 on the host (table entries of the forward sign; the
// inverse conjugates them)
inline std::vector<double> fft_twiddles(int n) {
  std::vector<double> w(2 * (size_t)n);
  const double pi = 3.14159265358979323846;
  for (int t = 0; t < n; ++t) {
    // reduce the angle to the first octant's symmetric form for accuracy
    const double a = 2.0 * pi * (double)t / (double)n;
    w[2 * t] = std::cos(a);
    w[2 * t + 1] = -std::sin(a);
  }
  return w;
}

// R-point DFT of v in place (sign -1 forward, +1 inverse), odd prime R from the table: the
// R-th roots are table entries t = k n / R.
template <typename T, int SIGN, int R>
__device__ __forceinline__ void dft_small(Cx<T>* v, const Cx<T>* __restrict__ tw, int n) {
  Cx<T> out[R];
  const int step = n / R;
#pragma unroll
  for (int s = 0; s < R; ++s) {
    Cx<T> acc = v[0];
#pragma unroll
    for (int q = 1; q < R; ++q) {
      Cx<T> w = tw[((q * s) % R) * step];
      if (SIGN > 0) w.y = -w.y;
      acc = cadd(acc, cmul(v[q], w));
    }
    out[s] = acc;
  }
#pragma unroll
  for (int s = 0; s < R; ++s) v[s] = out[s];
}

template <typename T, int SIGN, int R>
__device__ __forceinline__ void butterfly(Cx<T>* v, const Cx<T>* __restrict__ tw, int n) {
  if (R == 2) {
    const Cx<T> a = v[0], b = v[1];
    v[0] = cadd(a, b);
    v[1] = csub(a, b);
  } else if (R == 4) {
    const Cx<T> t0 = cadd(v[0], v[2]), t1 = csub(v[0], v[2]);
    const Cx<T> t2 = cadd(v[1], v[3]), t3 = csub(v[1], v[3]);
    v[0] = cadd(t0, t2);
    v[2] = csub(t0, t2);
    // X1 = t1 + sign i t3, X3 = t1 - sign i t3
    const Cx<T> it3 = cmuli<T, SIGN>(t3);
    v[1] = cadd(t1, it3);
    v[3] = csub(t1, it3);
  } else if (R == 3) {
    const T h = (T)0.86602540378443864676;  // sqrt(3) / 2
    const Cx<T> t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    const Cx<T> m = {v[0].x - (T)0.5 * t1.x, v[0].y - (T)0.5 * t1.y};
    const Cx<T> s = cmuli<T, SIGN>(Cx<T>{h * t2.x, h * t2.y});
    v[0] = cadd(v[0], t1);
    v[1] = cadd(m, s);
    v[2] = csub(m, s);
  } else {
    dft_small<T, SIGN, R>(v, tw, n);
  }
}

// One stage of radix R over ncol columns: src -> dst (both LDS, column-interleaved).
template <typename T, int SIGN, int R>
__device__ inline void fft_stage(const Cx<T>* src, Cx<T>* dst, int ncol, int n, int Ns,
                                 const Cx<T>* __restrict__ tw) {
  const int nb = n / R;
  const int tstep = n / (Ns * R);
  for (int q = threadIdx.x; q < ncol * nb; q += blockDim.x) {
    const int col = q % ncol, j = q / ncol;
    const int k = j % Ns;
    Cx<T> v[R];
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[(j + r * nb) * ncol + col];
    if (Ns > 1) {
#pragma unroll
      for (int r = 1; r < R; ++r) {
        Cx<T> w = tw[k * r * tstep];
        if (SIGN > 0) w.y = -w.y;
        v[r] = cmul(v[r], w);
      }
    }
    butterfly<T, SIGN, R>(v, tw, n);
    const int base = (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[(base + r * Ns) * ncol + col] = v[r];
  }
}

// Transforms the ncol columns held in a (scratch b); returns the buffer holding the result.
template <typename T, int SIGN>
__device__ inline Cx<T>* fft_lds(Cx<T>* a, Cx<T>* b, int ncol, const FftPlan& pl,
                                 const Cx<T>* __restrict__ tw) {
  int Ns = 1;
  for (int s = 0; s < pl.nst; ++s) {
    const int R = pl.rad[s];
    switch (R) {
      case 4: fft_stage<T, SIGN, 4>(a, b, ncol, pl.n, Ns, tw); break;
      case 2: fft_stage<T, SIGN, 2>(a, b, ncol, pl.n, Ns, tw); break;
      case 3: fft_stage<T, SIGN, 3>(a, b, ncol, pl.n, Ns, tw); break;
      case 5: fft_stage<T, SIGN, 5>(a, b, ncol, pl.n, Ns, tw); break;
      case 7: fft_stage<T, SIGN, 7>(a, b, ncol, pl.n, Ns, tw); break;
      case 11: fft_stage<T, SIGN, 11>(a, b, ncol, pl.n, Ns, tw); break;
      default: fft_stage<T, SIGN, 13>(a, b, ncol, pl.n, Ns, tw); break;
    }
    __syncthreads();
    Cx<T>* t = a;
    a = b;
    b = t;
    Ns *= R;
  }
  return a;
}
