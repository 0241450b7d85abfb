// Wavefront-wide scans of first-order affine recurrences on DPP lane moves (gfx9 data-parallel
// primitives: row shifts inside 16-lane rows, the row-15 / row-31 broadcasts, whole-wavefront
// shifts by one lane).  A lane holds the composite of its rows' maps; the scans compose those
// composites across the 64 lanes with VALU lane moves instead of LDS-routed permutes, so a
// step costs a few cycles rather than an LDS round trip.  Line solves (one tridiagonal system per
// wavefront, rows spread over the lanes) issue twelve such steps per line, one after another.
//
// A map y -> A + B y is held as (A, B); the identity is (0, 1).  A lane whose source lane lies
// outside the move's range keeps the identity passed as the old value, so the combine needs no
// lane test.  A double moves as its two 32-bit halves, a float as one.
#pragma once
#include <hip/hip_runtime.h>

namespace wave {

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp(double src, double old) {
  const long long s = __double_as_longlong(src), o = __double_as_longlong(old);
  const int lo = __builtin_amdgcn_update_dpp((int)o, (int)s, CTRL, ROW_MASK, 0xF, false);
  const int hi =
      __builtin_amdgcn_update_dpp((int)(o >> 32), (int)(s >> 32), CTRL, ROW_MASK, 0xF, false);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

template <int CTRL, int ROW_MASK>
__device__ __forceinline__ float dpp(float src, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(src), CTRL,
                                                    ROW_MASK, 0xF, false));
}

__device__ __forceinline__ double lane_value(double v, int lane) {
  const long long s = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)s, lane);
  const int hi = __builtin_amdgcn_readlane((int)(s >> 32), lane);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__device__ __forceinline__ float lane_value(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

// DPP controls
constexpr int kRowShr = 0x110;    // + n: lane reads lane - n of its row
constexpr int kRowShl = 0x100;    // + n: lane reads lane + n of its row
constexpr int kRowBcast15 = 0x142;
constexpr int kRowBcast31 = 0x143;
constexpr int kWaveShr1 = 0x138;  // lane reads lane - 1 (lane 0: none)
constexpr int kWaveShl1 = 0x130;  // lane reads lane + 1 (lane 63: none)

// forward step over N right-hand sides sharing B: (A_n, B) <- (A_n, B) o (predecessor)
template <int N, int CTRL, int ROW_MASK, typename T>
__device__ __forceinline__ void fstep(T (&A)[N], T& B) {
  T Ao[N];
#pragma unroll
  for (int q = 0; q < N; ++q) Ao[q] = dpp<CTRL, ROW_MASK>(A[q], (T)0);
  const T Bo = dpp<CTRL, ROW_MASK>(B, (T)1);
#pragma unroll
  for (int q = 0; q < N; ++q) A[q] = A[q] + B * Ao[q];
  B = B * Bo;
}

// backward step: (D_n, C) <- (D_n, C) o (successor composite)
template <int N, int CTRL, typename T>
__device__ __forceinline__ void bstep(T (&D)[N], T& C) {
  T Do[N];
#pragma unroll
  for (int q = 0; q < N; ++q) Do[q] = dpp<CTRL, 0xF>(D[q], (T)0);
  const T Co = dpp<CTRL, 0xF>(C, (T)1);
#pragma unroll
  for (int q = 0; q < N; ++q) D[q] = D[q] + C * Do[q];
  C = C * Co;
}

// Exclusive prefix of lane maps applied to y0: lane l gets (f_{l-1} o ... o f_0)(y0), lane 0 y0.
// (A, B) enter as the lane's own composite and leave as the inclusive prefix.
template <int N, typename T>
__device__ __forceinline__ void prefix_apply(T (&A)[N], T& B, const T (&y0)[N], T (&out)[N]) {
  fstep<N, kRowShr + 1, 0xF>(A, B);
  fstep<N, kRowShr + 2, 0xF>(A, B);
  fstep<N, kRowShr + 4, 0xF>(A, B);
  fstep<N, kRowShr + 8, 0xF>(A, B);
  fstep<N, kRowBcast15, 0xA>(A, B);  // rows 1, 3 take lanes 15, 47
  fstep<N, kRowBcast31, 0xC>(A, B);  // rows 2, 3 take lane 31
  const T Be = dpp<kWaveShr1, 0xF>(B, (T)1);
#pragma unroll
  for (int q = 0; q < N; ++q) out[q] = dpp<kWaveShr1, 0xF>(A[q], (T)0) + Be * y0[q];
}

// Exclusive suffix of lane maps applied to xn: lane l gets (g_{l+1} o ... o g_63)(xn), lane 63
// xn.  Inside a row by row shifts; across rows from the row composites held by lanes 16, 32, 48.
template <int N, typename T>
__device__ __forceinline__ void suffix_apply(T (&D)[N], T& C, const T (&xn)[N], T (&out)[N]) {
  bstep<N, kRowShl + 1>(D, C);
  bstep<N, kRowShl + 2>(D, C);
  bstep<N, kRowShl + 4>(D, C);
  bstep<N, kRowShl + 8>(D, C);
  const T C1 = lane_value(C, 16), C2 = lane_value(C, 32), C3 = lane_value(C, 48);
  const int row = (int)(threadIdx.x & 63) >> 4;
  // composites of the rows after row r: r = 2: row 3; r = 1: rows 2-3; r = 0: rows 1-3
  const T C23 = C2 * C3, C13 = C1 * C23;
  const T Ct = row == 0 ? C13 : (row == 1 ? C23 : (row == 2 ? C3 : (T)1));
#pragma unroll
  for (int q = 0; q < N; ++q) {
    const T D1 = lane_value(D[q], 16), D2 = lane_value(D[q], 32), D3 = lane_value(D[q], 48);
    const T D23 = D2 + C2 * D3;
    const T D13 = D1 + C1 * D23;
    const T Dt = row == 0 ? D13 : (row == 1 ? D23 : (row == 2 ? D3 : (T)0));
    D[q] = D[q] + C * Dt;
  }
  C = C * Ct;
  const T Ce = dpp<kWaveShl1, 0xF>(C, (T)1);
#pragma unroll
  for (int q = 0; q < N; ++q) out[q] = dpp<kWaveShl1, 0xF>(D[q], (T)0) + Ce * xn[q];
}

// one right-hand side
template <typename T>
__device__ __forceinline__ T prefix_apply(T& A, T& B, T y0) {
  T a[1] = {A}, y[1] = {y0}, o[1];
  prefix_apply<1>(a, B, y, o);
  A = a[0];
  return o[0];
}

template <typename T>
__device__ __forceinline__ T suffix_apply(T& D, T& C, T xn) {
  T d[1] = {D}, x[1] = {xn}, o[1];
  suffix_apply<1>(d, C, x, o);
  D = d[0];
  return o[0];
}

}  // namespace wave
