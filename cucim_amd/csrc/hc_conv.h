// Register-blocked 1-D correlation cores over LDS-resident lines.
//
// Both cores accumulate acc[j] += sum_t w[t] * x[j + t] with taps in
// increasing order and a fused multiply-add per tap, which is the reference's
// `sum += cast<W>(x) * wval` accumulation order.  Taps are consumed four at a
// time; a group whose four weights are all nonzero runs branch-free, a group
// holding a zero weight skips those taps one by one (the reference skips zero
// weights, _vendored/_ndimage_filters_core.py:308-309, which matters only for
// non-finite input).  Tap counts are padded to a multiple of 4 with zeros.
#pragma once

#include "hc_common.h"

namespace hc {

__device__ __forceinline__ void ld4(const float* p, float (&r)[4]) {
  float4 t = *reinterpret_cast<const float4*>(p);
  r[0] = t.x; r[1] = t.y; r[2] = t.z; r[3] = t.w;
}
__device__ __forceinline__ void ld4(const double* p, double (&r)[4]) {
  double2 a = *reinterpret_cast<const double2*>(p);
  double2 b = *reinterpret_cast<const double2*>(p + 2);
  r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y;
}

template <typename W, int P>
__device__ __forceinline__ void taps4(W (&acc)[P], const W* v, bool full, W w0, W w1, W w2,
                                      W w3) {
  if (full) {
#pragma unroll
    for (int j = 0; j < P; ++j) acc[j] = fma_(v[j], w0, acc[j]);
#pragma unroll
    for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 1], w1, acc[j]);
#pragma unroll
    for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 2], w2, acc[j]);
#pragma unroll
    for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 3], w3, acc[j]);
  } else {
    if (nz(w0)) {
#pragma unroll
      for (int j = 0; j < P; ++j) acc[j] = fma_(v[j], w0, acc[j]);
    }
    if (nz(w1)) {
#pragma unroll
      for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 1], w1, acc[j]);
    }
    if (nz(w2)) {
#pragma unroll
      for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 2], w2, acc[j]);
    }
    if (nz(w3)) {
#pragma unroll
      for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + 3], w3, acc[j]);
    }
  }
}

// Four consecutive outputs of a contiguous LDS row.  x is 16-byte aligned and
// x[0 .. K+7] must be readable (the next window is fetched one group ahead).
// tp.K is a multiple of 4.
template <typename W>
__device__ __forceinline__ void conv_row4(W (&acc)[4], const W* __restrict__ x,
                                          const TapPack<W>& tp) {
  W a[4], b[4], c[4], v[8];
  ld4(x, a);
  ld4(x + 4, b);
  const int K = tp.K;
  for (int t = 0; t < K; t += 4) {
    ld4(x + t + 8, c);
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[m] = a[m]; v[m + 4] = b[m]; }
    taps4<W, 4>(acc, v, group_full(tp, t), tp.w[t], tp.w[t + 1], tp.w[t + 2], tp.w[t + 3]);
#pragma unroll
    for (int m = 0; m < 4; ++m) { a[m] = b[m]; b[m] = c[m]; }
  }
}

// Two kernels of equal length over the same row window (first-derivative and
// smoothing passes of one Hessian line share their input reads).
template <typename W>
__device__ __forceinline__ void conv_row4x2(W (&accA)[4], W (&accB)[4], const W* __restrict__ x,
                                            const TapPack<W>& ta, const TapPack<W>& tb) {
  W a[4], b[4], c[4], v[8];
  ld4(x, a);
  ld4(x + 4, b);
  const int K = ta.K;
  for (int t = 0; t < K; t += 4) {
    ld4(x + t + 8, c);
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[m] = a[m]; v[m + 4] = b[m]; }
    taps4<W, 4>(accA, v, group_full(ta, t), ta.w[t], ta.w[t + 1], ta.w[t + 2], ta.w[t + 3]);
    taps4<W, 4>(accB, v, group_full(tb, t), tb.w[t], tb.w[t + 1], tb.w[t + 2], tb.w[t + 3]);
#pragma unroll
    for (int m = 0; m < 4; ++m) { a[m] = b[m]; b[m] = c[m]; }
  }
}

// Sixteen consecutive outputs along a column of an LDS tile with row stride
// S (elements, compile time so every read is an immediate offset).  Reads
// x[(0 .. K+15) * S].
template <typename W, int S>
__device__ __forceinline__ void conv_col16(W (&acc)[16], const W* __restrict__ x,
                                           const TapPack<W>& tp) {
  const int K = tp.K;
  for (int t = 0; t < K; t += 4) {
    W v[19];
    const W* xt = x + t * S;
#pragma unroll
    for (int m = 0; m < 19; ++m) v[m] = xt[m * S];
    taps4<W, 16>(acc, v, group_full(tp, t), tp.w[t], tp.w[t + 1], tp.w[t + 2], tp.w[t + 3]);
  }
}

template <typename W, int S>
__device__ __forceinline__ void conv_col16x2(W (&accA)[16], W (&accB)[16],
                                             const W* __restrict__ x, const TapPack<W>& ta,
                                             const TapPack<W>& tb) {
  const int K = ta.K;
  for (int t = 0; t < K; t += 4) {
    W v[19];
    const W* xt = x + t * S;
#pragma unroll
    for (int m = 0; m < 19; ++m) v[m] = xt[m * S];
    taps4<W, 16>(accA, v, group_full(ta, t), ta.w[t], ta.w[t + 1], ta.w[t + 2], ta.w[t + 3]);
    taps4<W, 16>(accB, v, group_full(tb, t), tb.w[t], tb.w[t + 1], tb.w[t + 2], tb.w[t + 3]);
  }
}

}  // namespace hc
