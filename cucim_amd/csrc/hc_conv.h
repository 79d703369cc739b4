// Register-blocked 1-D correlation cores over LDS-resident lines.
//
// Every core accumulates acc[j] += sum_t w[t] * x[j + t] with the taps in
// increasing order and one fused multiply-add per tap: the reference's
// `sum += cast<W>(x) * wval` order.  The reference skips zero weights
// (_vendored/_ndimage_filters_core.py:308-309).  With the accumulator
// starting at +0, a multiply-add with a zero weight leaves it unchanged for
// every finite x (x*0 is +-0 and the sum is never -0), so the skip only
// matters when x is inf or NaN.  The cores therefore come in two forms:
// SKIP=false multiplies every tap (branch-free; used when the staged tile is
// all finite, checked per work-group), SKIP=true tests each group of four
// taps against a host-computed "all nonzero" mask and skips zero taps one by
// one.  Tap counts are padded to a multiple of 4 with zeros.
//
// The weights are read from LDS (stage_taps at kernel start) rather than the
// kernel-argument segment, so that the window and weight reads of the next
// group are issued before the current group's multiply-adds: scalar memory
// returns out of order and would force a full wait every group.
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

// LDS copy of a kernel: K padded taps followed by 4 zeros (the pipelined
// reads run one group past the end).
template <typename W>
__device__ __forceinline__ void stage_taps(W* __restrict__ dst, const TapPack<W>& tp, int tid,
                                           int nthreads) {
  for (int k = tid; k < tp.K + 4; k += nthreads) dst[k] = k < tp.K ? tp.w[k] : (W)0;
}

template <typename W>
__device__ __forceinline__ W uniform(W w);
template <>
__device__ __forceinline__ float uniform<float>(float w) {
  return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(w)));
}
template <>
__device__ __forceinline__ double uniform<double>(double w) {
  const long long b = __double_as_longlong(w);
  const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
  const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
  return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// 64-bit "all four taps nonzero" mask of a tap pack (groups 0..63), loaded
// once before a loop.
template <typename W>
__device__ __forceinline__ uint64_t full_mask(const TapPack<W>& tp) {
  uint64_t m = tp.full[0];
  if (sizeof(tp.full) / sizeof(tp.full[0]) > 1) m |= (uint64_t)tp.full[1] << 32;
  return m;
}

// One group of 4 taps for P outputs: v[0 .. P+2] window, w[0..3] weights.
template <typename W, int P, bool SKIP>
__device__ __forceinline__ void group4(W (&acc)[P], const W* v, const W (&w)[4], uint64_t mask,
                                       int t) {
  if (!SKIP || ((mask >> (t >> 2)) & 1ull)) {
#pragma unroll
    for (int k = 0; k < 4; ++k)
#pragma unroll
      for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + k], w[k], acc[j]);
  } else {
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const W wk = uniform<W>(w[k]);
      if (nz(wk)) {
#pragma unroll
        for (int j = 0; j < P; ++j) acc[j] = fma_(v[j + k], wk, acc[j]);
      }
    }
  }
}

// ------------------------------------------------------------- row core --
// Four consecutive outputs of a contiguous LDS row.  x is 16-byte aligned and
// x[0 .. K+7] must be readable; wl: LDS taps (K + 4 entries).  The loop is
// unrolled three groups deep so the window registers rotate by renaming.
template <typename W, int NK, bool SKIP>
__device__ __forceinline__ void conv_row4n(W (&acc)[NK][4], const W* __restrict__ x,
                                           const W* const (&wl)[NK], const TapPack<W>* const (&tp)[NK],
                                           int K) {
  W r0[4], r1[4], r2[4];
  W w0[NK][4], w1[NK][4], w2[NK][4];
  uint64_t mk[NK];
#pragma unroll
  for (int n = 0; n < NK; ++n) mk[n] = SKIP ? full_mask(*tp[n]) : 0ull;
  auto blk = [&](const W (&A)[4], const W (&B)[4], const W (&ww)[NK][4], int t) {
    W v[8];
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[m] = A[m]; v[m + 4] = B[m]; }
#pragma unroll
    for (int n = 0; n < NK; ++n) group4<W, 4, SKIP>(acc[n], v, ww[n], mk[n], t);
  };
  ld4(x, r0);
  ld4(x + 4, r1);
#pragma unroll
  for (int n = 0; n < NK; ++n) ld4(wl[n], w0[n]);
  int t = 0;
  for (; t + 12 <= K; t += 12) {
    ld4(x + t + 8, r2);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, w1[n]);
    blk(r0, r1, w0, t);
    ld4(x + t + 12, r0);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 8, w2[n]);
    blk(r1, r2, w1, t + 4);
    ld4(x + t + 16, r1);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 12, w0[n]);
    blk(r2, r0, w2, t + 8);
  }
  if (t < K) {
    ld4(x + t + 8, r2);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, w1[n]);
    blk(r0, r1, w0, t);
    if (t + 4 < K) blk(r1, r2, w1, t + 4);
  }
}

template <bool SKIP, typename W>
__device__ __forceinline__ void conv_row4(W (&acc)[4], const W* __restrict__ x,
                                          const W* __restrict__ wl, const TapPack<W>& tp) {
  W a[1][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) a[0][j] = acc[j];
  const W* const wls[1] = {wl};
  const TapPack<W>* const tps[1] = {&tp};
  conv_row4n<W, 1, SKIP>(a, x, wls, tps, tp.K);
#pragma unroll
  for (int j = 0; j < 4; ++j) acc[j] = a[0][j];
}

// Two kernels of equal length over the same row window.
template <bool SKIP, typename W>
__device__ __forceinline__ void conv_row4x2(W (&accA)[4], W (&accB)[4], const W* __restrict__ x,
                                            const W* __restrict__ wa, const TapPack<W>& ta,
                                            const W* __restrict__ wb, const TapPack<W>& tb) {
  W a[2][4];
#pragma unroll
  for (int j = 0; j < 4; ++j) { a[0][j] = accA[j]; a[1][j] = accB[j]; }
  const W* const wls[2] = {wa, wb};
  const TapPack<W>* const tps[2] = {&ta, &tb};
  conv_row4n<W, 2, SKIP>(a, x, wls, tps, ta.K);
#pragma unroll
  for (int j = 0; j < 4; ++j) { accA[j] = a[0][j]; accB[j] = a[1][j]; }
}

// ---------------------------------------------------------- column core --
// Sixteen consecutive outputs along a column of an LDS tile with row stride
// S (compile time, so every read is an immediate offset).  Reads
// x[(0 .. K+18) * S]; the next group's window is fetched before the current
// group's multiply-adds (two-deep unroll, roles swap by renaming).
template <typename W, int S, int NK, bool SKIP>
__device__ __forceinline__ void conv_col16n(W (&acc)[NK][16], const W* __restrict__ x,
                                            const W* const (&wl)[NK],
                                            const TapPack<W>* const (&tp)[NK], int K) {
  W va[19], vb[19];
  W wa[NK][4], wb[NK][4];
  uint64_t mk[NK];
#pragma unroll
  for (int n = 0; n < NK; ++n) mk[n] = SKIP ? full_mask(*tp[n]) : 0ull;
  auto ldwin = [&](W (&v)[19], int t) {
    const W* xt = x + t * S;
#pragma unroll
    for (int m = 0; m < 19; ++m) v[m] = xt[m * S];
  };
  auto blk = [&](const W (&v)[19], const W (&ww)[NK][4], int t) {
#pragma unroll
    for (int n = 0; n < NK; ++n) group4<W, 16, SKIP>(acc[n], v, ww[n], mk[n], t);
  };
  ldwin(va, 0);
#pragma unroll
  for (int n = 0; n < NK; ++n) ld4(wl[n], wa[n]);
  int t = 0;
  for (; t + 8 <= K; t += 8) {
    ldwin(vb, t + 4);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, wb[n]);
    blk(va, wa, t);
    ldwin(va, t + 8);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 8, wa[n]);
    blk(vb, wb, t + 4);
  }
  if (t < K) blk(va, wa, t);
}

// P consecutive outputs (P = 8 or 16) along an LDS column: the general form
// of the column core; window P + 3 values per group.
template <typename W, int S, int NK, int P, bool SKIP>
__device__ __forceinline__ void conv_colP(W (&acc)[NK][P], const W* __restrict__ x,
                                          const W* const (&wl)[NK],
                                          const TapPack<W>* const (&tp)[NK], int K) {
  constexpr int V = P + 3;
  W va[V], vb[V];
  W wa[NK][4], wb[NK][4];
  uint64_t mk[NK];
#pragma unroll
  for (int n = 0; n < NK; ++n) mk[n] = SKIP ? full_mask(*tp[n]) : 0ull;
  auto ldwin = [&](W (&v)[V], int t) {
    const W* xt = x + t * S;
#pragma unroll
    for (int m = 0; m < V; ++m) v[m] = xt[m * S];
  };
  auto blk = [&](const W (&v)[V], const W (&ww)[NK][4], int t) {
#pragma unroll
    for (int n = 0; n < NK; ++n) group4<W, P, SKIP>(acc[n], v, ww[n], mk[n], t);
  };
  ldwin(va, 0);
#pragma unroll
  for (int n = 0; n < NK; ++n) ld4(wl[n], wa[n]);
  int t = 0;
  for (; t + 8 <= K; t += 8) {
    ldwin(vb, t + 4);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, wb[n]);
    blk(va, wa, t);
    ldwin(va, t + 8);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 8, wa[n]);
    blk(vb, wb, t + 4);
  }
  if (t < K) blk(va, wa, t);
}

// 8 consecutive outputs along an LDS row: acc[n][j] += sum_k w_n[k] x[j + k].
// The window x[t .. t + 11] of each group of 4 taps lives in three of four
// rotating 4-value registers (one 16-byte LDS read per group), the weights
// in two rotating sets; reads x[0 .. K + 11] and w[0 .. K + 3].
template <typename W, int NK, bool SKIP>
__device__ __forceinline__ void conv_row8n(W (&acc)[NK][8], const W* __restrict__ x,
                                           const W* const (&wl)[NK],
                                           const TapPack<W>* const (&tp)[NK], int K) {
  W ra[4], rb[4], rc[4], rd[4];
  W wa[NK][4], wb[NK][4];
  uint64_t mk[NK];
#pragma unroll
  for (int n = 0; n < NK; ++n) mk[n] = SKIP ? full_mask(*tp[n]) : 0ull;
  auto blk = [&](const W (&A)[4], const W (&B)[4], const W (&C)[4], const W (&ww)[NK][4],
                 int t) {
    W v[12];
#pragma unroll
    for (int m = 0; m < 4; ++m) { v[m] = A[m]; v[m + 4] = B[m]; v[m + 8] = C[m]; }
#pragma unroll
    for (int n = 0; n < NK; ++n) group4<W, 8, SKIP>(acc[n], v, ww[n], mk[n], t);
  };
  ld4(x, ra);
  ld4(x + 4, rb);
  ld4(x + 8, rc);
#pragma unroll
  for (int n = 0; n < NK; ++n) ld4(wl[n], wa[n]);
  int t = 0;
  for (; t + 16 <= K; t += 16) {
    ld4(x + t + 12, rd);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, wb[n]);
    blk(ra, rb, rc, wa, t);
    ld4(x + t + 16, ra);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 8, wa[n]);
    blk(rb, rc, rd, wb, t + 4);
    ld4(x + t + 20, rb);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 12, wb[n]);
    blk(rc, rd, ra, wa, t + 8);
    ld4(x + t + 24, rc);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 16, wa[n]);
    blk(rd, ra, rb, wb, t + 12);
  }
  // 0 .. 3 groups left; ra, rb, rc hold x[t .. t + 11], wa holds w[t .. t + 3]
  for (; t < K; t += 4) {
    ld4(x + t + 12, rd);
#pragma unroll
    for (int n = 0; n < NK; ++n) ld4(wl[n] + t + 4, wb[n]);
    blk(ra, rb, rc, wa, t);
#pragma unroll
    for (int m = 0; m < 4; ++m) { ra[m] = rb[m]; rb[m] = rc[m]; rc[m] = rd[m]; }
#pragma unroll
    for (int n = 0; n < NK; ++n)
#pragma unroll
      for (int m = 0; m < 4; ++m) wa[n][m] = wb[n][m];
  }
}

template <bool SKIP, typename W>
__device__ __forceinline__ void conv_row8(W (&acc)[8], const W* __restrict__ x,
                                          const W* __restrict__ wl, const TapPack<W>& tp) {
  W a[1][8];
#pragma unroll
  for (int j = 0; j < 8; ++j) a[0][j] = acc[j];
  const W* const wls[1] = {wl};
  const TapPack<W>* const tps[1] = {&tp};
  conv_row8n<W, 1, SKIP>(a, x, wls, tps, tp.K);
#pragma unroll
  for (int j = 0; j < 8; ++j) acc[j] = a[0][j];
}

// P = 4 or 8 consecutive outputs along an LDS row.
template <bool SKIP, int P, typename W>
__device__ __forceinline__ void conv_rowP(W (&acc)[P], const W* __restrict__ x,
                                          const W* __restrict__ wl, const TapPack<W>& tp) {
  if constexpr (P == 8)
    conv_row8<SKIP>(acc, x, wl, tp);
  else
    conv_row4<SKIP>(acc, x, wl, tp);
}

template <bool SKIP, typename W, int S>
__device__ __forceinline__ void conv_col16(W (&acc)[16], const W* __restrict__ x,
                                           const W* __restrict__ wl, const TapPack<W>& tp) {
  W a[1][16];
#pragma unroll
  for (int j = 0; j < 16; ++j) a[0][j] = acc[j];
  const W* const wls[1] = {wl};
  const TapPack<W>* const tps[1] = {&tp};
  conv_col16n<W, S, 1, SKIP>(a, x, wls, tps, tp.K);
#pragma unroll
  for (int j = 0; j < 16; ++j) acc[j] = a[0][j];
}

template <bool SKIP, typename W, int S>
__device__ __forceinline__ void conv_col16x2(W (&accA)[16], W (&accB)[16],
                                             const W* __restrict__ x, const W* __restrict__ wa,
                                             const TapPack<W>& ta, const W* __restrict__ wb,
                                             const TapPack<W>& tb) {
  W a[2][16];
#pragma unroll
  for (int j = 0; j < 16; ++j) { a[0][j] = accA[j]; a[1][j] = accB[j]; }
  const W* const wls[2] = {wa, wb};
  const TapPack<W>* const tps[2] = {&ta, &tb};
  conv_col16n<W, S, 2, SKIP>(a, x, wls, tps, ta.K);
#pragma unroll
  for (int j = 0; j < 16; ++j) { accA[j] = a[0][j]; accB[j] = a[1][j]; }
}

}  // namespace hc

namespace hc {
// Dispatch a templated body on the work-group's "tile holds a non-finite
// value" flag: body(std::true_type{}) or body(std::false_type{}).
template <class F>
__device__ __forceinline__ void with_skip(bool skip, F&& body) {
  if (skip) body(std::true_type{});
  else body(std::false_type{});
}
}  // namespace hc
