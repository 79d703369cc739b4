// Register-blocked 2-D separable passes (float32), shared by the fused 2-D
// filter and the batched 2-D tile pipeline.
//
// A work-group of 256 lanes owns a strip of 256 staged columns (one per lane)
// and TY output rows.  The vertical pass runs in registers: each lane loads
// its column's TY + K - 1 input values straight from global memory and forms
// TY outputs with taps known at compile time (the weights sit in SGPRs).  The
// horizontal pass reads the TY x 256 intermediate from LDS, 8 outputs per
// item.  Both passes accumulate every output as the reference does: one
// float32 fused multiply-add per non-zero tap, in tap order, from +0; zero
// taps are skipped with a wave-uniform branch, so non-finite inputs behave as
// in the reference's kernels.
#pragma once

#include "hc_common.h"

namespace hc {

template <int K>
struct RegTaps {
  float w[K];
};

constexpr int kRegCols = 256;              // staged columns per strip
constexpr int kRegP = 8;                   // horizontal outputs per item
constexpr int kRegUStride = kRegCols + 4;  // LDS row stride of the intermediate

// Non-zero test on the bit pattern (+0 and -0 are zero), scalar for
// wave-uniform weights.
__device__ __forceinline__ bool wnz(float w) { return (__float_as_uint(w) << 1) != 0u; }

// acc[i] = sum_k w[k] * v[i + k] for i < TY.
template <int K, int TY>
__device__ __forceinline__ void reg_vpass(float (&acc)[TY], const float (&v)[TY + K - 1],
                                          const RegTaps<K>& t) {
#pragma unroll
  for (int i = 0; i < TY; ++i) acc[i] = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float w = t.w[k];
    if (wnz(w)) {
#pragma unroll
      for (int i = 0; i < TY; ++i) acc[i] = __builtin_fmaf(v[i + k], w, acc[i]);
    }
  }
}

// acc[j] = sum_k w[k] * row[j + k] for j < kRegP; row 16-byte aligned in LDS.
template <int K>
__device__ __forceinline__ void reg_hpass(float (&acc)[kRegP], const float* __restrict__ row,
                                          const RegTaps<K>& t) {
  constexpr int NC = (kRegP + K - 1 + 3) / 4;
  float u[4 * NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const float4 q = *reinterpret_cast<const float4*>(row + 4 * c);
    u[4 * c] = q.x;
    u[4 * c + 1] = q.y;
    u[4 * c + 2] = q.z;
    u[4 * c + 3] = q.w;
  }
#pragma unroll
  for (int j = 0; j < kRegP; ++j) acc[j] = 0.f;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const float w = t.w[k];
    if (wnz(w)) {
#pragma unroll
      for (int j = 0; j < kRegP; ++j) acc[j] = __builtin_fmaf(u[j + k], w, acc[j]);
    }
  }
}

// Useful output columns of a strip for kernel length K: the horizontal
// items read 4 * ceil((kRegP + K - 1) / 4) staged values from their start.
__host__ __device__ constexpr int reg_strip_outputs(int K) {
  return ((kRegCols - 4 * ((kRegP + K - 1 + 3) / 4)) / kRegP + 1) * kRegP;
}

// Kernel lengths with a compiled register-blocked variant.
constexpr int kRegKs[] = {3, 5, 7, 9, 13, 17, 21, 25, 33};
constexpr int kRegMaxK = 33;

// Fused two-pass 2-D filter of a float32 image into a float32 output (the
// register-blocked kernel).  Returns HC_OK, or -1 when the arguments are
// outside what this kernel handles (the caller then uses the general tile
// kernel).
int launch_sep2d_reg(const hc_array* in, const hc_array* out, const hc_taps* t0,
                     const hc_taps* t1, int mode0, int mode1, double cval, hipStream_t st);

}  // namespace hc
