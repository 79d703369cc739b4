// Shared device/host helpers for the cucim_amd HIP kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>
#include <hip/hip_fp16.h>
#include <stdint.h>
#include <string>

#include "../../include/cucim_amd.h"

namespace hc {

// ---------------------------------------------------------------- errors --
void set_error(const std::string& msg);
int fail(int code, const std::string& msg);
int hip_check(hipError_t e, const char* what);

// ------------------------------------------------------------ tuning --
// Current value of a tile-shape knob (hc_tune_knob); 0 = built-in heuristic.
int tune_get(int knob);

// ------------------------------------------------------------- tracing --
// When enabled with hc_timing_enable(1), every kernel launch is bracketed by
// a pair of hipEvents on its stream and tagged with the kernel's name;
// hc_timing_collect() returns the elapsed times.
int timing_begin(const char* name, hipStream_t st);
int launch_error(hipError_t e, const char* name, size_t smem);
void timing_end(int slot, hipStream_t st);

// Launch with a clean error state, so a failure is attributed to this launch.
#define HC_LAUNCH(name, st, kernel, grid, block, smem, ...)              \
  do {                                                                   \
    (void)hipGetLastError();                                             \
    const int _slot = hc::timing_begin(name, st);                        \
    hipLaunchKernelGGL(kernel, grid, block, smem, st, __VA_ARGS__);      \
    const hipError_t _le = hipGetLastError();                            \
    hc::timing_end(_slot, st);                                           \
    if (_le != hipSuccess) return hc::launch_error(_le, name, smem);     \
  } while (0)

#define HC_HIP(call)                                   \
  do {                                                 \
    hipError_t _e = (call);                            \
    if (_e != hipSuccess) return hc::hip_check(_e, #call); \
  } while (0)

// ------------------------------------------------- separate rounding --
// Separately rounded multiply / add: the reference forms these in separate
// kernels (square in place, then accumulate), so they must not contract into
// a fused multiply-add (the pragma removes the contraction permission from
// these operations, which survives inlining).
__device__ __forceinline__ float mul_rn(float a, float b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ double mul_rn(double a, double b) {
#pragma clang fp contract(off)
  return a * b;
}
__device__ __forceinline__ float add_rn(float a, float b) {
#pragma clang fp contract(off)
  return a + b;
}
__device__ __forceinline__ double add_rn(double a, double b) {
#pragma clang fp contract(off)
  return a + b;
}

// --------------------------------------------------------------- dtypes --
inline int dtype_size(int dt) {
  switch (dt) {
    case HC_BOOL: case HC_INT8: case HC_UINT8: return 1;
    case HC_INT16: case HC_UINT16: case HC_FLOAT16: case HC_BFLOAT16: return 2;
    case HC_INT32: case HC_UINT32: case HC_FLOAT32: return 4;
    case HC_INT64: case HC_UINT64: case HC_FLOAT64: return 8;
    default: return 0;
  }
}

// Load element i of a typed buffer, converted to the accumulation type W
// (the reference's cast<W>(x): a C conversion, round-to-nearest for floats).
template <typename W>
__device__ __forceinline__ W load_as(const void* __restrict__ base, int64_t i, int dt) {
  switch (dt) {
    case HC_FLOAT32: return (W)((const float*)base)[i];
    case HC_FLOAT64: return (W)((const double*)base)[i];
    case HC_UINT8: return (W)((const uint8_t*)base)[i];
    case HC_INT8: return (W)((const int8_t*)base)[i];
    case HC_UINT16: return (W)((const uint16_t*)base)[i];
    case HC_INT16: return (W)((const int16_t*)base)[i];
    case HC_INT32: return (W)((const int32_t*)base)[i];
    case HC_UINT32: return (W)((const uint32_t*)base)[i];
    case HC_INT64: return (W)((const int64_t*)base)[i];
    case HC_UINT64: return (W)((const uint64_t*)base)[i];
    case HC_FLOAT16: return (W)__half2float(((const __half*)base)[i]);
    case HC_BOOL: return (W)(((const uint8_t*)base)[i] != 0);
    default: return (W)0;
  }
}

// Typed load: the fast path of the staging loops when the element type is
// already the accumulation type.
template <typename W>
__device__ __forceinline__ W load_w(const void* __restrict__ base, int64_t i) {
  return reinterpret_cast<const W*>(base)[i];
}

template <typename W>
__host__ __device__ __forceinline__ constexpr int w_dtype() {
  return sizeof(W) == 4 ? HC_FLOAT32 : HC_FLOAT64;
}

// Number of global loads each lane keeps in flight while staging a tile in
// LDS (loads are issued as a batch before their LDS stores).
constexpr int kStageBatch = 8;

template <typename W>
__device__ __forceinline__ bool finite(W v) {
  return v - v == (W)0;  // false for inf and NaN
}

// Stage `total` elements of a tile into LDS with kStageBatch global loads in
// flight per lane.  map(e, slot) returns the global element index of tile
// element e (or -1 for the constant-mode value) and sets its LDS slot (or -1
// to skip it).  TYPED: the buffer holds W already.  Returns whether this lane
// staged a non-finite value.
template <typename W, bool TYPED, class Map>
__device__ __forceinline__ bool stage_lds(W* __restrict__ lds, int total, int tid, int nthreads,
                                          const void* __restrict__ in, int dt, W cval,
                                          bool negate, Map map,
                                          const double* __restrict__ center = nullptr) {
  // center (device, optional): the value subtracted from every element
  // (after the negation, as c - x; the constant padding is not shifted)
  const bool ctr = center != nullptr;
  const W c = ctr ? (W)*center : (W)0;
  bool bad = false;
  for (int e0 = tid; e0 < total; e0 += nthreads * kStageBatch) {
    W v[kStageBatch];
    int slot[kStageBatch];
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u) {
      const int e = e0 + u * nthreads;
      slot[u] = -1;
      v[u] = cval;
      if (e < total) {
        const int64_t g = map(e, slot[u]);
        if (g >= 0) {
          const W x = TYPED ? load_w<W>(in, g) : load_as<W>(in, g, dt);
          v[u] = ctr ? (negate ? c - x : x - c) : (negate ? -x : x);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < kStageBatch; ++u)
      if (slot[u] >= 0) {
        lds[slot[u]] = v[u];
        bad |= !finite(v[u]);
      }
  }
  return bad;
}

template <typename W, class Map>
__device__ __forceinline__ bool stage_any(W* __restrict__ lds, int total, int tid, int nthreads,
                                          const void* __restrict__ in, int dt, W cval,
                                          bool negate, Map map,
                                          const double* __restrict__ center = nullptr) {
  if (dt == w_dtype<W>())
    return stage_lds<W, true>(lds, total, tid, nthreads, in, dt, cval, negate, map, center);
  return stage_lds<W, false>(lds, total, tid, nthreads, in, dt, cval, negate, map, center);
}

// Saturating truncation toward zero (float -> integer conversion of the GPU;
// NaN -> 0).
template <typename T, typename W>
__device__ __forceinline__ T sat_trunc(W v, W lo, W hi, T tlo, T thi) {
  if (!(v == v)) return (T)0;
  if (v <= lo) return tlo;
  if (v >= hi) return thi;
  return (T)v;
}

// cast<B>(A) of _vendored/_ndimage_filters_core.py:227-244: float -> unsigned
// is a >= 0 ? (B)a : -(B)(-a); other conversions are C casts (truncating,
// saturating on the GPU); float -> bool is a != 0.
template <typename W>
__device__ __forceinline__ void store_cast(void* __restrict__ base, int64_t i, int dt, W v) {
  switch (dt) {
    case HC_FLOAT32: ((float*)base)[i] = (float)v; return;
    case HC_FLOAT64: ((double*)base)[i] = (double)v; return;
    case HC_FLOAT16: ((__half*)base)[i] = __float2half_rn((float)v); return;
    case HC_BOOL: ((uint8_t*)base)[i] = (v != (W)0) ? 1 : 0; return;
    case HC_INT8:
      ((int8_t*)base)[i] = sat_trunc<int8_t, W>(v, (W)-128, (W)127, (int8_t)-128, (int8_t)127);
      return;
    case HC_INT16:
      ((int16_t*)base)[i] = sat_trunc<int16_t, W>(v, (W)-32768, (W)32767, (int16_t)-32768,
                                                  (int16_t)32767);
      return;
    case HC_INT32:
      ((int32_t*)base)[i] = sat_trunc<int32_t, W>(v, (W)-2147483648.0, (W)2147483647.0,
                                                  (int32_t)INT32_MIN, (int32_t)INT32_MAX);
      return;
    case HC_INT64:
      ((int64_t*)base)[i] = sat_trunc<int64_t, W>(v, (W)-9223372036854775808.0,
                                                  (W)9223372036854775807.0, INT64_MIN,
                                                  INT64_MAX);
      return;
    default: break;
  }
  // unsigned: a >= 0 ? (B)a : -(B)(-a), modular negation
  bool neg = !(v >= (W)0);
  W a = neg ? -v : v;
  switch (dt) {
    case HC_UINT8: {
      uint8_t m = sat_trunc<uint8_t, W>(a, (W)0, (W)255, 0, 255);
      ((uint8_t*)base)[i] = neg ? (uint8_t)(0u - m) : m;
      return;
    }
    case HC_UINT16: {
      uint16_t m = sat_trunc<uint16_t, W>(a, (W)0, (W)65535, 0, 65535);
      ((uint16_t*)base)[i] = neg ? (uint16_t)(0u - m) : m;
      return;
    }
    case HC_UINT32: {
      uint32_t m = sat_trunc<uint32_t, W>(a, (W)0, (W)4294967295.0, 0u, 0xffffffffu);
      ((uint32_t*)base)[i] = neg ? (uint32_t)(0u - m) : m;
      return;
    }
    case HC_UINT64: {
      uint64_t m = sat_trunc<uint64_t, W>(a, (W)0, (W)18446744073709551615.0, (uint64_t)0,
                                          ~(uint64_t)0);
      ((uint64_t*)base)[i] = neg ? (uint64_t)(0ull - m) : m;
      return;
    }
    default: return;
  }
}

// Round a value through dtype dt and back to W: the intermediate of a
// multi-pass filter is stored in the output dtype (_run_1d_filters).
template <typename W>
__device__ __forceinline__ W roundtrip(W v, int dt) {
  switch (dt) {
    case HC_FLOAT32: return (W)(float)v;
    case HC_FLOAT64: return (W)(double)v;
    default: {
      union { uint64_t u; double d; } buf;
      buf.u = 0;
      store_cast<W>(&buf, 0, dt, v);
      return load_as<W>(&buf, 0, dt);
    }
  }
}

// ---------------------------------------------------------- boundaries --
// Boundary extension of position p on an axis of length n,
// _vendored/_ndimage_util.py:186-308.  Returns -1 for constant mode outside
// the array.  In-range positions return unchanged.
__device__ __host__ __forceinline__ int64_t remap(int64_t p, int64_t n, int mode) {
  if (p >= 0 && p < n) return p;
  switch (mode) {
    case HC_MODE_REFLECT: {  // i<0 -> -1-i; i %= 2n; i = min(i, 2n-1-i)
      if (p < 0) p = -1 - p;
      p %= 2 * n;
      int64_t q = 2 * n - 1 - p;
      return p < q ? p : q;
    }
    case HC_MODE_MIRROR: {  // n==1 -> 0; i<0 -> -i; 1 + (i-1) % (2n-2); min(i, 2n-2-i)
      if (n == 1) return 0;
      if (p < 0) p = -p;
      p = 1 + (p - 1) % (2 * n - 2);
      int64_t q = 2 * n - 2 - p;
      return p < q ? p : q;
    }
    case HC_MODE_NEAREST:
      return p < 0 ? 0 : n - 1;
    case HC_MODE_WRAP: {
      p %= n;
      return p < 0 ? p + n : p;
    }
    default:
      return -1;
  }
}

// Work-groups are dispatched round-robin over the 8 XCDs (block b runs on XCD
// b % 8).  Remap block indices so that each XCD takes a contiguous range of
// tiles: neighbouring tiles, which share cache lines and halo rows, then sit
// behind the same L2.
__device__ __forceinline__ int64_t xcd_swizzle(int64_t b, int64_t nb) {
  const int64_t nx = 8;
  if (nb < 2 * nx) return b;
  const int64_t per = nb / nx;
  const int64_t full = per * nx;
  if (b >= full) return b;
  return (b % nx) * per + b / nx;
}

// ------------------------------------------------------------- weights --
// Kernel taps passed by value in the kernel-argument segment: the kernels
// index them with wave-uniform indices, which compile to scalar loads.
constexpr int kInlineTapBytes = 1024;

template <typename W>
struct TapPack {
  int K;       // padded tap count (multiple of 4 for tiled kernels)
  int anchor;  // tap k reads position i + k - anchor
  // bit g set: taps 4g .. 4g+3 are all nonzero (scalar test in the cores)
  uint32_t full[(kInlineTapBytes / sizeof(W) + 127) / 128];
  W w[kInlineTapBytes / sizeof(W)];
};

template <typename W>
__device__ __forceinline__ bool group_full(const TapPack<W>& tp, int t) {
  const int g = t >> 2;
  return (tp.full[g >> 5] >> (g & 31)) & 1u;
}

template <typename W>
constexpr int inline_cap() { return kInlineTapBytes / (int)sizeof(W); }

// Bit test that a float is nonzero, evaluated on the scalar unit when the
// value is wave-uniform (+0 and -0 are both zero).
__device__ __forceinline__ bool nz(float w) { return (__float_as_uint(w) << 1) != 0u; }
__device__ __forceinline__ bool nz(double w) {
  return ((uint64_t)__double_as_longlong(w) << 1) != 0ull;
}

__device__ __forceinline__ float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// Host-side helpers shared by the entry points.
struct Lines {
  int64_t outer;  // product of dims before axis
  int64_t N;      // axis length
  int64_t inner;  // product of dims after axis
};

bool is_c_contiguous(const hc_array* a);
Lines lines_of(const hc_array* a, int axis);
int64_t numel(const hc_array* a);
bool same_shape(const hc_array* a, const hc_array* b);

template <typename W>
int make_taps(const hc_taps* t, int pad_to, TapPack<W>* out);

// Eigenvalues of n x n symmetric matrices (n != 2, 3), k_eig_nd.hip.
int sym_eigvals_nd(const hc_array* S, int nS, const hc_array* eig, int sort, int abs_sort,
                   hipStream_t st);

}  // namespace hc
