// Pointwise and small-stencil kernels of the path: rgb2gray, img_as_float
// conversion, finite-difference gradient, the fused 2-D Sobel/Prewitt
// magnitude, mask erosion, running maximum and strided copy-with-cast.
//
// Reference (paths relative to python/cucim/src/cucim/skimage/):
//   color/colorconv.py:1037-1095, :160-180   rgb2gray
//   util/dtype.py:219-377                    _convert (int -> float)
//   _shared/_gradient.py                     gradient
//   filters/edges.py:75-85, 148-241          _mask_filter_result, _generic_edge_filter
#include <algorithm>
#include <cstring>

#include "hc_common.h"

namespace hc {

namespace {
int64_t grid_for(int64_t n) { return std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 256 * 64)); }

double imax_of(int dt) {
  switch (dt) {
    case HC_UINT8: return 255.0;
    case HC_INT8: return 127.0;
    case HC_UINT16: return 65535.0;
    case HC_INT16: return 32767.0;
    case HC_UINT32: return 4294967295.0;
    case HC_INT32: return 2147483647.0;
    case HC_UINT64: return 18446744073709551615.0;
    case HC_INT64: return 9223372036854775807.0;
    default: return 1.0;
  }
}
bool is_int(int dt) { return dt >= HC_INT8 && dt <= HC_UINT64; }
bool is_signed_int(int dt) {
  return dt == HC_INT8 || dt == HC_INT16 || dt == HC_INT32 || dt == HC_INT64;
}
}  // namespace


// one tap, skipped when the weight is zero (as the reference skips it)
template <typename F>
__device__ __forceinline__ F tap(F acc, F x, F w) { return nz(w) ? fma_(x, w, acc) : acc; }

// img_as_float-style conversion of one element to the computation type C:
// uint -> C(x) * C(1/imax); int -> max(that, -1); bool/float -> C(x).
template <typename C>
__device__ __forceinline__ C to_float_elem(const void* in, int64_t i, int dt, C scale, int sgn) {
  C v = load_as<C>(in, i, dt);
  if (scale != (C)0) {
    v = v * scale;
    if (sgn) v = v < (C)-1 ? (C)-1 : v;
  }
  return v;
}

// ------------------------------------------------------------- rgb2gray --
// gray = 0.2125 R + 0.7154 G + 0.0721 B with double literals, summed left to
// right in double without contraction, stored in the output float type.
template <typename C>
__global__ __launch_bounds__(256) void k_rgb2gray(const void* __restrict__ in, int dt,
                                                  void* __restrict__ out, int out_dt, int64_t n,
                                                  C scale, int sgn) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const C r = to_float_elem<C>(in, 3 * i, dt, scale, sgn);
    const C g = to_float_elem<C>(in, 3 * i + 1, dt, scale, sgn);
    const C b = to_float_elem<C>(in, 3 * i + 2, dt, scale, sgn);
    // computation-type values are first rounded to the float image type
    double rd, gd, bd;
    if (out_dt == HC_FLOAT32) {
      rd = (double)(float)r; gd = (double)(float)g; bd = (double)(float)b;
    } else {
      rd = (double)r; gd = (double)g; bd = (double)b;
    }
    double y = add_rn(add_rn(mul_rn(0.2125, rd), mul_rn(0.7154, gd)), mul_rn(0.0721, bd));
    if (out_dt == HC_FLOAT32) ((float*)out)[i] = (float)y;
    else ((double*)out)[i] = y;
  }
}

// uint8 RGB -> float32 gray, 1024 pixels per wave and step: the wave loads
// its 3072 interleaved bytes with three coalesced 16-byte loads per lane into
// its own LDS slice, then each lane converts 4 consecutive pixels (three
// aligned words) per round and writes them as one 16-byte store, so loads
// and stores both cover contiguous spans.  The per-pixel arithmetic is that
// of k_rgb2gray<float> (x * (float)(1/255) in float, then the double sum).
typedef unsigned int u32x4_t __attribute__((ext_vector_type(4)));
typedef float f32x4v_t __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_rgb2gray_u8(const u32x4_t* __restrict__ in,
                                                     f32x4v_t* __restrict__ out, int64_t nwork,
                                                     float scale) {
  __shared__ __align__(16) unsigned int S[4][768];  // 3072 bytes per wave
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  u32x4_t* Sv = reinterpret_cast<u32x4_t*>(S[wv]);
  for (int64_t w = (int64_t)blockIdx.x * 4 + wv; w < nwork; w += (int64_t)gridDim.x * 4) {
    const u32x4_t* src = in + 192 * w;  // 1024 pixels = 192 16-byte chunks
#pragma unroll
    for (int k = 0; k < 3; ++k) Sv[lane + 64 * k] = src[lane + 64 * k];
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int j = lane + 64 * k;  // group of pixels 4 j .. 4 j + 3
      const unsigned int wd[3] = {S[wv][3 * j], S[wv][3 * j + 1], S[wv][3 * j + 2]};
      float g[4];
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        float c[3];
#pragma unroll
        for (int ch = 0; ch < 3; ++ch) {
          const int byte = 3 * q + ch;
          c[ch] = (float)((wd[byte >> 2] >> (8 * (byte & 3))) & 255u) * scale;
        }
        const double y = add_rn(add_rn(mul_rn(0.2125, (double)c[0]),
                                       mul_rn(0.7154, (double)c[1])),
                                mul_rn(0.0721, (double)c[2]));
        g[q] = (float)y;
      }
      __builtin_nontemporal_store(f32x4v_t{g[0], g[1], g[2], g[3]}, out + 256 * w + j);
    }
    __builtin_amdgcn_wave_barrier();  // the slice is rewritten by the next step
  }
}

template <typename C>
__global__ __launch_bounds__(256) void k_to_float(const void* __restrict__ in, int dt,
                                                  void* __restrict__ out, int out_dt, int64_t n,
                                                  C scale, int sgn) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const C v = to_float_elem<C>(in, i, dt, scale, sgn);
    if (out_dt == HC_FLOAT32) ((float*)out)[i] = (float)v;
    else ((double*)out)[i] = (double)v;
  }
}

// ------------------------------------------------------------- gradient --
template <typename F>
__global__ __launch_bounds__(256) void k_gradient(const F* __restrict__ f, F* __restrict__ o,
                                                  int64_t n, int64_t N, int64_t inner) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int64_t p = (i / inner) % N;
    F v;
    if (p == 0) v = f[i + inner] - f[i];
    else if (p == N - 1) v = f[i] - f[i - inner];
    else v = (f[i + inner] - f[i - inner]) / (F)2.0;
    o[i] = v;
  }
}

// ------------------------------------------------------ edge magnitude --
// 2-D _generic_edge_filter with axis=None: for edge axis 0 convolve with the
// edge kernel along axis 0 then the smoothing kernel along axis 1, square;
// the same with the axes swapped; add; sqrt.  Every 1-D pass result is
// rounded to F (separate arrays upstream).  cval is not forwarded upstream,
// so constant mode pads with 0.  Weights are correlation weights (already
// reversed from the convolution weights).
constexpr int kEdgeTX = 64, kEdgeTY = 16;

template <typename F>
struct EdgeArgs {
  const void* in;
  int in_dt;
  F* out;
  int64_t H, Wd;
  int mode;
  F scale;
  int sgn;
  int nty, ntx;
  F e[3], s[3];  // correlation weights
};

template <typename F>
__global__ __launch_bounds__(256) void k_edge2d(const EdgeArgs<F> a) {
  __shared__ F T[kEdgeTY + 2][kEdgeTX + 3];
  __shared__ F EY[kEdgeTY][kEdgeTX + 3];  // edge along axis 0, cols -1..TX
  __shared__ F EX[kEdgeTY + 2][kEdgeTX + 1];  // edge along axis 1, rows -1..TY
  const int64_t tile = blockIdx.x;
  const int64_t ty = tile / a.ntx, tx = tile - ty * a.ntx;
  const int64_t y0 = ty * kEdgeTY, x0 = tx * kEdgeTX;
  const int tid = threadIdx.x;
  const bool cst = a.mode == HC_MODE_CONSTANT;
  // image tile with one halo row/column on each side
  for (int e = tid; e < (kEdgeTY + 2) * (kEdgeTX + 2); e += 256) {
    const int r = e / (kEdgeTX + 2), c = e - r * (kEdgeTX + 2);
    const int64_t y = remap(y0 + r - 1, a.H, a.mode), x = remap(x0 + c - 1, a.Wd, a.mode);
    T[r][c] = (y < 0 || x < 0) ? (F)0 : to_float_elem<F>(a.in, y * a.Wd + x, a.in_dt, a.scale, a.sgn);
  }
  __syncthreads();
  // first passes (edge kernel)
  for (int e = tid; e < kEdgeTY * (kEdgeTX + 2); e += 256) {
    const int r = e / (kEdgeTX + 2), c = e - r * (kEdgeTX + 2);
    F v = (F)0;
    v = tap(v, T[r][c], a.e[0]);
    v = tap(v, T[r + 1][c], a.e[1]);
    v = tap(v, T[r + 2][c], a.e[2]);
    const int64_t x = x0 + c - 1;
    if (cst && (x < 0 || x >= a.Wd)) v = (F)0;
    EY[r][c] = v;
  }
  for (int e = tid; e < (kEdgeTY + 2) * kEdgeTX; e += 256) {
    const int r = e / kEdgeTX, c = e - r * kEdgeTX;
    F v = (F)0;
    v = tap(v, T[r][c], a.e[0]);
    v = tap(v, T[r][c + 1], a.e[1]);
    v = tap(v, T[r][c + 2], a.e[2]);
    const int64_t y = y0 + r - 1;
    if (cst && (y < 0 || y >= a.H)) v = (F)0;
    EX[r][c] = v;
  }
  __syncthreads();
  for (int e = tid; e < kEdgeTY * kEdgeTX; e += 256) {
    const int r = e / kEdgeTX, c = e - r * kEdgeTX;
    const int64_t y = y0 + r, x = x0 + c;
    if (y >= a.H || x >= a.Wd) continue;
    F A = (F)0, B = (F)0;
    A = tap(A, EY[r][c], a.s[0]);
    A = tap(A, EY[r][c + 1], a.s[1]);
    A = tap(A, EY[r][c + 2], a.s[2]);
    B = tap(B, EX[r][c], a.s[0]);
    B = tap(B, EX[r + 1][c], a.s[1]);
    B = tap(B, EX[r + 2][c], a.s[2]);
    a.out[y * a.Wd + x] = sqrt(add_rn(mul_rn(A, A), mul_rn(B, B)));
  }
}

// The same magnitude for a float input of the output's type, as a window
// sliding down columns.  A wave owns a strip of 64 * VN columns (VN = 16 /
// sizeof(F) consecutive columns per lane, one 16-byte load per row) and a run
// of `rows` output rows; it reads rows y0 - 1 .. y0 + rows.  The columns left
// and right of a lane's VN come from the neighbouring lanes (cross-lane
// permute); the strip's two outer columns are loaded by lanes 0 and 63.  Rows
// are fetched kEdgeU at a time, the next group's loads issued before the
// current group is used.  Each output takes the taps in the same order as
// k_edge2d, so the two kernels agree bit for bit.  Strips whose columns and
// neighbours lie inside the image (and whose rows are 16-byte aligned) use
// vector loads and streaming vector stores; the first and last strip, or
// every strip when the width is not a multiple of VN, map each column
// through the boundary mode and store per element.
constexpr int kEdgeU = 4;

template <typename F>
struct EdgeVecT;
template <>
struct EdgeVecT<float> {
  typedef float T __attribute__((ext_vector_type(4)));
  static constexpr int N = 4;
};
template <>
struct EdgeVecT<double> {
  typedef double T __attribute__((ext_vector_type(2)));
  static constexpr int N = 2;
};

template <typename F>
struct EdgeColArgs {
  const F* in;
  F* out;
  int64_t H, Wd;
  int mode;
  int nstrip, rows;
  int64_t nwork;  // waves: nstrip * ceil(H / rows)
  int vec;        // width a multiple of VN and both arrays 16-byte aligned
  F e[3], s[3];
};

// Three taps in order.  Z >= 0: bit k set means weight k is zero, known at
// compile time (the edge weights [w, 0, -w] of Sobel, Scharr and Prewitt with
// nonzero smoothing weights); Z < 0: zero weights are tested at run time.
template <int Z, typename F>
__device__ __forceinline__ F tap3(F a, F b, F c, const F* w) {
  F v = (F)0;
  if (Z < 0) {
    v = tap(v, a, w[0]);
    v = tap(v, b, w[1]);
    return tap(v, c, w[2]);
  }
  if (!(Z & 1)) v = fma_(a, w[0], v);
  if (!(Z & 2)) v = fma_(b, w[1], v);
  if (!(Z & 4)) v = fma_(c, w[2], v);
  return v;
}

// remap() for positions at most one step outside the axis (p in [-1, n]),
// as selects, without the general path's divisions.
__device__ __forceinline__ int64_t remap_near(int64_t p, int64_t n, int mode) {
  const bool lo = p < 0, hi = p >= n;
  const bool mir = mode == HC_MODE_MIRROR, wrp = mode == HC_MODE_WRAP;
  const int64_t at_lo = wrp ? n - 1 : (mir && n > 1 ? 1 : 0);
  const int64_t at_hi = wrp ? 0 : (mir ? (n > 1 ? n - 2 : 0) : n - 1);
  const int64_t r = lo ? at_lo : (hi ? at_hi : p);
  return mode == HC_MODE_CONSTANT && (lo || hi) ? -1 : r;
}

// One row's loads, issued without branches or lane masks so that several
// rows stay in flight: the VN columns of the lane and, in raw[VN], the
// strip's outer column (lanes 0..31 the left one, 32..63 the right one, so
// the load touches two cache lines).  Rows that read as zero (constant mode
// outside the image, rows past y = H that feed no stored output) load row 0
// and are zeroed where they are used.
template <typename F, bool VEC>
__device__ __forceinline__ bool edge_fetch(const EdgeColArgs<F>& a, int64_t y, int64_t x0,
                                           int64_t xl, int lane, const int* cidx, F* raw) {
  constexpr int VN = EdgeVecT<F>::N;
  const int64_t yy = remap_near(y, a.H, a.mode);
  const bool zero = y > a.H || yy < 0;
  const F* row = a.in + (zero ? 0 : yy) * a.Wd;
  if (VEC) {
    const typename EdgeVecT<F>::T v = *reinterpret_cast<const typename EdgeVecT<F>::T*>(row + xl);
#pragma unroll
    for (int k = 0; k < VN; ++k) raw[k] = v[k];
    raw[VN] = row[lane < 32 ? x0 - 1 : x0 + 64 * VN];
  } else {
#pragma unroll
    for (int k = 0; k <= VN; ++k) raw[k] = row[cidx[k] < 0 ? 0 : cidx[k]];
  }
  return zero;
}

template <typename F, bool VEC, int ZE, int ZS>
__device__ __forceinline__ void edge_strip(const EdgeColArgs<F>& a, int64_t y0, int64_t x0,
                                           int lane) {
  constexpr int VN = EdgeVecT<F>::N;
  constexpr int U = VEC ? kEdgeU : 2;  // rows per fetch group; divides rows
  const int64_t xl = x0 + (int64_t)lane * VN;
  const bool cst = a.mode == HC_MODE_CONSTANT;
  // Border strips: column indices mapped once; -1 reads as zero (constant
  // mode outside the image, or a column past x = Wd that no stored output
  // reads).  cidx[VN] is the strip's outer column for lanes 0 and 63.
  int cidx[VN + 1];
#pragma unroll
  for (int k = 0; k <= VN; ++k) cidx[k] = 0;
  if (!VEC) {
#pragma unroll
    for (int k = 0; k < VN; ++k) {
      const int64_t x = xl + k;
      cidx[k] = x <= a.Wd ? (int)remap_near(x, a.Wd, a.mode) : -1;
    }
    const int64_t xe = lane == 0 ? x0 - 1 : x0 + 64 * VN;
    cidx[VN] = (lane == 0 || lane == 63) && xe <= a.Wd ? (int)remap_near(xe, a.Wd, a.mode) : -1;
  }
  F tm[VN + 2], t0[VN + 2], exm[VN], ex0[VN];
#pragma unroll
  for (int k = 0; k < VN + 2; ++k) tm[k] = t0[k] = (F)0;
#pragma unroll
  for (int k = 0; k < VN; ++k) exm[k] = ex0[k] = (F)0;

  // row y's values are in (raw, zero); output row y - 1 when `out`
  auto step = [&](const F* raw, bool zero, int64_t y, bool out) {
    F c[VN + 1];
#pragma unroll
    for (int k = 0; k <= VN; ++k) {
      bool z = zero;
      if (!VEC) z = z || cidx[k] < 0;
      c[k] = z ? (F)0 : raw[k];
    }
    // columns xl - 1 .. xl + VN of row y
    F tp[VN + 2];
#pragma unroll
    for (int k = 0; k < VN; ++k) tp[k + 1] = c[k];
    const F lft = __shfl_up(c[VN - 1], 1);
    const F rgt = __shfl_down(c[0], 1);
    tp[0] = lane == 0 ? c[VN] : lft;
    tp[VN + 1] = lane == 63 ? c[VN] : rgt;
    // edge along axis 1 of row y (zero outside the image in constant mode)
    F exp_[VN];
    const bool yout = cst && (y < 0 || y >= a.H);
#pragma unroll
    for (int k = 0; k < VN; ++k)
      exp_[k] = yout ? (F)0 : tap3<ZE>(tp[k], tp[k + 1], tp[k + 2], a.e);
    if (out) {
      const int64_t yo = y - 1;
      F ey[VN + 2];
#pragma unroll
      for (int j = 0; j < VN + 2; ++j) {
        F v = tap3<ZE>(tm[j], t0[j], tp[j], a.e);
        if (!VEC && cst) {
          const int64_t x = xl - 1 + j;
          if (x < 0 || x >= a.Wd) v = (F)0;
        }
        ey[j] = v;
      }
      F res[VN];
#pragma unroll
      for (int k = 0; k < VN; ++k) {
        const F A = tap3<ZS>(ey[k], ey[k + 1], ey[k + 2], a.s);
        const F B = tap3<ZS>(exm[k], ex0[k], exp_[k], a.s);
        res[k] = sqrt(add_rn(mul_rn(A, A), mul_rn(B, B)));
      }
      if (yo < a.H) {
        F* orow = a.out + yo * a.Wd;
        if (VEC) {
          typename EdgeVecT<F>::T o;
#pragma unroll
          for (int k = 0; k < VN; ++k) o[k] = res[k];
          __builtin_nontemporal_store(o, reinterpret_cast<typename EdgeVecT<F>::T*>(orow + xl));
        } else {
#pragma unroll
          for (int k = 0; k < VN; ++k)
            if (xl + k < a.Wd) orow[xl + k] = res[k];
        }
      }
    }
#pragma unroll
    for (int k = 0; k < VN + 2; ++k) {
      tm[k] = t0[k];
      t0[k] = tp[k];
    }
#pragma unroll
    for (int k = 0; k < VN; ++k) {
      exm[k] = ex0[k];
      ex0[k] = exp_[k];
    }
  };

  // Rows y0 - 1 and y0 start the window; groups of U rows y0 + 1 .. y0 + rows
  // follow, each group's loads issued before the previous group is used.  Two
  // register sets alternate (no copies between them, which would make the
  // compiler wait for loads in flight), and the fetches are unconditional
  // inside the loop (a conditional fetch makes it wait at the join).
  F pre[2][VN + 1], ga[U][VN + 1], gb[U][VN + 1];
  bool zpre[2], za[U], zb[U];
  auto fetch = [&](int g, F (*buf)[VN + 1], bool* z) {
#pragma unroll
    for (int i = 0; i < U; ++i)
      z[i] = edge_fetch<F, VEC>(a, y0 + 1 + g * U + i, x0, xl, lane, cidx, buf[i]);
  };
  auto run = [&](int g, F (*buf)[VN + 1], const bool* z) {
#pragma unroll
    for (int i = 0; i < U; ++i) step(buf[i], z[i], y0 + 1 + g * U + i, true);
  };
#pragma unroll
  for (int i = 0; i < 2; ++i) zpre[i] = edge_fetch<F, VEC>(a, y0 - 1 + i, x0, xl, lane, cidx, pre[i]);
  fetch(0, ga, za);
  step(pre[0], zpre[0], y0 - 1, false);
  step(pre[1], zpre[1], y0, false);
  const int nb = a.rows / U;  // at least 2
  int g = 0;
  for (; g + 2 < nb; g += 2) {
    fetch(g + 1, gb, zb);
    run(g, ga, za);
    fetch(g + 2, ga, za);
    run(g + 1, gb, zb);
  }
  if (g + 1 < nb) {
    fetch(g + 1, gb, zb);
    run(g, ga, za);
    run(g + 1, gb, zb);
  } else {
    run(g, ga, za);
  }
}

template <typename F, int ZE, int ZS>
__global__ __launch_bounds__(256) void k_edge2d_cols(const EdgeColArgs<F> a) {
  constexpr int VN = EdgeVecT<F>::N;
  const int64_t g = (int64_t)blockIdx.x * 4 + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (g >= a.nwork) return;  // whole waves; the kernel has no barrier
  const int lane = threadIdx.x & 63;
  const int64_t chunk = g / a.nstrip, strip = g - chunk * a.nstrip;
  const int64_t x0 = strip * 64 * VN, y0 = chunk * a.rows;
  if (a.vec && x0 >= 1 && x0 + 64 * VN < a.Wd)
    edge_strip<F, true, ZE, ZS>(a, y0, x0, lane);
  else
    edge_strip<F, false, ZE, ZS>(a, y0, x0, lane);
}

// ------------------------------------------------------ mask erosion --
template <typename F>
__global__ __launch_bounds__(256) void k_mask_eroded(F* __restrict__ out,
                                                     const uint8_t* __restrict__ mask, int64_t n,
                                                     int ndim, int64_t sh0, int64_t sh1,
                                                     int64_t sh2, int64_t sh3) {
  const int64_t sh[4] = {sh0, sh1, sh2, sh3};
  int nb = 1;
  for (int d = 0; d < ndim; ++d) nb *= 3;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    int64_t pos[4] = {0, 0, 0, 0};
    int64_t rem = i;
    for (int d = ndim - 1; d >= 0; --d) {
      pos[d] = rem % sh[d];
      rem /= sh[d];
    }
    bool keep = true;
    for (int k = 0; k < nb && keep; ++k) {
      int kk = k;
      int64_t off = 0;
      for (int d = ndim - 1; d >= 0; --d) {
        const int64_t q = pos[d] + (kk % 3) - 1;
        kk /= 3;
        if (q < 0 || q >= sh[d]) { keep = false; break; }
        int64_t stride = 1;
        for (int e = d + 1; e < ndim; ++e) stride *= sh[e];
        off += q * stride;
      }
      if (keep && mask[off] == 0) keep = false;
    }
    out[i] = out[i] * (keep ? (F)1 : (F)0);
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_maximum(F* __restrict__ a, const F* __restrict__ b, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const F x = a[i], y = b[i];
    a[i] = (x != x) ? x : ((y != y) ? y : (x > y ? x : y));
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_square_acc(F* __restrict__ acc, const F* __restrict__ v,
                                                    int64_t n, int first) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const F s = mul_rn(v[i], v[i]);
    acc[i] = first ? s : add_rn(acc[i], s);
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_binary(const F* a, const F* b, F* out, int64_t n,
                                                int op) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const F x = a[i], y = b[i];
    out[i] = op == HC_BIN_MUL ? mul_rn(x, y) : add_rn(x, op == HC_BIN_ADD ? y : -y);
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_div_scalar(F* __restrict__ a, int64_t n, F d) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    a[i] = a[i] / d;
}

// Sparse correlation (filters/_sparse.py:55-118): out[o] = sum over the
// kernel's nonzero taps, in order, of w_k * P[o + idx_k], where P is the image
// padded by pad_d on both sides of each axis (index remapping, constant mode
// pads with 0).  Each product and each add is rounded once in F, as the
// reference's `out += val * view` steps are; w = 1 and w = -1 take the view
// (or its negation) without a multiply.  One launch handles a chunk of taps;
// later chunks continue the sum from out.
constexpr int kSparseTaps = 64;

struct SparseArgs {
  const void* in;
  int in_dt;
  void* out;
  int ndim;
  int64_t n;                       // output elements
  int64_t oshape[HC_MAX_NDIM];     // output shape
  int64_t ishape[HC_MAX_NDIM];     // input shape
  int64_t istride[HC_MAX_NDIM];    // input element strides
  int pad[HC_MAX_NDIM];            // padding before each axis
  int mode;
  int ntaps, first;                // first != 0: out starts at this chunk's tap 0
  int16_t idx[kSparseTaps][HC_MAX_NDIM];
  double w[kSparseTaps];
};

template <typename F>
__device__ __forceinline__ F sparse_term(double w, F v) {
  if (w == 1.0) return v;
  if (w == -1.0) return -v;
  return mul_rn((F)w, v);
}

template <typename F>
__global__ __launch_bounds__(256) void k_correlate_sparse(const SparseArgs a) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    int64_t o[HC_MAX_NDIM];
    int64_t r = i;
    for (int d = a.ndim - 1; d >= 0; --d) {
      o[d] = r % a.oshape[d];
      r /= a.oshape[d];
    }
    F s = a.first ? (F)0 : reinterpret_cast<F*>(a.out)[i];
    for (int k = 0; k < a.ntaps; ++k) {
      int64_t off = 0;
      bool inside = true;
      for (int d = 0; d < a.ndim; ++d) {
        const int64_t q = remap(o[d] + a.idx[k][d] - a.pad[d], a.ishape[d], a.mode);
        inside &= q >= 0;
        off += (q >= 0 ? q : 0) * a.istride[d];
      }
      const F v = inside ? load_as<F>(a.in, off, a.in_dt) : (F)0;
      const F t = sparse_term<F>(a.w[k], v);
      s = (a.first && k == 0) ? t : add_rn(s, t);
    }
    reinterpret_cast<F*>(a.out)[i] = s;
  }
}

// Integer element ops wrap modulo 2^bits like the in-place integer ufuncs:
// U is the unsigned type of the element width, W an unsigned type at least
// 32 bits wide (no promotion to signed int, so no overflow).
template <typename U, typename W>
__device__ __forceinline__ U wrap_op(U x, U y, int op) {
  const W a = (W)x, b = (W)y;
  return (U)(op == HC_BIN_MUL ? a * b : (op == HC_BIN_ADD ? a + b : a - b));
}

template <typename U, typename W>
__global__ __launch_bounds__(256) void k_binary_int(const U* a, const U* b, U* out, int64_t n,
                                                    int op) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    out[i] = wrap_op<U, W>(a[i], b[i], op);
}

template <typename U, typename W>
__global__ __launch_bounds__(256) void k_square_acc_int(U* __restrict__ acc,
                                                        const U* __restrict__ v, int64_t n,
                                                        int first) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const U s = wrap_op<U, W>(v[i], v[i], HC_BIN_MUL);
    acc[i] = first ? s : wrap_op<U, W>(acc[i], s, HC_BIN_ADD);
  }
}

template <typename F>
__global__ __launch_bounds__(256) void k_sqrt(F* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    a[i] = sqrt(a[i]);
}

// sqrt of an integer array in place with an unsafe cast back (sqrt(out, out,
// casting='unsafe')): computed in the float type the ufunc picks for the
// input (F: float for 8/16-bit, double for 32/64-bit; for 8-bit the ufunc's
// float16 result truncates to the same integer as float's, since sqrt(k^2-1)
// stays more than 1/32 below k), truncated toward zero; a negative input
// gives NaN, stored as 0.
template <typename T, typename F>
__global__ __launch_bounds__(256) void k_sqrt_int(T* __restrict__ a, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const F r = sqrt((F)a[i]);
    a[i] = (r != r) ? (T)0 : (T)r;
  }
}

struct CopyArgs {
  const void* in;
  void* out;
  int in_dt, out_dt;
  int ndim;
  int64_t n;
  int64_t shape[HC_MAX_NDIM];
  int64_t istride[HC_MAX_NDIM];
  int64_t ostride[HC_MAX_NDIM];
};

// kind: 0 = same dtype (raw element copy), 1 = integer/bool -> integer/bool
// (C integer conversion, modular), 2 = through the float type W.
__device__ __forceinline__ uint64_t load_bits(const void* p, int64_t i, int sz) {
  switch (sz) {
    case 1: return ((const uint8_t*)p)[i];
    case 2: return ((const uint16_t*)p)[i];
    case 4: return ((const uint32_t*)p)[i];
    default: return ((const uint64_t*)p)[i];
  }
}
__device__ __forceinline__ void store_bits(void* p, int64_t i, int sz, uint64_t v) {
  switch (sz) {
    case 1: ((uint8_t*)p)[i] = (uint8_t)v; return;
    case 2: ((uint16_t*)p)[i] = (uint16_t)v; return;
    case 4: ((uint32_t*)p)[i] = (uint32_t)v; return;
    default: ((uint64_t*)p)[i] = v; return;
  }
}
__device__ __forceinline__ int64_t load_int(const void* p, int64_t i, int dt) {
  switch (dt) {
    case HC_BOOL: return ((const uint8_t*)p)[i] != 0;
    case HC_INT8: return ((const int8_t*)p)[i];
    case HC_UINT8: return ((const uint8_t*)p)[i];
    case HC_INT16: return ((const int16_t*)p)[i];
    case HC_UINT16: return ((const uint16_t*)p)[i];
    case HC_INT32: return ((const int32_t*)p)[i];
    case HC_UINT32: return ((const uint32_t*)p)[i];
    default: return ((const int64_t*)p)[i];  // int64 / uint64 bit pattern
  }
}

template <typename W>
__global__ __launch_bounds__(256) void k_copy_cast(const CopyArgs a, int kind, int isz, int osz) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    int64_t rem = i, io = 0, oo = 0;
#pragma unroll
    for (int d = HC_MAX_NDIM - 1; d >= 0; --d) {
      if (d < a.ndim) {
        const int64_t p = rem % a.shape[d];
        rem /= a.shape[d];
        io += p * a.istride[d];
        oo += p * a.ostride[d];
      }
    }
    if (kind == 0) {
      store_bits(a.out, oo, osz, load_bits(a.in, io, isz));
    } else if (kind == 1) {
      const int64_t v = load_int(a.in, io, a.in_dt);
      if (a.out_dt == HC_BOOL) store_bits(a.out, oo, 1, v != 0);
      else store_bits(a.out, oo, osz, (uint64_t)v);
    } else {
      store_cast<W>(a.out, oo, a.out_dt, load_as<W>(a.in, io, a.in_dt));
    }
  }
}

// Contiguous same-dtype copy, 16 bytes per lane per access, 4 accesses in
// flight per lane; n16 = byte count / 16.
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

__global__ __launch_bounds__(256) void k_copy_vec(const u32x4* __restrict__ in,
                                                  u32x4* __restrict__ out, int64_t n16) {
  const int64_t stride = (int64_t)gridDim.x * 1024;
  for (int64_t base = (int64_t)blockIdx.x * 1024 + threadIdx.x; base < n16; base += stride) {
    u32x4 v[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + 256 * u;
      if (i < n16) v[u] = __builtin_nontemporal_load(in + i);
    }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t i = base + 256 * u;
      if (i < n16) __builtin_nontemporal_store(v[u], out + i);
    }
  }
}

// Rows per wave: 8 unless the HC_TUNE_EDGE_ROWS knob says otherwise.  Short
// runs measure fastest at 2048^2 and 8192^2 alike (the more waves, the more
// rows in flight; the halo rows are mostly served by L2).
template <typename F>
int launch_edge_cols(const hc_array* in, const hc_array* out, const double* edge3,
                     const double* smooth3, int mode, hipStream_t st) {
  constexpr int VN = EdgeVecT<F>::N;
  EdgeColArgs<F> a;
  std::memset(&a, 0, sizeof(a));
  a.in = (const F*)in->data;
  a.out = (F*)out->data;
  a.H = in->shape[0];
  a.Wd = in->shape[1];
  a.mode = mode;
  const int64_t nstrip = (a.Wd + 64 * VN - 1) / (64 * VN);
  int rows = tune_get(HC_TUNE_EDGE_ROWS);
  if (rows == 0) rows = 8;
  const int64_t nwork = nstrip * ((a.H + rows - 1) / rows);
  if (a.Wd >= INT32_MAX || (nwork + 3) / 4 > INT32_MAX)
    return fail(HC_ERR_VALUE, "image too large for the edge kernel's grid");
  a.nstrip = (int)nstrip;
  a.rows = rows;
  a.nwork = nwork;
  a.vec = (a.Wd % VN == 0) && ((uintptr_t)a.in % 16 == 0) && ((uintptr_t)a.out % 16 == 0);
  for (int k = 0; k < 3; ++k) {
    a.e[k] = (F)edge3[k];
    a.s[k] = (F)smooth3[k];
  }
  int ze = 0, zs = 0;
  for (int k = 0; k < 3; ++k) {
    ze |= a.e[k] == (F)0 ? 1 << k : 0;
    zs |= a.s[k] == (F)0 ? 1 << k : 0;
  }
  const dim3 grid((unsigned)((nwork + 3) / 4));
  if (ze == 2 && zs == 0)
    HC_LAUNCH("k_edge2d_cols", st, (k_edge2d_cols<F, 2, 0>), grid, dim3(256), 0, a);
  else
    HC_LAUNCH("k_edge2d_cols", st, (k_edge2d_cols<F, -1, -1>), grid, dim3(256), 0, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_rgb2gray(const hc_array* in, const hc_array* out, void* stream) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null array");
  if (in->ndim < 1 || in->shape[in->ndim - 1] != 3)
    return fail(HC_ERR_VALUE, "the input array must have size 3 along `channel_axis`");
  if (!is_c_contiguous(in) || !is_c_contiguous(out))
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  if (out->ndim != in->ndim - 1) return fail(HC_ERR_VALUE, "output shape not correct");
  for (int d = 0; d < out->ndim; ++d)
    if (out->shape[d] != in->shape[d]) return fail(HC_ERR_VALUE, "output shape not correct");
  const int64_t n = numel(out);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  const int dt = in->dtype;
  const int sgn = is_signed_int(dt);
  // computation type: first of (out, float32, float64) at least as wide as the input
  const int isz = dtype_size(dt);
  const bool c64 = is_int(dt) ? (out->dtype == HC_FLOAT64 || isz > 4) : (dt == HC_FLOAT64 || out->dtype == HC_FLOAT64);
  if (c64) {
    const double sc = is_int(dt) ? 1.0 / imax_of(dt) : 0.0;
    HC_LAUNCH("k_rgb2gray", st, k_rgb2gray<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, in->data, dt, out->data, out->dtype, n, sc, sgn);
  } else {
    const float sc = is_int(dt) ? (float)(1.0 / imax_of(dt)) : 0.0f;
    int64_t done = 0;
    if (dt == HC_UINT8 && out->dtype == HC_FLOAT32 &&
        (((uintptr_t)in->data | (uintptr_t)out->data) & 15) == 0 && n >= 1024) {
      // runs of 1024 pixels, 16-byte aligned in both arrays
      const int64_t nw = n / 1024;
      const int64_t blocks = std::max<int64_t>(1, std::min<int64_t>((nw + 3) / 4, 256 * 16));
      HC_LAUNCH("k_rgb2gray_u8", st, k_rgb2gray_u8, dim3((unsigned)blocks), dim3(256), 0,
                (const u32x4_t*)in->data, (f32x4v_t*)out->data, nw, sc);
      done = nw * 1024;
    }
    if (done < n)
      HC_LAUNCH("k_rgb2gray", st, k_rgb2gray<float>, dim3((unsigned)grid_for(n - done)), dim3(256),
                0, (const uint8_t*)in->data + 3 * done * dtype_size(dt), dt,
                (float*)out->data + done, out->dtype, n - done, sc, sgn);
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_to_float(const hc_array* in, const hc_array* out, void* stream) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(in, out) || !is_c_contiguous(in) || !is_c_contiguous(out))
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  const int64_t n = numel(in);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  const int dt = in->dtype;
  const int sgn = is_signed_int(dt);
  const int isz = dtype_size(dt);
  const bool c64 = is_int(dt) ? (out->dtype == HC_FLOAT64 || isz > 4) : (out->dtype == HC_FLOAT64 || dt == HC_FLOAT64);
  if (c64) {
    const double sc = is_int(dt) ? 1.0 / imax_of(dt) : 0.0;
    HC_LAUNCH("k_to_float", st, k_to_float<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, in->data, dt, out->data, out->dtype, n, sc, sgn);
  } else {
    const float sc = is_int(dt) ? (float)(1.0 / imax_of(dt)) : 0.0f;
    HC_LAUNCH("k_to_float", st, k_to_float<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, in->data, dt, out->data, out->dtype, n, sc, sgn);
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_gradient(const hc_array* in, const hc_array* out, int axis, void* stream) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(in, out) || !is_c_contiguous(in) || !is_c_contiguous(out) ||
      in->dtype != out->dtype)
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes and dtypes");
  if (in->dtype != HC_FLOAT32 && in->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "gradient needs float32/float64");
  if (axis < 0) axis += in->ndim;
  if (axis < 0 || axis >= in->ndim) return fail(HC_ERR_VALUE, "invalid axis");
  if (in->shape[axis] < 2)
    return fail(HC_ERR_VALUE,
                "Shape of array too small to calculate a numerical gradient, at least 2 "
                "elements are required.");
  const Lines ln = lines_of(in, axis);
  const int64_t n = numel(in);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (in->dtype == HC_FLOAT64)
    HC_LAUNCH("k_gradient", st, k_gradient<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (const double*)in->data, (double*)out->data, n, ln.N, ln.inner);
  else
    HC_LAUNCH("k_gradient", st, k_gradient<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (const float*)in->data, (float*)out->data, n, ln.N, ln.inner);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_edge_magnitude(const hc_array* in, const hc_array* out, const double* edge3,
                      const double* smooth3, int axis_mask, int mode, const hc_array* work,
                      void* stream) {
  (void)work;
  if (!in || !out || !edge3 || !smooth3) return fail(HC_ERR_VALUE, "null argument");
  if (in->ndim != 2 || axis_mask != 3)
    return fail(HC_ERR_UNSUPPORTED, "fused edge magnitude handles 2-D inputs over both axes");
  if (!same_shape(in, out) || !is_c_contiguous(in) || !is_c_contiguous(out))
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  if (mode < HC_MODE_REFLECT || mode > HC_MODE_CONSTANT)
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  const int64_t H = in->shape[0], Wd = in->shape[1];
  if (H == 0 || Wd == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  const int dt = in->dtype;
  if (dt == out->dtype) {
    if (dt == HC_FLOAT32)
      return launch_edge_cols<float>(in, out, edge3, smooth3, mode, st);
    return launch_edge_cols<double>(in, out, edge3, smooth3, mode, st);
  }
  const int ntx = (int)((Wd + kEdgeTX - 1) / kEdgeTX), nty = (int)((H + kEdgeTY - 1) / kEdgeTY);
  const int64_t ntiles = (int64_t)ntx * nty;
  if (out->dtype == HC_FLOAT64) {
    EdgeArgs<double> a;
    std::memset(&a, 0, sizeof(a));
    a.in = in->data; a.in_dt = dt; a.out = (double*)out->data; a.H = H; a.Wd = Wd;
    a.mode = mode; a.scale = is_int(dt) ? 1.0 / imax_of(dt) : 0.0; a.sgn = is_signed_int(dt);
    a.ntx = ntx; a.nty = nty;
    for (int k = 0; k < 3; ++k) { a.e[k] = edge3[k]; a.s[k] = smooth3[k]; }
    HC_LAUNCH("k_edge2d", st, k_edge2d<double>, dim3((unsigned)ntiles), dim3(256), 0, a);
  } else {
    EdgeArgs<float> a;
    std::memset(&a, 0, sizeof(a));
    a.in = in->data; a.in_dt = dt; a.out = (float*)out->data; a.H = H; a.Wd = Wd;
    a.mode = mode; a.scale = is_int(dt) ? (float)(1.0 / imax_of(dt)) : 0.0f;
    a.sgn = is_signed_int(dt);
    a.ntx = ntx; a.nty = nty;
    for (int k = 0; k < 3; ++k) { a.e[k] = (float)edge3[k]; a.s[k] = (float)smooth3[k]; }
    HC_LAUNCH("k_edge2d", st, k_edge2d<float>, dim3((unsigned)ntiles), dim3(256), 0, a);
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_mask_eroded(const hc_array* out, const hc_array* mask, void* stream) {
  if (!out || !mask) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(out, mask) || !is_c_contiguous(out) || !is_c_contiguous(mask))
    return fail(HC_ERR_VALUE, "mask must be C-contiguous with the output's shape");
  if (mask->dtype != HC_BOOL && mask->dtype != HC_UINT8)
    return fail(HC_ERR_VALUE, "mask must be bool or uint8");
  if (out->ndim > 4) return fail(HC_ERR_UNSUPPORTED, "mask erosion supports up to 4-D");
  const int64_t n = numel(out);
  if (n == 0) return HC_OK;
  int64_t sh[4] = {1, 1, 1, 1};
  for (int d = 0; d < out->ndim; ++d) sh[d] = out->shape[d];
  hipStream_t st = (hipStream_t)stream;
  if (out->dtype == HC_FLOAT64)
    HC_LAUNCH("k_mask_eroded", st, k_mask_eroded<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (double*)out->data, (const uint8_t*)mask->data, n, out->ndim, sh[0], sh[1], sh[2], sh[3]);
  else if (out->dtype == HC_FLOAT32)
    HC_LAUNCH("k_mask_eroded", st, k_mask_eroded<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (float*)out->data, (const uint8_t*)mask->data, n, out->ndim, sh[0], sh[1], sh[2], sh[3]);
  else
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_maximum_inplace(const hc_array* a, const hc_array* b, void* stream) {
  if (!a || !b) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(a, b) || !is_c_contiguous(a) || !is_c_contiguous(b) || a->dtype != b->dtype)
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes and dtypes");
  const int64_t n = numel(a);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (a->dtype == HC_FLOAT64)
    HC_LAUNCH("k_maximum", st, k_maximum<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (double*)a->data, (const double*)b->data, n);
  else if (a->dtype == HC_FLOAT32)
    HC_LAUNCH("k_maximum", st, k_maximum<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (float*)a->data, (const float*)b->data, n);
  else
    return fail(HC_ERR_VALUE, "maximum needs float32/float64");
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_square_acc(const hc_array* acc, const hc_array* v, int first, void* stream) {
  if (!acc || !v) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(acc, v) || !is_c_contiguous(acc) || !is_c_contiguous(v) || acc->dtype != v->dtype)
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes and dtypes");
  const int64_t n = numel(acc);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (acc->dtype == HC_FLOAT64)
    HC_LAUNCH("k_square_acc", st, k_square_acc<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (double*)acc->data, (const double*)v->data, n, first);
  else if (acc->dtype == HC_FLOAT32)
    HC_LAUNCH("k_square_acc", st, k_square_acc<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (float*)acc->data, (const float*)v->data, n, first);
  else {
    void* A = acc->data;
    const void* V = v->data;
    switch (dtype_size(acc->dtype) * (acc->dtype <= HC_UINT64 && acc->dtype != HC_BOOL)) {
      case 1: HC_LAUNCH("k_square_acc_int", st, (k_square_acc_int<uint8_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (uint8_t*)A, (const uint8_t*)V, n, first); break;
      case 2: HC_LAUNCH("k_square_acc_int", st, (k_square_acc_int<uint16_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (uint16_t*)A, (const uint16_t*)V, n, first); break;
      case 4: HC_LAUNCH("k_square_acc_int", st, (k_square_acc_int<uint32_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (uint32_t*)A, (const uint32_t*)V, n, first); break;
      case 8: HC_LAUNCH("k_square_acc_int", st, (k_square_acc_int<uint64_t, uint64_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (uint64_t*)A, (const uint64_t*)V, n, first); break;
      default: return fail(HC_ERR_VALUE, "square_acc needs an integer or float32/float64 array");
    }
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_binary(const hc_array* a, const hc_array* b, const hc_array* out, int op, void* stream) {
  if (!a || !b || !out) return fail(HC_ERR_VALUE, "null array");
  if (op != HC_BIN_SUB && op != HC_BIN_MUL && op != HC_BIN_ADD)
    return fail(HC_ERR_VALUE, "unknown binary op");
  if (!same_shape(a, b) || !same_shape(a, out) || !is_c_contiguous(a) || !is_c_contiguous(b) ||
      !is_c_contiguous(out) || a->dtype != b->dtype || a->dtype != out->dtype)
    return fail(HC_ERR_VALUE, "arrays must be C-contiguous with equal shapes and dtypes");
  const int64_t n = numel(a);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (a->dtype == HC_FLOAT64)
    HC_LAUNCH("k_binary", st, k_binary<double>, dim3((unsigned)grid_for(n)), dim3(256), 0,
              (const double*)a->data, (const double*)b->data, (double*)out->data, n, op);
  else if (a->dtype == HC_FLOAT32)
    HC_LAUNCH("k_binary", st, k_binary<float>, dim3((unsigned)grid_for(n)), dim3(256), 0,
              (const float*)a->data, (const float*)b->data, (float*)out->data, n, op);
  else {
    const void* A = a->data;
    const void* B = b->data;
    void* O = out->data;
    switch (dtype_size(a->dtype) * (a->dtype <= HC_UINT64 && a->dtype != HC_BOOL)) {
      case 1: HC_LAUNCH("k_binary_int", st, (k_binary_int<uint8_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (const uint8_t*)A, (const uint8_t*)B, (uint8_t*)O, n, op); break;
      case 2: HC_LAUNCH("k_binary_int", st, (k_binary_int<uint16_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (const uint16_t*)A, (const uint16_t*)B, (uint16_t*)O, n, op); break;
      case 4: HC_LAUNCH("k_binary_int", st, (k_binary_int<uint32_t, uint32_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (const uint32_t*)A, (const uint32_t*)B, (uint32_t*)O, n, op); break;
      case 8: HC_LAUNCH("k_binary_int", st, (k_binary_int<uint64_t, uint64_t>), dim3((unsigned)grid_for(n)), dim3(256), 0, (const uint64_t*)A, (const uint64_t*)B, (uint64_t*)O, n, op); break;
      default: return fail(HC_ERR_VALUE, "binary ops need integer or float32/float64 arrays");
    }
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_div_scalar(const hc_array* a, double d, void* stream) {
  if (!a) return fail(HC_ERR_VALUE, "null array");
  if (!is_c_contiguous(a)) return fail(HC_ERR_VALUE, "array must be C-contiguous");
  const int64_t n = numel(a);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (a->dtype == HC_FLOAT64)
    HC_LAUNCH("k_div_scalar", st, k_div_scalar<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (double*)a->data, n, d);
  else if (a->dtype == HC_FLOAT32)
    HC_LAUNCH("k_div_scalar", st, k_div_scalar<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (float*)a->data, n, (float)d);
  else
    return fail(HC_ERR_VALUE, "division by a scalar needs float32/float64");
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_correlate_sparse(const hc_array* in, const hc_array* out, const int32_t* kshape,
                        const int32_t* tap_idx, const double* tap_w, int ntaps, int valid,
                        int mode, void* stream) {
  if (!in || !out || !kshape || (ntaps > 0 && (!tap_idx || !tap_w)))
    return fail(HC_ERR_VALUE, "null argument");
  if (in->ndim != out->ndim || in->ndim < 1 || in->ndim > HC_MAX_NDIM)
    return fail(HC_ERR_VALUE, "input and output ranks differ or are out of range");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  if (!is_c_contiguous(out)) return fail(HC_ERR_VALUE, "output must be C-contiguous");
  if (ntaps < 1) return fail(HC_ERR_VALUE, "at least one tap (the corner) is needed");
  if (!valid && (mode < HC_MODE_REFLECT || mode > HC_MODE_CONSTANT))
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  const int nd = in->ndim;
  SparseArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in->data;
  a.in_dt = in->dtype;
  a.out = out->data;
  a.ndim = nd;
  a.mode = valid ? HC_MODE_CONSTANT : mode;
  for (int d = 0; d < nd; ++d)
    if (kshape[d] < 1 || kshape[d] > 32767) return fail(HC_ERR_VALUE, "kernel shape out of range");
  for (int k = 0; k < ntaps; ++k)
    for (int d = 0; d < nd; ++d) {
      const int32_t v = tap_idx[(int64_t)k * nd + d];
      if (v < 0 || v >= kshape[d]) return fail(HC_ERR_VALUE, "tap index outside the kernel");
    }
  for (int d = 0; d < nd; ++d) {
    a.ishape[d] = in->shape[d];
    a.istride[d] = in->strides[d];
    a.oshape[d] = out->shape[d];
    // padded modes pad kshape // 2 on both sides (odd kernels keep the
    // input's shape); 'valid' reads without padding and shrinks by kshape - 1
    const int64_t want =
        valid ? std::max<int64_t>(0, in->shape[d] - kshape[d] + 1) : in->shape[d];
    if (!valid && kshape[d] % 2 == 0) return fail(HC_ERR_VALUE, "even kernel size");
    if (out->shape[d] != want) return fail(HC_ERR_VALUE, "output shape does not match");
    a.pad[d] = valid ? 0 : kshape[d] / 2;
  }
  a.n = numel(out);
  if (a.n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  for (int k0 = 0; k0 < ntaps; k0 += kSparseTaps) {
    a.ntaps = std::min(kSparseTaps, ntaps - k0);
    a.first = k0 == 0;
    for (int k = 0; k < a.ntaps; ++k) {
      for (int d = 0; d < nd; ++d) a.idx[k][d] = (int16_t)tap_idx[(int64_t)(k0 + k) * nd + d];
      a.w[k] = tap_w[k0 + k];
    }
    if (out->dtype == HC_FLOAT64)
      HC_LAUNCH("k_correlate_sparse", st, k_correlate_sparse<double>, dim3((unsigned)grid_for(a.n)), dim3(256), 0, a);
    else
      HC_LAUNCH("k_correlate_sparse", st, k_correlate_sparse<float>, dim3((unsigned)grid_for(a.n)), dim3(256), 0, a);
    HC_HIP(hipGetLastError());
  }
  return HC_OK;
}

int hc_sqrt_inplace(const hc_array* a, void* stream) {
  if (!a) return fail(HC_ERR_VALUE, "null array");
  if (!is_c_contiguous(a)) return fail(HC_ERR_VALUE, "array must be C-contiguous");
  const int64_t n = numel(a);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (a->dtype == HC_FLOAT64)
    HC_LAUNCH("k_sqrt", st, k_sqrt<double>, dim3((unsigned)grid_for(n)), dim3(256), 0, (double*)a->data, n);
  else if (a->dtype == HC_FLOAT32)
    HC_LAUNCH("k_sqrt", st, k_sqrt<float>, dim3((unsigned)grid_for(n)), dim3(256), 0, (float*)a->data, n);
  else {
    void* A = a->data;
    const dim3 g((unsigned)grid_for(n));
    switch (a->dtype) {
      case HC_INT8: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<int8_t, float>), g, dim3(256), 0, (int8_t*)A, n); break;
      case HC_UINT8: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<uint8_t, float>), g, dim3(256), 0, (uint8_t*)A, n); break;
      case HC_INT16: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<int16_t, float>), g, dim3(256), 0, (int16_t*)A, n); break;
      case HC_UINT16: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<uint16_t, float>), g, dim3(256), 0, (uint16_t*)A, n); break;
      case HC_INT32: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<int32_t, double>), g, dim3(256), 0, (int32_t*)A, n); break;
      case HC_UINT32: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<uint32_t, double>), g, dim3(256), 0, (uint32_t*)A, n); break;
      case HC_INT64: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<int64_t, double>), g, dim3(256), 0, (int64_t*)A, n); break;
      case HC_UINT64: HC_LAUNCH("k_sqrt_int", st, (k_sqrt_int<uint64_t, double>), g, dim3(256), 0, (uint64_t*)A, n); break;
      default: return fail(HC_ERR_VALUE, "sqrt needs an integer or float32/float64 array");
    }
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_copy_cast(const hc_array* in, const hc_array* out, void* stream) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null array");
  if (!same_shape(in, out)) return fail(HC_ERR_VALUE, "shapes differ");
  if (in->ndim > HC_MAX_NDIM) return fail(HC_ERR_VALUE, "unsupported ndim");
  CopyArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = in->data;
  a.out = out->data;
  a.in_dt = in->dtype;
  a.out_dt = out->dtype;
  a.ndim = in->ndim;
  a.n = numel(in);
  for (int d = 0; d < in->ndim; ++d) {
    a.shape[d] = in->shape[d];
    a.istride[d] = in->strides[d];
    a.ostride[d] = out->strides[d];
  }
  if (a.n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  const bool iin = in->dtype <= HC_UINT64, iout = out->dtype <= HC_UINT64;  // ints and bool
  const int kind = in->dtype == out->dtype ? 0 : (iin && iout ? 1 : 2);
  const int isz = dtype_size(in->dtype), osz = dtype_size(out->dtype);
  if (kind == 0 && is_c_contiguous(in) && is_c_contiguous(out)) {
    const int64_t bytes = a.n * isz;
    if (bytes % 16 == 0 && ((uintptr_t)in->data | (uintptr_t)out->data) % 16 == 0) {
      const int64_t n16 = bytes / 16;
      const int64_t want = (n16 + 1023) / 1024;
      const unsigned grid = (unsigned)(want < 4096 ? want : 4096);
      HC_LAUNCH("k_copy_vec", st, k_copy_vec, dim3(grid), dim3(256), 0,
                (const u32x4*)in->data, (u32x4*)out->data, n16);
      return HC_OK;
    }
  }
  const bool f64 = in->dtype == HC_FLOAT64 || out->dtype == HC_FLOAT64 || isz >= 4;
  if (f64)
    HC_LAUNCH("k_copy_cast", st, k_copy_cast<double>, dim3((unsigned)grid_for(a.n)), dim3(256), 0, a, kind, isz, osz);
  else
    HC_LAUNCH("k_copy_cast", st, k_copy_cast<float>, dim3((unsigned)grid_for(a.n)), dim3(256), 0, a, kind, isz, osz);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

}  // extern "C"
