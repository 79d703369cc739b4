// correlate1d / correlate / fused 2-D separable kernels for gfx950.
//
// Reference behaviour (paths relative to python/cucim/src/cucim/skimage/):
//   _vendored/_ndimage_filters.py:271-344   1-D dispatch (shared_memory vs elementwise)
//   _vendored/_ndimage_filters_core.py:247-408  elementwise n-D kernel
//   filters/_separable_filtering.py:977-1070    shared-memory line kernel
//   _vendored/_ndimage_util.py:186-308       boundary modes
//
// Kernels:
//   k_gather         one work-item per output, any strides, any kernel length
//                    (the 'elementwise' algorithm and the universal fallback)
//   k_tiled_contig   axis is the contiguous one: each wave stages a
//                    1024-output line segment plus halo in LDS and computes
//                    4 outputs per lane from 16-byte LDS reads
//   k_tiled_strided  axis is strided: a work-group stages 64 lines x
//                    (64 outputs + halo) in LDS; each lane computes 16
//                    consecutive outputs of its line
//   k_sep2d          two passes of a 2-D array in one launch through an LDS
//                    tile; the intermediate is rounded to the output dtype
#include <algorithm>
#include <cstring>

#include "hc_conv.h"
#include "hc_stencil.h"

namespace hc {

// ------------------------------------------------------------ gather --
template <typename W>
struct GatherArgs {
  const void* in;
  void* out;
  int in_dt, out_dt;
  int ndim;
  int all_const;
  int64_t total;
  int64_t shape[HC_MAX_NDIM];
  int64_t istride[HC_MAX_NDIM];
  int32_t wsh[HC_MAX_NDIM];   // footprint per dim (1 = not filtered)
  int32_t off[HC_MAX_NDIM];   // wsh//2 + origin
  int32_t mode[HC_MAX_NDIM];
  int64_t ntaps;
  const W* wdev;  // taps when ntaps exceeds the inline capacity
  W cval;
  TapPack<W> tp;
};

template <typename W>
__global__ __launch_bounds__(256) void k_gather(const GatherArgs<W> a) {
  typedef const __attribute__((address_space(4))) W* cW;
  const cW wd = (cW)a.wdev;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.total;
       i += (int64_t)gridDim.x * 256) {
    int64_t pos[HC_MAX_NDIM];
    int64_t rem = i;
#pragma unroll
    for (int d = HC_MAX_NDIM - 1; d >= 0; --d) {
      pos[d] = 0;
      if (d < a.ndim) {
        const int64_t s = a.shape[d];
        pos[d] = rem % s;
        rem /= s;
      }
    }
    int32_t k[HC_MAX_NDIM];
#pragma unroll
    for (int d = 0; d < HC_MAX_NDIM; ++d) k[d] = 0;
    W acc = (W)0;
    for (int64_t t = 0; t < a.ntaps; ++t) {
      const W w = wd ? wd[t] : a.tp.w[t];
      if (nz(w)) {
        int64_t offs = 0;
        bool outside = false;
#pragma unroll
        for (int d = 0; d < HC_MAX_NDIM; ++d) {
          if (d < a.ndim) {
            int64_t q = pos[d] + k[d] - a.off[d];
            q = remap(q, a.shape[d], a.mode[d]);
            if (q < 0) { outside = true; q = 0; }
            offs += q * a.istride[d];
          }
        }
        const W v = outside ? a.cval : load_as<W>(a.in, offs, a.in_dt);
        acc = fma_(v, w, acc);
      }
      // advance the C-order tap counter
#pragma unroll
      for (int d = HC_MAX_NDIM - 1; d >= 0; --d) {
        if (d < a.ndim) {
          if (++k[d] < a.wsh[d]) break;
          k[d] = 0;
        }
      }
    }
    store_cast<W>(a.out, i, a.out_dt, acc);
  }
}

// ------------------------------------------------------ tiled kernels --
template <typename W>
struct TiledArgs {
  const void* in;
  void* out;
  int in_dt, out_dt;
  int64_t nlines;  // number of lines
  int64_t N;       // line length
  int64_t inner;   // element stride along a line (1 = contiguous)
  int64_t nseg;    // segments per line
  int seg;         // contiguous: outputs per wave segment (multiple of 256)
  int mode;
  W cval;
  TapPack<W> tp;   // K padded to a multiple of 4
};

constexpr int kStrCols = 64;     // strided: lines per work-group
constexpr int kStrStride = 65;   // strided: LDS row stride (elements)
constexpr int kStrTA = 64;       // strided: outputs per line per work-group
constexpr int kStrWaveBlocks = kStrTA / 64;  // 16-output blocks per wave

// 16-byte loads each lane keeps in flight while staging.
constexpr int kVecBatch = 8;
// 16-byte chunks a lane loads per round when it stages a row segment: a
// 1024-output segment and its halo are about 5 chunks per lane
constexpr int kRowBatch = 5;

// 16-byte vectors of the compute type.
template <typename W> struct Vec16;
template <> struct Vec16<float> {
  typedef float T __attribute__((ext_vector_type(4)));
  static constexpr int n = 4;
};
template <> struct Vec16<double> {
  typedef double T __attribute__((ext_vector_type(2)));
  static constexpr int n = 2;
};

// ext[s] = row[p0 + s] (boundary-extended) for s in [0, L), one wave.  Input
// chunks of Vec16 elements aligned to the row start are read with one
// 16-byte load when they lie inside the row, element-wise through the
// boundary remap otherwise.  Requires row 16-byte aligned.
template <typename W>
__device__ __forceinline__ bool stage_row_vec(W* __restrict__ ext, int L, int lane,
                                              const W* __restrict__ row, int64_t N, int64_t p0,
                                              int mode, W cval) {
  constexpr int VN = Vec16<W>::n;
  typedef typename Vec16<W>::T V;
  const int64_t a0 = p0 >= 0 ? (p0 / VN) * VN : -((-p0 + VN - 1) / VN) * VN;
  const int shift = (int)(p0 - a0);
  const int nc = (shift + L + VN - 1) / VN;
  // the row holds a whole number of 16-byte chunks (N a multiple of VN,
  // N >= VN): gl is the start of its last one
  const int64_t gl = N - VN;
  bool bad = false;
  for (int c0 = lane; c0 < nc; c0 += 64 * kRowBatch) {
    // every load of the batch is issued before any is used: unconditional,
    // its chunk index clamped into the row (a chunk that is not wholly
    // inside is replaced below); a load under a branch would be waited for
    // before the next one is issued
    V t[kRowBatch];
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int64_t g = a0 + (int64_t)(c0 + 64 * u) * VN;
      const int64_t gc = g < 0 ? 0 : (g > gl ? gl : g);
      t[u] = *reinterpret_cast<const V*>(row + gc);
    }
#pragma unroll
    for (int u = 0; u < kRowBatch; ++u) {
      const int c = c0 + 64 * u;
      if (c >= nc) continue;
      const int64_t g = a0 + (int64_t)c * VN;
      W v[VN];
      if (g >= 0 && g + VN <= N) {
#pragma unroll
        for (int j = 0; j < VN; ++j) v[j] = t[u][j];
      } else {
#pragma unroll
        for (int j = 0; j < VN; ++j) {
          const int64_t r = remap(g + j, N, mode);
          v[j] = r >= 0 ? row[r] : cval;
        }
      }
#pragma unroll
      for (int j = 0; j < VN; ++j) {
        const int sl = c * VN + j - shift;
        if (sl >= 0 && sl < L) {
          ext[sl] = v[j];
          bad |= !finite(v[j]);
        }
      }
    }
  }
  return bad;
}

// VEC: input and output have the compute dtype, 16-byte aligned rows
// (line length a multiple of 16 bytes): vector loads and stores.
template <typename W, bool VEC>
__global__ __launch_bounds__(256) void k_tiled_contig(const TiledArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  W* smem = reinterpret_cast<W*>(smem_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int L = a.seg + a.tp.K + 4;  // conv_row4 reads 8 past the last group
  W* ext = smem + wv * L;
  W* wl = smem + 4 * L;
  stage_taps<W>(wl, a.tp, threadIdx.x, 256);
  const int64_t nb = gridDim.x;
  const int64_t seg = xcd_swizzle(blockIdx.x, nb) * 4 + wv;
  const bool active = seg < a.nlines * a.nseg;
  int64_t lbase = 0, s0 = 0;
  bool bad = false;
  if (active) {
    const int64_t line = seg / a.nseg;
    s0 = (seg - line * a.nseg) * a.seg;
    lbase = line * a.N;
    const int64_t p0 = s0 - a.tp.anchor;
    const int64_t N = a.N, lb = lbase;
    const int mode = a.mode;
    if (VEC)
      bad = stage_row_vec<W>(ext, L, lane, reinterpret_cast<const W*>(a.in) + lb, N, p0, mode,
                             a.cval);
    else
      bad = stage_any<W>(ext, L, lane, 64, a.in, a.in_dt, a.cval, false,
                         [=](int e, int& slot) -> int64_t {
                           slot = e;
                           const int64_t r = remap(p0 + e, N, mode);
                           return r < 0 ? -1 : lb + r;
                         });
  }
  const bool skip = __syncthreads_or(bad);
  if (!active) return;
  with_skip(skip, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
#pragma unroll 1
    for (int g = 0; g < a.seg / 256; ++g) {
      const int o0 = 256 * g + 4 * lane;
      const int64_t o = s0 + o0;
      if (o >= a.N) break;
      W acc[4] = {(W)0, (W)0, (W)0, (W)0};
      conv_row4<SK>(acc, ext + o0, wl, a.tp);
      if (VEC && o + 4 <= a.N) {
        W* op = reinterpret_cast<W*>(a.out) + lbase + o;
        // streaming stores: the output is not read again by this launch
        if (sizeof(W) == 4) {
          typename Vec16<float>::T t = {(float)acc[0], (float)acc[1], (float)acc[2], (float)acc[3]};
          __builtin_nontemporal_store(t, reinterpret_cast<typename Vec16<float>::T*>(op));
        } else {
          typename Vec16<double>::T t0 = {(double)acc[0], (double)acc[1]};
          typename Vec16<double>::T t1 = {(double)acc[2], (double)acc[3]};
          __builtin_nontemporal_store(t0, reinterpret_cast<typename Vec16<double>::T*>(op));
          __builtin_nontemporal_store(t1, reinterpret_cast<typename Vec16<double>::T*>(op) + 1);
        }
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j)
          if (o + j < a.N) store_cast<W>(a.out, lbase + o + j, a.out_dt, acc[j]);
      }
    }
  });
}

// VEC: input and output have the compute dtype, the element stride of the
// line (inner) is a multiple of 16 bytes and the data 16-byte aligned: each
// 16-byte load brings one position of Vec16 neighbouring lines.
template <typename W, bool VEC>
__global__ __launch_bounds__(256) void k_tiled_strided(const TiledArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  W* ext = reinterpret_cast<W*>(smem_raw);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int64_t nlb = (a.nlines + kStrCols - 1) / kStrCols;
  const int64_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t lb = tile % nlb, sg = tile / nlb;
  const int64_t L = lb * kStrCols + lane;
  const bool lok = L < a.nlines;
  const int64_t lo = lok ? L / a.inner : 0, li = lok ? L - lo * a.inner : 0;
  const int64_t lbase = lo * a.N * a.inner + li;
  const int64_t a0 = sg * kStrTA;
  const int R = kStrTA + a.tp.K + 4;  // conv_col16 prefetches one group ahead
  W* wl = ext + R * kStrStride;
  stage_taps<W>(wl, a.tp, threadIdx.x, 256);
  const int64_t p0 = a0 - a.tp.anchor;
  bool bad;
  {
    const int64_t N = a.N, inner = a.inner;
    const int mode = a.mode;
    if (VEC) {
      constexpr int VN = Vec16<W>::n, CPR = 64 / VN;  // chunks per LDS row
      typedef typename Vec16<W>::T V;
      const W* in = reinterpret_cast<const W*>(a.in);
      const int64_t Lt = lb * kStrCols;
      const int total = R * CPR;
      // 256 is a multiple of CPR: the thread's chunk column is fixed, so its
      // line base is computed once
      static_assert(256 % CPR == 0, "chunk column per thread");
      const int cc = (threadIdx.x % CPR) * VN;
      const int64_t Lc = Lt + cc;
      const bool cok = Lc < a.nlines;
      const int64_t lo2 = cok ? Lc / inner : 0;
      const W* colp = in + lo2 * N * inner + (Lc - lo2 * inner);
      bad = false;
      for (int e0 = threadIdx.x; e0 < total; e0 += 256 * kVecBatch) {
        W v[kVecBatch][VN];
#pragma unroll
        for (int u = 0; u < kVecBatch; ++u) {
          const int e = e0 + 256 * u;
          const int q = e / CPR;
#pragma unroll
          for (int j = 0; j < VN; ++j) v[u][j] = a.cval;
          if (e < total && cok) {
            const int64_t pq = p0 + q;
            const int64_t r = (pq >= 0 && pq < N) ? pq : remap(pq, N, mode);
            if (r >= 0) {
              const V t = *reinterpret_cast<const V*>(colp + r * inner);
#pragma unroll
              for (int j = 0; j < VN; ++j) v[u][j] = t[j];
            }
          }
        }
#pragma unroll
        for (int u = 0; u < kVecBatch; ++u) {
          const int e = e0 + 256 * u;
          if (e < total) {
            const int q = e / CPR;
#pragma unroll
            for (int j = 0; j < VN; ++j) {
              ext[q * kStrStride + cc + j] = v[u][j];
              bad |= !finite(v[u][j]);
            }
          }
        }
      }
    } else {
      bad = stage_any<W>(ext, R * 64, threadIdx.x, 256, a.in, a.in_dt, a.cval, false,
                         [=](int e, int& slot) -> int64_t {
                           const int q = e >> 6, c = e & 63;  // c == lane
                           slot = q * kStrStride + c;
                           if (!lok) return -1;
                           const int64_t r = remap(p0 + q, N, mode);
                           return r < 0 ? -1 : lbase + r * inner;
                         });
    }
  }
  const bool skip = __syncthreads_or(bad);
  if (!lok) return;
#pragma unroll 1
  for (int blk = 0; blk < kStrWaveBlocks; ++blk) {
    const int rb = 16 * (wv + 4 * blk);  // first output row of the block
    const int64_t n0 = a0 + rb;
    if (n0 >= a.N) return;
    with_skip(skip, [&](auto sk) {
      constexpr bool SK = decltype(sk)::value;
      W acc[16];
#pragma unroll
      for (int j = 0; j < 16; ++j) acc[j] = (W)0;
      conv_col16<SK, W, kStrStride>(acc, ext + rb * kStrStride + lane, wl, a.tp);
      if (VEC) {
        W* op = reinterpret_cast<W*>(a.out) + lbase;
#pragma unroll
        for (int j = 0; j < 16; ++j)
          if (n0 + j < a.N) __builtin_nontemporal_store(acc[j], op + (n0 + j) * a.inner);
      } else {
#pragma unroll
        for (int j = 0; j < 16; ++j)
          if (n0 + j < a.N) store_cast<W>(a.out, lbase + (n0 + j) * a.inner, a.out_dt, acc[j]);
      }
    });
  }
}

// ------------------------------------------------------ fused 2-D --
constexpr int kSepCols = 128;  // LDS tile width (elements)

template <typename W>
struct Sep2dArgs {
  const void* in;
  void* out;
  int in_dt, out_dt;
  int64_t H, Wd;
  int TX;
  int64_t ntx, ntiles;
  int mode0, mode1;
  W cval;
  TapPack<W> t0, t1;
};

// VEC: typed input/output, rows 16-byte aligned: the tile is staged with
// 16-byte loads (aligned to the row start) and written with 16-byte stores.
template <typename W, int TY, bool VEC>
__global__ __launch_bounds__(256) void k_sep2d(const Sep2dArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  W* A = reinterpret_cast<W*>(smem_raw);
  const int RY = TY + a.t0.K + 4;  // conv_col16 prefetches one group ahead
  W* U = A + RY * kSepCols;
  W* w0l = U + TY * kSepCols;
  W* w1l = w0l + a.t0.K + 4;
  stage_taps<W>(w0l, a.t0, threadIdx.x, 256);
  stage_taps<W>(w1l, a.t1, threadIdx.x, 256);
  const int CX = a.TX + a.t1.K;
  const int nq = a.TX / 4;
  for (int64_t tile = blockIdx.x; tile < a.ntiles; tile += gridDim.x) {
    const int64_t ty = tile / a.ntx, tx = tile - ty * a.ntx;
    const int64_t y0 = ty * TY, x0 = tx * a.TX;
    // stage the input tile with both axes boundary-extended
    bool bad;
    {
      const int64_t H = a.H, Wd = a.Wd, ya = y0 - a.t0.anchor, xa = x0 - a.t1.anchor;
      const int m0 = a.mode0, m1 = a.mode1;
      if (VEC) {
        constexpr int VN = Vec16<W>::n;
        typedef typename Vec16<W>::T V;
        const W* in = reinterpret_cast<const W*>(a.in);
        const int64_t ax0 = xa >= 0 ? (xa / VN) * VN : -((-xa + VN - 1) / VN) * VN;
        const int shift = (int)(xa - ax0);
        const int cpr = (shift + CX + VN - 1) / VN;  // chunks per tile row
        const int total = RY * cpr;
        bad = false;
        for (int e0 = threadIdx.x; e0 < total; e0 += 256 * kVecBatch) {
          W v[kVecBatch][VN];
#pragma unroll
          for (int u = 0; u < kVecBatch; ++u) {
            const int e = e0 + 256 * u;
#pragma unroll
            for (int j = 0; j < VN; ++j) v[u][j] = a.cval;
            if (e < total) {
              const int q = e / cpr, cc = e - q * cpr;
              const int64_t y = remap(ya + q, H, m0);
              const int64_t gx = ax0 + (int64_t)cc * VN;
              if (y >= 0) {
                const W* row = in + y * Wd;
                if (gx >= 0 && gx + VN <= Wd) {
                  const V t = *reinterpret_cast<const V*>(row + gx);
#pragma unroll
                  for (int j = 0; j < VN; ++j) v[u][j] = t[j];
                } else {
#pragma unroll
                  for (int j = 0; j < VN; ++j) {
                    const int64_t x = remap(gx + j, Wd, m1);
                    if (x >= 0) v[u][j] = row[x];
                  }
                }
              }
            }
          }
#pragma unroll
          for (int u = 0; u < kVecBatch; ++u) {
            const int e = e0 + 256 * u;
            if (e < total) {
              const int q = e / cpr, cc = e - q * cpr;
#pragma unroll
              for (int j = 0; j < VN; ++j) {
                const int c = cc * VN + j - shift;
                if (c >= 0 && c < CX) {
                  A[q * kSepCols + c] = v[u][j];
                  bad |= !finite(v[u][j]);
                }
              }
            }
          }
        }
      } else {
        bad = stage_any<W>(A, RY * kSepCols, threadIdx.x, 256, a.in, a.in_dt, a.cval, false,
                           [=](int e, int& slot) -> int64_t {
                             const int q = e >> 7, c = e & (kSepCols - 1);
                             slot = c < CX ? e : -1;
                             const int64_t y = remap(ya + q, H, m0), x = remap(xa + c, Wd, m1);
                             return (y < 0 || x < 0) ? -1 : y * Wd + x;
                           });
      }
    }
    const bool skip0 = __syncthreads_or(bad);
    // pass 1 (axis 0) over every staged column, rounded to the output dtype
    bool bad1 = false;
    with_skip(skip0, [&](auto sk) {
      constexpr bool SK = decltype(sk)::value;
      for (int it = threadIdx.x; it < (TY / 16) * kSepCols; it += 256) {
        const int ch = it / kSepCols, c = it - ch * kSepCols;
        if (c >= CX) continue;
        W acc[16];
#pragma unroll
        for (int j = 0; j < 16; ++j) acc[j] = (W)0;
        conv_col16<SK, W, kSepCols>(acc, A + (16 * ch) * kSepCols + c, w0l, a.t0);
        const int64_t xg = x0 - a.t1.anchor + c;
        const bool xout = (xg < 0 || xg >= a.Wd) && a.mode1 == HC_MODE_CONSTANT;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
          const W u = xout ? a.cval : roundtrip<W>(acc[j], a.out_dt);
          bad1 |= !finite(u);
          U[(16 * ch + j) * kSepCols + c] = u;
        }
      }
    });
    const bool skip1 = __syncthreads_or(bad1);
    // pass 2 (axis 1) along the rows of the intermediate
    with_skip(skip1, [&](auto sk) {
      constexpr bool SK = decltype(sk)::value;
      for (int it = threadIdx.x; it < TY * nq; it += 256) {
        const int i = it / nq, o = (it - i * nq) * 4;
        const int64_t y = y0 + i, x = x0 + o;
        if (y >= a.H || x >= a.Wd) continue;
        W acc[4] = {(W)0, (W)0, (W)0, (W)0};
        conv_row4<SK>(acc, U + i * kSepCols + o, w1l, a.t1);
        if (VEC && x + 4 <= a.Wd) {
          W* op = reinterpret_cast<W*>(a.out) + y * a.Wd + x;
          if (sizeof(W) == 4) {
            typename Vec16<float>::T t = {(float)acc[0], (float)acc[1], (float)acc[2],
                                          (float)acc[3]};
            *reinterpret_cast<typename Vec16<float>::T*>(op) = t;
          } else {
            typename Vec16<double>::T t0 = {(double)acc[0], (double)acc[1]};
            typename Vec16<double>::T t1 = {(double)acc[2], (double)acc[3]};
            reinterpret_cast<typename Vec16<double>::T*>(op)[0] = t0;
            reinterpret_cast<typename Vec16<double>::T*>(op)[1] = t1;
          }
        } else {
#pragma unroll
          for (int j = 0; j < 4; ++j)
            if (x + j < a.Wd) store_cast<W>(a.out, y * a.Wd + x + j, a.out_dt, acc[j]);
        }
      }
    });
    __syncthreads();
  }
}

// --------------------------------------------------------- host side --
namespace {

bool valid_dtype(int dt) { return dt >= HC_BOOL && dt <= HC_FLOAT64; }
bool valid_mode(int m) { return m >= HC_MODE_REFLECT && m <= HC_MODE_CONSTANT; }

int check_in_out(const hc_array* in, const hc_array* out) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null array");
  if (in->ndim < 1 || in->ndim > HC_MAX_NDIM) return fail(HC_ERR_VALUE, "unsupported ndim");
  if (!valid_dtype(in->dtype) || !valid_dtype(out->dtype))
    return fail(HC_ERR_UNSUPPORTED, "unsupported dtype");
  if (!same_shape(in, out)) return fail(HC_ERR_VALUE, "output shape not correct");
  if (!is_c_contiguous(out)) return fail(HC_ERR_VALUE, "output must be C-contiguous");
  return HC_OK;
}

template <typename W>
int launch_gather_1d(const hc_array* in, const hc_array* out, const hc_taps* taps, int axis,
                     int mode, double cval, hipStream_t st) {
  GatherArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  a.in = in->data;
  a.out = out->data;
  a.in_dt = in->dtype;
  a.out_dt = out->dtype;
  a.ndim = in->ndim;
  a.total = numel(in);
  for (int d = 0; d < in->ndim; ++d) {
    a.shape[d] = in->shape[d];
    a.istride[d] = in->strides[d];
    a.wsh[d] = 1;
    a.off[d] = 0;
    a.mode[d] = mode;
  }
  a.wsh[axis] = (int32_t)taps->K;
  a.off[axis] = taps->anchor;
  a.ntaps = taps->K;
  a.cval = (W)cval;
  if (taps->K <= inline_cap<W>()) {
    a.tp.K = (int)taps->K;
    for (int64_t k = 0; k < taps->K; ++k) a.tp.w[k] = (W)taps->w[k];
    a.wdev = nullptr;
  } else {
    if (!taps->w_dev) return fail(HC_ERR_VALUE, "long kernel needs a device copy of the weights");
    a.wdev = (const W*)taps->w_dev;
  }
  if (a.total == 0) return HC_OK;
  const int64_t blocks = std::min<int64_t>((a.total + 255) / 256, 1 << 20);
  HC_LAUNCH("k_gather", st, k_gather<W>, dim3((unsigned)blocks), dim3(256), 0, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

// Outputs per wave segment of the contiguous-axis kernel: the tuned value
// (hc_tune_set(HC_TUNE_CONTIG_SEG, ...)) or, by default, the longest of
// 1024 / 512 / 256 that still gives 16 waves per CU.
int contig_segment(int64_t N, int64_t nlines) {
  const int t = tune_get(HC_TUNE_CONTIG_SEG);
  if (t > 0) return t;
  int seg = 1024;
  while (seg > 256 && nlines * ((N + seg - 1) / seg) < 16 * 256) seg /= 2;
  return seg;
}

template <typename W>
int launch_tiled_1d(const hc_array* in, const hc_array* out, const hc_taps* taps, int axis,
                    int mode, double cval, hipStream_t st) {
  TiledArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  int rc = make_taps<W>(taps, 4, &a.tp);
  if (rc != HC_OK) return rc;
  const Lines ln = lines_of(in, axis);
  a.in = in->data;
  a.out = out->data;
  a.in_dt = in->dtype;
  a.out_dt = out->dtype;
  a.N = ln.N;
  a.inner = ln.inner;
  a.mode = mode;
  a.cval = (W)cval;
  if (numel(in) == 0) return HC_OK;
  if (ln.inner == 1) {
    const bool vec_contig = in->dtype == w_dtype<W>() && out->dtype == w_dtype<W>() &&
                            (ln.N * (int64_t)sizeof(W)) % 16 == 0 &&
                            (((uintptr_t)in->data | (uintptr_t)out->data) % 16) == 0 &&
                            is_c_contiguous(in) && is_c_contiguous(out);
    a.nlines = ln.outer;
    a.seg = contig_segment(ln.N, a.nlines);
    a.nseg = (ln.N + a.seg - 1) / a.seg;
    const int64_t nsegs = a.nlines * a.nseg;
    const int64_t blocks = (nsegs + 3) / 4;
    if (blocks > 0x7fffffff) return fail(HC_ERR_VALUE, "array too large");
    const size_t smem = ((size_t)4 * (a.seg + a.tp.K + 4) + a.tp.K + 4) * sizeof(W);
    if (smem > 160 * 1024) return fail(HC_ERR_RESOURCE, "segment does not fit in LDS");
    if (smem > 64 * 1024) {
      const hipError_t e = hipFuncSetAttribute(
          vec_contig ? (const void*)k_tiled_contig<W, true> : (const void*)k_tiled_contig<W, false>,
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(k_tiled_contig)");
    }
    if (vec_contig)
      HC_LAUNCH("k_tiled_contig", st, (k_tiled_contig<W, true>), dim3((unsigned)blocks),
                dim3(256), smem, a);
    else
      HC_LAUNCH("k_tiled_contig", st, (k_tiled_contig<W, false>), dim3((unsigned)blocks),
                dim3(256), smem, a);
  } else {
    a.nlines = ln.outer * ln.inner;
    a.nseg = (ln.N + kStrTA - 1) / kStrTA;
    const int64_t nlb = (a.nlines + kStrCols - 1) / kStrCols;
    const int64_t blocks = nlb * a.nseg;
    if (blocks > 0x7fffffff) return fail(HC_ERR_VALUE, "array too large");
    const size_t smem =
        ((size_t)(kStrTA + a.tp.K + 4) * kStrStride + a.tp.K + 4) * sizeof(W);
    const bool vec = in->dtype == w_dtype<W>() && out->dtype == w_dtype<W>() &&
                     (ln.inner * (int64_t)sizeof(W)) % 16 == 0 &&
                     (((uintptr_t)in->data | (uintptr_t)out->data) % 16) == 0 &&
                     is_c_contiguous(in) && is_c_contiguous(out);
    if (smem > 160 * 1024) return fail(HC_ERR_RESOURCE, "line tile does not fit in LDS");
    if (smem > 64 * 1024) {
      const hipError_t e = hipFuncSetAttribute(
          vec ? (const void*)k_tiled_strided<W, true> : (const void*)k_tiled_strided<W, false>,
          hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
      if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(k_tiled_strided)");
    }
    if (vec)
      HC_LAUNCH("k_tiled_strided", st, (k_tiled_strided<W, true>), dim3((unsigned)blocks),
                dim3(256), smem, a);
    else
      HC_LAUNCH("k_tiled_strided", st, (k_tiled_strided<W, false>), dim3((unsigned)blocks),
                dim3(256), smem, a);
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

template <typename W>
int launch_gather_nd(const hc_array* in, const hc_array* out, const void* w_dev,
                     const int64_t* wshape, const int32_t* offsets, const int32_t* modes,
                     double cval, hipStream_t st) {
  GatherArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  a.in = in->data;
  a.out = out->data;
  a.in_dt = in->dtype;
  a.out_dt = out->dtype;
  a.ndim = in->ndim;
  a.total = numel(in);
  a.ntaps = 1;
  a.all_const = 1;
  for (int d = 0; d < in->ndim; ++d) {
    a.shape[d] = in->shape[d];
    a.istride[d] = in->strides[d];
    if (wshape[d] < 1) return fail(HC_ERR_VALUE, "filter weights array has incorrect shape");
    a.wsh[d] = (int32_t)wshape[d];
    a.off[d] = offsets[d];
    if (offsets[d] < 0 || offsets[d] >= wshape[d]) return fail(HC_ERR_VALUE, "invalid origin");
    if (!valid_mode(modes[d])) return fail(HC_ERR_VALUE, "boundary mode not supported");
    a.mode[d] = modes[d];
    a.ntaps *= wshape[d];
  }
  if (!w_dev) return fail(HC_ERR_VALUE, "weights device pointer is null");
  a.wdev = (const W*)w_dev;
  a.cval = (W)cval;
  if (a.total == 0) return HC_OK;
  const int64_t blocks = std::min<int64_t>((a.total + 255) / 256, 1 << 20);
  HC_LAUNCH("k_gather", st, k_gather<W>, dim3((unsigned)blocks), dim3(256), 0, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

template <typename W, int TY, bool VEC>
int launch_sep2d_ty(const Sep2dArgs<W>& a, hipStream_t st, size_t smem) {
  if (smem > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute((const void*)k_sep2d<W, TY, VEC>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)smem);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      return fail(HC_ERR_RESOURCE, std::string("LDS request refused: ") + hipGetErrorString(e));
    }
  }
  const int64_t grid = std::min<int64_t>(a.ntiles, 256 * 8);
  HC_LAUNCH("k_sep2d", st, (k_sep2d<W, TY, VEC>), dim3((unsigned)grid), dim3(256), smem, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

template <typename W>
int launch_sep2d(const hc_array* in, const hc_array* out, const hc_taps* t0, const hc_taps* t1,
                 int mode0, int mode1, double cval, hipStream_t st) {
  Sep2dArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  int rc = make_taps<W>(t0, 4, &a.t0);
  if (rc != HC_OK) return rc;
  rc = make_taps<W>(t1, 4, &a.t1);
  if (rc != HC_OK) return rc;
  if (a.t1.K > kSepCols - 20 || a.t0.K > 128)
    return fail(HC_ERR_RESOURCE, "kernels too long for the fused 2-D tile");
  a.in = in->data;
  a.out = out->data;
  a.in_dt = in->dtype;
  a.out_dt = out->dtype;
  a.H = in->shape[0];
  a.Wd = in->shape[1];
  a.TX = kSepCols - a.t1.K - 4;  // row reads run 8 past the last group
  a.mode0 = mode0;
  a.mode1 = mode1;
  a.cval = (W)cval;
  if (a.H == 0 || a.Wd == 0) return HC_OK;
  int TY = tune_get(HC_TUNE_SEP2D_TY);
  if (TY != 16 && TY != 32 && TY != 64) TY = sizeof(W) == 4 ? 32 : 16;
  a.ntx = (a.Wd + a.TX - 1) / a.TX;
  const int64_t nty = (a.H + TY - 1) / TY;
  a.ntiles = a.ntx * nty;
  const size_t smem =
      ((size_t)(TY + a.t0.K + 4 + TY) * kSepCols + a.t0.K + a.t1.K + 8) * sizeof(W);
  if (smem > 160 * 1024) return fail(HC_ERR_RESOURCE, "kernels too long for the fused 2-D tile");
  const bool vec = in->dtype == w_dtype<W>() && out->dtype == w_dtype<W>() &&
                   (a.Wd * (int64_t)sizeof(W)) % 16 == 0 &&
                   (((uintptr_t)in->data | (uintptr_t)out->data) % 16) == 0;
  if (vec) {
    if (TY == 16) return launch_sep2d_ty<W, 16, true>(a, st, smem);
    if (TY == 32) return launch_sep2d_ty<W, 32, true>(a, st, smem);
    return launch_sep2d_ty<W, 64, true>(a, st, smem);
  }
  if (TY == 16) return launch_sep2d_ty<W, 16, false>(a, st, smem);
  if (TY == 32) return launch_sep2d_ty<W, 32, false>(a, st, smem);
  return launch_sep2d_ty<W, 64, false>(a, st, smem);
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int hc_correlate1d(const hc_array* in, const hc_array* out, const hc_taps* taps, int axis,
                   int mode, double cval, int algorithm, void* stream) {
  int rc = check_in_out(in, out);
  if (rc != HC_OK) return rc;
  if (!taps) return fail(HC_ERR_VALUE, "null taps");
  if (axis < 0) axis += in->ndim;
  if (axis < 0 || axis >= in->ndim) return fail(HC_ERR_VALUE, "invalid axis");
  if (!valid_mode(mode)) return fail(HC_ERR_VALUE, "boundary mode not supported");
  if (taps->K < 1) return fail(HC_ERR_VALUE, "incorrect filter size");
  if (taps->anchor < 0 || taps->anchor >= taps->K) return fail(HC_ERR_VALUE, "invalid origin");
  hipStream_t st = (hipStream_t)stream;
  const bool f64 = taps->w_f64 != 0;
  const bool contig = is_c_contiguous(in);
  const int64_t Kp = (taps->K + 3) / 4 * 4;
  const bool fits = Kp <= (f64 ? inline_cap<double>() : inline_cap<float>());
  if (algorithm == HC_ALGO_TILED) {
    if (!contig) return fail(HC_ERR_VALUE, "tiled kernels need a C-contiguous input");
    if (!fits) return fail(HC_ERR_RESOURCE, "kernel too long for the LDS-tiled line kernels");
  }
  const bool tiled = algorithm != HC_ALGO_GATHER && contig && fits;
  if (tiled)
    return f64 ? launch_tiled_1d<double>(in, out, taps, axis, mode, cval, st)
               : launch_tiled_1d<float>(in, out, taps, axis, mode, cval, st);
  return f64 ? launch_gather_1d<double>(in, out, taps, axis, mode, cval, st)
             : launch_gather_1d<float>(in, out, taps, axis, mode, cval, st);
}

int hc_correlate_nd(const hc_array* in, const hc_array* out, const void* w_dev,
                    const int64_t* wshape, const int32_t* offsets, const int32_t* modes, int w_f64,
                    double cval, void* stream) {
  int rc = check_in_out(in, out);
  if (rc != HC_OK) return rc;
  if (!wshape || !offsets || !modes) return fail(HC_ERR_VALUE, "null argument");
  hipStream_t st = (hipStream_t)stream;
  return w_f64 ? launch_gather_nd<double>(in, out, w_dev, wshape, offsets, modes, cval, st)
               : launch_gather_nd<float>(in, out, w_dev, wshape, offsets, modes, cval, st);
}

int hc_separable2d(const hc_array* in, const hc_array* out, const hc_taps* taps0,
                   const hc_taps* taps1, int mode0, int mode1, double cval, void* stream) {
  int rc = check_in_out(in, out);
  if (rc != HC_OK) return rc;
  if (in->ndim != 2) return fail(HC_ERR_VALUE, "hc_separable2d needs a 2-D array");
  if (!is_c_contiguous(in)) return fail(HC_ERR_VALUE, "input must be C-contiguous");
  if (!valid_mode(mode0) || !valid_mode(mode1))
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  if (!taps0 || !taps1) return fail(HC_ERR_VALUE, "null taps");
  if (taps0->w_f64 != taps1->w_f64) return fail(HC_ERR_VALUE, "weights dtypes differ");
  hipStream_t st = (hipStream_t)stream;
  if (taps0->K < 1 || taps1->K < 1) return fail(HC_ERR_VALUE, "incorrect filter size");
  if (taps0->anchor < 0 || taps0->anchor >= taps0->K || taps1->anchor < 0 ||
      taps1->anchor >= taps1->K)
    return fail(HC_ERR_VALUE, "invalid origin");
  rc = launch_sep2d_reg(in, out, taps0, taps1, mode0, mode1, cval, st);
  if (rc != -1) return rc;
  return taps0->w_f64 ? launch_sep2d<double>(in, out, taps0, taps1, mode0, mode1, cval, st)
                      : launch_sep2d<float>(in, out, taps0, taps1, mode0, mode1, cval, st);
}

}  // extern "C"
