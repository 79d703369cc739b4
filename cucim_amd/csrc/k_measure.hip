// Reductions and labelling after the vesselness filter (gfx950): the image
// range and histogram behind filters/thresholding.py:298-475 (threshold_otsu,
// threshold_yen via exposure/exposure.py:213-345 histogram), the foreground /
// background means of threshold_li (:614-782), and the connected-component
// labelling of measure/_label.py:21-124.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>

#include "hc_common.h"

namespace hc {

namespace {

constexpr int kRedThreads = 256;
constexpr int kRedMaxBlocks = 1024;
constexpr int kHistLdsBins = 4096;  // LDS-privatised counters up to this many bins

// min / max keys: double for floats (NaN-propagating), int64 / uint64 for
// integers, so that every integer value is represented exactly.
template <typename K>
__device__ __forceinline__ K kmin(K a, K b) {
  if constexpr (std::is_floating_point<K>::value) {
    if (a != a || b != b) return (K)NAN;
  }
  return b < a ? b : a;
}
template <typename K>
__device__ __forceinline__ K kmax(K a, K b) {
  if constexpr (std::is_floating_point<K>::value) {
    if (a != a || b != b) return (K)NAN;
  }
  return b > a ? b : a;
}

template <typename K>
__device__ __forceinline__ void block_minmax(K& lo, K& hi, K* s_lo, K* s_hi) {
  const int tid = threadIdx.x;
  s_lo[tid] = lo;
  s_hi[tid] = hi;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (tid < o) {
      s_lo[tid] = kmin(s_lo[tid], s_lo[tid + o]);
      s_hi[tid] = kmax(s_hi[tid], s_hi[tid + o]);
    }
    __syncthreads();
  }
  lo = s_lo[0];
  hi = s_hi[0];
}

template <typename T, typename K>
__global__ __launch_bounds__(kRedThreads) void k_minmax_part(const T* __restrict__ x, int64_t n,
                                                             K* __restrict__ part) {
  __shared__ K s_lo[kRedThreads], s_hi[kRedThreads];
  K lo = (K)x[0], hi = lo;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const K v = (K)x[i];
    lo = kmin(lo, v);
    hi = kmax(hi, v);
  }
  block_minmax(lo, hi, s_lo, s_hi);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = lo;
    part[2 * blockIdx.x + 1] = hi;
  }
}

template <typename K>
__global__ __launch_bounds__(kRedThreads) void k_minmax_final(const K* __restrict__ part, int nb,
                                                              K* __restrict__ out) {
  __shared__ K s_lo[kRedThreads], s_hi[kRedThreads];
  K lo = part[0], hi = part[1];
  for (int b = threadIdx.x; b < nb; b += kRedThreads) {
    lo = kmin(lo, part[2 * b]);
    hi = kmax(hi, part[2 * b + 1]);
  }
  block_minmax(lo, hi, s_lo, s_hi);
  if (threadIdx.x == 0) {
    out[0] = lo;
    out[1] = hi;
  }
}

__global__ void k_fill_i32(int32_t* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0;
}

__global__ void k_copy_i32(const int32_t* src, int32_t* dst) {
  if (threadIdx.x == 0) dst[0] = src[0];
}

__global__ void k_fill_u64(unsigned long long* p, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = 0ull;
}

// Bin of x: integer mode (edges == nullptr) x - lo for x in [lo, lo + nbins);
// float mode the i with edges[i] <= x < edges[i + 1], the last bin closed
// (numpy.histogram's rule), values outside [edges[0], edges[nbins]] and NaN
// dropped.
template <typename T>
__device__ __forceinline__ int64_t bin_of(T x, const double* edges, int64_t nbins, int64_t lo) {
  if constexpr (std::is_floating_point<T>::value) {
    const double v = (double)x;
    if (!(v >= edges[0] && v <= edges[nbins])) return -1;
    int64_t a = 0, b = nbins;  // invariant: edges[a] <= v, edges[b] > v or b = nbins
    while (b - a > 1) {
      const int64_t m = (a + b) >> 1;
      if (edges[m] <= v) a = m;
      else b = m;
    }
    return a;
  } else {
    const int64_t b = (int64_t)x - lo;
    return (b >= 0 && b < nbins) ? b : -1;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_hist_lds(const T* __restrict__ x, int64_t n,
                                                  const double* __restrict__ edges,
                                                  int64_t nbins, int64_t lo,
                                                  unsigned long long* __restrict__ counts) {
  __shared__ unsigned int h[kHistLdsBins];
  __shared__ double e[kHistLdsBins + 1];
  for (int i = threadIdx.x; i < nbins; i += 256) h[i] = 0u;
  const bool fmode = edges != nullptr;
  if (fmode)
    for (int i = threadIdx.x; i <= nbins; i += 256) e[i] = edges[i];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int64_t b = bin_of<T>(x[i], fmode ? e : nullptr, nbins, lo);
    if (b >= 0) atomicAdd(&h[b], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < nbins; i += 256)
    if (h[i]) atomicAdd(&counts[i], (unsigned long long)h[i]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_hist_global(const T* __restrict__ x, int64_t n,
                                                     const double* __restrict__ edges,
                                                     int64_t nbins, int64_t lo,
                                                     unsigned long long* __restrict__ counts) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int64_t b = bin_of<T>(x[i], edges, nbins, lo);
    if (b >= 0) atomicAdd(&counts[b], 1ull);
  }
}

// Per-block sums for threshold_li's float iteration: over the finite values
// v = x - vmin (rounded in T as the reference's `image -= image_min`),
// foreground v > t and background v <= t: sum and count of each.
template <typename T>
__global__ __launch_bounds__(kRedThreads) void k_li_part(const T* __restrict__ x, int64_t n,
                                                         double vmin, double t,
                                                         double* __restrict__ part) {
  __shared__ double s[4][kRedThreads];
  double sf = 0, cf = 0, sb = 0, cb = 0;
  for (int64_t i = (int64_t)blockIdx.x * kRedThreads + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kRedThreads) {
    const T xi = x[i];
    if (!isfinite((double)xi)) continue;
    const double v = (double)(T)(xi - (T)vmin);
    if (v > t) {
      sf += v;
      cf += 1;
    } else {
      sb += v;
      cb += 1;
    }
  }
  const int tid = threadIdx.x;
  s[0][tid] = sf;
  s[1][tid] = cf;
  s[2][tid] = sb;
  s[3][tid] = cb;
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (tid < o)
      for (int k = 0; k < 4; ++k) s[k][tid] += s[k][tid + o];
    __syncthreads();
  }
  if (tid < 4) part[4 * blockIdx.x + tid] = s[tid][0];
}

__global__ __launch_bounds__(kRedThreads) void k_li_final(const double* __restrict__ part, int nb,
                                                          double* __restrict__ out) {
  __shared__ double s[4][kRedThreads];
  const int tid = threadIdx.x;
  double a[4] = {0, 0, 0, 0};
  for (int b = tid; b < nb; b += kRedThreads)
    for (int k = 0; k < 4; ++k) a[k] += part[4 * b + k];
  for (int k = 0; k < 4; ++k) s[k][tid] = a[k];
  __syncthreads();
  for (int o = kRedThreads / 2; o > 0; o >>= 1) {
    if (tid < o)
      for (int k = 0; k < 4; ++k) s[k][tid] += s[k][tid + o];
    __syncthreads();
  }
  if (tid < 4) out[tid] = s[tid][0];
}

// ---------------------------------------------------------------- label --
// Union-find over the voxel grid (one int32 parent per voxel, stored in the
// output array).  Foreground voxels (key != background) start as their own
// roots; every foreground voxel unites with each foreground neighbour of equal
// key across the structure's "backward" offsets (the first nonzero coordinate
// of the offset is -1; the other half of the structure is covered by the
// neighbour's own pass).  Roots only ever point to smaller indices, so each
// component's root is its smallest linear index, and ranking the roots in
// index order numbers the components in raster order of their first voxel.
constexpr int kLabelDirs = 64;  // offsets per launch

struct LabelArgs {
  const void* key;
  int key_dt;
  int64_t bg;                    // background key
  int32_t* y;                    // parents, then labels
  int64_t n;
  int ndim;
  int64_t shape[HC_MAX_NDIM];
  int64_t stride[HC_MAX_NDIM];   // element strides of the C-contiguous grid
  int ndirs;
  int8_t dirs[kLabelDirs][HC_MAX_NDIM];
};

// Integer key of element i: integers as int64 (uint64 bit patterns kept),
// floats truncated toward zero as an intp conversion does.
__device__ __forceinline__ int64_t key_of(const void* p, int64_t i, int dt) {
  switch (dt) {
    case HC_BOOL: case HC_UINT8: return ((const uint8_t*)p)[i];
    case HC_INT8: return ((const int8_t*)p)[i];
    case HC_UINT16: return ((const uint16_t*)p)[i];
    case HC_INT16: return ((const int16_t*)p)[i];
    case HC_UINT32: return ((const uint32_t*)p)[i];
    case HC_INT32: return ((const int32_t*)p)[i];
    case HC_UINT64: case HC_INT64: return ((const int64_t*)p)[i];
    case HC_FLOAT32: {
      const float v = ((const float*)p)[i];
      return v != v ? 0 : (int64_t)v;
    }
    case HC_FLOAT64: {
      const double v = ((const double*)p)[i];
      return v != v ? 0 : (int64_t)v;
    }
    default: return 0;
  }
}

__device__ __forceinline__ int32_t ld_parent(const int32_t* y, int32_t i) {
  return __hip_atomic_load(y + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Root of x with path halving (each visited voxel is pointed at its
// grandparent; parents only decrease, so the writes keep every tree intact).
__device__ __forceinline__ int32_t uf_find(int32_t* y, int32_t x) {
  int32_t cur = ld_parent(y, x);
  if (cur != x) {
    int32_t prev = x, next;
    while (cur > (next = ld_parent(y, cur))) {
      __hip_atomic_store(y + prev, next, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      prev = cur;
      cur = next;
    }
  }
  return cur;
}

// Root of x without writes: used once every union is done, when each voxel
// then stores its root.  (Path halving there could overwrite a voxel's stored
// root with an older, non-root ancestor read before that store.)
__device__ __forceinline__ int32_t uf_root(const int32_t* y, int32_t x) {
  int32_t cur = x, next;
  while ((next = ld_parent(y, cur)) != cur) cur = next;
  return cur;
}

__device__ __forceinline__ void uf_unite(int32_t* y, int32_t a, int32_t b) {
  int32_t ra = uf_find(y, a), rb = uf_find(y, b);
  while (ra != rb) {
    if (ra < rb) {  // hook the larger root under the smaller
      const int32_t t = ra;
      ra = rb;
      rb = t;
    }
    const int32_t old = atomicCAS(y + ra, ra, rb);
    if (old == ra) break;
    ra = uf_find(y, old);
  }
}

__global__ __launch_bounds__(256) void k_label_init(const LabelArgs a) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256)
    a.y[i] = key_of(a.key, i, a.key_dt) != a.bg ? (int32_t)i : -1;
}

__global__ __launch_bounds__(256) void k_label_unite(const LabelArgs a) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    const int64_t ki = key_of(a.key, i, a.key_dt);
    if (ki == a.bg) continue;
    int64_t c[HC_MAX_NDIM];
    int64_t r = i;
    for (int d = a.ndim - 1; d >= 0; --d) {
      c[d] = r % a.shape[d];
      r /= a.shape[d];
    }
    for (int k = 0; k < a.ndirs; ++k) {
      int64_t j = i;
      bool ok = true;
      for (int d = 0; d < a.ndim; ++d) {
        const int64_t q = c[d] + a.dirs[k][d];
        ok &= q >= 0 && q < a.shape[d];
        j += a.dirs[k][d] * a.stride[d];
      }
      if (!ok || key_of(a.key, j, a.key_dt) != ki) continue;
      uf_unite(a.y, (int32_t)i, (int32_t)j);
    }
  }
}

// Ranks of the roots: per 2048-voxel block an exclusive count of the roots
// before each voxel (stored for every voxel; only roots' entries are read),
// the block's root total in part[b].
constexpr int kScanItems = 8;
constexpr int kScanBlock = 256 * kScanItems;

__global__ __launch_bounds__(256) void k_label_flatten_scan(int32_t* __restrict__ y, int64_t n,
                                                            int32_t* __restrict__ rank,
                                                            int32_t* __restrict__ part) {
  __shared__ int32_t s[256];
  const int64_t base = (int64_t)blockIdx.x * kScanBlock + (int64_t)threadIdx.x * kScanItems;
  int32_t f[kScanItems];
  int32_t tot = 0;
#pragma unroll
  for (int u = 0; u < kScanItems; ++u) {
    const int64_t i = base + u;
    int32_t isroot = 0;
    if (i < n) {
      const int32_t p = y[i];
      if (p >= 0) {
        const int32_t root = uf_root(y, (int32_t)i);
        // compressed: every voxel points at its root (only this thread
        // writes y[i] in this kernel; others only read it, as an ancestor)
        __hip_atomic_store(y + i, root, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        isroot = root == (int32_t)i;
      }
    }
    f[u] = tot;
    tot += isroot;
  }
  const int tid = threadIdx.x;
  s[tid] = tot;
  __syncthreads();
  for (int o = 1; o < 256; o <<= 1) {  // inclusive scan of the thread totals
    const int32_t v = tid >= o ? s[tid - o] : 0;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  const int32_t before = tid ? s[tid - 1] : 0;
#pragma unroll
  for (int u = 0; u < kScanItems; ++u)
    if (base + u < n) rank[base + u] = before + f[u];
  if (tid == 255) part[blockIdx.x] = s[255];
}

// Exclusive scan of the block totals (one work-group; each thread takes a
// contiguous run), the component count into part[nb].
__global__ __launch_bounds__(1024) void k_label_scan_parts(int32_t* __restrict__ part, int nb) {
  __shared__ int32_t s[1024];
  const int tid = threadIdx.x;
  const int per = (nb + 1023) / 1024;
  const int b0 = tid * per, b1 = min(nb, b0 + per);
  int32_t t = 0;
  for (int b = b0; b < b1; ++b) t += part[b];
  s[tid] = t;
  __syncthreads();
  for (int o = 1; o < 1024; o <<= 1) {
    const int32_t v = tid >= o ? s[tid - o] : 0;
    __syncthreads();
    s[tid] += v;
    __syncthreads();
  }
  int32_t run = tid ? s[tid - 1] : 0;
  for (int b = b0; b < b1; ++b) {
    const int32_t v = part[b];
    part[b] = run;
    run += v;
  }
  if (tid == 1023) part[nb] = s[1023];
}

__global__ __launch_bounds__(256) void k_label_final(int32_t* __restrict__ y, int64_t n,
                                                     const int32_t* __restrict__ rank,
                                                     const int32_t* __restrict__ part) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int32_t r = y[i];
    y[i] = r < 0 ? 0 : rank[r] + part[r / kScanBlock] + 1;
  }
}

// Scratch of the allocation-free hc_minmax_noalloc: one (min, max) pair of
// 8-byte keys per partial block.  Each GPU holds its own copy until the
// process ends; host code reaches it only through the address
// hipGetSymbolAddress returns on that device.
constexpr int kNoAllocScratchBytes = kRedMaxBlocks * 16;
__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// The current device's address of `scratch`, looked up on the first call
// per device (with that device current) and kept.
int scratch_ptr(void** out) {
  int dev = 0;
  HC_HIP(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) return fail(HC_ERR_UNSUPPORTED, "device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    HC_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return HC_OK;
}

int red_blocks(int64_t n) {
  return (int)std::max<int64_t>(1, std::min<int64_t>(kRedMaxBlocks, (n + 2047) / 2048));
}

}  // namespace

}  // namespace hc

using namespace hc;

extern "C" {

int hc_minmax(const hc_array* in, void* result_dev, void* work_dev, int64_t work_bytes,
              void* stream) {
  if (!in || !result_dev || !work_dev) return fail(HC_ERR_VALUE, "null argument");
  if (!is_c_contiguous(in)) return fail(HC_ERR_VALUE, "input must be C-contiguous");
  const int64_t n = numel(in);
  if (n == 0) return fail(HC_ERR_VALUE, "zero-size array has no minimum");
  const int nb = red_blocks(n);
  if (work_bytes < (int64_t)nb * 16) return fail(HC_ERR_VALUE, "work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  const void* x = in->data;
#define HC_MINMAX(T, K)                                                                      \
  do {                                                                                       \
    HC_LAUNCH("k_minmax_part", st, (k_minmax_part<T, K>), dim3(nb), dim3(kRedThreads), 0,   \
              (const T*)x, n, (K*)work_dev);                                                 \
    HC_LAUNCH("k_minmax_final", st, k_minmax_final<K>, dim3(1), dim3(kRedThreads), 0,       \
              (const K*)work_dev, nb, (K*)result_dev);                                       \
  } while (0)
  switch (in->dtype) {
    case HC_FLOAT32: HC_MINMAX(float, double); break;
    case HC_FLOAT64: HC_MINMAX(double, double); break;
    case HC_BOOL: case HC_UINT8: HC_MINMAX(uint8_t, uint64_t); break;
    case HC_UINT16: HC_MINMAX(uint16_t, uint64_t); break;
    case HC_UINT32: HC_MINMAX(uint32_t, uint64_t); break;
    case HC_UINT64: HC_MINMAX(uint64_t, uint64_t); break;
    case HC_INT8: HC_MINMAX(int8_t, int64_t); break;
    case HC_INT16: HC_MINMAX(int16_t, int64_t); break;
    case HC_INT32: HC_MINMAX(int32_t, int64_t); break;
    case HC_INT64: HC_MINMAX(int64_t, int64_t); break;
    default: return fail(HC_ERR_VALUE, "minmax: unsupported dtype");
  }
#undef HC_MINMAX
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_minmax_noalloc(const hc_array* in, void* result_dev, void* stream) {
  if (!in || !result_dev) return fail(HC_ERR_VALUE, "null argument");
  if (numel(in) > 0 && red_blocks(numel(in)) * 16 > kNoAllocScratchBytes)
    return fail(HC_ERR_VALUE, "partials exceed the static scratch");
  void* p = nullptr;
  const int rc = scratch_ptr(&p);
  if (rc != HC_OK) return rc;
  return hc_minmax(in, result_dev, p, kNoAllocScratchBytes, stream);
}

int hc_histogram(const hc_array* in, int64_t nbins, const double* edges_dev, int64_t lo,
                 uint64_t* counts_dev, void* stream) {
  if (!in || !counts_dev) return fail(HC_ERR_VALUE, "null argument");
  if (!is_c_contiguous(in)) return fail(HC_ERR_VALUE, "input must be C-contiguous");
  if (nbins < 1) return fail(HC_ERR_VALUE, "nbins must be positive");
  const bool fl = in->dtype == HC_FLOAT32 || in->dtype == HC_FLOAT64;
  if (fl != (edges_dev != nullptr))
    return fail(HC_ERR_VALUE, "float images take bin edges, integer images an offset");
  const int64_t n = numel(in);
  hipStream_t st = (hipStream_t)stream;
  unsigned long long* c = (unsigned long long*)counts_dev;
  HC_LAUNCH("k_fill_u64", st, k_fill_u64, dim3((unsigned)std::min<int64_t>(1024, (nbins + 255) / 256)), dim3(256), 0, c, nbins);
  if (n == 0) return HC_OK;
  const bool lds = nbins <= kHistLdsBins;
  const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>(2048, (n + 4095) / 4096)));
  const void* x = in->data;
#define HC_HIST(T)                                                                          \
  do {                                                                                      \
    if (lds)                                                                                \
      HC_LAUNCH("k_hist_lds", st, k_hist_lds<T>, g, dim3(256), 0, (const T*)x, n, edges_dev, \
                nbins, lo, c);                                                              \
    else                                                                                    \
      HC_LAUNCH("k_hist_global", st, k_hist_global<T>, g, dim3(256), 0, (const T*)x, n,     \
                edges_dev, nbins, lo, c);                                                   \
  } while (0)
  switch (in->dtype) {
    case HC_FLOAT32: HC_HIST(float); break;
    case HC_FLOAT64: HC_HIST(double); break;
    case HC_BOOL: case HC_UINT8: HC_HIST(uint8_t); break;
    case HC_UINT16: HC_HIST(uint16_t); break;
    case HC_UINT32: HC_HIST(uint32_t); break;
    case HC_UINT64: HC_HIST(uint64_t); break;
    case HC_INT8: HC_HIST(int8_t); break;
    case HC_INT16: HC_HIST(int16_t); break;
    case HC_INT32: HC_HIST(int32_t); break;
    case HC_INT64: HC_HIST(int64_t); break;
    default: return fail(HC_ERR_VALUE, "histogram: unsupported dtype");
  }
#undef HC_HIST
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int64_t hc_label_work_elems(int64_t n) {
  // a rank per voxel, the block totals and the component count
  return n < 0 ? -1 : n + (n + kScanBlock - 1) / kScanBlock + 1;
}

int hc_label(const hc_array* in, int64_t background, int connectivity, const hc_array* out,
             int32_t* work_dev, int64_t work_elems, int32_t* count_dev, void* stream) {
  if (!in || !out || !work_dev) return fail(HC_ERR_VALUE, "null argument");
  const int nd = in->ndim;
  if (nd < 1 || nd > HC_MAX_NDIM) return fail(HC_ERR_VALUE, "rank out of range");
  if (!same_shape(in, out) || !is_c_contiguous(in) || !is_c_contiguous(out) ||
      out->dtype != HC_INT32)
    return fail(HC_ERR_VALUE, "labels: int32, C-contiguous, the input's shape");
  if (connectivity < 1 || connectivity > nd)
    return fail(HC_ERR_VALUE, "Connectivity below 1 or above ndim is illegal.");
  const int64_t n = numel(in);
  if (n >= ((int64_t)1 << 31) - 1) return fail(HC_ERR_UNSUPPORTED, "more than 2^31 - 2 voxels");
  const int64_t nb = (n + kScanBlock - 1) / kScanBlock;
  if (work_elems < n + nb + 1) return fail(HC_ERR_VALUE, "work buffer too small");
  if (in->dtype > HC_FLOAT64 || in->dtype == HC_FLOAT16)
    return fail(HC_ERR_VALUE, "labels: unsupported dtype");
  hipStream_t st = (hipStream_t)stream;
  int32_t* y = (int32_t*)out->data;
  int32_t* rank = work_dev;
  int32_t* part = work_dev + n;
  if (n == 0) {
    if (count_dev) HC_LAUNCH("k_fill_i32", st, k_fill_i32, dim3(1), dim3(64), 0, count_dev, (int64_t)1);
    return HC_OK;
  }
  LabelArgs a;
  std::memset(&a, 0, sizeof(a));
  a.key = in->data;
  a.key_dt = in->dtype;
  a.bg = background;
  a.y = y;
  a.n = n;
  a.ndim = nd;
  int64_t s = 1;
  for (int d = nd - 1; d >= 0; --d) {
    a.shape[d] = in->shape[d];
    a.stride[d] = s;
    s *= in->shape[d];
  }
  const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 8192)));
  HC_LAUNCH("k_label_init", st, k_label_init, g, dim3(256), 0, a);
  // backward offsets of generate_binary_structure(nd, connectivity): at most
  // `connectivity` nonzero coordinates, the first nonzero one -1
  int total = 1;
  for (int d = 0; d < nd; ++d) total *= 3;
  for (int code = 0; code < total; ++code) {
    int8_t o[HC_MAX_NDIM];
    int c = code, nz = 0, first = 0;
    for (int d = nd - 1; d >= 0; --d) {
      o[d] = (int8_t)(c % 3 - 1);
      c /= 3;
    }
    for (int d = 0; d < nd; ++d)
      if (o[d]) {
        if (!nz) first = o[d];
        ++nz;
      }
    if (nz == 0 || nz > connectivity || first != -1) continue;
    std::memcpy(a.dirs[a.ndirs], o, sizeof(o));
    if (++a.ndirs == kLabelDirs) {
      HC_LAUNCH("k_label_unite", st, k_label_unite, g, dim3(256), 0, a);
      a.ndirs = 0;
    }
  }
  if (a.ndirs) HC_LAUNCH("k_label_unite", st, k_label_unite, g, dim3(256), 0, a);
  HC_LAUNCH("k_label_flatten_scan", st, k_label_flatten_scan, dim3((unsigned)nb), dim3(256), 0, y, n, rank, part);
  HC_LAUNCH("k_label_scan_parts", st, k_label_scan_parts, dim3(1), dim3(1024), 0, part, (int)nb);
  HC_LAUNCH("k_label_final", st, k_label_final, g, dim3(256), 0, y, n, rank, part);
  if (count_dev)
    HC_LAUNCH("k_copy_i32", st, k_copy_i32, dim3(1), dim3(64), 0, part + nb, count_dev);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_li_means(const hc_array* in, double vmin, double t, double* out4_dev, void* work_dev,
                int64_t work_bytes, void* stream) {
  if (!in || !out4_dev || !work_dev) return fail(HC_ERR_VALUE, "null argument");
  if (!is_c_contiguous(in)) return fail(HC_ERR_VALUE, "input must be C-contiguous");
  const int64_t n = numel(in);
  const int nb = red_blocks(n);
  if (work_bytes < (int64_t)nb * 32) return fail(HC_ERR_VALUE, "work buffer too small");
  hipStream_t st = (hipStream_t)stream;
  if (in->dtype == HC_FLOAT32)
    HC_LAUNCH("k_li_part", st, k_li_part<float>, dim3(nb), dim3(kRedThreads), 0, (const float*)in->data, n, vmin, t, (double*)work_dev);
  else if (in->dtype == HC_FLOAT64)
    HC_LAUNCH("k_li_part", st, k_li_part<double>, dim3(nb), dim3(kRedThreads), 0, (const double*)in->data, n, vmin, t, (double*)work_dev);
  else
    return fail(HC_ERR_VALUE, "threshold_li's float iteration needs float32/float64");
  HC_LAUNCH("k_li_final", st, k_li_final, dim3(1), dim3(kRedThreads), 0, (const double*)work_dev, nb, out4_dev);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

}  // extern "C"
