// Sato tubeness, Meijering neuriteness and the hybrid Hessian filter's
// final step for gfx950 (reference: python/cucim/src/cucim/skimage/
// filters/ridges.py:50-225 and :479-565).
//
// Both filters take the Hessian elements of one sigma (hessian_matrix with
// Gaussian derivatives), form the closed-form eigenvalues sorted descending
// (feature/corner.py:824-856, hessian_matrix_eigvals), zero the ones below
// 1e-12 in magnitude, and fold the per-voxel response into the running
// maximum over sigmas (cp.maximum, NaN-propagating; the maximum starts at 0).
//
//   k_sato       v = prod(max(l_i, 0) for all but the smallest)^(1/(ndim-1))
//                    * sigma^2
//   k_meij_vals  v_i = l_i + alpha * sum_{j != i} l_j; v = max(v_k, 0) for the
//                first k of largest |v_k|; the maximum of v over the image is
//                reduced into scratch (non-negative floats and NaN order like
//                their bit patterns, so an unsigned atomic max is the max)
//   k_meij_fold  out = max(out, v / m if m > 0 else v)
//   k_fill_le    a[a <= t] = value  (the hybrid Hessian filter's last line)
#include <algorithm>

#include "hc_common.h"
#include "hc_eig.h"

namespace hc {

template <typename W>
struct Bits;
template <>
struct Bits<float> {
  typedef unsigned int U;
  __device__ static U of(float v) { return __float_as_uint(v); }
  __device__ static float to(U u) { return __uint_as_float(u); }
};
template <>
struct Bits<double> {
  typedef unsigned long long U;
  __device__ static U of(double v) { return (U)__double_as_longlong(v); }
  __device__ static double to(U u) { return __longlong_as_double((long long)u); }
};

template <typename W>
struct SArgs {
  const W* s[6];
  int64_t n;
};

template <typename W, int ND>
__device__ __forceinline__ void eig_desc(const SArgs<W>& a, int64_t i, W (&l)[3]) {
  if (ND == 3) {
    // (S0..S5) = (M00, M01, M02, M11, M12, M22) -> (a, d, f, b, e, c)
    eig33<W>(a.s[0][i], a.s[3][i], a.s[5][i], a.s[1][i], a.s[4][i], a.s[2][i], 1, 0, l[0],
             l[1], l[2]);
  } else {
    eig22<W>(a.s[0][i], a.s[1][i], a.s[2][i], 1, 0, l[0], l[1]);
    l[2] = (W)0;
  }
#pragma unroll
  for (int k = 0; k < ND; ++k)
    if (fabs((double)l[k]) < 1e-12) l[k] = (W)0;
}

template <typename W>
__device__ __forceinline__ W clip0(W v) {
  return (v > (W)0 || v != v) ? v : (W)0;  // +0 for v <= 0 (no -0), NaN kept
}

template <typename W, int ND>
__global__ __launch_bounds__(256) void k_sato(const SArgs<W> a, W* __restrict__ out, W sig2,
                                              int first) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    W l[3];
    eig_desc<W, ND>(a, i, l);
    W v;
    if (ND == 3) {
      const W p = clip0(l[0]) * clip0(l[1]);
      v = pow(p, (W)0.5);
    } else {
      v = clip0(l[0]);  // x ** 1.0
    }
    v = v * sig2;
    out[i] = nanmax(first ? (W)0 : out[i], v);
  }
}

template <typename W, int ND>
__global__ __launch_bounds__(256) void k_meij_vals(const SArgs<W> a, W* __restrict__ vals,
                                                   W alpha,
                                                   typename Bits<W>::U* __restrict__ mx) {
  typedef typename Bits<W>::U U;
  __shared__ U red[256];
  U m = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    W l[3];
    eig_desc<W, ND>(a, i, l);
    W best = (W)0, babs = (W)-1;
    bool nan_seen = false;
#pragma unroll
    for (int r = 0; r < ND; ++r) {
      W acc = (W)0;
#pragma unroll
      for (int c = 0; c < ND; ++c) acc = fma_(r == c ? (W)1 : alpha, l[c], acc);
      const W aa = fabs(acc);
      // argmax of |v|: the first maximum, a NaN counting as the maximum
      if (!nan_seen && (aa != aa || aa > babs)) {
        best = acc;
        babs = aa;
        nan_seen = aa != aa;
      }
    }
    const W v = clip0(best);
    vals[i] = v;
    const U b = Bits<W>::of(v);
    m = b > m ? b : m;
  }
  red[threadIdx.x] = m;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s && red[threadIdx.x + s] > red[threadIdx.x])
      red[threadIdx.x] = red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicMax(mx, red[0]);
}

template <typename W>
__global__ __launch_bounds__(256) void k_meij_fold(const W* __restrict__ vals, W* __restrict__ out,
                                                   int64_t n,
                                                   const typename Bits<W>::U* __restrict__ mx,
                                                   int first) {
  const W m = Bits<W>::to(*mx);
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    W v = vals[i];
    if (m > (W)0) v = v / m;
    out[i] = nanmax(first ? (W)0 : out[i], v);
  }
}

template <typename U>
__global__ void k_zero_word(U* p) {
  if (threadIdx.x == 0 && blockIdx.x == 0) p[0] = (U)0;
}

template <typename W>
__global__ __launch_bounds__(256) void k_fill_le(W* __restrict__ a, int64_t n, W t, W value) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    if (a[i] <= t) a[i] = value;
}

namespace {

int check_hessian(const hc_array* H, int nH, const hc_array* out) {
  if (!H || !out) return fail(HC_ERR_VALUE, "null argument");
  if (nH != 3 && nH != 6) return fail(HC_ERR_VALUE, "2-D or 3-D Hessian expected");
  const int dt = out->dtype;
  if (dt != HC_FLOAT32 && dt != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "output must be float32 or float64");
  if (!is_c_contiguous(out)) return fail(HC_ERR_VALUE, "output must be C-contiguous");
  for (int k = 0; k < nH; ++k)
    if (!same_shape(&H[k], out) || !is_c_contiguous(&H[k]) || H[k].dtype != dt)
      return fail(HC_ERR_VALUE, "Hessian elements must be C-contiguous arrays like the output");
  return HC_OK;
}

template <typename W>
SArgs<W> sargs(const hc_array* H, int nH) {
  SArgs<W> a;
  for (int k = 0; k < 6; ++k) a.s[k] = (const W*)H[k < nH ? k : 0].data;
  a.n = numel(&H[0]);
  return a;
}

int64_t grid_of(int64_t n) { return std::min<int64_t>((n + 255) / 256, 256 * 64); }

template <typename W>
int sato_t(const hc_array* H, int nH, const hc_array* out, double sig2, int first,
           hipStream_t st) {
  const SArgs<W> a = sargs<W>(H, nH);
  if (a.n == 0) return HC_OK;
  W* o = (W*)out->data;
  if (nH == 6)
    HC_LAUNCH("k_sato", st, (k_sato<W, 3>), dim3((unsigned)grid_of(a.n)), dim3(256), 0, a, o,
              (W)sig2, first);
  else
    HC_LAUNCH("k_sato", st, (k_sato<W, 2>), dim3((unsigned)grid_of(a.n)), dim3(256), 0, a, o,
              (W)sig2, first);
  return HC_OK;
}

template <typename W>
int meij_t(const hc_array* H, int nH, const hc_array* out, const hc_array* tmp, double alpha,
           int first, void* scratch, hipStream_t st) {
  typedef typename Bits<W>::U U;
  const SArgs<W> a = sargs<W>(H, nH);
  if (a.n == 0) return HC_OK;
  U* mx = (U*)scratch;
  W* v = (W*)tmp->data;
  HC_LAUNCH("k_zero_word", st, k_zero_word<U>, dim3(1), dim3(64), 0, mx);
  if (nH == 6)
    HC_LAUNCH("k_meij_vals", st, (k_meij_vals<W, 3>), dim3((unsigned)grid_of(a.n)), dim3(256),
              0, a, v, (W)alpha, mx);
  else
    HC_LAUNCH("k_meij_vals", st, (k_meij_vals<W, 2>), dim3((unsigned)grid_of(a.n)), dim3(256),
              0, a, v, (W)alpha, mx);
  HC_LAUNCH("k_meij_fold", st, k_meij_fold<W>, dim3((unsigned)grid_of(a.n)), dim3(256), 0, v,
            (W*)out->data, a.n, mx, first);
  return HC_OK;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int hc_sato_fold(const hc_array* H, int nH, const hc_array* out, double sigma, int first_sigma,
                 void* stream) {
  int rc = check_hessian(H, nH, out);
  if (rc != HC_OK) return rc;
  const double sig2 = sigma * sigma;
  hipStream_t st = (hipStream_t)stream;
  return out->dtype == HC_FLOAT64 ? sato_t<double>(H, nH, out, sig2, first_sigma, st)
                                  : sato_t<float>(H, nH, out, sig2, first_sigma, st);
}

int hc_meijering_fold(const hc_array* H, int nH, const hc_array* out, const hc_array* tmp,
                      double alpha, int first_sigma, void* scratch, void* stream) {
  int rc = check_hessian(H, nH, out);
  if (rc != HC_OK) return rc;
  if (!tmp || !scratch) return fail(HC_ERR_VALUE, "null argument");
  if (!same_shape(tmp, out) || !is_c_contiguous(tmp) || tmp->dtype != out->dtype)
    return fail(HC_ERR_VALUE, "tmp must be a C-contiguous array like the output");
  hipStream_t st = (hipStream_t)stream;
  return out->dtype == HC_FLOAT64
             ? meij_t<double>(H, nH, out, tmp, alpha, first_sigma, scratch, st)
             : meij_t<float>(H, nH, out, tmp, alpha, first_sigma, scratch, st);
}

int hc_fill_le(const hc_array* a, double threshold, double value, void* stream) {
  if (!a) return fail(HC_ERR_VALUE, "null argument");
  if (!is_c_contiguous(a)) return fail(HC_ERR_VALUE, "array must be C-contiguous");
  const int64_t n = numel(a);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  if (a->dtype == HC_FLOAT64)
    HC_LAUNCH("k_fill_le", st, k_fill_le<double>, dim3((unsigned)grid_of(n)), dim3(256), 0,
              (double*)a->data, n, threshold, value);
  else if (a->dtype == HC_FLOAT32)
    HC_LAUNCH("k_fill_le", st, k_fill_le<float>, dim3((unsigned)grid_of(n)), dim3(256), 0,
              (float*)a->data, n, (float)threshold, (float)value);
  else
    return fail(HC_ERR_VALUE, "array must be float32 or float64");
  return HC_OK;
}

}  // extern "C"
