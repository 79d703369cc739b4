// Gaussian kernels built inside the library (host code), and the entry points
// that take sigmas instead of kernels: gaussian_filter, hessian_matrix and
// frangi for C callers that do not run the reference's Python kernel code.
//
// The kernel is _gaussian_kernel1d of _vendored/_ndimage_filters.py:689-717,
// evaluated in float64 the way numpy evaluates that function: phi(x) =
// exp((-0.5 / sigma^2) * x^2) over x = -r..r, divided by its pairwise sum
// (numpy's summation order: 8 interleaved partial sums in blocks of at most
// 128, halves above); for order n the coefficient recurrence
// q <- (D + P) q and q(x) = sum_k x^k q_k with fused multiply-adds in index
// order (the float64 matrix-vector products numpy hands to BLAS); the kernel
// is (q phi) reversed.  The library's exp is not numpy's: the float64 taps
// can differ from the Python path's in the last bit (float32 taps then
// agree but for a 2^-28-per-tap chance).
#include <cmath>
#include <cstring>
#include <vector>

#include "hc_common.h"

extern "C" {
int hc_correlate1d(const hc_array* in, const hc_array* out, const hc_taps* taps, int axis,
                   int mode, double cval, int algorithm, void* stream);
int hc_separable2d(const hc_array* in, const hc_array* out, const hc_taps* taps0,
                   const hc_taps* taps1, int mode0, int mode1, double cval, void* stream);
int hc_copy_cast(const hc_array* in, const hc_array* out, void* stream);
int hc_max_inline_taps(int w_f64);
}

namespace hc {
namespace {

double pairwise_sum(const double* a, int64_t n) {
  if (n < 8) {
    double r = 0.0;
    for (int64_t i = 0; i < n; ++i) r += a[i];
    return r;
  }
  if (n <= 128) {
    double r[8];
    for (int j = 0; j < 8; ++j) r[j] = a[j];
    int64_t i = 8;
    for (; i < n - (n % 8); i += 8)
      for (int j = 0; j < 8; ++j) r[j] += a[i + j];
    double res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]));
    for (; i < n; ++i) res += a[i];
    return res;
  }
  int64_t n2 = n / 2;
  n2 -= n2 % 8;
  return pairwise_sum(a, n2) + pairwise_sum(a + n2, n - n2);
}

void kernel1d(double sigma, int order, int64_t radius, std::vector<double>& out) {
  const int64_t n = 2 * radius + 1;
  const double sigma2 = sigma * sigma;
  const double c = -0.5 / sigma2;
  std::vector<double> phi((size_t)n);
  for (int64_t m = 0; m < n; ++m) {
    const double x = (double)(m - radius);
    phi[(size_t)m] = std::exp(c * (x * x));
  }
  const double s = pairwise_sum(phi.data(), n);
  for (auto& v : phi) v /= s;
  out.assign((size_t)n, 0.0);
  if (order == 0) {
    out = phi;
    return;
  }
  // q <- (D + P) q, D[i][i+1] = i + 1, P[i][i-1] = 1 / -sigma^2
  const int no = order + 1;
  std::vector<double> q((size_t)no, 0.0), t((size_t)no);
  q[0] = 1.0;
  const double p = 1.0 / -sigma2;
  for (int it = 0; it < order; ++it) {
    for (int i = 0; i < no; ++i) {
      double acc = 0.0;
      for (int j = 0; j < no; ++j) {
        const double Q = (j == i + 1 ? (double)(i + 1) : 0.0) + (j == i - 1 ? p : 0.0);
        acc = std::fma(Q, q[(size_t)j], acc);
      }
      t[(size_t)i] = acc;
    }
    q = t;
  }
  for (int64_t m = 0; m < n; ++m) {
    const int64_t x = m - radius;
    double acc = 0.0, xk = 1.0;
    int64_t xi = 1;
    for (int k = 0; k < no; ++k) {
      xk = (double)xi;
      acc = std::fma(xk, q[(size_t)k], acc);
      xi *= x;
    }
    out[(size_t)(n - 1 - m)] = acc * phi[(size_t)m];
  }
}

// Round to the weights dtype (float32 unless f64).
inline double to_w(double v, bool f64) { return f64 ? v : (double)(float)v; }

// Taps of one axis kept alive for the duration of a call.
struct AxisKernel {
  std::vector<double> w;
  int anchor = 0;
};

// Hessian kernels of one axis (feature/corner.py:218-232): sigma / sqrt(2),
// radius int(truncate * s + 0.5), order 0 and 1 rounded to the dtype and
// trimmed to their common nonzero support.
void hessian_axis(double sigma, double truncate, bool f64, AxisKernel& k0, AxisKernel& k1) {
  const double sq1_2 = 1.0 / std::sqrt(2.0);
  const double ss = sq1_2 * sigma;
  const int64_t r = (int64_t)(truncate * ss + 0.5);
  std::vector<double> w0, w1;
  kernel1d(ss, 0, r, w0);
  kernel1d(ss, 1, r, w1);
  for (auto& v : w0) v = to_w(v, f64);
  for (auto& v : w1) v = to_w(v, f64);
  int64_t lo = -1, hi = -1;
  for (int64_t i = 0; i < (int64_t)w0.size(); ++i)
    if (w0[(size_t)i] != 0.0 || w1[(size_t)i] != 0.0) {
      if (lo < 0) lo = i;
      hi = i + 1;
    }
  if (lo < 0) {
    k0.w.assign(1, 0.0);
    k1.w.assign(1, 0.0);
    k0.anchor = k1.anchor = 0;
    return;
  }
  k0.w.assign(w0.begin() + lo, w0.begin() + hi);
  k1.w.assign(w1.begin() + lo, w1.begin() + hi);
  k0.anchor = k1.anchor = (int)(r - lo);
}

hc_taps taps_of(const AxisKernel& k, bool f64) {
  hc_taps t;
  std::memset(&t, 0, sizeof(t));
  t.w = k.w.data();
  t.K = (int64_t)k.w.size();
  t.anchor = k.anchor;
  t.w_f64 = f64 ? 1 : 0;
  t.w_dev = nullptr;
  return t;
}

bool weights_f64(int in_dt) {
  // promote(input dtype, float32): float64 for float64 and 32/64-bit integers
  return in_dt == HC_FLOAT64 || in_dt == HC_INT32 || in_dt == HC_UINT32 || in_dt == HC_INT64 ||
         in_dt == HC_UINT64;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int hc_gaussian_kernel1d(double sigma, int order, int64_t radius, double* out) {
  if (!out) return fail(HC_ERR_VALUE, "null output");
  if (order < 0) return fail(HC_ERR_VALUE, "order must be non-negative");
  if (radius < 0) return fail(HC_ERR_VALUE, "radius must be non-negative");
  std::vector<double> w;
  kernel1d(sigma, order, radius, w);
  std::memcpy(out, w.data(), w.size() * sizeof(double));
  return HC_OK;
}

int hc_gaussian_nd(const hc_array* in, const hc_array* out, const double* sigma,
                   const int* order, const int64_t* radius, const int* mode, double cval,
                   double truncate, const hc_array* work, void* stream) {
  if (!in || !out || !sigma || !order || !mode) return fail(HC_ERR_VALUE, "null argument");
  const int nd = in->ndim;
  if (nd < 1 || nd > HC_MAX_NDIM) return fail(HC_ERR_VALUE, "rank out of range");
  const bool f64 = weights_f64(in->dtype);
  std::vector<AxisKernel> ks;
  std::vector<int> axes;
  for (int d = 0; d < nd; ++d) {
    if (sigma[d] < 0) return fail(HC_ERR_VALUE, "sigma must be non-negative");
    if (order[d] < 0) return fail(HC_ERR_VALUE, "order must be non-negative");
    if (sigma[d] < 1e-15) continue;  // _ndimage_filters.py:646-653
    const int64_t r =
        (radius && radius[d] >= 0) ? radius[d] : (int64_t)(truncate * sigma[d] + 0.5);
    if (r <= 0) continue;
    AxisKernel k;
    kernel1d(sigma[d], order[d], r, k.w);
    for (auto& v : k.w) v = to_w(v, f64);
    k.anchor = (int)r;
    // leading / trailing zero taps contribute nothing (zero weights are
    // skipped), so they are trimmed as the Python layer does
    int64_t lo = -1, hi = -1;
    for (int64_t i = 0; i < (int64_t)k.w.size(); ++i)
      if (k.w[(size_t)i] != 0.0) {
        if (lo < 0) lo = i;
        hi = i + 1;
      }
    if (lo < 0) {
      k.w.assign(1, 0.0);
      k.anchor = 0;
    } else {
      k.w = std::vector<double>(k.w.begin() + lo, k.w.begin() + hi);
      k.anchor -= (int)lo;
    }
    if ((int64_t)k.w.size() > hc_max_inline_taps(f64 ? 1 : 0))
      return fail(HC_ERR_RESOURCE, "kernel longer than the inline tap limit (pass a device "
                                   "copy through hc_correlate1d)");
    ks.push_back(k);
    axes.push_back(d);
  }
  const int np = (int)axes.size();
  if (np == 0) return hc_copy_cast(in, out, stream);
  std::vector<hc_taps> t(np);
  for (int i = 0; i < np; ++i) t[(size_t)i] = taps_of(ks[(size_t)i], f64);
  if (np == 1) return hc_correlate1d(in, out, &t[0], axes[0], mode[axes[0]], cval, HC_ALGO_AUTO, stream);
  if (np == 2 && nd == 2 && in->dtype == HC_FLOAT32 && out->dtype == HC_FLOAT32 && !f64 &&
      t[0].K <= 33 && t[1].K <= 33) {
    const int rc = hc_separable2d(in, out, &t[0], &t[1], mode[0], mode[1], cval, stream);
    if (rc != HC_ERR_RESOURCE) return rc;
  }
  if (!work || !same_shape(work, out) || work->dtype != out->dtype || !is_c_contiguous(work))
    return fail(HC_ERR_VALUE, "more than one pass needs a work array like the output");
  // ping-pong between work and out so that the last pass lands in out
  const hc_array* cur = in;
  for (int i = 0; i < np; ++i) {
    const hc_array* dst = ((np - 1 - i) % 2 == 0) ? out : work;
    const int rc = hc_correlate1d(cur, dst, &t[(size_t)i], axes[(size_t)i],
                                  mode[axes[(size_t)i]], cval, HC_ALGO_AUTO, stream);
    if (rc != HC_OK) return rc;
    cur = dst;
  }
  return HC_OK;
}

int hc_hessian_gaussian(const hc_array* in, const hc_array* H, int nH, const hc_taps* taps_s,
                        const hc_taps* taps_d, int mode, double cval, int order_xy,
                        void* scratch, void* stream);
int hc_frangi(const hc_array* in, const hc_array* out, const hc_taps* taps_s,
              const hc_taps* taps_d, int nsig, double alpha, double beta, double gamma,
              int has_gamma, int black_ridges, int mode, double cval, const hc_array* work,
              int nwork, void* scratch, const double* center, void* stream);

int hc_hessian_matrix(const hc_array* in, const hc_array* H, int nH, const double* sigma,
                      int mode, double cval, int order_xy, void* scratch, void* stream) {
  if (!in || !sigma) return fail(HC_ERR_VALUE, "null argument");
  const int nd = in->ndim;
  if (nd < 1 || nd > HC_MAX_NDIM) return fail(HC_ERR_VALUE, "rank out of range");
  const bool f64 = in->dtype == HC_FLOAT64;
  bool all_gt1 = true;
  for (int d = 0; d < nd; ++d) all_gt1 &= sigma[d] > 1;
  const double truncate = all_gt1 ? 8.0 : 100.0;
  std::vector<AxisKernel> k0((size_t)nd), k1((size_t)nd);
  std::vector<hc_taps> ts((size_t)nd), td((size_t)nd);
  for (int d = 0; d < nd; ++d) {
    hessian_axis(sigma[d], truncate, f64, k0[(size_t)d], k1[(size_t)d]);
    ts[(size_t)d] = taps_of(k0[(size_t)d], f64);
    td[(size_t)d] = taps_of(k1[(size_t)d], f64);
  }
  return hc_hessian_gaussian(in, H, nH, ts.data(), td.data(), mode, cval, order_xy, scratch,
                             stream);
}

int hc_frangi_sigmas(const hc_array* in, const hc_array* out, const double* sigmas, int nsig,
                     double alpha, double beta, double gamma, int has_gamma, int black_ridges,
                     int mode, double cval, const hc_array* work, int nwork, void* scratch,
                     void* stream) {
  if (!in || !out || (nsig > 0 && !sigmas)) return fail(HC_ERR_VALUE, "null argument");
  if (nsig < 1) return fail(HC_ERR_VALUE, "sigmas must contain at least one value");
  const int nd = in->ndim;
  if (nd != 2 && nd != 3) return fail(HC_ERR_VALUE, "frangi handles 2-D and 3-D images");
  const bool f64 = out->dtype == HC_FLOAT64;
  std::vector<AxisKernel> k0((size_t)nsig * nd), k1((size_t)nsig * nd);
  std::vector<hc_taps> ts((size_t)nsig * nd), td((size_t)nsig * nd);
  for (int s = 0; s < nsig; ++s) {
    const double truncate = sigmas[s] > 1 ? 8.0 : 100.0;
    for (int d = 0; d < nd; ++d) {
      const size_t i = (size_t)s * nd + d;
      hessian_axis(sigmas[s], truncate, f64, k0[i], k1[i]);
      ts[i] = taps_of(k0[i], f64);
      td[i] = taps_of(k1[i], f64);
    }
  }
  return hc_frangi(in, out, ts.data(), td.data(), nsig, alpha, beta, gamma, has_gamma,
                   black_ridges, mode, cval, work, nwork, scratch, nullptr, stream);
}

}  // extern "C"
