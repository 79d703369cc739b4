// Host-side self-test of the C ABI, compiled (host code only) with
// AddressSanitizer and UndefinedBehaviorSanitizer by test_host_sanitize.py.
//
// It exercises what the library does on the host before any device work:
// Gaussian tap construction, batch planning, knob and tracing state, and
// the argument checks of the entry points (each case below is rejected
// before a HIP call).  Lines "K <sigma> <order> <radius> <hex values>" and
// "P <items> <item0...>" are compared with the production library by the
// Python side; a line "FAIL ..." or a sanitizer report fails the test.
#include <cinttypes>
#include <cstdio>
#include <cstring>
#include <initializer_list>
#include <string>
#include <vector>

#include "cucim_amd.h"

namespace {

int g_fails = 0;

void expect(bool ok, const char* what, int line) {
  if (!ok) {
    std::printf("FAIL line %d: %s (last error: %s)\n", line, what, hc_last_error());
    ++g_fails;
  }
}
#define EXPECT(c) expect((c), #c, __LINE__)

// an error return with a message
#define REJECTS(call)                                              \
  do {                                                             \
    const int rc_ = (call);                                        \
    expect(rc_ != HC_OK && std::strlen(hc_last_error()) > 0, #call, __LINE__); \
  } while (0)

hc_array arr(void* p, int dtype, std::initializer_list<int64_t> shape) {
  hc_array a;
  std::memset(&a, 0, sizeof(a));
  a.data = p;
  a.dtype = dtype;
  a.ndim = (int32_t)shape.size();
  int d = 0;
  for (int64_t s : shape) a.shape[d++] = s;
  int64_t st = 1;
  for (d = a.ndim - 1; d >= 0; --d) {
    a.strides[d] = st;
    st *= a.shape[d];
  }
  return a;
}

void kernels() {
  const double sigmas[] = {0.3, 0.5, 1.0, 1.5, 2.0, 3.7, 5.0, 9.0};
  for (double s : sigmas)
    for (int order = 0; order <= 3; ++order) {
      const int64_t radius = (int64_t)(4.0 * s + 0.5);
      std::vector<double> w((size_t)(2 * radius + 1));
      EXPECT(hc_gaussian_kernel1d(s, order, radius, w.data()) == HC_OK);
      std::printf("K %.17g %d %" PRId64, s, order, radius);
      for (double v : w) std::printf(" %a", v);
      std::printf("\n");
    }
  double one[1];
  EXPECT(hc_gaussian_kernel1d(1.0, 0, 0, one) == HC_OK);
  REJECTS(hc_gaussian_kernel1d(1.0, -1, 3, one));
  REJECTS(hc_gaussian_kernel1d(1.0, 0, -1, one));
  REJECTS(hc_gaussian_kernel1d(1.0, 0, 3, nullptr));
}

void planning() {
  // uneven image shapes, including empty ones
  const int64_t shapes[][2] = {{512, 512}, {1, 1}, {0, 7}, {33, 1000}, {1023, 17}, {5, 0}};
  const int n = (int)(sizeof(shapes) / sizeof(shapes[0]));
  for (int ty : {16, 32})
    for (int K : {1, 9, 17, 33}) {
      std::vector<hc_tile2d> t((size_t)n);
      std::memset(t.data(), 0, t.size() * sizeof(hc_tile2d));
      for (int i = 0; i < n; ++i) {
        t[(size_t)i].H = shapes[i][0];
        t[(size_t)i].W = shapes[i][1];
      }
      const int64_t total = hc_batch2d_plan(t.data(), n, K, ty, nullptr, 0);
      EXPECT(total >= 0);
      std::vector<int32_t> items((size_t)(total > 0 ? total : 1));
      EXPECT(hc_batch2d_plan(t.data(), n, K, ty, items.data(), total) == total);
      // a table one entry short is refused, not overrun
      if (total > 0) EXPECT(hc_batch2d_plan(t.data(), n, K, ty, items.data(), total - 1) == -1);
      std::printf("P %d %d %" PRId64, K, ty, total);
      for (const auto& x : t) std::printf(" %" PRId64, x.item0);
      int64_t prev = 0;
      for (int64_t i = 0; i < total; ++i) {
        EXPECT(items[(size_t)i] >= prev && items[(size_t)i] < n);
        prev = items[(size_t)i];
      }
      std::printf("\n");
    }
  hc_tile2d bad;
  std::memset(&bad, 0, sizeof(bad));
  bad.H = -1;
  EXPECT(hc_batch2d_plan(&bad, 1, 9, 16, nullptr, 0) == -1);
  EXPECT(hc_batch2d_plan(nullptr, 1, 9, 16, nullptr, 0) == -1);
  EXPECT(hc_batch2d_plan(&bad, 1, 34, 16, nullptr, 0) == -1);
  EXPECT(hc_batch2d_plan(&bad, 1, 9, 24, nullptr, 0) == -1);
}

void state() {
  EXPECT(hc_version() >= 1);
  EXPECT(hc_max_inline_taps(0) > 0 && hc_max_inline_taps(1) > 0);
  EXPECT(hc_max_resident_line(0) > 0);
  for (int k : {-1, (int)HC_TUNE_NKNOBS, 1 << 30}) {
    REJECTS(hc_tune_set(k, 0));
  }
  REJECTS(hc_tune_set(HC_TUNE_CONTIG_SEG, 300));
  REJECTS(hc_tune_set(HC_TUNE_SEP2D_TY, 12));
  REJECTS(hc_tune_set(HC_TUNE_EDGE_ROWS, 7));
  EXPECT(hc_tune_set(HC_TUNE_CONTIG_SEG, 512) == HC_OK);
  EXPECT(hc_tune_get(HC_TUNE_CONTIG_SEG) == 512);
  EXPECT(hc_tune_set(HC_TUNE_CONTIG_SEG, 0) == HC_OK);
  // nothing recorded: collecting into small buffers writes terminated text
  EXPECT(hc_timing_enable(0) == HC_OK);
  char names[4];
  double ms[1];
  EXPECT(hc_timing_collect(names, (int)sizeof(names), ms, 1) == 0);
  EXPECT(hc_timing_collect(nullptr, 0, nullptr, 0) == 0);
  EXPECT(hc_label_work_elems(0) >= 1);
  EXPECT(hc_label_work_elems(1000) >= 1001);
  EXPECT(hc_cnn3d_packed_size(24, 24) > 0);
}

void argument_checks() {
  float dummy[4];
  void* p = dummy;
  hc_array a = arr(p, HC_FLOAT32, {8, 8});
  hc_array b = arr(p, HC_FLOAT32, {8, 9});
  hc_array c3 = arr(p, HC_FLOAT32, {4, 4, 4});
  hc_array i32 = arr(p, HC_INT32, {8, 8});
  hc_array bad_dt = arr(p, 99, {8, 8});
  hc_array deep = arr(p, HC_FLOAT32, {1});
  deep.ndim = HC_MAX_NDIM + 1;
  double w[3] = {0.25, 0.5, 0.25};
  hc_taps t;
  std::memset(&t, 0, sizeof(t));
  t.w = w;
  t.K = 3;
  t.anchor = 1;
  hc_taps t0 = t;
  t0.K = 0;
  hc_taps ta = t;
  ta.anchor = 3;

  REJECTS(hc_correlate1d(nullptr, &a, &t, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &b, &t, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, nullptr, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, &t, 2, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, &t, -3, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, &t, 0, 17, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, &t0, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&a, &a, &ta, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&bad_dt, &a, &t, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_correlate1d(&deep, &deep, &t, 0, HC_MODE_REFLECT, 0.0, 0, nullptr));
  REJECTS(hc_separable2d(&a, &b, &t, &t, HC_MODE_REFLECT, HC_MODE_REFLECT, 0.0, nullptr));

  const double sig[HC_MAX_NDIM] = {1.0, 1.0, 1.0};
  const double negsig[HC_MAX_NDIM] = {1.0, -1.0, 1.0};
  const int ord[HC_MAX_NDIM] = {0, 0, 0};
  const int negord[HC_MAX_NDIM] = {0, -2, 0};
  const int md[HC_MAX_NDIM] = {HC_MODE_REFLECT, HC_MODE_REFLECT, HC_MODE_REFLECT};
  REJECTS(hc_gaussian_nd(&a, &a, nullptr, ord, nullptr, md, 0.0, 4.0, nullptr, nullptr));
  REJECTS(hc_gaussian_nd(&a, &a, negsig, ord, nullptr, md, 0.0, 4.0, nullptr, nullptr));
  REJECTS(hc_gaussian_nd(&a, &a, sig, negord, nullptr, md, 0.0, 4.0, nullptr, nullptr));
  REJECTS(hc_gaussian_nd(&deep, &deep, sig, ord, nullptr, md, 0.0, 4.0, nullptr, nullptr));

  hc_array H3[3] = {a, a, a};
  REJECTS(hc_hessian_matrix(&a, H3, 2, sig, HC_MODE_REFLECT, 0.0, 0, nullptr, nullptr));
  REJECTS(hc_hessian_matrix(&c3, H3, 3, sig, HC_MODE_REFLECT, 0.0, 0, nullptr, nullptr));
  REJECTS(hc_frangi_sigmas(&a, &a, sig, 0, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3, 3,
                           nullptr, nullptr));
  REJECTS(hc_frangi_sigmas(&a, nullptr, sig, 1, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3,
                           3, nullptr, nullptr));
  REJECTS(hc_frangi_sigmas(&deep, &deep, sig, 1, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3,
                           3, nullptr, nullptr));
  REJECTS(hc_frangi(&a, &b, &t, &t, 1, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3, 3,
                    nullptr, nullptr, nullptr));
  REJECTS(hc_frangi(&a, &i32, &t, &t, 1, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3, 3,
                    nullptr, nullptr, nullptr));
  REJECTS(hc_frangi(&a, &a, nullptr, nullptr, 1, 0.5, 0.5, 0.0, 0, 1, HC_MODE_REFLECT, 0.0, H3,
                    3, nullptr, nullptr, nullptr));
  REJECTS(hc_frangi(&a, &a, &t, &t, 1, 0.5, 0.5, 0.0, 0, 1, 9, 0.0, H3, 3, nullptr, nullptr,
                    nullptr));

  uint64_t counts[4];
  double edges[5] = {0, 1, 2, 3, 4};
  REJECTS(hc_histogram(&a, 0, edges, 0, counts, nullptr));
  REJECTS(hc_histogram(&a, 4, nullptr, 0, counts, nullptr));
  REJECTS(hc_histogram(&i32, 4, edges, 0, counts, nullptr));
  REJECTS(hc_histogram(nullptr, 4, edges, 0, counts, nullptr));

  int32_t work[4];
  hc_array lab = arr(p, HC_INT32, {8, 8});
  REJECTS(hc_label(&a, 0, 0, &lab, work, 4, nullptr, nullptr));
  REJECTS(hc_label(&a, 0, 3, &lab, work, 4, nullptr, nullptr));
  REJECTS(hc_label(&a, 0, 1, &a, work, 4, nullptr, nullptr));
  REJECTS(hc_label(&a, 0, 1, &lab, work, 4, nullptr, nullptr));  // work too small
  REJECTS(hc_label(&a, 0, 1, &lab, nullptr, 4, nullptr, nullptr));

  REJECTS(hc_minmax(&a, p, p, 8, nullptr));  // work below one partial result
  REJECTS(hc_binary(&a, &b, &a, HC_BIN_SUB, nullptr));
  REJECTS(hc_binary(&a, &a, &a, 7, nullptr));
}

}  // namespace

int main() {
  kernels();
  planning();
  state();
  argument_checks();
  std::printf("DONE %d\n", g_fails);
  return g_fails == 0 ? 0 : 1;
}
