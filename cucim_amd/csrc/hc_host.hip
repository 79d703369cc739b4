// Host-side helpers of the C ABI: error state, array checks, tap packing.
#include <atomic>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "hc_common.h"

namespace hc {

namespace {
thread_local std::string g_last_error;
}

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int launch_error(hipError_t e, const char* name, size_t smem) {
  g_last_error = std::string("launch of ") + name + " failed (" + std::to_string(smem) +
                 " bytes LDS): " + hipGetErrorString(e);
  return HC_ERR_HIP;
}

int hip_check(hipError_t e, const char* what) {
  g_last_error = std::string("HIP error in ") + what + ": " + hipGetErrorString(e);
  return HC_ERR_HIP;
}

bool is_c_contiguous(const hc_array* a) {
  int64_t expect = 1;
  for (int d = a->ndim - 1; d >= 0; --d) {
    if (a->shape[d] != 1 && a->strides[d] != expect) return false;
    expect *= a->shape[d];
  }
  return true;
}

Lines lines_of(const hc_array* a, int axis) {
  Lines l{1, a->shape[axis], 1};
  for (int d = 0; d < axis; ++d) l.outer *= a->shape[d];
  for (int d = axis + 1; d < a->ndim; ++d) l.inner *= a->shape[d];
  return l;
}

int64_t numel(const hc_array* a) {
  int64_t n = 1;
  for (int d = 0; d < a->ndim; ++d) n *= a->shape[d];
  return n;
}

bool same_shape(const hc_array* a, const hc_array* b) {
  if (a->ndim != b->ndim) return false;
  for (int d = 0; d < a->ndim; ++d)
    if (a->shape[d] != b->shape[d]) return false;
  return true;
}

template <typename W>
int make_taps(const hc_taps* t, int pad_to, TapPack<W>* out) {
  if (t == nullptr || t->w == nullptr || t->K < 1)
    return fail(HC_ERR_VALUE, "incorrect filter size");
  if (t->anchor < 0 || t->anchor >= t->K) return fail(HC_ERR_VALUE, "invalid origin");
  const int64_t Kp = (t->K + pad_to - 1) / pad_to * pad_to;
  if (Kp > inline_cap<W>())
    return fail(HC_ERR_RESOURCE, "kernel too long for the LDS-tiled line kernels");
  std::memset(out, 0, sizeof(*out));
  out->K = (int)Kp;
  out->anchor = t->anchor;
  for (int64_t k = 0; k < t->K; ++k) out->w[k] = (W)t->w[k];
  for (int g = 0; g < (int)(Kp / 4); ++g) {
    bool all = true;
    for (int k = 0; k < 4; ++k) all = all && out->w[4 * g + k] != (W)0;
    if (all) out->full[g >> 5] |= 1u << (g & 31);
  }
  return HC_OK;
}

std::atomic<int> g_tune[HC_TUNE_NKNOBS];
int tune_get(int knob) { return g_tune[knob].load(std::memory_order_relaxed); }

namespace {
struct TimingRec {
  const char* name;
  hipEvent_t a, b;
};
std::mutex g_tmu;
bool g_timing = false;
std::vector<TimingRec> g_recs;
std::vector<hipEvent_t> g_pool;

hipEvent_t take_event() {
  if (!g_pool.empty()) {
    hipEvent_t e = g_pool.back();
    g_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}
}  // namespace

int timing_begin(const char* name, hipStream_t st) {
  if (!g_timing) return -1;
  std::lock_guard<std::mutex> lk(g_tmu);
  TimingRec r{name, take_event(), take_event()};
  if (!r.a || !r.b) return -1;
  (void)hipEventRecord(r.a, st);
  g_recs.push_back(r);
  return (int)g_recs.size() - 1;
}

void timing_end(int slot, hipStream_t st) {
  if (slot < 0) return;
  std::lock_guard<std::mutex> lk(g_tmu);
  if (slot < (int)g_recs.size()) (void)hipEventRecord(g_recs[slot].b, st);
}

template int make_taps<float>(const hc_taps*, int, TapPack<float>*);
template int make_taps<double>(const hc_taps*, int, TapPack<double>*);

}  // namespace hc

extern "C" {

const char* hc_last_error(void) { return hc::g_last_error.c_str(); }

int hc_version(void) { return 1; }

int hc_tune_set(int knob, int value) {
  if (knob < 0 || knob >= HC_TUNE_NKNOBS) return hc::fail(HC_ERR_VALUE, "unknown tuning knob");
  if (knob == HC_TUNE_CONTIG_SEG && value != 0 &&
      (value < 256 || value > 4096 || value % 256 != 0))
    return hc::fail(HC_ERR_VALUE, "segment must be a multiple of 256 in [256, 4096]");
  if (knob == HC_TUNE_HESSIAN_VALU && value != 0 && value != 1)
    return hc::fail(HC_ERR_VALUE, "the Hessian vector-ALU switch takes 0 or 1");
  if (knob == HC_TUNE_SEP2D_TY && value != 0 && value != 8 && value != 16 && value != 32 &&
      value != 64)
    return hc::fail(HC_ERR_VALUE, "tile rows must be 8, 16, 32 or 64");
  if (knob == HC_TUNE_EDGE_ROWS && value != 0 && value != 8 && value != 16 && value != 32 &&
      value != 64 && value != 128)
    return hc::fail(HC_ERR_VALUE, "edge rows per wave must be 8, 16, 32, 64 or 128");
  if (knob == HC_TUNE_MFMA_WAVES && value != 0 && value != 4 && value != 8)
    return hc::fail(HC_ERR_VALUE, "matrix-core Hessian waves per work-group must be 4 or 8");
  hc::g_tune[knob].store(value);
  return HC_OK;
}

int hc_tune_get(int knob) {
  if (knob < 0 || knob >= HC_TUNE_NKNOBS) return -1;
  return hc::g_tune[knob].load();
}

int hc_timing_enable(int on) {
  std::lock_guard<std::mutex> lk(hc::g_tmu);
  hc::g_timing = on != 0;
  return HC_OK;
}

int hc_timing_collect(char* names, int names_len, double* ms, int max_records) {
  std::lock_guard<std::mutex> lk(hc::g_tmu);
  int n = 0, pos = 0;
  for (auto& r : hc::g_recs) {
    if (n < max_records) {
      float t = 0.f;
      if (hipEventSynchronize(r.b) != hipSuccess ||
          hipEventElapsedTime(&t, r.a, r.b) != hipSuccess)
        t = -1.f;
      ms[n] = t;
      if (names && pos < names_len) {
        const int len = (int)std::strlen(r.name);
        if (pos + len + 1 <= names_len) {
          std::memcpy(names + pos, r.name, len);
          names[pos + len] = '\n';
          pos += len + 1;
        }
      }
      ++n;
    }
    hc::g_pool.push_back(r.a);
    hc::g_pool.push_back(r.b);
  }
  if (names && pos < names_len) names[pos] = 0;
  hc::g_recs.clear();
  return n;
}

int hc_max_inline_taps(int w_f64) {
  return w_f64 ? hc::inline_cap<double>() : hc::inline_cap<float>();
}

}  // extern "C"
