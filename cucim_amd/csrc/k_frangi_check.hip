// Sampled plain-Frangi check of the surrogate (gfx950).
//
// For every sampled line of a volume (one work-group each), the plain 3-D
// Frangi vesselness (python/cucim/src/cucim/skimage/filters/ridges.py:322-476,
// black_ridges as given, gamma fixed) is evaluated exactly on up to 64 voxels
// along axis 0 and compared with the surrogate's output there: the line's
// sum of squared differences, sum of squared plain values and sample count.
//
// Per sigma the Hessian element H_ab = G_b(G_a(I)) is, along each axis, a
// correlation with the composite 1-D kernel k1 * k2 (full convolution of the
// first and second pass kernels, anchor a1 + a2) as long as no pass reaches
// the volume's border; the host places the sampled voxels where that holds.
// With the composite kernels kx, ky, kz (each of SS, SD, DD):
//   Cx_k(z', y')  = sum_w kx_k[w] I(z', y', x0 + w - A)        (per wave, per z')
//   Cy_e(z')      = sum_v ky_e[v] Cx_{kx_e}(z', y0 + v - A)    (6 elements)
//   H_e(z)        = sum_u kz_e[u] Cy_e(z + u - A)
// then the closed-form eigenvalues sorted by |value| and the vesselness of
// the plain kernels (hc_eig.h), maximum over sigmas.
#include <algorithm>

#include "hc_common.h"
#include "hc_eig.h"

namespace hc {

constexpr int kChkSamples = 64;
constexpr int kChkMaxL = 128;   // composite kernel length bound
constexpr int kChkMaxZ = kChkSamples + kChkMaxL;

struct CheckArgs {
  const float* vol;      // (D, H, W)
  const float* sur;      // surrogate output (D, H, W)
  int64_t D, H, W;
  const int32_t* lines;  // per brick: x0, y0, za, zb
  int nb;
  const float* kc;       // per sigma: SS, SD, DD, each L floats
  const int32_t* kinfo;  // per sigma: offset into kc, L, A
  int nsig;
  float sign;            // -1: black_ridges=False (the image is negated)
  double alpha_sq, beta_sq, gamma_sq;
  float* plain_out;      // per line: kChkSamples values (NaN where unused)
  float* err_out;        // per line: sum (sur - plain)^2, sum plain^2, samples
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}

__global__ __launch_bounds__(256) void k_frangi_check(const CheckArgs a) {
  __shared__ float Cx[4][kChkMaxL][3];
  __shared__ float Cy[kChkMaxZ][6];
  __shared__ float res[kChkSamples];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int x0 = a.lines[4 * b], y0 = a.lines[4 * b + 1];
  const int za = a.lines[4 * b + 2], zb = a.lines[4 * b + 3];
  // a line whose windows would leave the volume (or that is longer than the
  // sample buffer) is not evaluated: ns = 0 marks it as without samples
  int ns = zb - za;
  if (ns < 0 || ns > kChkSamples) ns = 0;
  for (int s = 0; s < a.nsig; ++s) {
    const int L = a.kinfo[3 * s + 1], A = a.kinfo[3 * s + 2];
    const bool inside = L >= 1 && L <= kChkMaxL && A >= 0 && za - A >= 0 &&
                        (int64_t)za - A + ns + L - 1 <= a.D && y0 - A >= 0 &&
                        (int64_t)y0 - A + L <= a.H && x0 - A >= 0 &&
                        (int64_t)x0 - A + L <= a.W;
    if (!inside) ns = 0;
  }
  if (tid < kChkSamples) res[tid] = 0.f;
  __syncthreads();
  // element e: (z kernel, y kernel, x kernel) indices into {SS, SD, DD}
  const int kz_of[6] = {2, 1, 1, 0, 0, 0};
  const int ky_of[6] = {0, 1, 0, 2, 1, 0};
  const int kx_of[6] = {0, 0, 1, 0, 1, 2};
  for (int s = 0; s < a.nsig && ns > 0; ++s) {
    const int off = a.kinfo[3 * s], L = a.kinfo[3 * s + 1], A = a.kinfo[3 * s + 2];
    const float* kk = a.kc + off;  // kk[k * L + i]
    const int nz = ns + L - 1;     // z' = za - A + j, j < nz
    for (int j = wv; j < nz; j += 4) {
      const int64_t zq = za - A + j;
      // Cx for every y row of the window
      for (int v = 0; v < L; ++v) {
        const int64_t yq = y0 - A + v;
        const float* row = a.vol + (zq * a.H + yq) * a.W + (x0 - A);
        float p0 = 0.f, p1 = 0.f, p2 = 0.f;
        for (int w = lane; w < L; w += 64) {
          const float x = row[w];
          p0 = fmaf(kk[w], x, p0);
          p1 = fmaf(kk[L + w], x, p1);
          p2 = fmaf(kk[2 * L + w], x, p2);
        }
        p0 = wave_sum(p0);
        p1 = wave_sum(p1);
        p2 = wave_sum(p2);
        if (lane == 0) {
          Cx[wv][v][0] = p0;
          Cx[wv][v][1] = p1;
          Cx[wv][v][2] = p2;
        }
      }
      // Cy (LDS writes of lane 0 are visible to the wave after the fence)
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
      __builtin_amdgcn_wave_barrier();
#pragma unroll
      for (int e = 0; e < 6; ++e) {
        float p = 0.f;
        const float* ky = kk + ky_of[e] * L;
        for (int v = lane; v < L; v += 64) p = fmaf(ky[v], Cx[wv][v][kx_of[e]], p);
        p = wave_sum(p);
        if (lane == 0) Cy[j][e] = p;
      }
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // the samples
    if (tid < ns) {
      float Hm[6];
#pragma unroll
      for (int e = 0; e < 6; ++e) {
        const float* kz = kk + kz_of[e] * L;
        float h = 0.f;
        for (int u = 0; u < L; ++u) h = fmaf(kz[u], Cy[tid + u][e], h);
        Hm[e] = a.sign * h;
      }
      float l1, l2, l3;
      // (H00, H01, H02, H11, H12, H22) -> (a, d, f, b, e, c)
      eig33<float>(Hm[0], Hm[3], Hm[5], Hm[1], Hm[4], Hm[2], 0, 1, l1, l2, l3);
      const float rg = frangi_rg3<float>(l1, l2, l3);
      const float v = frangi_v3<float>(l1, l2, l3, rg, a.alpha_sq, a.beta_sq, a.gamma_sq);
      res[tid] = s == 0 ? v : nanmax(res[tid], v);
    }
    __syncthreads();
  }
  // per-brick sampled error against the surrogate
  if (wv == 0) {
    float dn = 0.f, nn = 0.f;
    if (lane < ns) {
      const float p = res[lane];
      const float q = a.sur[((int64_t)(za + lane) * a.H + y0) * a.W + x0];
      dn = (q - p) * (q - p);
      nn = p * p;
      a.plain_out[(int64_t)b * kChkSamples + lane] = p;
    } else {
      a.plain_out[(int64_t)b * kChkSamples + lane] = __builtin_nanf("");
    }
    dn = wave_sum(dn);
    nn = wave_sum(nn);
    if (lane == 0) {
      a.err_out[3 * (int64_t)b] = dn;
      a.err_out[3 * (int64_t)b + 1] = nn;
      a.err_out[3 * (int64_t)b + 2] = (float)(ns > 0 ? ns : 0);
    }
  }
}

}  // namespace hc

using namespace hc;

extern "C" {

int hc_frangi_check(const hc_array* vol, const hc_array* sur, const int32_t* lines_dev, int nb,
                    const float* kc_dev, const int32_t* kinfo_dev, const int32_t* kinfo_host,
                    int nsig, int black_ridges, double alpha, double beta, double gamma,
                    float* plain_out, float* err_out, void* stream) {
  if (!vol || !sur || !lines_dev || !kc_dev || !kinfo_dev || !kinfo_host || !plain_out ||
      !err_out)
    return fail(HC_ERR_VALUE, "null argument");
  if (vol->dtype != HC_FLOAT32 || vol->ndim != 3 || !is_c_contiguous(vol) ||
      !same_shape(vol, sur) || sur->dtype != HC_FLOAT32 || !is_c_contiguous(sur))
    return fail(HC_ERR_VALUE, "volume and surrogate output must be C-contiguous 3-D float32");
  if (nb < 0 || nsig < 1) return fail(HC_ERR_VALUE, "bad counts");
  for (int s = 0; s < nsig; ++s)
    if (kinfo_host[3 * s + 1] < 1 || kinfo_host[3 * s + 1] > kChkMaxL)
      return fail(HC_ERR_RESOURCE, "composite kernel longer than 128 taps");
  if (gamma <= 0) return fail(HC_ERR_VALUE, "the check needs a fixed gamma");
  if (nb == 0) return HC_OK;
  CheckArgs a;
  a.vol = (const float*)vol->data;
  a.sur = (const float*)sur->data;
  a.D = vol->shape[0];
  a.H = vol->shape[1];
  a.W = vol->shape[2];
  a.lines = lines_dev;
  a.nb = nb;
  a.kc = kc_dev;
  a.kinfo = kinfo_dev;
  a.nsig = nsig;
  a.sign = black_ridges ? 1.f : -1.f;
  a.alpha_sq = 2 * alpha * alpha;
  a.beta_sq = 2 * beta * beta;
  a.gamma_sq = 2 * gamma * gamma;
  a.plain_out = plain_out;
  a.err_out = err_out;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_frangi_check", st, k_frangi_check, dim3((unsigned)nb), dim3(256), 0, a);
  return HC_OK;
}

}  // extern "C"
