// 3-D convolutional surrogate of the multi-scale Frangi stack (gfx950).
//
// The network (5 layers of 3x3x3 "valid" convolutions, 24 channels, GELU
// between layers) has no counterpart in the reference; it stands in for
// filters/ridges.py:322-476 (python/cucim/src/cucim/skimage/) and its output
// is compared with the plain kernels'.
//
// Activations are channels-last bfloat16 volumes (D, H, W, C); weights are
// bfloat16, accumulation float32 on the matrix cores:
//
//   k_conv3d_mfma<CI, CO, ACT, OUT>  implicit GEMM, M = output voxels (16
//       consecutive x per tile), N = output channels (16 per tile), K =
//       27 taps x CI input channels in chunks of 32 (tap-major, channel-minor;
//       CI = 24 puts every 8-element K group of a lane inside one tap, so an
//       A fragment is one 16-byte LDS read).  v_mfma_f32_16x16x32_bf16:
//       lane l holds A[l % 16][8 (l / 16) + j], B[8 (l / 16) + j][l % 16] and
//       D[4 (l / 16) + r][l % 16].  Weights arrive pre-packed in that B order
//       (chunk, n-tile, lane, 8 values) and are staged in LDS once per
//       work-group.  A work-group computes a 2 x 8 x 16 block of outputs
//       (4 waves x 4 M-tiles).
//   k_pad_bf16        reflect / nearest / ... padding of a float32 volume into
//                     the bfloat16 single-channel input of the first layer
//   training kernels  (float32): GELU backward, input gradient (full
//                     correlation with the transposed kernel), weight / bias
//                     gradients (per-block partial sums, atomically added),
//                     masked MSE gradient, Adam step, bf16 packing.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hc_common.h"

namespace hc {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int kBY = 8, kBX = 16;        // output block rows / columns
constexpr int kTY = kBY + 2, kTX = kBX + 2;  // input tile rows / columns

__device__ __forceinline__ float bf2f(__bf16 v) { return (float)v; }
__device__ __forceinline__ __bf16 f2bf(float v) { return (__bf16)v; }  // RNE

// GELU in its tanh form (torch.nn.GELU(approximate="tanh")):
// 0.5 x (1 + tanh(z / 2)) = x * sigmoid(z), z = 2 sqrt(2/pi) (x + 0.044715 x^3),
// with the hardware exp and reciprocal (the network is trained with this
// same function and its derivative).
__device__ __forceinline__ float gelu_z(float x) {
  return 1.5957691216057308f * fmaf(0.044715f * x, x * x, x);
}
__device__ __forceinline__ float gelu(float x) {
  return x * __builtin_amdgcn_rcpf(1.0f + __expf(-gelu_z(x)));
}
__device__ __forceinline__ float gelu_grad(float x) {
  const float s = __builtin_amdgcn_rcpf(1.0f + __expf(-gelu_z(x)));
  const float dz = 1.5957691216057308f * fmaf(3.0f * 0.044715f, x * x, 1.0f);
  return fmaf(x * s * (1.0f - s), dz, s);
}

enum { ACT_NONE = 0, ACT_GELU = 1, ACT_CLAMP01 = 2 };

struct ConvArgs {
  const __bf16* in;   // (D, H, W, CI)
  void* out;          // (D-2, H-2, W-2, CO): bf16, or float32 when OUT_F32
  float* pre;         // optional float32 pre-activation (training), or null
  const bf16x8* wpk;  // packed weights: [chunk][ntile][64 lanes]
  const float* bias;  // CO
  int D, H, W;        // input extents
  int nbx, nby;       // blocks along x, y
  int64_t nblocks;
};

template <int CI>
struct ConvGeom {
  static constexpr int K = 27 * CI;
  static constexpr int NCH = (K + 31) / 32;  // K chunks of 32
};

// Block b of the output grid (x fastest) -> origin.
template <int BZ>
__device__ __forceinline__ void block_origin(const ConvArgs& a, int64_t b, int& x0, int& y0,
                                             int& z0) {
  x0 = (int)(b % a.nbx) * kBX;
  y0 = (int)((b / a.nbx) % a.nby) * kBY;
  z0 = (int)(b / ((int64_t)a.nbx * a.nby)) * BZ;
}

// BZ output planes per block, two waves per plane (4 rows of 16 voxels each).
template <int CI, int CO, int ACT, bool OUT_F32, int BZ>
__global__ __launch_bounds__(128 * BZ) void k_conv3d_mfma(const ConvArgs a) {
  constexpr int NTH = 128 * BZ;
  constexpr int kBZ = BZ, kTZ = BZ + 2;
  constexpr int NCH = ConvGeom<CI>::NCH;
  constexpr int NT = (CO + 15) / 16;
  constexpr int TVOX = kTZ * kTY * kTX;
  constexpr int TILE = TVOX * CI;                   // bf16 elements
  constexpr bool VEC = CI % 8 == 0;
  constexpr int TV = VEC ? TVOX * (CI / 8) : 1;     // 16-byte vectors of a tile
  constexpr int NPF = VEC ? (TV + NTH - 1) / NTH : 1;  // prefetch registers per lane
  constexpr bool STAGE_OUT = !OUT_F32 && CO % 8 == 0;
  extern __shared__ __align__(16) unsigned char smem[];
  bf16x8* Wl = reinterpret_cast<bf16x8*>(smem);   // NCH * NT * 64
  __bf16* T = reinterpret_cast<__bf16*>(smem + NCH * NT * 64 * 16);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  for (int i = tid; i < NCH * NT * 64; i += NTH) Wl[i] = a.wpk[i];
  const int64_t G = gridDim.x;
  const int64_t slot = xcd_swizzle(blockIdx.x, G);
  // input tile of block b: global -> registers (VEC) or straight to LDS
  auto fetch = [&](int64_t b, bf16x8 (&r)[NPF]) {
    int x0, y0, z0;
    block_origin<BZ>(a, b, x0, y0, z0);
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int e = tid + NTH * u;
      bf16x8 val = {};
      if (e < TV) {
        constexpr int V = VEC ? CI / 8 : 1;
        const int v = e % V, vox = e / V;
        const int tx = vox % kTX, ty = (vox / kTX) % kTY, tz = vox / (kTX * kTY);
        const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
        if (gx < a.W && gy < a.H && gz < a.D)
          val = reinterpret_cast<const bf16x8*>(
              a.in + (((int64_t)gz * a.H + gy) * a.W + gx) * CI)[v];
      }
      r[u] = val;
    }
  };
  auto put = [&](const bf16x8 (&r)[NPF]) {
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int e = tid + NTH * u;
      if (e < TV) reinterpret_cast<bf16x8*>(T)[e] = r[u];
    }
  };
  auto stage_scalar = [&](int64_t b) {
    int x0, y0, z0;
    block_origin<BZ>(a, b, x0, y0, z0);
    for (int e = tid; e < TILE; e += NTH) {
      const int c = e % CI, vox = e / CI;
      const int tx = vox % kTX, ty = (vox / kTX) % kTY, tz = vox / (kTX * kTY);
      const int gx = x0 + tx, gy = y0 + ty, gz = z0 + tz;
      __bf16 val = (__bf16)0.0f;
      if (gx < a.W && gy < a.H && gz < a.D)
        val = a.in[(((int64_t)gz * a.H + gy) * a.W + gx) * CI + c];
      T[e] = val;
    }
  };
  bf16x8 pf[NPF];
  if (VEC && slot < a.nblocks) fetch(slot, pf);
  const int OD = a.D - 2, OH = a.H - 2, OW = a.W - 2;
  const int oz = wv >> 1, oy0 = (wv & 1) * 4;
  const int xi = lane & 15, kg = lane >> 4;  // A row (voxel x), K group
  for (int64_t base = 0; base < a.nblocks; base += G) {
    const int64_t blk = base + slot;
    if (blk >= a.nblocks) break;  // uniform per work-group
    int x0, y0, z0;
    block_origin<BZ>(a, blk, x0, y0, z0);
    __syncthreads();  // the previous block is done with T
    if (VEC) {
      put(pf);
      if (blk + G < a.nblocks) fetch(blk + G, pf);  // in flight during the MFMA loop
    } else {
      stage_scalar(blk);
    }
    __syncthreads();
    f32x4 acc[4][NT];
#pragma unroll
    for (int m = 0; m < 4; ++m)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[m][t] = f32x4{0.f, 0.f, 0.f, 0.f};
    // fragments of K chunk c: B from the packed weights, A from the tile
    auto frag = [&](int c, bf16x8 (&av)[4], bf16x8 (&bw)[NT]) {
#pragma unroll
      for (int t = 0; t < NT; ++t) bw[t] = Wl[(c * NT + t) * 64 + lane];
      if (VEC) {
        const int k = 32 * c + 8 * kg;  // first K index of this lane's group
        const int tap = k / CI, ci0 = k - tap * CI;
        const bool valid = tap < 27;
        const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          const int vox = ((oz + dz) * kTY + (oy0 + m + dy)) * kTX + (xi + dx);
          bf16x8 v = {};
          if (valid) v = *reinterpret_cast<const bf16x8*>(T + vox * CI + ci0);
          av[m] = v;
        }
      } else {
        // CI == 1: a lane's 8 K values are 8 consecutive taps of one voxel
#pragma unroll
        for (int m = 0; m < 4; ++m) {
          bf16x8 v = {};
#pragma unroll
          for (int j = 0; j < 8; ++j) {
            const int tap = 32 * c + 8 * kg + j;
            if (tap < 27) {
              const int dz = tap / 9, dy = (tap / 3) % 3, dx = tap % 3;
              v[j] = T[((oz + dz) * kTY + (oy0 + m + dy)) * kTX + (xi + dx)];
            }
          }
          av[m] = v;
        }
      }
    };
    auto mma = [&](const bf16x8 (&av)[4], const bf16x8 (&bw)[NT]) {
#pragma unroll
      for (int m = 0; m < 4; ++m)
#pragma unroll
        for (int t = 0; t < NT; ++t)
          acc[m][t] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(av[m], bw[t], acc[m][t], 0, 0, 0);
    };
    // two fragment sets: the loads of chunk c + 1 are in flight while the
    // matrix instructions of chunk c run
    bf16x8 a0[4], b0[NT], a1[4], b1[NT];
    frag(0, a0, b0);
#pragma unroll 1
    for (int c = 0; c < NCH; c += 2) {
      if (c + 1 < NCH) frag(c + 1, a1, b1);
      mma(a0, b0);
      if (c + 2 < NCH) frag(c + 2, a0, b0);
      if (c + 1 < NCH) mma(a1, b1);
    }
    // epilogue: D[4 (l/16) + r][l % 16] = (x = 4 kg + r, channel 16 t + l % 16)
    const int gz = z0 + oz;
    __bf16* O = T;  // output block staging (kBZ * kBY * kBX voxels x CO)
    if (STAGE_OUT) __syncthreads();  // every wave is done reading T
#pragma unroll
    for (int m = 0; m < 4; ++m) {
      const int gy = y0 + oy0 + m;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int co = 16 * t + (lane & 15);
        if (co >= CO) continue;
        const float bias = a.bias[co];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int lx = 4 * kg + r, gx = x0 + lx;
          const bool inside = gx < OW && gy < OH && gz < OD;
          const float p = acc[m][t][r] + bias;
          float y = p;
          if (ACT == ACT_GELU) y = gelu(p);
          if (ACT == ACT_CLAMP01) y = fminf(fmaxf(p, 0.f), 1.f);
          const int64_t o = (((int64_t)gz * OH + gy) * OW + gx) * CO + co;
          if (inside && a.pre) a.pre[o] = p;
          if (STAGE_OUT) {
            O[((oz * kBY + oy0 + m) * kBX + lx) * CO + co] = f2bf(y);
          } else if (inside) {
            if (OUT_F32)
              reinterpret_cast<float*>(a.out)[o] = y;
            else
              reinterpret_cast<__bf16*>(a.out)[o] = f2bf(y);
          }
        }
      }
    }
    if (STAGE_OUT) {
      __syncthreads();
      constexpr int VPV = CO / 8;  // 16-byte vectors per voxel
      for (int e = tid; e < kBZ * kBY * kBX * VPV; e += NTH) {
        const int v = e % VPV, vox = e / VPV;
        const int lx = vox % kBX, ly = (vox / kBX) % kBY, lz = vox / (kBX * kBY);
        const int gx = x0 + lx, gy = y0 + ly, gzz = z0 + lz;
        if (gx < OW && gy < OH && gzz < OD)
          reinterpret_cast<bf16x8*>(reinterpret_cast<__bf16*>(a.out) +
                                    (((int64_t)gzz * OH + gy) * OW + gx) * CO)[v] =
              reinterpret_cast<const bf16x8*>(O)[e];
      }
    }
  }
}

// --------------------------------------------------------------- padding --
// out (D + 2p, H + 2p, W + 2p) bf16 = in (D, H, W) float32 boundary-extended
__global__ __launch_bounds__(256) void k_pad_bf16(const float* __restrict__ in,
                                                  __bf16* __restrict__ out, int D, int H, int W,
                                                  int p, int mode, float cval) {
  const int PD = D + 2 * p, PH = H + 2 * p, PW = W + 2 * p;
  const int64_t n = (int64_t)PD * PH * PW;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int x = (int)(i % PW), y = (int)((i / PW) % PH), z = (int)(i / ((int64_t)PW * PH));
    const int64_t sx = remap(x - p, W, mode), sy = remap(y - p, H, mode),
                  sz = remap(z - p, D, mode);
    float v = cval;
    if (sx >= 0 && sy >= 0 && sz >= 0) v = in[(sz * H + sy) * W + sx];
    out[i] = f2bf(v);
  }
}

// ------------------------------------------------------------- training --
// g (out-shaped, float32) *= gelu'(pre)
__global__ __launch_bounds__(256) void k_gelu_bwd(float* __restrict__ g,
                                                  const float* __restrict__ pre, int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256)
    g[i] *= gelu_grad(pre[i]);
}

// Input gradient of a valid 3x3x3 convolution: dx[p][ci] =
// sum_{tap, co} dy[p - off(tap)][co] * w[tap][ci][co] over valid p - off.
// w: float32 [27][CI][CO].
__global__ __launch_bounds__(256) void k_conv3d_dgrad(const float* __restrict__ dy,
                                                      const float* __restrict__ w,
                                                      float* __restrict__ dx, int D, int H,
                                                      int W, int CI, int CO) {
  const int OD = D - 2, OH = H - 2, OW = W - 2;
  const int64_t n = (int64_t)D * H * W * CI;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const int ci = (int)(i % CI);
    const int64_t v = i / CI;
    const int x = (int)(v % W), y = (int)((v / W) % H), z = (int)(v / ((int64_t)W * H));
    float s = 0.f;
    for (int dz = 0; dz < 3; ++dz) {
      const int oz = z - dz;
      if (oz < 0 || oz >= OD) continue;
      for (int dyy = 0; dyy < 3; ++dyy) {
        const int oy = y - dyy;
        if (oy < 0 || oy >= OH) continue;
        for (int dxx = 0; dxx < 3; ++dxx) {
          const int ox = x - dxx;
          if (ox < 0 || ox >= OW) continue;
          const int tap = (dz * 3 + dyy) * 3 + dxx;
          const float* g = dy + (((int64_t)oz * OH + oy) * OW + ox) * CO;
          const float* wr = w + ((int64_t)tap * CI + ci) * CO;
          for (int co = 0; co < CO; ++co) s = fmaf(g[co], wr[co], s);
        }
      }
    }
    dx[i] = s;
  }
}

// Weight gradient: dw[tap][ci][co] += sum_p x[p + off(tap)][ci] * dy[p][co],
// one work-group per slab of output voxels; db[co] += sum_p dy[p][co].
// x: bf16 activations (the forward input), dy: float32.
__global__ __launch_bounds__(256) void k_conv3d_wgrad(const __bf16* __restrict__ x,
                                                      const float* __restrict__ dy,
                                                      float* __restrict__ dw,
                                                      float* __restrict__ db, int D, int H,
                                                      int W, int CI, int CO, int64_t per_block) {
  const int OD = D - 2, OH = H - 2, OW = W - 2;
  const int64_t nvox = (int64_t)OD * OH * OW;
  const int64_t v0 = (int64_t)blockIdx.x * per_block;
  const int64_t v1 = v0 + per_block < nvox ? v0 + per_block : nvox;
  const int nw = 27 * CI * CO;
  for (int j = threadIdx.x; j < nw + CO; j += 256) {
    float s = 0.f;
    if (j < nw) {
      const int co = j % CO, ci = (j / CO) % CI, tap = j / (CO * CI);
      const int dz = tap / 9, dyy = (tap / 3) % 3, dxx = tap % 3;
      for (int64_t v = v0; v < v1; ++v) {
        const int ox = (int)(v % OW), oy = (int)((v / OW) % OH),
                  oz = (int)(v / ((int64_t)OW * OH));
        const float xv =
            bf2f(x[(((int64_t)(oz + dz) * H + (oy + dyy)) * W + (ox + dxx)) * CI + ci]);
        s = fmaf(xv, dy[v * CO + co], s);
      }
      atomicAdd(dw + j, s);
    } else {
      const int co = j - nw;
      for (int64_t v = v0; v < v1; ++v) s += dy[v * CO + co];
      atomicAdd(db + co, s);
    }
  }
}

// MSE gradient over the interior: g = 2 (y - t) / count where mask, else 0;
// loss partial sums atomically added to *loss.
__global__ __launch_bounds__(256) void k_mse_grad(const float* __restrict__ y,
                                                  const float* __restrict__ t,
                                                  float* __restrict__ g, int64_t n, float inv_n,
                                                  float* __restrict__ loss) {
  __shared__ float red[256];
  float s = 0.f;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const float d = y[i] - t[i];
    g[i] = 2.f * d * inv_n;
    s += d * d;
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if (threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) atomicAdd(loss, red[0] * inv_n);
}

// Adam on float32 master weights.
__global__ __launch_bounds__(256) void k_adam(float* __restrict__ p, const float* __restrict__ g,
                                              float* __restrict__ m, float* __restrict__ v,
                                              int64_t n, float lr, float b1, float b2, float eps,
                                              float c1, float c2) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const float gi = g[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    p[i] -= lr * (mi / c1) / (sqrtf(vi / c2) + eps);
  }
}

// Pack float32 weights [27][CI][CO] into the bf16 B-fragment order.
__global__ __launch_bounds__(256) void k_pack_weights(const float* __restrict__ w,
                                                      bf16x8* __restrict__ pk, int CI, int CO) {
  const int K = 27 * CI, NCH = (K + 31) / 32, NT = (CO + 15) / 16;
  for (int i = blockIdx.x * 256 + threadIdx.x; i < NCH * NT * 64; i += gridDim.x * 256) {
    const int lane = i % 64, t = (i / 64) % NT, c = i / (64 * NT);
    const int n = 16 * t + (lane & 15);
    bf16x8 v;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int k = 32 * c + 8 * (lane >> 4) + j;
      float f = 0.f;
      if (k < K && n < CO) f = w[(int64_t)k * CO + n];  // k = tap * CI + ci
      v[j] = f2bf(f);
    }
    pk[i] = v;
  }
}

namespace {

int64_t grid_of(int64_t n) { return n <= 0 ? 1 : std::min<int64_t>((n + 255) / 256, 256 * 64); }

template <int CI, int CO, int ACT, bool F32, int BZ>
int conv_launch(ConvArgs a, hipStream_t st) {
  constexpr int NCH = ConvGeom<CI>::NCH, NT = (CO + 15) / 16;
  a.nbx = (a.W - 2 + kBX - 1) / kBX;
  a.nby = (a.H - 2 + kBY - 1) / kBY;
  const int nbz = (a.D - 2 + BZ - 1) / BZ;
  a.nblocks = (int64_t)a.nbx * a.nby * nbz;
  if (a.nblocks > 0x7fffffff) return fail(HC_ERR_VALUE, "volume too large");
  // weights, then the input tile (also the output block's staging area)
  const size_t tile =
      std::max<size_t>((size_t)(BZ + 2) * kTY * kTX * CI, (size_t)BZ * kBY * kBX * CO);
  const size_t smem = (size_t)NCH * NT * 64 * 16 + tile * 2;
  if (smem > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute((const void*)k_conv3d_mfma<CI, CO, ACT, F32, BZ>,
                                             hipFuncAttributeMaxDynamicSharedMemorySize,
                                             (int)smem);
    if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(k_conv3d_mfma)");
  }
  // persistent: as many work-groups as fit the CUs' LDS, each walking blocks
  // with the next block's input prefetched during the matrix loop
  const int per_cu = std::max<int>(1, (int)((160 * 1024) / smem));
  const int64_t grid = std::min<int64_t>(a.nblocks, (int64_t)256 * per_cu);
  HC_LAUNCH("k_conv3d_mfma", st, (k_conv3d_mfma<CI, CO, ACT, F32, BZ>), dim3((unsigned)grid),
            dim3(128 * BZ), smem, a);
  return HC_OK;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int hc_cnn3d_packed_size(int ci, int co) {
  if (ci < 1 || co < 1) return -1;
  return ((27 * ci + 31) / 32) * ((co + 15) / 16) * 64 * 16;
}

int hc_cnn3d_pack(const float* w_dev, void* packed_dev, int ci, int co, void* stream) {
  if (!w_dev || !packed_dev) return fail(HC_ERR_VALUE, "null argument");
  if (ci < 1 || co < 1) return fail(HC_ERR_VALUE, "channel counts must be positive");
  const int n = hc_cnn3d_packed_size(ci, co) / 16;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_pack_weights", st, k_pack_weights, dim3((unsigned)grid_of(n)), dim3(256), 0, w_dev,
            (bf16x8*)packed_dev, ci, co);
  return HC_OK;
}

int hc_cnn3d_conv(const hc_array* in, const hc_array* out, const void* packed,
                  const float* bias_dev, int act, float* pre_dev, void* stream) {
  if (!in || !out || !packed || !bias_dev) return fail(HC_ERR_VALUE, "null argument");
  if (in->dtype != HC_BFLOAT16 || in->ndim != 4 || !is_c_contiguous(in))
    return fail(HC_ERR_VALUE, "input must be a C-contiguous (D, H, W, C) bfloat16 array");
  if (out->ndim != 4 || !is_c_contiguous(out))
    return fail(HC_ERR_VALUE, "output must be a C-contiguous (D-2, H-2, W-2, C) array");
  const int ci = (int)in->shape[3], co = (int)out->shape[3];
  for (int d = 0; d < 3; ++d)
    if (out->shape[d] != in->shape[d] - 2 || in->shape[d] < 3)
      return fail(HC_ERR_VALUE, "valid convolution: output extents are input extents - 2");
  if (((uintptr_t)in->data | (uintptr_t)packed) % 16)
    return fail(HC_ERR_VALUE, "input and packed weights must be 16-byte aligned");
  ConvArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = (const __bf16*)in->data;
  a.out = out->data;
  a.pre = pre_dev;
  a.wpk = (const bf16x8*)packed;
  a.bias = bias_dev;
  a.D = (int)in->shape[0];
  a.H = (int)in->shape[1];
  a.W = (int)in->shape[2];
  hipStream_t st = (hipStream_t)stream;
  const bool f32 = out->dtype == HC_FLOAT32;
  if (!f32 && out->dtype != HC_BFLOAT16) return fail(HC_ERR_VALUE, "output must be bf16 or f32");
  if (ci == 1 && co == 24 && act == ACT_GELU && !f32)
    return conv_launch<1, 24, ACT_GELU, false, 2>(a, st);
  if (ci == 24 && co == 24 && act == ACT_GELU && !f32)
    return conv_launch<24, 24, ACT_GELU, false, 4>(a, st);
  if (ci == 24 && co == 1 && act == ACT_CLAMP01 && f32)
    return conv_launch<24, 1, ACT_CLAMP01, true, 4>(a, st);
  if (ci == 24 && co == 1 && act == ACT_NONE && f32)
    return conv_launch<24, 1, ACT_NONE, true, 4>(a, st);
  return fail(HC_ERR_UNSUPPORTED, "layer shape / activation not compiled");
}

int hc_cnn3d_pad(const hc_array* in, const hc_array* out, int pad, int mode, double cval,
                 void* stream) {
  if (!in || !out) return fail(HC_ERR_VALUE, "null argument");
  if (in->dtype != HC_FLOAT32 || in->ndim != 3 || !is_c_contiguous(in))
    return fail(HC_ERR_VALUE, "input must be a C-contiguous 3-D float32 array");
  if (out->dtype != HC_BFLOAT16 || !is_c_contiguous(out) || out->ndim < 3)
    return fail(HC_ERR_VALUE, "output must be a C-contiguous bfloat16 array");
  for (int d = 0; d < 3; ++d)
    if (out->shape[d] != in->shape[d] + 2 * pad)
      return fail(HC_ERR_VALUE, "output extents must be input + 2 pad");
  if (out->ndim == 4 && out->shape[3] != 1) return fail(HC_ERR_VALUE, "one channel expected");
  const int64_t n = numel(out);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_pad_bf16", st, k_pad_bf16, dim3((unsigned)grid_of(n)), dim3(256), 0,
            (const float*)in->data, (__bf16*)out->data, (int)in->shape[0], (int)in->shape[1],
            (int)in->shape[2], pad, mode, (float)cval);
  return HC_OK;
}

int hc_cnn3d_gelu_bwd(float* g, const float* pre, int64_t n, void* stream) {
  if (!g || !pre || n < 0) return fail(HC_ERR_VALUE, "bad argument");
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_gelu_bwd", st, k_gelu_bwd, dim3((unsigned)grid_of(n)), dim3(256), 0, g, pre, n);
  return HC_OK;
}

int hc_cnn3d_dgrad(const float* dy, const float* w, float* dx, int D, int H, int W, int ci,
                   int co, void* stream) {
  if (!dy || !w || !dx || D < 3 || H < 3 || W < 3 || ci < 1 || co < 1)
    return fail(HC_ERR_VALUE, "bad argument");
  const int64_t n = (int64_t)D * H * W * ci;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_conv3d_dgrad", st, k_conv3d_dgrad, dim3((unsigned)grid_of(n)), dim3(256), 0, dy,
            w, dx, D, H, W, ci, co);
  return HC_OK;
}

int hc_cnn3d_wgrad(const void* x_bf16, const float* dy, float* dw, float* db, int D, int H,
                   int W, int ci, int co, void* stream) {
  if (!x_bf16 || !dy || !dw || !db || D < 3 || H < 3 || W < 3 || ci < 1 || co < 1)
    return fail(HC_ERR_VALUE, "bad argument");
  const int64_t nvox = (int64_t)(D - 2) * (H - 2) * (W - 2);
  const int64_t blocks = std::min<int64_t>(nvox, 2048);
  const int64_t per = (nvox + blocks - 1) / blocks;
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_conv3d_wgrad", st, k_conv3d_wgrad, dim3((unsigned)blocks), dim3(256), 0,
            (const __bf16*)x_bf16, dy, dw, db, D, H, W, ci, co, per);
  return HC_OK;
}

int hc_cnn3d_mse_grad(const float* y, const float* t, float* g, int64_t n, float* loss,
                      void* stream) {
  if (!y || !t || !g || !loss || n <= 0) return fail(HC_ERR_VALUE, "bad argument");
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_mse_grad", st, k_mse_grad, dim3((unsigned)std::min<int64_t>(grid_of(n), 1024)),
            dim3(256), 0, y, t, g, n, 1.0f / (float)n, loss);
  return HC_OK;
}

int hc_cnn3d_adam(float* p, const float* g, float* m, float* v, int64_t n, double lr,
                  double b1, double b2, double eps, int step, void* stream) {
  if (!p || !g || !m || !v || n < 0 || step < 1) return fail(HC_ERR_VALUE, "bad argument");
  if (n == 0) return HC_OK;
  const float c1 = (float)(1.0 - std::pow(b1, step)), c2 = (float)(1.0 - std::pow(b2, step));
  hipStream_t st = (hipStream_t)stream;
  HC_LAUNCH("k_adam", st, k_adam, dim3((unsigned)grid_of(n)), dim3(256), 0, p, g, m, v, n,
            (float)lr, (float)b1, (float)b2, (float)eps, c1, c2);
  return HC_OK;
}

}  // extern "C"
