// Register-blocked fused 2-D separable filter (float32) for gfx950.
//
// Reference behaviour: two calls of the correlate1d line kernel in
// _run_1d_filters (python/cucim/src/cucim/skimage/_vendored/
// _ndimage_filters_core.py:124-163), axis 0 then axis 1, float32 weights,
// float32 intermediate; boundary modes of _vendored/_ndimage_util.py:186-308.
//
// k_sep2d_reg<K, TY>: one work-group per TY x TX output tile (TX from
// reg_strip_outputs(K)); lane c owns staged column x0 - anchor1 + c.
#include <cstring>

#include "hc_stencil.h"

namespace hc {

struct Reg2dArgs {
  const float* in;
  float* out;
  int64_t H, W;
  int64_t ntx, ntiles;
  int TX;
  int a0, a1;  // anchors of the vertical / horizontal kernel
  int mode0, mode1;
  float cval;
};

template <int K, int TY>
__global__ __launch_bounds__(256) void k_sep2d_reg(const Reg2dArgs a, const RegTaps<K> t0,
                                                   const RegTaps<K> t1) {
  __shared__ __align__(16) float U[TY * kRegUStride];
  const int64_t tile = xcd_swizzle(blockIdx.x, gridDim.x);
  const int64_t ty = tile / a.ntx, tx = tile - ty * a.ntx;
  const int64_t y0 = ty * TY, x0 = tx * a.TX;
  const int c = threadIdx.x;
  // staged column; -1: outside the image in constant mode (intermediate = cval)
  const int64_t xr = remap(x0 - a.a1 + c, a.W, a.mode1);
  const int64_t ya = y0 - a.a0;
  float v[TY + K - 1];
  if (ya >= 0 && ya + TY + K - 1 <= a.H) {
    // every staged row inside the image (the tile is not at the top or
    // bottom border): consecutive rows, no boundary mapping
    const float* col = a.in + ya * a.W + (xr >= 0 ? xr : 0);
#pragma unroll
    for (int j = 0; j < TY + K - 1; ++j) v[j] = col[(int64_t)j * a.W];
    if (xr < 0) {
#pragma unroll
      for (int j = 0; j < TY + K - 1; ++j) v[j] = a.cval;
    }
  } else {
#pragma unroll
    for (int j = 0; j < TY + K - 1; ++j) {
      const int64_t yr = remap(ya + j, a.H, a.mode0);
      v[j] = (yr >= 0 && xr >= 0) ? a.in[yr * a.W + xr] : a.cval;
    }
  }
  float acc[TY];
  reg_vpass<K, TY>(acc, v, t0);
#pragma unroll
  for (int i = 0; i < TY; ++i) U[i * kRegUStride + c] = xr < 0 ? a.cval : acc[i];
  __syncthreads();
  const int nq = a.TX / kRegP;
  for (int it = threadIdx.x; it < TY * nq; it += 256) {
    const int i = it / nq, o = (it - i * nq) * kRegP;
    const int64_t y = y0 + i, x = x0 + o;
    if (y >= a.H || x >= a.W) continue;
    float r[kRegP];
    reg_hpass<K>(r, U + i * kRegUStride + o, t1);
    float* op = a.out + y * a.W + x;
    if (x + kRegP <= a.W && ((uintptr_t)op & 15) == 0) {
      // streaming stores: the output is not read again by this launch
      typedef float f4v __attribute__((ext_vector_type(4)));
      __builtin_nontemporal_store(f4v{r[0], r[1], r[2], r[3]}, reinterpret_cast<f4v*>(op));
      __builtin_nontemporal_store(f4v{r[4], r[5], r[6], r[7]}, reinterpret_cast<f4v*>(op) + 1);
    } else {
#pragma unroll
      for (int j = 0; j < kRegP; ++j)
        if (x + j < a.W) op[j] = r[j];
    }
  }
}

namespace {

template <int K>
void fill_taps(RegTaps<K>* r, const hc_taps* t) {
  std::memset(r, 0, sizeof(*r));
  for (int64_t k = 0; k < t->K; ++k) r->w[k] = (float)t->w[k];
}

template <int K, int TY>
int launch_reg(Reg2dArgs a, const hc_taps* t0, const hc_taps* t1, hipStream_t st) {
  RegTaps<K> r0, r1;
  fill_taps<K>(&r0, t0);
  fill_taps<K>(&r1, t1);
  a.TX = reg_strip_outputs(K);
  a.ntx = (a.W + a.TX - 1) / a.TX;
  const int64_t nty = (a.H + TY - 1) / TY;
  a.ntiles = a.ntx * nty;
  if (a.ntiles > 0x7fffffff) return fail(HC_ERR_VALUE, "array too large");
  HC_LAUNCH("k_sep2d_reg", st, (k_sep2d_reg<K, TY>), dim3((unsigned)a.ntiles), dim3(256), 0,
            a, r0, r1);
  return HC_OK;
}

template <int K>
int launch_reg_ty(const Reg2dArgs& a, const hc_taps* t0, const hc_taps* t1, hipStream_t st) {
  // Default: 8 rows up to 4 Mpixel (2048^2: more work-groups for a launch
  // that is one wave of the chip, 18 vs 19 us), 16 above (8, 16 and 32 rows
  // measure within noise of each other at 8192^2)
  int ty = tune_get(HC_TUNE_SEP2D_TY);
  if (ty != 8 && ty != 16 && ty != 32) ty = a.H * a.W <= (int64_t(1) << 22) ? 8 : 16;
  if (ty == 8) return launch_reg<K, 8>(a, t0, t1, st);
  return ty == 32 ? launch_reg<K, 32>(a, t0, t1, st) : launch_reg<K, 16>(a, t0, t1, st);
}

}  // namespace

int launch_sep2d_reg(const hc_array* in, const hc_array* out, const hc_taps* t0,
                     const hc_taps* t1, int mode0, int mode1, double cval, hipStream_t st) {
  if (in->dtype != HC_FLOAT32 || out->dtype != HC_FLOAT32 || t0->w_f64 || t1->w_f64) return -1;
  if (!is_c_contiguous(in) || !is_c_contiguous(out) || in->ndim != 2) return -1;
  const int64_t K = t0->K > t1->K ? t0->K : t1->K;
  if (K > kRegMaxK) return -1;
  Reg2dArgs a;
  std::memset(&a, 0, sizeof(a));
  a.in = (const float*)in->data;
  a.out = (float*)out->data;
  a.H = in->shape[0];
  a.W = in->shape[1];
  a.a0 = t0->anchor;
  a.a1 = t1->anchor;
  a.mode0 = mode0;
  a.mode1 = mode1;
  a.cval = (float)cval;
  if (a.H == 0 || a.W == 0) return HC_OK;
  if (K <= 3) return launch_reg_ty<3>(a, t0, t1, st);
  if (K <= 5) return launch_reg_ty<5>(a, t0, t1, st);
  if (K <= 7) return launch_reg_ty<7>(a, t0, t1, st);
  if (K <= 9) return launch_reg_ty<9>(a, t0, t1, st);
  if (K <= 13) return launch_reg_ty<13>(a, t0, t1, st);
  if (K <= 17) return launch_reg_ty<17>(a, t0, t1, st);
  if (K <= 21) return launch_reg_ty<21>(a, t0, t1, st);
  if (K <= 25) return launch_reg_ty<25>(a, t0, t1, st);
  return launch_reg_ty<33>(a, t0, t1, st);
}

// ------------------------------------------------------------------------
// Batched rgb2gray -> gaussian -> edge magnitude over uneven tiles.
//
// Reference behaviour, per tile (python/cucim/src/cucim/skimage/):
//   color/colorconv.py:1037-1095   gray = 0.2125 R + 0.7154 G + 0.0721 B of
//                                   img_as_float(rgb) (float32 for uint8),
//                                   summed in double, stored in float32
//   _shared/filters.py:21-166       gaussian: axis 0 then axis 1 passes
//   filters/edges.py:148-241        edge magnitude: edge pass then smoothing
//                                   pass per axis, sqrt(A*A + B*B)
// Every stage rounds to float32 exactly where the separate calls store
// their float32 outputs, so the fused result equals the three calls'.
//
// One work item = TY x TX output pixels of one tile.  The gaussian output G
// is formed with a one-pixel halo (TY + 2 rows, GX = TX + 2 columns) so the
// 3x3 edge stencils read it from LDS; halo rows / columns outside the image
// are replaced by the edge filter's boundary values (reflect, mirror,
// nearest: the remapped G value, which lies inside the item; constant: 0,
// as the reference does not forward cval to the edge filter).
namespace {

struct GgsArgs {
  const hc_tile2d* tiles;
  const int32_t* item_tile;
  int64_t n_items;
  int TX;      // output columns per item
  int ag;      // gaussian anchor
  int gmode;   // gaussian boundary mode
  int emode;   // edge-filter boundary mode (not wrap)
  float gcval;
  float scale;  // float32(1 / 255)
  float e[3], s[3];  // edge / smoothing correlation weights
};

// Image pointers read from the tile table are generic; as generic (flat)
// accesses every load would also count against the LDS counter, and the
// compiler then waits for all loads in flight before using any.  They are
// global memory, so they are cast to the global address space.
typedef const __attribute__((address_space(1))) uint8_t* gu8_ptr;
typedef __attribute__((address_space(1))) float* gf32_ptr;

// gray = float32(0.2125 r' + 0.7154 g' + 0.0721 b'), products and sums in
// double, r' = float32(r * float32(1/255)) (colorconv.py:1037-1095 after
// img_as_float).  Each double product depends on one byte only: tables of the
// 256 products per channel, filled per work-group in LDS, give the same bits
// with two double additions per pixel.
struct GrayLut {
  double r[256], g[256], b[256];
};

__device__ __forceinline__ void gray_lut_fill(GrayLut* t, float scale) {
  for (int i = threadIdx.x; i < 256; i += blockDim.x) {
    const double v = (double)((float)i * scale);
    t->r[i] = mul_rn(0.2125, v);
    t->g[i] = mul_rn(0.7154, v);
    t->b[i] = mul_rn(0.0721, v);
  }
}

__device__ __forceinline__ float gray_lut(gu8_ptr p, const GrayLut* t) {
  return (float)add_rn(add_rn(t->r[p[0]], t->g[p[1]]), t->b[p[2]]);
}

// one tap with the reference's zero-weight skip
__device__ __forceinline__ float etap(float acc, float x, float w) {
  return wnz(w) ? __builtin_fmaf(x, w, acc) : acc;
}

}  // namespace

template <int K, int TY>
__global__ __launch_bounds__(256) void k_tiles_ggs(const GgsArgs a, const RegTaps<K> tg) {
  constexpr int GR = TY + 2;  // gaussian rows with the edge halo
  // V: the vertical pass's output; the horizontal pass overwrites it in
  // place with G (the gaussian with its edge halo)
  __shared__ __align__(16) float V[GR * kRegUStride];
  float* const G = V;
  __shared__ GrayLut glut;
  gray_lut_fill(&glut, a.scale);
  __syncthreads();
  const int GX = a.TX + 2;
  const int c = threadIdx.x;
  const int64_t G_ = gridDim.x;
  const int64_t slot = xcd_swizzle(blockIdx.x, G_);
  for (int64_t base = 0; base < a.n_items; base += G_) {
    const int64_t it = base + slot;
    if (it >= a.n_items) break;  // uniform per work-group
    const hc_tile2d d = a.tiles[a.item_tile[it]];
    const int64_t H = d.H, W = d.W;
    const int64_t local = it - d.item0;
    const int64_t nbx = (W + a.TX - 1) / a.TX;
    const int64_t by = local / nbx, bx = local - by * nbx;
    const int64_t y0 = by * TY, x0 = bx * a.TX;
    const gu8_ptr in = (gu8_ptr)d.in;
    // 1. gray + vertical gaussian pass, one staged column per lane
    {
      const int64_t xr = remap(x0 - 1 - a.ag + c, W, a.gmode);
      const int64_t ya = y0 - 1 - a.ag;
      float v[GR + K - 1];
      if (ya >= 0 && ya + GR + K - 1 <= H) {
        // every staged row inside the tile: consecutive rows, no boundary
        // mapping
        const gu8_ptr col = in + 3 * (ya * W + (xr >= 0 ? xr : 0));
#pragma unroll
        for (int j = 0; j < GR + K - 1; ++j) v[j] = gray_lut(col + 3 * W * j, &glut);
        if (xr < 0) {
#pragma unroll
          for (int j = 0; j < GR + K - 1; ++j) v[j] = a.gcval;
        }
      } else {
#pragma unroll
        for (int j = 0; j < GR + K - 1; ++j) {
          const int64_t yr = remap(ya + j, H, a.gmode);
          v[j] = (yr >= 0 && xr >= 0) ? gray_lut(in + 3 * (yr * W + xr), &glut) : a.gcval;
        }
      }
      float acc[GR];
      reg_vpass<K, GR>(acc, v, tg);
#pragma unroll
      for (int i = 0; i < GR; ++i) V[i * kRegUStride + c] = xr < 0 ? a.gcval : acc[i];
    }
    __syncthreads();
    // 2. horizontal gaussian pass into G, in place: every item's inputs are
    // read before any item's outputs are written
    {
      constexpr int nq = reg_strip_outputs(K) / kRegP, NQ = (GR * nq + 255) / 256;
      float r[NQ][kRegP];
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        const int q = threadIdx.x + 256 * m;
        if (q < GR * nq) {
          const int i = q / nq, o = (q - i * nq) * kRegP;
          reg_hpass<K>(r[m], V + i * kRegUStride + o, tg);
        }
      }
      __syncthreads();
#pragma unroll
      for (int m = 0; m < NQ; ++m) {
        const int q = threadIdx.x + 256 * m;
        if (q < GR * nq) {
          const int i = q / nq, o = (q - i * nq) * kRegP;
#pragma unroll
          for (int j = 0; j < kRegP; ++j) G[i * kRegUStride + o + j] = r[m][j];
        }
      }
    }
    __syncthreads();
    // 3. edge-filter boundary of G (items touching the image border only)
    const bool top = y0 == 0, bottom = y0 + TY + 1 > H, left = x0 == 0,
               right = x0 + a.TX + 1 > W;
    if (top || bottom) {
      const bool cst = a.emode == HC_MODE_CONSTANT;
      for (int e = threadIdx.x; e < 2 * GX; e += 256) {
        const int which = e / GX, gc = e - which * GX;
        const int64_t y = which == 0 ? y0 - 1 : H;  // the halo row to replace
        const int64_t gr = y - (y0 - 1);
        if ((which == 0 && !top) || (which == 1 && (!bottom || gr > TY + 1))) continue;
        float val = 0.f;
        if (!cst) val = G[(remap(y, H, a.emode) - (y0 - 1)) * kRegUStride + gc];
        G[gr * kRegUStride + gc] = val;
      }
      __syncthreads();
    }
    if (left || right) {
      const bool cst = a.emode == HC_MODE_CONSTANT;
      for (int e = threadIdx.x; e < 2 * GR; e += 256) {
        const int which = e / GR, gr = e - which * GR;
        const int64_t x = which == 0 ? x0 - 1 : W;
        const int64_t gc = x - (x0 - 1);
        if ((which == 0 && !left) || (which == 1 && (!right || gc > GX - 1))) continue;
        float val = 0.f;
        if (!cst) val = G[gr * kRegUStride + (remap(x, W, a.emode) - (x0 - 1))];
        G[gr * kRegUStride + gc] = val;
      }
      __syncthreads();
    }
    // 4. edge magnitude, one output column per lane
    if (c < a.TX && x0 + c < W) {
      const gf32_ptr out = (gf32_ptr)d.out;
      // a 3x3 window slides down the lane's column: G rows i, i + 1, i + 2
      // (image rows y - 1 .. y + 1) at columns x - 1 .. x + 1, and the
      // horizontal edge response of each row, are carried to the next row
      const float* gc = G + c;
      auto edge_x = [&](const float (&g)[3]) {
        float u = 0.f;
        u = etap(u, g[0], a.e[0]);
        u = etap(u, g[1], a.e[1]);
        return etap(u, g[2], a.e[2]);  // edge along axis 1 of one row
      };
      float gm[3], g0[3], gp[3];
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        gm[q] = gc[q];
        g0[q] = gc[kRegUStride + q];
      }
      float exm = edge_x(gm), ex0 = edge_x(g0);
#pragma unroll 4
      for (int i = 0; i < TY; ++i) {
        const int64_t y = y0 + i;
        if (y >= H) break;
#pragma unroll
        for (int q = 0; q < 3; ++q) gp[q] = gc[(i + 2) * kRegUStride + q];
        const float exp_ = edge_x(gp);
        float ey[3];
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          float v = 0.f;
          v = etap(v, gm[q], a.e[0]);
          v = etap(v, g0[q], a.e[1]);
          ey[q] = etap(v, gp[q], a.e[2]);  // edge along axis 0 at column x - 1 + q
        }
        float A = 0.f, B = 0.f;
        A = etap(A, ey[0], a.s[0]);
        B = etap(B, exm, a.s[0]);
        A = etap(A, ey[1], a.s[1]);
        B = etap(B, ex0, a.s[1]);
        A = etap(A, ey[2], a.s[2]);
        B = etap(B, exp_, a.s[2]);
        __builtin_nontemporal_store(sqrtf(add_rn(mul_rn(A, A), mul_rn(B, B))),
                                    out + y * W + x0 + c);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
          gm[q] = g0[q];
          g0[q] = gp[q];
        }
        exm = ex0;
        ex0 = exp_;
      }
    }
    __syncthreads();  // V and G are reused by the next item
  }
}

namespace {

template <int K, int TY>
int launch_ggs(GgsArgs a, const hc_taps* g, hipStream_t st) {
  RegTaps<K> t;
  fill_taps<K>(&t, g);
  a.TX = reg_strip_outputs(K) - 2;
  const int64_t grid = a.n_items < 256 * 8 ? a.n_items : 256 * 8;
  HC_LAUNCH("k_tiles_ggs", st, (k_tiles_ggs<K, TY>), dim3((unsigned)grid), dim3(256), 0, a, t);
  return HC_OK;
}

template <int TY>
int launch_ggs_k(const GgsArgs& a, const hc_taps* g, hipStream_t st) {
  const int64_t K = g->K;
  if (K <= 5) return launch_ggs<5, TY>(a, g, st);
  if (K <= 9) return launch_ggs<9, TY>(a, g, st);
  if (K <= 13) return launch_ggs<13, TY>(a, g, st);
  if (K <= 17) return launch_ggs<17, TY>(a, g, st);
  if (K <= 25) return launch_ggs<25, TY>(a, g, st);
  return launch_ggs<33, TY>(a, g, st);
}

int ggs_k_compiled(int64_t K) {
  return K <= 5 ? 5 : K <= 9 ? 9 : K <= 13 ? 13 : K <= 17 ? 17 : K <= 25 ? 25 : 33;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int64_t hc_batch2d_plan(hc_tile2d* tiles, int n, int K, int ty, int32_t* item_tile,
                        int64_t cap) {
  if (!tiles || n < 0) return fail(HC_ERR_VALUE, "null tiles"), -1;
  if (K < 1 || K > kRegMaxK) return fail(HC_ERR_VALUE, "kernel length outside 1..33"), -1;
  if (ty != 16 && ty != 32) return fail(HC_ERR_VALUE, "tile rows must be 16 or 32"), -1;
  const int64_t TX = reg_strip_outputs(ggs_k_compiled(K)) - 2;
  int64_t total = 0;
  for (int t = 0; t < n; ++t) {
    if (tiles[t].H < 0 || tiles[t].W < 0) return fail(HC_ERR_VALUE, "negative tile shape"), -1;
    tiles[t].item0 = total;
    const int64_t items = ((tiles[t].H + ty - 1) / ty) * ((tiles[t].W + TX - 1) / TX);
    if (item_tile) {
      if (total + items > cap) return fail(HC_ERR_VALUE, "item table too small"), -1;
      for (int64_t i = 0; i < items; ++i) item_tile[total + i] = t;
    }
    total += items;
  }
  if (total > 0x7fffffff) return fail(HC_ERR_VALUE, "too many work items"), -1;
  return total;
}

int hc_batch2d_gray_gauss_sobel(const hc_tile2d* tiles_dev, int n, const int32_t* item_tile_dev,
                                int64_t n_items, const hc_taps* gauss, int gauss_mode,
                                double gauss_cval, const double* edge3, const double* smooth3,
                                int edge_mode, int ty, void* stream) {
  if (!tiles_dev || !item_tile_dev || !gauss || !edge3 || !smooth3)
    return fail(HC_ERR_VALUE, "null argument");
  if (n < 0 || n_items < 0) return fail(HC_ERR_VALUE, "negative count");
  if (gauss->w_f64) return fail(HC_ERR_VALUE, "gaussian weights must be float32");
  if (gauss->K < 1 || gauss->K > kRegMaxK)
    return fail(HC_ERR_RESOURCE, "gaussian kernel longer than 33 taps");
  if (gauss->anchor < 0 || gauss->anchor >= gauss->K) return fail(HC_ERR_VALUE, "invalid origin");
  if (gauss_mode < HC_MODE_REFLECT || gauss_mode > HC_MODE_CONSTANT ||
      edge_mode < HC_MODE_REFLECT || edge_mode > HC_MODE_CONSTANT)
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  if (edge_mode == HC_MODE_WRAP)
    return fail(HC_ERR_UNSUPPORTED, "wrap mode edges are not fused (halo outside the item)");
  if (ty != 16 && ty != 32) return fail(HC_ERR_VALUE, "tile rows must be 16 or 32");
  if (n_items == 0) return HC_OK;
  GgsArgs a;
  std::memset(&a, 0, sizeof(a));
  a.tiles = tiles_dev;
  a.item_tile = item_tile_dev;
  a.n_items = n_items;
  a.ag = gauss->anchor;
  a.gmode = gauss_mode;
  a.emode = edge_mode;
  a.gcval = (float)gauss_cval;
  a.scale = (float)(1.0 / 255.0);
  for (int k = 0; k < 3; ++k) {
    a.e[k] = (float)edge3[k];
    a.s[k] = (float)smooth3[k];
  }
  hipStream_t st = (hipStream_t)stream;
  return ty == 32 ? launch_ggs_k<32>(a, gauss, st) : launch_ggs_k<16>(a, gauss, st);
}

}  // extern "C"
