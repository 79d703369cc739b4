// Fused Hessian-by-Gaussian-derivatives, eigenvalues and Frangi vesselness.
//
// Reference (paths relative to python/cucim/src/cucim/skimage/):
//   feature/corner.py:161-254  _hessian_matrix_with_gaussian: gradients
//       g_a = gaussian_filter(I, order=e_a) then H_ab = gaussian_filter(g_a,
//       order=e_b); every 1-D pass re-extends its input at the border.
//   feature/corner.py:473-756  closed-form eigenvalues
//   filters/ridges.py:322-476  frangi (gamma from the first sigma, running max)
//
// Restructuring.  Along one axis x the reference applies two 1-D passes to
// each H_ab: first d (order-1 kernel) if x == a else s (order-0), then d if
// x == b else s.  Passes along different axes commute exactly, so H_ab is a
// product over axes of per-axis "chains" (first op, second op).  One launch
// per axis keeps whole lines resident in LDS, computes the first ops, re-
// extends their results with the boundary mode (exactly what the
// reference's second pass sees), applies the second ops and writes only the
// distinct chain results.  3-D: axis 0 (3 chains from the image), axis 1
// (6 chains), axis 2 (the six H elements, consumed in registers by the
// eigenvalue / Frangi epilogue).  Intermediates live in the caller's H (or
// work) arrays, in place, because each work-group reads its lines completely
// before writing the same lines.
#include <algorithm>
#include <type_traits>
#include <cstring>
#include <utility>
#include <vector>

#include "hc_conv.h"
#include "hc_eig.h"

namespace hc {

constexpr int kChC = 16;        // strided chain pass: lines per work-group
constexpr int kChS = kChC + 1;  // LDS row stride
constexpr int kFinIPT = 1;      // final pass: items per thread
constexpr int kFinP = 8;        // final pass: consecutive outputs per item
constexpr int kMaxLds = 160 * 1024;

struct ChainOut {
  void* out;
  int first;   // 0: smoothing (order 0), 1: derivative (order 1)
  int second;  // kernel of the second pass
};

struct ChainJob {
  const void* in;
  int in_dt;
  int nout;
  ChainOut o[3];
};

template <typename W>
struct ChainArgs {
  ChainJob job[3];
  int64_t N, inner, nlines;
  int E;        // LDS line length: roundup(N) + K
  int Nr;       // N rounded up to the output block
  int rows;     // contiguous variant: lines per work-group
  int mode;
  int negate;   // first pass of Frangi with black_ridges=False: image = -image
  // first pass, optional (device): a constant subtracted from the image (see
  // image_center); every Hessian element is unchanged by it in exact
  // arithmetic, the staged values are smaller
  const double* center;
  W cval;
  TapPack<W> S, D;  // same padded length and anchor
  int t0, Km;       // matrix-core variant: taps t0 .. t0 + Km - 1 hold every
                    // nonzero weight of S and D
  int shift;        // matrix-core variant: buffer row of line position 0
  int njobs;        // matrix-core variant: jobs (persistent grid)
};

// --------------------------------------------------- strided chain pass --
// Two LDS line buffers: the input lines are staged in A; the first passes
// (smoothing and/or derivative) run into registers, then, after a barrier,
// overwrite A (smoothing) and B (derivative); the second passes read A / B.
constexpr int kChP = 8;  // outputs per lane along a line
constexpr int kChMaxItems = 4;  // first-pass items a lane holds in registers

template <typename W, int NI>
__global__ __launch_bounds__(256) void k_chain_strided(const ChainArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  W* bufA = reinterpret_cast<W*>(smem_raw);
  const int E = a.E, Nr = a.Nr;
  W* bufB = bufA + E * kChS;
  W* wS = bufB + E * kChS;
  W* wD = wS + a.S.K + 4;
  const ChainJob& jb = a.job[blockIdx.y];
  const int anc = a.S.anchor;
  const int64_t L0 = xcd_swizzle(blockIdx.x, gridDim.x) * kChC;
  const int tid = threadIdx.x;
  stage_taps<W>(wS, a.S, tid, 256);
  stage_taps<W>(wD, a.D, tid, 256);
  bool need_s = false, need_d = false;
  for (int k = 0; k < jb.nout; ++k) {
    need_s |= jb.o[k].first == 0;
    need_d |= jb.o[k].first == 1;
  }
  // 1. stage whole lines (+ halo) of the job's input into A
  bool bad;
  {
    const int64_t N = a.N, inner = a.inner, nlines = a.nlines;
    const int mode = a.mode;
    bad = stage_any<W>(bufA, E * kChC, tid, 256, jb.in, jb.in_dt, a.cval, a.negate != 0,
                       [=](int e, int& slot) -> int64_t {
                         const int q = e / kChC, c = e - q * kChC;
                         slot = q * kChS + c;
                         const int64_t L = L0 + c;
                         if (L >= nlines) return -1;
                         const int64_t r = remap((int64_t)q - anc, N, mode);
                         if (r < 0) return -1;
                         const int64_t lo = L / inner, li = L - lo * inner;
                         return lo * N * inner + li + r * inner;
                       },
                       a.center);
  }
  const bool skip0 = __syncthreads_or(bad);
  // 2. first passes over positions [0, Nr) into registers (<= NI items/lane)
  const int nitems = (Nr / kChP) * kChC;
  W fs[NI][kChP], fd[NI][kChP];
  with_skip(skip0, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
#pragma unroll
    for (int i = 0; i < NI; ++i) {
      const int it = tid + 256 * i;
#pragma unroll
      for (int j = 0; j < kChP; ++j) { fs[i][j] = (W)0; fd[i][j] = (W)0; }
      if (it < nitems) {
        const int ch = it / kChC, c = it - ch * kChC;
        const W* x = bufA + (kChP * ch) * kChS + c;
        W acc[2][kChP];
#pragma unroll
        for (int j = 0; j < kChP; ++j) { acc[0][j] = (W)0; acc[1][j] = (W)0; }
        if (need_s && need_d) {
          const W* const wls[2] = {wS, wD};
          const TapPack<W>* const tps[2] = {&a.S, &a.D};
          conv_colP<W, kChS, 2, kChP, SK>(acc, x, wls, tps, a.S.K);
        } else {
          W acc1[1][kChP];
#pragma unroll
          for (int j = 0; j < kChP; ++j) acc1[0][j] = (W)0;
          const W* const wls[1] = {need_s ? wS : wD};
          const TapPack<W>* const tps[1] = {need_s ? &a.S : &a.D};
          conv_colP<W, kChS, 1, kChP, SK>(acc1, x, wls, tps, a.S.K);
#pragma unroll
          for (int j = 0; j < kChP; ++j) acc[need_s ? 0 : 1][j] = acc1[0][j];
        }
#pragma unroll
        for (int j = 0; j < kChP; ++j) { fs[i][j] = acc[0][j]; fd[i][j] = acc[1][j]; }
      }
    }
  });
  __syncthreads();  // every read of the staged input is done
  bool bad1 = false;
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int it = tid + 256 * i;
    if (it < nitems) {
      const int ch = it / kChC, c = it - ch * kChC;
#pragma unroll
      for (int j = 0; j < kChP; ++j) {
        const int q = anc + kChP * ch + j;
        bufA[q * kChS + c] = fs[i][j];
        bufB[q * kChS + c] = fd[i][j];
        bad1 |= !(finite(fs[i][j]) && finite(fd[i][j]));
      }
    }
  }
  const bool skip1 = __syncthreads_or(bad1) || skip0;
  // 3. boundary-extend the first-pass results
  const int nh = E - a.N;  // positions [0, anc) and [anc + N, E)
  for (int e = tid; e < nh * kChC; e += 256) {
    const int h = e / kChC, c = e - h * kChC;
    const int q = h < anc ? h : h + (int)a.N;
    const int64_t r = remap((int64_t)q - anc, a.N, a.mode);
    W vs = a.cval, vd = a.cval;
    if (r >= 0) {
      vs = bufA[(r + anc) * kChS + c];
      vd = bufB[(r + anc) * kChS + c];
    }
    bufA[q * kChS + c] = vs;
    bufB[q * kChS + c] = vd;
  }
  __syncthreads();
  // 4. second passes, sharing window reads between outputs of one source
  with_skip(skip1, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
    for (int it = tid; it < nitems; it += 256) {
      const int ch = it / kChC, c = it - ch * kChC;
      const int64_t L = L0 + c;
      if (L >= a.nlines) continue;
      const int64_t lo = L / a.inner, li = L - lo * a.inner;
      const int64_t base = lo * a.N * a.inner + li;
#pragma unroll 1
      for (int src = 0; src < 2; ++src) {
        int ks = -1, kd = -1;
        for (int k = 0; k < jb.nout; ++k) {
          if (jb.o[k].first != src) continue;
          if (jb.o[k].second == 0) ks = k; else kd = k;
        }
        if (ks < 0 && kd < 0) continue;
        const W* x = (src ? bufB : bufA) + (kChP * ch) * kChS + c;
        W r[2][kChP];
#pragma unroll
        for (int j = 0; j < kChP; ++j) { r[0][j] = (W)0; r[1][j] = (W)0; }
        if (ks >= 0 && kd >= 0) {
          const W* const wls[2] = {wS, wD};
          const TapPack<W>* const tps[2] = {&a.S, &a.D};
          conv_colP<W, kChS, 2, kChP, SK>(r, x, wls, tps, a.S.K);
        } else {
          W r1[1][kChP];
#pragma unroll
          for (int j = 0; j < kChP; ++j) r1[0][j] = (W)0;
          const W* const wls[1] = {ks >= 0 ? wS : wD};
          const TapPack<W>* const tps[1] = {ks >= 0 ? &a.S : &a.D};
          conv_colP<W, kChS, 1, kChP, SK>(r1, x, wls, tps, a.S.K);
#pragma unroll
          for (int j = 0; j < kChP; ++j) r[ks >= 0 ? 0 : 1][j] = r1[0][j];
        }
        W* os = ks >= 0 ? reinterpret_cast<W*>(jb.o[ks].out) : nullptr;
        W* od = kd >= 0 ? reinterpret_cast<W*>(jb.o[kd].out) : nullptr;
#pragma unroll
        for (int j = 0; j < kChP; ++j) {
          const int64_t n = kChP * ch + j;
          if (n < a.N) {
            if (os) os[base + n * a.inner] = r[0][j];
            if (od) od[base + n * a.inner] = r[1][j];
          }
        }
      }
    }
  });
}

// ------------------------------------- strided chain pass, matrix cores --
// The same two passes on the matrix cores (float32): a pass along the line,
// for 16 output positions p and the work-group's 16 lines c, is the product
//   Y[p][c] = sum_q T[p][q] X[q][c],  T[p][q] = w[q - p]  (Toeplitz),
// done with v_mfma_f32_16x16x4f32 over input blocks q = 16 (b + d) + 4 u + k:
// A[i][k] = w[16 d + 4 u + k - i] (from a zero-padded LDS copy of the taps),
// B[k][c] = X[q][c] (the LDS line buffer, conflict-free), D[i][c] -> lane l
// holds positions 4 (l / 16) + r of line l % 16.  The sums are rounded as
// the matrix core rounds them, not tap by tap; a work-group whose lines hold
// a non-finite value computes every output instead as the reference does
// (float32 fused multiply-add per non-zero tap, in tap order), in the same
// register layout.
//
// Only the span t0 .. t0 + Km - 1 of taps that holds every nonzero weight of
// both sets is used (Gaussian tails round to zero in float32).  Buffer row r
// holds line position r - shift, shift = anchor - t0, so output p reads rows
// p .. p + Km - 1.  A work-group computes kChainSPW consecutive stages (a
// tile of 16 lines of one job) and keeps the next stage's input in flight
// in registers while the current one is computed.
typedef float f32x4_t __attribute__((ext_vector_type(4)));

// Row stride (floats) of the matrix-core passes' line buffers: 16, so that a
// wave's B fragment (4 rows x 16 lines) is 64 consecutive words (no bank
// conflict) and two buffers of a 256-voxel line fit four work-groups per CU.
// The chain pass stages strided lines in this position-major layout (16
// adjacent lines are 16 consecutive words in memory too).
constexpr int kMfS = 16;

// The final pass stages contiguous lines line-major instead: line l's row r
// at X[l * LS + r], LS = 2 (mod 4) and >= E.  Then a lane's 4 consecutive
// positions of one line are two 8-byte words, and the staging stores (16
// lines at one position per 16 lanes), the B fragments (16 lines x 2 rows
// per 32 lanes) and the halo copies fall in distinct banks of the 32 that
// serve each half-wave.
__host__ __device__ constexpr int lm_stride(int E) { return E + ((6 - E % 4) % 4); }

// Length of the zero-padded tap arrays: the passes read wp[15 + 4 s + k - i]
// for steps s < ceil((Km + 15) / 4), k < 4, i < 16.
__host__ __device__ constexpr int mfma_wp(int Km) { return 4 * ((Km + 18) / 4) + 16; }

__device__ __forceinline__ f32x4_t mfma4(float a, float b, f32x4_t c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// The wave's index in the work-group, as a wave-uniform (scalar) value, so
// that tests on it are scalar branches rather than lane masks.
__device__ __forceinline__ int wave_id() {
  return __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
}

// Values per thread of one tile's interior held in flight: 16 lines of up to
// 16 NW NBW positions (NBW blocks of 16 per wave) over 64 NW threads.
__host__ __device__ constexpr int mfma_npf(int NBW) { return 4 * NBW; }

// Waves per SIMD the compiler budgets registers for: four blocks per wave
// need the larger share (the chain pass's two first-op results, the final
// pass's six Hessian inputs); two or one leave room for more waves.
__host__ __device__ constexpr int chain_wpe(int NBW) { return NBW >= 3 ? 5 : 6; }
__host__ __device__ constexpr int final_wpe(int NBW) { return NBW >= 3 ? 4 : 6; }

// Blocks b = wave + NW i (i < NBW) of 16 positions: acc[s][i] for the NS
// weight sets wp[s] (zero-padded: trimmed tap t at wp[15 + t]).  The A
// fragment of a step of 4 inputs serves every block of the wave.
template <int NBW, int NS, int NW>
__device__ __forceinline__ void chain_pass_mfma(f32x4_t (&acc)[NS][NBW], const float* X,
                                                const float* const (&wp)[NS], int nd,
                                                int nblk, int wv, int lane, int Km, int ps,
                                                int ls) {
  // steps of 4 inputs that the band of 16 outputs needs (the last 16-row
  // block is cut short)
  const int nsteps = (Km + 15 + 3) / 4;
  const int kk = lane >> 4, ii = lane & 15;
#pragma unroll
  for (int i = 0; i < NBW; ++i) {
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s][i] = f32x4_t{0.f, 0.f, 0.f, 0.f};
  }
  // blocks past the last are computed on the last one's rows (no branch in
  // the loop); their results are never stored
  int rb[NBW];
#pragma unroll
  for (int i = 0; i < NBW; ++i) rb[i] = (min(wv + NW * i, nblk - 1) * 16 + kk) * ps + ii * ls;
  // Software-pipelined over two operand sets: the operands of step st + 1
  // are read from LDS while the products of step st run, so no product
  // waits for its own reads (a read past the last step re-reads the last
  // step's operands, which are then unused).
  const int w0 = 15 + kk - ii, last = nsteps - 1;
  auto fetch = [&](float (&a)[NS], float (&b)[NBW], int st) {
    const int t = min(st, last);
#pragma unroll
    for (int s = 0; s < NS; ++s) a[s] = wp[s][w0 + 4 * t];
#pragma unroll
    for (int i = 0; i < NBW; ++i) b[i] = X[rb[i] + 4 * t * ps];
  };
  auto step = [&](const float (&a)[NS], const float (&b)[NBW]) {
#pragma unroll
    for (int i = 0; i < NBW; ++i) {
#pragma unroll
      for (int s = 0; s < NS; ++s) acc[s][i] = mfma4(a[s], b[i], acc[s][i]);
    }
  };
  float a0[NS], b0[NBW], a1[NS], b1[NBW];
  fetch(a0, b0, 0);
  int st = 0;
  for (; st + 1 < nsteps; st += 2) {
    fetch(a1, b1, st + 1);
    step(a0, b0);
    fetch(a0, b0, st + 2);
    step(a1, b1);
  }
  if (st < nsteps) step(a0, b0);  // an odd step count's last step
  (void)nd;
}

// The reference's arithmetic for the same outputs, in the same layout (the
// taps outside the span are zero, so the same products in the same order).
template <int NBW, int NS, int NW>
__device__ __forceinline__ void chain_pass_exact(f32x4_t (&acc)[NS][NBW], const float* X,
                                                 const float* const (&wp)[NS], int Km,
                                                 int nblk, int wv, int lane, int ps, int ls) {
  const int kk = lane >> 4, ii = lane & 15;
#pragma unroll
  for (int i = 0; i < NBW; ++i) {
#pragma unroll
    for (int s = 0; s < NS; ++s) acc[s][i] = f32x4_t{0.f, 0.f, 0.f, 0.f};
    const int b = wv + NW * i;
    if (b >= nblk) continue;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int p = 16 * b + 4 * kk + r;
#pragma unroll
      for (int s = 0; s < NS; ++s) {
        float v = 0.f;
        for (int t = 0; t < Km; ++t) {
          const float w = wp[s][15 + t];
          if (w != 0.f) v = fmaf(X[(p + t) * ps + ii * ls], w, v);
        }
        acc[s][i][r] = v;
      }
    }
  }
}

template <int NBW, int NS, int NW>
__device__ __forceinline__ void chain_pass(bool exact, f32x4_t (&acc)[NS][NBW], const float* X,
                                           const float* const (&wp)[NS], int Km, int nd,
                                           int nblk, int wv, int lane, int ps = kMfS,
                                           int ls = 1) {
  if (exact)
    chain_pass_exact<NBW, NS, NW>(acc, X, wp, Km, nblk, wv, lane, ps, ls);
  else
    chain_pass_mfma<NBW, NS, NW>(acc, X, wp, nd, nblk, wv, lane, Km, ps, ls);
}

__device__ __forceinline__ float load_in(const void* in, int dt, int64_t g, bool negate,
                                         bool ctr, float c) {
  const float y =
      dt == HC_FLOAT32 ? reinterpret_cast<const float*>(in)[g] : load_as<float>(in, g, dt);
  return ctr ? (negate ? c - y : y - c) : (negate ? -y : y);
}

// Zero-padded trimmed taps of both sets into LDS.
__device__ __forceinline__ void stage_trimmed(float* wpS, float* wpD, const TapPack<float>& S,
                                              const TapPack<float>& D, int t0, int Km, int WP,
                                              int tid, int nt) {
  for (int e = tid; e < WP; e += nt) {
    const int t = e - 15;
    wpS[e] = (t >= 0 && t < Km) ? S.w[t0 + t] : 0.f;
    wpD[e] = (t >= 0 && t < Km) ? D.w[t0 + t] : 0.f;
  }
}

// Source row of every halo row of the line buffer (rows of position outside
// [0, N)): shift + the boundary-mapped position, or -1 for the constant.
// The same for every column and buffer, so computed once per work-group.
__device__ __forceinline__ void halo_table(short* hsrc, int E, int64_t N, int shift, int mode,
                                           int tid, int nt) {
  const int nh = E - (int)N;
  for (int h = tid; h < nh; h += nt) {
    const int q = h < shift ? h : h + (int)N;
    const int64_t r = remap((int64_t)q - shift, N, mode);
    hsrc[h] = (short)(r >= 0 ? (int)r + shift : -1);
  }
}

// Boundary-extend the halo rows of X (and Y) from the rows inside.
__device__ __forceinline__ void fill_halo(float* X, float* Y, int E, int N, int shift,
                                          const short* hsrc, float cval, int tid, int nt,
                                          int ps = kMfS, int ls = 1) {
  const int nh = E - N;
  for (int e = tid; e < nh * kChC; e += nt) {
    const int h = e >> 4, c = e & 15;
    const int q = h < shift ? h : h + N;
    const int src = hsrc[h];
    X[q * ps + c * ls] = src >= 0 ? X[src * ps + c * ls] : cval;
    if (Y) Y[q * ps + c * ls] = src >= 0 ? Y[src * ps + c * ls] : cval;
  }
}

// Stages (tile, job) per work-group of the matrix-core chain pass.
constexpr int kChainSPW = 1;

template <int NBW, int NW>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(chain_wpe(NBW)))) void
k_chain_mfma(const ChainArgs<float> a) {
  constexpr int NPF = mfma_npf(NBW), NT = 64 * NW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int E = a.E, Nr = a.Nr, Km = a.Km, shift = a.shift;
  const int nd = (Km + 30) / 16;  // input blocks per output block
  const int WP = mfma_wp(Km);     // padded tap array length
  float* bufA = reinterpret_cast<float*>(smem_raw);  // the one line buffer
  float* wpS = bufA + E * kMfS;
  float* wpD = wpS + WP;
  short* hsrc = reinterpret_cast<short*>(wpD + WP);
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int kk = lane >> 4, ii = lane & 15;
  stage_trimmed(wpS, wpD, a.S, a.D, a.t0, Km, WP, tid, NT);
  halo_table(hsrc, E, a.N, shift, a.mode, tid, NT);
  const int64_t N = a.N, inner = a.inner, nlines = a.nlines;
  const int mode = a.mode, nblk = Nr / 16;
  const int64_t ntile = (nlines + kChC - 1) / kChC, nst = ntile * a.njobs;
  const int64_t G = gridDim.x;
  // this thread stages line c = tid % 16, positions tid / 16 + 4 NW u
  const int cs = tid & 15, ns0 = tid >> 4;
  auto load = [&](float (&v)[NPF], int64_t st) {
    const int job = (int)(st / ntile);
    const int64_t L = (st - job * ntile) * kChC + cs;
    const ChainJob& jb = a.job[job];
    const bool ok = L < nlines;
    const int64_t lo = ok ? L / inner : 0, li = ok ? L - lo * inner : 0;
    const int64_t base = lo * N * inner + li;
    if (jb.in_dt == HC_FLOAT32) {
      // All NPF loads are issued before any is waited for: every load is
      // unconditional and in bounds (index 0 stands in for what lies
      // outside the line or past the last line), and the zeroing is applied
      // to the loaded values afterwards.  Written as a conditional load, the
      // compiler emits a lane-masked load per element followed by its wait,
      // so the loads of a lane would run one after another.
      const float* in = reinterpret_cast<const float*>(jb.in);
      const bool neg = a.negate != 0;
      const bool ctr = a.center != nullptr;
      const float c = ctr ? (float)*a.center : 0.f;
#pragma unroll
      for (int u = 0; u < NPF; ++u) {
        const int n = ns0 + 4 * NW * u;
        const bool use = ok && n < N;
        v[u] = in[base + (use ? (int64_t)n : 0) * inner];
      }
#pragma unroll
      for (int u = 0; u < NPF; ++u) {
        const int n = ns0 + 4 * NW * u;
        const bool use = ok && n < N;
        const float y = v[u];
        v[u] = use ? (ctr ? (neg ? c - y : y - c) : (neg ? -y : y)) : 0.f;
      }
    } else {
#pragma unroll
      for (int u = 0; u < NPF; ++u) {
        const int n = ns0 + 4 * NW * u;
        v[u] = (ok && n < N) ? load_in(jb.in, jb.in_dt, base + n * inner, a.negate != 0,
                                       a.center != nullptr,
                                       a.center ? (float)*a.center : 0.f)
                             : 0.f;
      }
    }
  };
  const bool cval_bad = !finite(a.cval);
  float pf[NPF];
  const int64_t st0 = xcd_swizzle(blockIdx.x, G) * kChainSPW;
  load(pf, st0);
#pragma unroll
  for (int k = 0; k < kChainSPW; ++k) {
    const int64_t st = st0 + k;
    if (st >= nst) break;
    const int job = (int)(st / ntile);
    const int64_t L0 = (st - job * ntile) * kChC;
    const ChainJob& jb = a.job[job];
    bool bad = false;
#pragma unroll
    for (int u = 0; u < NPF; ++u) {
      const int n = ns0 + 4 * NW * u;
      if (n < N) {
        bufA[(shift + n) * kMfS + cs] = pf[u];
        bad |= !finite(pf[u]);
      }
    }
    const bool skip0 = __syncthreads_or(bad) || cval_bad;
    // the work-group's next stage is in flight during this one's passes
    if (st + 1 < nst && st + 1 < st0 + kChainSPW) load(pf, st + 1);
    fill_halo(bufA, nullptr, E, (int)N, shift, hsrc, a.cval, tid, NT);
    __syncthreads();
    bool need_s = false, need_d = false;
    for (int k = 0; k < jb.nout; ++k) {
      need_s |= jb.o[k].first == 0;
      need_d |= jb.o[k].first == 1;
    }
    // first passes into registers
    f32x4_t f[2][NBW];
    if (need_s && need_d) {
      const float* const wp[2] = {wpS, wpD};
      chain_pass<NBW, 2, NW>(skip0, f, bufA, wp, Km, nd, nblk, wv, lane);
    } else {
      f32x4_t g[1][NBW];
      const float* const wp[1] = {need_s ? wpS : wpD};
      chain_pass<NBW, 1, NW>(skip0, g, bufA, wp, Km, nd, nblk, wv, lane);
#pragma unroll
      for (int i = 0; i < NBW; ++i) {
        f[0][i] = g[0][i];
        f[1][i] = g[0][i];
      }
    }
    __syncthreads();  // every read of the staged input is done
    bool bad1 = false;
#pragma unroll
    for (int i = 0; i < NBW; ++i) {
      if (wv + NW * i >= nblk) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) bad1 |= !(finite(f[0][i][r]) && finite(f[1][i][r]));
    }
    const int64_t L = L0 + ii;
    const bool lok = L < nlines;
    const int64_t lo = lok ? L / inner : 0, li = lok ? L - lo * inner : 0;
    const int64_t base = lo * N * inner + li;
    // second passes, per first op: its results go through the one line
    // buffer in turn (the other op's results wait in registers)
    bool skip1 = skip0, first_src = true;
#pragma unroll
    for (int src = 0; src < 2; ++src) {
      int ks = -1, kd = -1;
      for (int k = 0; k < jb.nout; ++k) {
        if (jb.o[k].first != src) continue;
        if (jb.o[k].second == 0) ks = k; else kd = k;
      }
      if (ks < 0 && kd < 0) continue;
      if (!first_src) __syncthreads();  // every read of the previous results is done
#pragma unroll
      for (int i = 0; i < NBW; ++i) {
        const int b = wv + NW * i;
        if (b >= nblk) continue;
#pragma unroll
        for (int r = 0; r < 4; ++r) bufA[(shift + 16 * b + 4 * kk + r) * kMfS + ii] = f[src][i][r];
      }
      if (first_src)
        skip1 = __syncthreads_or(bad1) || skip0;
      else
        __syncthreads();
      first_src = false;
      fill_halo(bufA, nullptr, E, (int)N, shift, hsrc, a.cval, tid, NT);
      __syncthreads();
      f32x4_t y[2][NBW];
      if (ks >= 0 && kd >= 0) {
        const float* const wp[2] = {wpS, wpD};
        chain_pass<NBW, 2, NW>(skip1, y, bufA, wp, Km, nd, nblk, wv, lane);
      } else {
        f32x4_t g[1][NBW];
        const float* const wp[1] = {ks >= 0 ? wpS : wpD};
        chain_pass<NBW, 1, NW>(skip1, g, bufA, wp, Km, nd, nblk, wv, lane);
#pragma unroll
        for (int i = 0; i < NBW; ++i) {
          y[0][i] = g[0][i];
          y[1][i] = g[0][i];
        }
      }
      float* os = ks >= 0 ? reinterpret_cast<float*>(jb.o[ks].out) : nullptr;
      float* od = kd >= 0 ? reinterpret_cast<float*>(jb.o[kd].out) : nullptr;
      if (lok) {
#pragma unroll
        for (int i = 0; i < NBW; ++i) {
          const int b = wv + NW * i;
          if (b >= nblk) continue;
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int64_t n = 16 * b + 4 * kk + r;
            if (n < N) {
              if (os) os[base + n * inner] = y[0][i][r];
              if (od) od[base + n * inner] = y[1][i][r];
            }
          }
        }
      }
    }
    __syncthreads();  // the buffers are restaged for the next stage
  }
}

// ------------------------------------------------ contiguous chain pass --
// The first Hessian pass when it runs along a contiguous axis (2-D arrays
// whose axis-0 lines are too long for the strided pass).  Lines are rows.
template <typename W>
__global__ __launch_bounds__(256) void k_chain_contig(const ChainArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int E = a.E, Nr = a.Nr, R = a.rows;
  W* xin = reinterpret_cast<W*>(smem_raw);
  W* xs = xin + R * E;
  W* xd = xs + R * E;
  W* wS = xd + R * E;
  W* wD = wS + a.S.K + 4;
  const ChainJob& jb = a.job[blockIdx.y];
  const int anc = a.S.anchor;
  const int64_t row0 = (int64_t)blockIdx.x * R;
  const int tid = threadIdx.x;
  stage_taps<W>(wS, a.S, tid, 256);
  stage_taps<W>(wD, a.D, tid, 256);
  bool need_s = false, need_d = false;
  for (int k = 0; k < jb.nout; ++k) {
    need_s |= jb.o[k].first == 0;
    need_d |= jb.o[k].first == 1;
  }
  bool bad;
  {
    const int64_t N = a.N, nlines = a.nlines;
    const int mode = a.mode;
    bad = stage_any<W>(xin, R * E, tid, 256, jb.in, jb.in_dt, a.cval, a.negate != 0,
                       [=](int e, int& slot) -> int64_t {
                         const int rr = e / E, q = e - rr * E;
                         slot = e;
                         const int64_t row = row0 + rr;
                         if (row >= nlines) return -1;
                         const int64_t r = remap((int64_t)q - anc, N, mode);
                         return r < 0 ? -1 : row * N + r;
                       },
                       a.center);
  }
  const bool skip0 = __syncthreads_or(bad);
  const int nq = Nr / 4;
  bool bad1 = false;
  with_skip(skip0, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
    for (int it = tid; it < R * nq; it += 256) {
      const int rr = it / nq, o = (it - rr * nq) * 4;
      W as[4] = {(W)0, (W)0, (W)0, (W)0}, ad[4] = {(W)0, (W)0, (W)0, (W)0};
      const W* x = xin + rr * E + o;
      if (need_s && need_d) conv_row4x2<SK>(as, ad, x, wS, a.S, wD, a.D);
      else if (need_s) conv_row4<SK>(as, x, wS, a.S);
      else conv_row4<SK>(ad, x, wD, a.D);
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        xs[rr * E + anc + o + j] = as[j];
        xd[rr * E + anc + o + j] = ad[j];
        bad1 |= !(finite(as[j]) && finite(ad[j]));
      }
    }
  });
  const bool skip1 = __syncthreads_or(bad1) || skip0;
  const int nh = E - (int)a.N;
  for (int e = tid; e < R * nh; e += 256) {
    const int rr = e / nh, h = e - rr * nh;
    const int q = h < anc ? h : h + (int)a.N;
    const int64_t r = remap((int64_t)q - anc, a.N, a.mode);
    W vs = a.cval, vd = a.cval;
    if (r >= 0) {
      vs = xs[rr * E + r + anc];
      vd = xd[rr * E + r + anc];
    }
    xs[rr * E + q] = vs;
    xd[rr * E + q] = vd;
  }
  __syncthreads();
  with_skip(skip1, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
    for (int it = tid; it < R * nq; it += 256) {
      const int rr = it / nq, o = (it - rr * nq) * 4;
      const int64_t row = row0 + rr;
      if (row >= a.nlines) continue;
#pragma unroll 1
      for (int src = 0; src < 2; ++src) {
        int ks = -1, kd = -1;
        for (int k = 0; k < jb.nout; ++k) {
          if (jb.o[k].first != src) continue;
          if (jb.o[k].second == 0) ks = k; else kd = k;
        }
        if (ks < 0 && kd < 0) continue;
        const W* x = (src ? xd : xs) + rr * E + o;
        W r0[4] = {(W)0, (W)0, (W)0, (W)0}, r1[4] = {(W)0, (W)0, (W)0, (W)0};
        if (ks >= 0 && kd >= 0) conv_row4x2<SK>(r0, r1, x, wS, a.S, wD, a.D);
        else if (ks >= 0) conv_row4<SK>(r0, x, wS, a.S);
        else conv_row4<SK>(r1, x, wD, a.D);
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          if (o + j < a.N) {
            if (ks >= 0) reinterpret_cast<W*>(jb.o[ks].out)[row * a.N + o + j] = r0[j];
            if (kd >= 0) reinterpret_cast<W*>(jb.o[kd].out)[row * a.N + o + j] = r1[j];
          }
        }
      }
    }
  });
}

// ----------------------------------------------------------- final pass --
enum Epilogue { EPI_STORE = 0, EPI_FRANGI = 1 };

template <typename W>
struct FinalArgs {
  const W* in[6];
  int first[6], second[6];
  int nin;          // 3 (2-D) or 6 (3-D)
  int64_t N, nrows;
  int E, Nr, rows;
  int mode;
  W cval;
  TapPack<W> S, D;
  int epi;
  W* out[6];        // EPI_STORE targets (may alias in[])
  // EPI_FRANGI
  W* fout;          // vesselness running maximum
  W* eig[3];        // eigenvalue store (first sigma, gamma unknown)
  int store_eig;    // 1: write eig + reduce max r_g; 0: vesselness now
  int first_sigma;  // 1: fout = v, else fout = max(fout, v)
  double alpha_sq, beta_sq, gamma_sq;
  const double* gamma_sq_dev;  // when non-null, overrides gamma_sq
  void* rg_max;     // uint32 (float) / uint64 (double) bit pattern
  int t0, Km, shift;  // matrix-core variant: nonzero tap span of S and D,
                      // buffer row of line position 0
  int vec4;           // matrix-core variant: rows staged by 16-byte loads
};

template <typename W>
__device__ __forceinline__ void atomic_max_bits(void* p, W v);
template <>
__device__ __forceinline__ void atomic_max_bits<float>(void* p, float v) {
  unsigned int b = __float_as_uint(v);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned int x = __shfl_xor(b, o, 64);
    b = x > b ? x : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(reinterpret_cast<unsigned int*>(p), b);
}
template <>
__device__ __forceinline__ void atomic_max_bits<double>(void* p, double v) {
  unsigned long long b = (unsigned long long)__double_as_longlong(v);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const unsigned long long x = __shfl_xor(b, o, 64);
    b = x > b ? x : b;
  }
  if ((threadIdx.x & 63) == 0) atomicMax(reinterpret_cast<unsigned long long*>(p), b);
}

// One input of the final pass: stage its rows, first op, re-extend, second
// op into res (IPT items of 4 outputs per thread).
template <typename W, int IPT, int P, int J>
__device__ __forceinline__ void final_input(const FinalArgs<W>& a, W* xin, W* xf, const W* wS,
                                            const W* wD, int64_t row0, W (&res)[IPT][P]) {
  const int E = a.E, R = a.rows, anc = a.S.anchor, tid = threadIdx.x;
  const int nq = a.Nr / P, nitems = R * nq;
  bool bad;
  {
    const int64_t N = a.N, nrows = a.nrows;
    const int mode = a.mode;
    bad = stage_lds<W, true>(xin, R * E, tid, 256, a.in[J], w_dtype<W>(), a.cval, false,
                             [=](int e, int& slot) -> int64_t {
                               const int rr = e / E, q = e - rr * E;
                               slot = e;
                               const int64_t row = row0 + rr;
                               if (row >= nrows) return -1;
                               const int64_t r = remap((int64_t)q - anc, N, mode);
                               return r < 0 ? -1 : row * N + r;
                             });
  }
  const bool skip0 = __syncthreads_or(bad);
  const TapPack<W>& t1 = a.first[J] ? a.D : a.S;
  const W* w1 = a.first[J] ? wD : wS;
  bool bad1 = false;
  with_skip(skip0, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + 256 * i;
      if (it < nitems) {
        const int rr = it / nq, o = (it - rr * nq) * P;
        W acc[P];
#pragma unroll
        for (int m = 0; m < P; ++m) acc[m] = (W)0;
        conv_rowP<SK, P>(acc, xin + rr * E + o, w1, t1);
#pragma unroll
        for (int m = 0; m < P; ++m) {
          xf[rr * E + anc + o + m] = acc[m];
          bad1 |= !finite(acc[m]);
        }
      }
    }
  });
  const bool skip1 = __syncthreads_or(bad1) || skip0;
  const int nh = E - (int)a.N;
  for (int e = tid; e < R * nh; e += 256) {
    const int rr = e / nh, h = e - rr * nh;
    const int q = h < anc ? h : h + (int)a.N;
    const int64_t r = remap((int64_t)q - anc, a.N, a.mode);
    xf[rr * E + q] = r >= 0 ? xf[rr * E + r + anc] : a.cval;
  }
  __syncthreads();
  const TapPack<W>& t2 = a.second[J] ? a.D : a.S;
  const W* w2 = a.second[J] ? wD : wS;
  with_skip(skip1, [&](auto sk) {
    constexpr bool SK = decltype(sk)::value;
#pragma unroll
    for (int i = 0; i < IPT; ++i) {
      const int it = tid + 256 * i;
#pragma unroll
      for (int m = 0; m < P; ++m) res[i][m] = (W)0;
      if (it < nitems) {
        const int rr = it / nq, o = (it - rr * nq) * P;
        conv_rowP<SK, P>(res[i], xf + rr * E + o, w2, t2);
      }
    }
  });
  __syncthreads();
}

template <typename W, int NIN, int IPT, int P, size_t... Js>
__device__ __forceinline__ void final_all(const FinalArgs<W>& a, W* xin, W* xf, const W* wS,
                                          const W* wD, int64_t row0, W (&res)[NIN][IPT][P],
                                          std::index_sequence<Js...>) {
  (final_input<W, IPT, P, (int)Js>(a, xin, xf, wS, wD, row0, res[Js]), ...);
}

template <typename W, int NIN, int IPT, int P>
__global__ __launch_bounds__(256) void k_final_contig(const FinalArgs<W> a) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  const int E = a.E, R = a.rows;
  W* xin = reinterpret_cast<W*>(smem_raw);
  W* xf = xin + R * E;
  W* wS = xf + R * E;
  W* wD = wS + a.S.K + 4;
  const int64_t row0 = xcd_swizzle(blockIdx.x, gridDim.x) * R;
  const int tid = threadIdx.x;
  const int nq = a.Nr / P;
  const int nitems = R * nq;
  stage_taps<W>(wS, a.S, tid, 256);
  stage_taps<W>(wD, a.D, tid, 256);
  W res[NIN][IPT][P];
  final_all<W, NIN, IPT, P>(a, xin, xf, wS, wD, row0, res, std::make_index_sequence<NIN>{});
  // epilogue: every item holds all Hessian elements of its P voxels
  double gsq = a.gamma_sq;
  if (a.epi == EPI_FRANGI && a.gamma_sq_dev) gsq = *a.gamma_sq_dev;
  const FrangiNegRecip nrc = frangi_neg_recip(a.alpha_sq, a.beta_sq, gsq);
  W rgmax = (W)0;
#pragma unroll
  for (int i = 0; i < IPT; ++i) {
    const int it = tid + 256 * i;
    if (it >= nitems) continue;
    const int rr = it / nq, o = (it - rr * nq) * P;
    const int64_t row = row0 + rr;
    if (row >= a.nrows) continue;
#pragma unroll
    for (int m = 0; m < P; ++m) {
      if (o + m >= a.N) continue;
      const int64_t idx = row * a.N + o + m;
      if (a.epi == EPI_STORE) {
#pragma unroll
        for (int j = 0; j < NIN; ++j) a.out[j][idx] = res[j][i][m];
        continue;
      }
      if (NIN == 6) {
        W l1, l2, l3;
        // H order (H00, H01, H02, H11, H12, H22) = (a, d, f, b, e, c)
        eig33<W>(res[0][i][m], res[3 % NIN][i][m], res[5 % NIN][i][m], res[1][i][m],
                 res[4 % NIN][i][m], res[2][i][m], 0, 1, l1, l2, l3);
        const W rg = frangi_rg3<W>(l1, l2, l3);
        if (a.store_eig) {
          a.eig[0][idx] = l1;
          a.eig[1][idx] = l2;
          a.eig[2][idx] = l3;
          rgmax = nanmax<W>(rgmax, rg);
        } else {
          const W v = frangi_v3r<W>(l1, l2, l3, rg, nrc);
          a.fout[idx] = a.first_sigma ? v : nanmax<W>(a.fout[idx], v);
        }
      } else {
        W l1, l2;
        eig22<W>(res[0][i][m], res[1][i][m], res[2][i][m], 0, 1, l1, l2);
        const W rg = frangi_rg2<W>(l1, l2);
        if (a.store_eig) {
          a.eig[0][idx] = l1;
          a.eig[1][idx] = l2;
          rgmax = nanmax<W>(rgmax, rg);
        } else {
          const W v = frangi_v2r<W>(l1, l2, rg, nrc);
          a.fout[idx] = a.first_sigma ? v : nanmax<W>(a.fout[idx], v);
        }
      }
    }
  }
  if (a.epi == EPI_FRANGI && a.store_eig) atomic_max_bits<W>(a.rg_max, rgmax);
}

// The final pass on the matrix cores (float32): tiles of 16 rows, each row
// staged whole and transposed (position-major, row-minor) so that the
// Toeplitz passes of the strided chain apply unchanged; every lane ends up
// with all Hessian elements of 4 NBW voxels of one row, and the epilogue is
// the one of k_final_contig.  The tile's next input is in flight in
// registers while the current one is computed.
struct FinalGeom {
  int E, Km, shift, nd, nblk, LS;
  int64_t N, nrows, ntile, G;
};

// Stage input J's rows of the tile into the line-major buffer X (row r of
// line l at X[l * LS + r], interior from row shift): thread t takes line
// l = t % 16 and the 4-value chunks c = t / 16 + (NT / 16) k, as 16-byte
// loads and two 8-byte stores when vec (rows of a multiple of 4 values,
// 16-byte aligned), else value by value.  Rows past the last are staged as
// zeros.  Every load is issued before any is used (the index of a masked
// one is clamped into the row and its value replaced afterwards).  Returns
// whether a staged value is not finite.
template <int NBW, int NW>
__device__ __forceinline__ bool final_stage(const FinalArgs<float>& a, const FinalGeom& g,
                                            float* X, int LS, int J, int64_t tile, int tid) {
  typedef float F4 __attribute__((ext_vector_type(4)));
  typedef float F2 __attribute__((ext_vector_type(2)));
  constexpr int CS = 4 * NW;  // chunk slots per pass over the work-group
  const int l = tid & 15, c0 = tid >> 4, n = (int)g.N;
  const int64_t row = tile * kChC + l;
  const bool rok = row < g.nrows;
  const float* rp = a.in[J] + (rok ? row : 0) * g.N;
  float* xl = X + l * LS + g.shift;
  bool bad = false;
  if (a.vec4) {
    // N <= 16 NW NBW, so a line has at most NBW CS chunks
    F4 v[NBW];
#pragma unroll
    for (int k = 0; k < NBW; ++k) {
      const int c = c0 + CS * k;
      v[k] = reinterpret_cast<const F4*>(rp)[4 * c < n ? c : 0];
    }
#pragma unroll
    for (int k = 0; k < NBW; ++k) {
      const int c = c0 + CS * k;
      if (4 * c >= n) continue;
      const F4 w = rok ? v[k] : F4{0.f, 0.f, 0.f, 0.f};
      bad |= !(finite(w[0]) && finite(w[1]) && finite(w[2]) && finite(w[3]));
      *reinterpret_cast<F2*>(xl + 4 * c) = F2{w[0], w[1]};
      *reinterpret_cast<F2*>(xl + 4 * c + 2) = F2{w[2], w[3]};
    }
  } else {
    float v[4 * NBW];
#pragma unroll
    for (int k = 0; k < 4 * NBW; ++k) {
      const int q = c0 + CS * k;
      v[k] = rp[q < n ? q : 0];
    }
#pragma unroll
    for (int k = 0; k < 4 * NBW; ++k) {
      const int q = c0 + CS * k;
      if (q >= n) continue;
      const float w = rok ? v[k] : 0.f;
      bad |= !finite(w);
      xl[q] = w;
    }
  }
  return bad;
}

// Input J of the tile: stage it, first op, re-extend, second op into res.
template <int NIN, int NBW, int NW, int J>
__device__ __forceinline__ void final_input_mfma(const FinalArgs<float>& a, const FinalGeom& g,
                                                 float* xin, float* xf, const float* wpS,
                                                 const float* wpD, const short* hsrc,
                                                 int64_t tile, f32x4_t (&res)[NBW]) {
  typedef float F2 __attribute__((ext_vector_type(2)));
  const int E = g.E, shift = g.shift, LS = g.LS;
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int kk = lane >> 4, ii = lane & 15;
  constexpr int NT = 64 * NW;
  const bool bad = final_stage<NBW, NW>(a, g, xin, LS, J, tile, tid);
  const bool skip0 = __syncthreads_or(bad) || !finite(a.cval);
  fill_halo(xin, nullptr, E, (int)g.N, shift, hsrc, a.cval, tid, NT, 1, LS);
  __syncthreads();
  f32x4_t f[1][NBW];
  {
    const float* const wp[1] = {a.first[J] ? wpD : wpS};
    chain_pass<NBW, 1, NW>(skip0, f, xin, wp, g.Km, g.nd, g.nblk, wv, lane, 1, LS);
  }
  bool bad1 = false;
#pragma unroll
  for (int i = 0; i < NBW; ++i) {
    if (wv + NW * i >= g.nblk) continue;
    // positions 16 b + 4 kk + r (r < 4) of line ii: two 8-byte words
    float* d = xf + ii * LS + shift + 16 * (wv + NW * i) + 4 * kk;
    *reinterpret_cast<F2*>(d) = F2{f[0][i][0], f[0][i][1]};
    *reinterpret_cast<F2*>(d + 2) = F2{f[0][i][2], f[0][i][3]};
#pragma unroll
    for (int r = 0; r < 4; ++r) bad1 |= !finite(f[0][i][r]);
  }
  const bool skip1 = __syncthreads_or(bad1) || skip0;
  fill_halo(xf, nullptr, E, (int)g.N, shift, hsrc, a.cval, tid, NT, 1, LS);
  __syncthreads();
  {
    const float* const wp[1] = {a.second[J] ? wpD : wpS};
    chain_pass<NBW, 1, NW>(skip1, f, xf, wp, g.Km, g.nd, g.nblk, wv, lane, 1, LS);
  }
#pragma unroll
  for (int i = 0; i < NBW; ++i) res[i] = f[0][i];
  // xin is next written after the barrier at the start of the next input,
  // which every wave reaches only after its first pass here; xf is next
  // written after the first pass of the next input, behind that barrier too
}

template <int NIN, int NBW, int NW, size_t... Js>
__device__ __forceinline__ void final_all_mfma(const FinalArgs<float>& a, const FinalGeom& g,
                                               float* xin, float* xf, const float* wpS,
                                               const float* wpD, const short* hsrc,
                                               int64_t tile, f32x4_t (&res)[NIN][NBW],
                                               std::index_sequence<Js...>) {
  (final_input_mfma<NIN, NBW, NW, (int)Js>(a, g, xin, xf, wpS, wpD, hsrc, tile, res[Js]), ...);
}

// EP: the epilogue, fixed at compile time so that each variant keeps only
// its own arguments in scalar registers: 0 store the Hessian (EPI_STORE),
// 1 Frangi storing eigenvalues and reducing max r_g (first sigma, gamma
// unknown), 2 Frangi vesselness into the running maximum.
template <int NIN, int NBW, int NW, int EP>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(final_wpe(NBW)))) void
k_final_mfma(const FinalArgs<float> a) {
  constexpr int NT = 64 * NW;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  FinalGeom g;
  g.E = a.E;
  g.Km = a.Km;
  g.shift = a.shift;
  g.nd = (a.Km + 30) / 16;
  g.nblk = a.Nr / 16;
  g.LS = lm_stride(a.E);
  g.N = a.N;
  g.nrows = a.nrows;
  g.ntile = (a.nrows + kChC - 1) / kChC;
  g.G = gridDim.x;
  const int WP = mfma_wp(g.Km);
  // two line-major buffers of 16 lines; the interior of a line (row shift)
  // starts at an even word, so that its 8-byte words are aligned
  float* smem_f = reinterpret_cast<float*>(smem_raw);
  float* xin = smem_f + (g.shift & 1);
  float* xf = xin + kChC * g.LS;
  float* wpS = smem_f + 2 + 2 * kChC * g.LS;
  float* wpD = wpS + WP;
  short* hsrc = reinterpret_cast<short*>(wpD + WP);
  const int tid = threadIdx.x, lane = tid & 63, wv = wave_id();
  const int kk = lane >> 4, ii = lane & 15;
  stage_trimmed(wpS, wpD, a.S, a.D, a.t0, a.Km, WP, tid, NT);
  halo_table(hsrc, g.E, g.N, g.shift, a.mode, tid, NT);
  double gsq = a.gamma_sq;
  if (EP != 0 && a.gamma_sq_dev) gsq = *a.gamma_sq_dev;
  const FrangiNegRecip nrc = frangi_neg_recip(a.alpha_sq, a.beta_sq, gsq);
  const int64_t N = g.N;
  float rgmax = 0.f;
  const int64_t tile = xcd_swizzle(blockIdx.x, g.G);
  {
    f32x4_t res[NIN][NBW];
    final_all_mfma<NIN, NBW, NW>(a, g, xin, xf, wpS, wpD, hsrc, tile, res,
                                 std::make_index_sequence<NIN>{});
    // epilogue (k_final_contig's)
    const int64_t row = tile * kChC + ii;
    if (row < g.nrows) {
#pragma unroll
      for (int v = 0; v < 4 * NBW; ++v) {
        const int i = v >> 2;
        if (wv + NW * i >= g.nblk) break;
        const int64_t p = 16 * (wv + NW * i) + 4 * kk + (v & 3);
        if (p >= N) continue;
        float h[NIN];
#pragma unroll
        for (int j = 0; j < NIN; ++j) h[j] = res[j][i][v & 3];
        const int64_t idx = row * N + p;
        if (EP == 0) {
#pragma unroll
          for (int j = 0; j < NIN; ++j) a.out[j][idx] = h[j];
          continue;
        }
        if (NIN == 6) {
          float l1, l2, l3;
          // H order (H00, H01, H02, H11, H12, H22) = (a, d, f, b, e, c)
          eig33<float>(h[0], h[3 % NIN], h[5 % NIN], h[1], h[4 % NIN], h[2], 0, 1, l1, l2, l3);
          const float rg = frangi_rg3<float>(l1, l2, l3);
          if (EP == 1) {
            a.eig[0][idx] = l1;
            a.eig[1][idx] = l2;
            a.eig[2][idx] = l3;
            rgmax = nanmax<float>(rgmax, rg);
          } else {
            const float vv = frangi_v3r<float>(l1, l2, l3, rg, nrc);
            a.fout[idx] = a.first_sigma ? vv : nanmax<float>(a.fout[idx], vv);
          }
        } else {
          float l1, l2;
          eig22<float>(h[0], h[1], h[2], 0, 1, l1, l2);
          const float rg = frangi_rg2<float>(l1, l2);
          if (EP == 1) {
            a.eig[0][idx] = l1;
            a.eig[1][idx] = l2;
            rgmax = nanmax<float>(rgmax, rg);
          } else {
            const float vv = frangi_v2r<float>(l1, l2, rg, nrc);
            a.fout[idx] = a.first_sigma ? vv : nanmax<float>(a.fout[idx], vv);
          }
        }
      }
    }
  }
  if (EP == 1) atomic_max_bits<float>(a.rg_max, rgmax);
}

// ---------------------------------------------------- small kernels --
__global__ void k_frangi_reset(unsigned long long* rg) {
  if (threadIdx.x == 0 && blockIdx.x == 0) rg[0] = 0ull;
}

template <typename W>
__global__ void k_frangi_gamma(const unsigned long long* rg, double* gamma_sq) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double s;
  if (sizeof(W) == 4) s = (double)__uint_as_float((unsigned int)(rg[0] & 0xffffffffull));
  else s = __longlong_as_double((long long)rg[0]);
  gamma_sq[0] = gamma_sq_from_max(s);
}

template <typename W, int NDIM>
__global__ __launch_bounds__(256) void k_frangi_from_eig(const W* __restrict__ e0,
                                                         const W* __restrict__ e1,
                                                         const W* __restrict__ e2,
                                                         W* __restrict__ out, int64_t n,
                                                         double alpha_sq, double beta_sq,
                                                         const double* gamma_sq_dev) {
  const double gsq = *gamma_sq_dev;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    if (NDIM == 3) {
      const W l1 = e0[i], l2 = e1[i], l3 = e2[i];
      out[i] = frangi_v3<W>(l1, l2, l3, frangi_rg3<W>(l1, l2, l3), alpha_sq, beta_sq, gsq);
    } else {
      const W l1 = e0[i], l2 = e1[i];
      out[i] = frangi_v2<W>(l1, l2, frangi_rg2<W>(l1, l2), beta_sq, gsq);
    }
  }
}

// vesselness from stored eigenvalues, with the running maximum
template <typename W, int NDIM>
__global__ __launch_bounds__(256) void k_frangi_eig_max(const W* __restrict__ e0,
                                                        const W* __restrict__ e1,
                                                        const W* __restrict__ e2,
                                                        W* __restrict__ out, int64_t n,
                                                        double alpha_sq, double beta_sq,
                                                        double gamma_sq,
                                                        const double* gamma_sq_dev, int first) {
  const double gsq = gamma_sq_dev ? *gamma_sq_dev : gamma_sq;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    W v;
    if (NDIM == 3) {
      const W l1 = e0[i], l2 = e1[i], l3 = e2[i];
      v = frangi_v3<W>(l1, l2, l3, frangi_rg3<W>(l1, l2, l3), alpha_sq, beta_sq, gsq);
    } else {
      const W l1 = e0[i], l2 = e1[i];
      v = frangi_v2<W>(l1, l2, frangi_rg2<W>(l1, l2), beta_sq, gsq);
    }
    out[i] = first ? v : nanmax<W>(out[i], v);
  }
}

// max over voxels of r_g = sum lambda^2 (bit-pattern max, NaN wins)
template <typename W, int NDIM>
__global__ __launch_bounds__(256) void k_rg_max(const W* __restrict__ e0, const W* __restrict__ e1,
                                                const W* __restrict__ e2, int64_t n, void* rg) {
  W m = (W)0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    const W r = NDIM == 3 ? frangi_rg3<W>(e0[i], e1[i], e2[i]) : frangi_rg2<W>(e0[i], e1[i]);
    m = nanmax<W>(m, r);
  }
  atomic_max_bits<W>(rg, m);
}

template <typename W>
__global__ __launch_bounds__(256) void k_sym_eig(const W* __restrict__ s0, const W* __restrict__ s1,
                                                 const W* __restrict__ s2, const W* __restrict__ s3,
                                                 const W* __restrict__ s4, const W* __restrict__ s5,
                                                 W* __restrict__ e0, W* __restrict__ e1,
                                                 W* __restrict__ e2, int64_t n, int nS,
                                                 int descending, int abs_sort) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    if (nS == 6) {
      W l1, l2, l3;
      // (S0..S5) = (M00, M01, M02, M11, M12, M22) -> (a, d, f, b, e, c)
      eig33<W>(s0[i], s3[i], s5[i], s1[i], s4[i], s2[i], descending, abs_sort, l1, l2, l3);
      e0[i] = l1;
      e1[i] = l2;
      e2[i] = l3;
    } else {
      W l1, l2;
      eig22<W>(s0[i], s1[i], s2[i], descending, abs_sort, l1, l2);
      e0[i] = l1;
      e1[i] = l2;
    }
  }
}

// ------------------------------------------------------------ planning --
namespace {

struct Pair {
  int a, b;  // first derivative axis, second derivative axis
};

// outputs in combinations_with_replacement order of the (possibly reversed)
// axes (corner.py:246-253)
int hessian_pairs(int ndim, int order_xy, Pair* p) {
  int ax[3];
  for (int d = 0; d < ndim; ++d) ax[d] = order_xy ? ndim - 1 - d : d;
  int n = 0;
  for (int i = 0; i < ndim; ++i)
    for (int j = i; j < ndim; ++j) p[n++] = Pair{ax[i], ax[j]};
  return n;
}

inline int chain_first(const Pair& p, int x) { return p.a == x ? 1 : 0; }
inline int chain_second(const Pair& p, int x) { return p.b == x ? 1 : 0; }

template <typename W>
int lds_strided(int E) { return 2 * E * kChS * (int)sizeof(W); }

template <typename W>
int pick_rows(int E, int Nr, int nbuf, int max_items) {
  // lines per work-group: as many as fit max_items 4-output items and ~64 KiB
  const int nq = Nr / 4;
  int r = std::max(1, max_items / nq);
  while (r > 1 && nbuf * r * E * (int)sizeof(W) > 64 * 1024) --r;
  return r;
}

template <typename W>
struct AxisTaps {
  TapPack<W> S, D;
};

template <typename W>
int axis_taps(const hc_taps* ts, const hc_taps* td, AxisTaps<W>* out) {
  if (ts->K != td->K || ts->anchor != td->anchor)
    return fail(HC_ERR_VALUE, "order-0 and order-1 kernels must have equal length and anchor");
  int rc = make_taps<W>(ts, 4, &out->S);
  if (rc != HC_OK) return rc;
  rc = make_taps<W>(td, 4, &out->D);
  if (rc != HC_OK) return rc;
  if (out->S.K > 128) return fail(HC_ERR_RESOURCE, "kernel too long for the fused Hessian");
  return HC_OK;
}

// Matrix-core passes: on unless the HC_TUNE_HESSIAN_VALU knob is set.
inline bool chain_mfma_enabled() { return tune_get(HC_TUNE_HESSIAN_VALU) == 0; }

// Geometry of the matrix-core passes over lines of N: the nonzero span of
// both tap sets (zero taps at the ends add exact zeros for finite inputs;
// non-finite tiles take the tap-by-tap path), the buffer shift and length,
// the LDS bytes.  False when they do not apply.
struct MfmaGeomHost {
  int t0, Km, shift, Nr, E, nd, lds;
};

inline bool mfma_geom(const TapPack<float>& S, const TapPack<float>& D, int64_t N, int nbuf,
                      MfmaGeomHost& g, bool line_major = false) {
  if (!chain_mfma_enabled() || N < 1 || N > 512) return false;
  int lo = S.K, hi = -1;
  for (int t = 0; t < S.K; ++t)
    if (S.w[t] != 0.f || D.w[t] != 0.f) {
      lo = std::min(lo, t);
      hi = t;
    }
  if (hi < 0) return false;
  g.t0 = lo;
  g.Km = hi - lo + 1;
  g.shift = S.anchor - lo;
  if (g.shift < 0 || g.shift > g.Km + 14) return false;
  g.Nr = (int)((N + 15) / 16 * 16);
  // rows the passes touch: the matrix-core steps read up to row
  // Nr - 17 + 4 ceil((Km + 15) / 4), the tap-by-tap path up to Nr + Km - 2,
  // the first-pass results go to rows shift .. shift + Nr - 1
  const int nsteps = (g.Km + 18) / 4;
  g.E = std::max(std::max(g.Nr + 4 * nsteps - 16, g.Nr + g.Km - 1), g.shift + g.Nr);
  g.nd = (g.Km + 30) / 16;
  // nbuf line buffers (line-major: 16 lines of lm_stride(E) and two words
  // of alignment room), the padded taps, the halo source rows (16-bit, one
  // per row outside [shift, shift + N))
  const int buf = line_major ? 2 + nbuf * kChC * lm_stride(g.E) : nbuf * g.E * kMfS;
  g.lds = (buf + 2 * mfma_wp(g.Km)) * (int)sizeof(float) + ((int)(g.E - N) + 1) / 2 * 4;
  return g.lds <= kMaxLds;
}

// Waves per work-group of the matrix-core passes: 4 for lines of at most
// 256, 8 up to 512 (a wave holds at most 4 blocks of 16 outputs); the
// HC_TUNE_MFMA_WAVES knob set to 8 spreads lines of 129..256 over 8 waves
// of 2 blocks each (fewer accumulators per wave, more waves per CU for the
// same LDS).
inline int mfma_waves(int64_t N) {
  if (N > 64 * 4) return 8;
  return tune_get(HC_TUNE_MFMA_WAVES) == 8 && N > 128 ? 8 : 4;
}

template <int NBW, int NW>
int launch_chain_mfma_nbw(ChainArgs<float>& a, const MfmaGeomHost& g, hipStream_t st) {
  const void* f = (const void*)k_chain_mfma<NBW, NW>;
  if (g.lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, g.lds);
    if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(k_chain_mfma)");
  }
  const int64_t work = (a.nlines + kChC - 1) / kChC * a.njobs;
  const int64_t G = (work + kChainSPW - 1) / kChainSPW;
  HC_LAUNCH("k_chain_mfma", st, (k_chain_mfma<NBW, NW>), dim3((unsigned)G), dim3(64 * NW), g.lds,
            a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

// Matrix-core strided chain pass (float32, lines of at most 512); returns -1
// when it does not apply.
inline int launch_chain_mfma(ChainArgs<float>& a, hipStream_t st) {
  const int nblk = (int)((a.N + 15) / 16);
  MfmaGeomHost g;
  if (!mfma_geom(a.S, a.D, a.N, 1, g)) return -1;
  a.t0 = g.t0;
  a.Km = g.Km;
  a.shift = g.shift;
  a.Nr = g.Nr;
  a.E = g.E;
  if (mfma_waves(a.N) == 8)
    return nblk <= 16 ? launch_chain_mfma_nbw<2, 8>(a, g, st) : launch_chain_mfma_nbw<4, 8>(a, g, st);
  if (nblk <= 4) return launch_chain_mfma_nbw<1, 4>(a, g, st);
  if (nblk <= 6) return launch_chain_mfma_nbw<3, 2>(a, g, st);
  if (nblk <= 8) return launch_chain_mfma_nbw<2, 4>(a, g, st);
  if (nblk <= 12) return launch_chain_mfma_nbw<3, 4>(a, g, st);
  return launch_chain_mfma_nbw<4, 4>(a, g, st);
}

template <typename W>
int run_chain_pass(const ChainJob* jobs, int njobs, const hc_array* shape_src, int axis,
                   const AxisTaps<W>& tp, int mode, double cval, int negate, hipStream_t st,
                   const double* center = nullptr) {
  ChainArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  for (int j = 0; j < njobs; ++j) a.job[j] = jobs[j];
  const Lines ln = lines_of(shape_src, axis);
  a.N = ln.N;
  a.inner = ln.inner;
  a.mode = mode;
  a.negate = negate;
  a.center = center;
  a.cval = (W)cval;
  a.S = tp.S;
  a.D = tp.D;
  if (ln.inner > 1 && std::is_same<W, float>::value) {
    ChainArgs<float>& af = *reinterpret_cast<ChainArgs<float>*>(&a);
    af.nlines = ln.outer * ln.inner;
    af.njobs = njobs;
    const int rc = launch_chain_mfma(af, st);
    if (rc >= 0) return rc;
  }
  if (ln.inner > 1) {
    a.nlines = ln.outer * ln.inner;
    a.Nr = (int)((ln.N + kChP - 1) / kChP * kChP);
    a.E = a.Nr + a.S.K + 4;  // the column core prefetches one group ahead
    const int nitems = (a.Nr / kChP) * kChC;
    if (nitems > 256 * kChMaxItems)
      return fail(HC_ERR_RESOURCE, "line too long for the fused Hessian");
    const int lds = lds_strided<W>(a.E) + 2 * (a.S.K + 4) * (int)sizeof(W);
    if (lds > kMaxLds) return fail(HC_ERR_RESOURCE, "line too long for the fused Hessian");
    const int64_t blocks = (a.nlines + kChC - 1) / kChC;
    if (nitems <= 512)
      HC_LAUNCH("k_chain_strided", st, (k_chain_strided<W, 2>), dim3((unsigned)blocks, njobs),
                dim3(256), lds, a);
    else
      HC_LAUNCH("k_chain_strided", st, (k_chain_strided<W, kChMaxItems>),
                dim3((unsigned)blocks, njobs), dim3(256), lds, a);
  } else {
    a.nlines = ln.outer;
    a.Nr = (int)((ln.N + 3) / 4 * 4);
    a.E = a.Nr + a.S.K + 4;
    a.rows = pick_rows<W>(a.E, a.Nr, 3, 1024);
    const int lds = (3 * a.rows * a.E + 2 * (a.S.K + 4)) * (int)sizeof(W);
    if (lds > kMaxLds) return fail(HC_ERR_RESOURCE, "line too long for the fused Hessian");
    const int64_t blocks = (a.nlines + a.rows - 1) / a.rows;
    HC_LAUNCH("k_chain_contig", st, k_chain_contig<W>, dim3((unsigned)blocks, njobs), dim3(256), lds, a);
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

template <typename W>
int prepare_final(FinalArgs<W>* a, const hc_array* shape_src, const AxisTaps<W>& tp, int mode,
                  double cval) {
  const int axis = shape_src->ndim - 1;
  const Lines ln = lines_of(shape_src, axis);
  a->N = ln.N;
  a->nrows = ln.outer;
  a->Nr = (int)((ln.N + kFinP - 1) / kFinP * kFinP);
  a->E = a->Nr + tp.S.K + 4;  // the row core reads K + 11 past an item's start
  if (a->Nr / kFinP > kFinIPT * 256)
    return fail(HC_ERR_RESOURCE, "last axis too long for the fused Hessian epilogue");
  a->rows = pick_rows<W>(a->E, a->Nr, 2, kFinIPT * 256 * kFinP / 4);
  if (2 * a->rows * a->E * (int)sizeof(W) > kMaxLds)
    return fail(HC_ERR_RESOURCE, "last axis too long for the fused Hessian epilogue");
  a->mode = mode;
  a->cval = (W)cval;
  a->S = tp.S;
  a->D = tp.D;
  return HC_OK;
}

template <int NIN, int NBW, int NW, int EP>
int launch_final_mfma_ep(const FinalArgs<float>& a, const MfmaGeomHost& g, hipStream_t st) {
  const void* f = (const void*)k_final_mfma<NIN, NBW, NW, EP>;
  if (g.lds > 64 * 1024) {
    const hipError_t e = hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, g.lds);
    if (e != hipSuccess) return hip_check(e, "hipFuncSetAttribute(k_final_mfma)");
  }
  const int64_t G = (a.nrows + kChC - 1) / kChC;
  HC_LAUNCH("k_final_mfma", st, (k_final_mfma<NIN, NBW, NW, EP>), dim3((unsigned)G),
            dim3(64 * NW), g.lds, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

template <int NIN, int NBW, int NW>
int launch_final_mfma_nbw(const FinalArgs<float>& a, const MfmaGeomHost& g, hipStream_t st) {
  if (a.epi == EPI_STORE) return launch_final_mfma_ep<NIN, NBW, NW, 0>(a, g, st);
  if (a.store_eig) return launch_final_mfma_ep<NIN, NBW, NW, 1>(a, g, st);
  return launch_final_mfma_ep<NIN, NBW, NW, 2>(a, g, st);
}

// Matrix-core final pass for float32 rows of at most 512 voxels; returns -1
// when it does not apply.
template <int NIN>
int launch_final_mfma(const FinalArgs<float>& a0, hipStream_t st) {
  const int nblk = (int)((a0.N + 15) / 16);
  MfmaGeomHost g;
  if (!mfma_geom(a0.S, a0.D, a0.N, 2, g, true)) return -1;
  FinalArgs<float> a = a0;
  a.t0 = g.t0;
  a.Km = g.Km;
  a.shift = g.shift;
  a.Nr = g.Nr;
  a.E = g.E;
  // rows of a multiple of 4 values from 16-byte aligned arrays: 16-byte loads
  a.vec4 = a.N % 4 == 0;
  for (int j = 0; j < NIN; ++j) a.vec4 &= ((uintptr_t)a.in[j] & 15) == 0;
  if (mfma_waves(a.N) == 8)
    return nblk <= 16 ? launch_final_mfma_nbw<NIN, 2, 8>(a, g, st)
                      : launch_final_mfma_nbw<NIN, 4, 8>(a, g, st);
  if (nblk <= 4) return launch_final_mfma_nbw<NIN, 1, 4>(a, g, st);
  if (nblk <= 6) return launch_final_mfma_nbw<NIN, 3, 2>(a, g, st);
  if (nblk <= 8) return launch_final_mfma_nbw<NIN, 2, 4>(a, g, st);
  if (nblk <= 12) return launch_final_mfma_nbw<NIN, 3, 4>(a, g, st);
  return launch_final_mfma_nbw<NIN, 4, 4>(a, g, st);
}

template <typename W>
int launch_final(const FinalArgs<W>& a, hipStream_t st) {
  if (std::is_same<W, float>::value) {
    const FinalArgs<float>& af = *reinterpret_cast<const FinalArgs<float>*>(&a);
    const int rc = a.nin == 6 ? launch_final_mfma<6>(af, st) : launch_final_mfma<3>(af, st);
    if (rc >= 0) return rc;
  }
  const int lds = (2 * a.rows * a.E + 2 * (a.S.K + 4)) * (int)sizeof(W);
  const int64_t blocks = (a.nrows + a.rows - 1) / a.rows;
  if (a.nin == 6)
    HC_LAUNCH("k_final_contig", st, (k_final_contig<W, 6, kFinIPT, kFinP>),
              dim3((unsigned)blocks), dim3(256), lds, a);
  else
    HC_LAUNCH("k_final_contig", st, (k_final_contig<W, 3, kFinIPT, kFinP>),
              dim3((unsigned)blocks), dim3(256), lds, a);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

// Check whether every pass of the fused schedule fits; returns HC_OK or
// HC_ERR_RESOURCE without launching anything.
template <typename W>
int plan_fits(const hc_array* img, const AxisTaps<W>* tp) {
  const int nd = img->ndim;
  for (int x = 0; x < nd - 1; ++x) {
    const int Nr = (int)((img->shape[x] + kChP - 1) / kChP * kChP);
    if (lds_strided<W>(Nr + tp[x].S.K + 4) + 2 * (tp[x].S.K + 4) * (int)sizeof(W) > kMaxLds ||
        (Nr / kChP) * kChC > 256 * kChMaxItems)
      return fail(HC_ERR_RESOURCE, "line too long for the fused Hessian");
  }
  const int64_t N = img->shape[nd - 1];
  const int Nr = (int)((N + kFinP - 1) / kFinP * kFinP);
  if (Nr / kFinP > kFinIPT * 256)
    return fail(HC_ERR_RESOURCE, "last axis too long for the fused Hessian epilogue");
  const int E = Nr + tp[nd - 1].S.K + 4;
  if (2 * E * (int)sizeof(W) > kMaxLds)
    return fail(HC_ERR_RESOURCE, "last axis too long for the fused Hessian epilogue");
  return HC_OK;
}

// ------------------------------------------------------ image centering --
// Every Hessian element is a composite of two first-derivative kernels along
// its axes (antisymmetric, zero sum), so it annihilates constants: H(I - c)
// = H(I) exactly, for every mode that extends the image by its own values
// (not 'constant').  The first pass stages I - c with c the image's
// midrange, which halves the magnitude of the staged values (smaller
// rounding in every later pass) and makes an image and its 255-complement
// (black vs white ridges of 8-bit data) stage identical values.  c is
// reduced on the device into the scratch (bytes 32..63), so no host
// synchronisation is needed; with NaN or infinite values c = 0.
__device__ __forceinline__ unsigned long long order_key(double v) {
  const unsigned long long b = (unsigned long long)__double_as_longlong(v);
  return (b >> 63) ? ~b : (b | 0x8000000000000000ull);
}
__device__ __forceinline__ double key_value(unsigned long long k) {
  return __longlong_as_double((long long)((k >> 63) ? (k & 0x7fffffffffffffffull) : ~k));
}

__global__ void k_center_reset(unsigned long long* s) {
  if (threadIdx.x == 0) {
    s[0] = ~0ull;  // min key
    s[1] = 0ull;   // max key
    s[2] = 0ull;   // non-finite seen
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_center_reduce(const void* __restrict__ x, int dt,
                                                       int64_t n, unsigned long long* s) {
  __shared__ double slo[256], shi[256];
  __shared__ int sbad;
  if (threadIdx.x == 0) sbad = 0;
  __syncthreads();
  double lo = INFINITY, hi = -INFINITY;
  bool bad = false;
  auto take = [&](double v) {
    if (!isfinite(v)) {
      bad = true;
    } else {
      lo = v < lo ? v : lo;
      hi = v > hi ? v : hi;
    }
  };
  int64_t i0 = 0;
  if constexpr (std::is_same<T, float>::value) {
    // 16-byte loads, four per lane in flight, over the 16-byte aligned part
    typedef float F4 __attribute__((ext_vector_type(4)));
    const float* xf = reinterpret_cast<const float*>(x);
    const int64_t head = std::min<int64_t>(n, ((16 - ((uintptr_t)xf & 15)) & 15) / 4);
    const int64_t nv = (n - head) / 4;
    const F4* xv = reinterpret_cast<const F4*>(xf + head);
    const int64_t stride = (int64_t)gridDim.x * 256;
    for (int64_t c = (int64_t)blockIdx.x * 256 + threadIdx.x; c < nv; c += 4 * stride) {
      F4 t[4];
#pragma unroll
      // a repeated vector (past the end) changes neither minimum nor maximum
      for (int u = 0; u < 4; ++u) t[u] = c + u * stride < nv ? xv[c + u * stride] : xv[c];
#pragma unroll
      for (int u = 0; u < 4; ++u)
#pragma unroll
        for (int j = 0; j < 4; ++j) take((double)t[u][j]);
    }
    // the unaligned head and the tail (fewer than 4 values each)
    const int64_t t0 = head + 4 * nv;
    if (blockIdx.x == 0 && threadIdx.x < head) take((double)xf[threadIdx.x]);
    if (blockIdx.x == 0 && threadIdx.x < n - t0) take((double)xf[t0 + threadIdx.x]);
    i0 = n;
  }
  for (int64_t i = i0 + (int64_t)blockIdx.x * 256 + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * 256) {
    double v;
    if constexpr (std::is_same<T, void>::value)
      v = load_as<double>(x, i, dt);
    else
      v = (double)reinterpret_cast<const T*>(x)[i];
    take(v);
  }
  if (bad) sbad = 1;
  const int tid = threadIdx.x;
  slo[tid] = lo;
  shi[tid] = hi;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (tid < o) {
      slo[tid] = slo[tid + o] < slo[tid] ? slo[tid + o] : slo[tid];
      shi[tid] = shi[tid + o] > shi[tid] ? shi[tid + o] : shi[tid];
    }
    __syncthreads();
  }
  if (tid == 0) {
    if (slo[0] <= shi[0]) {
      atomicMin(s, order_key(slo[0]));
      atomicMax(s + 1, order_key(shi[0]));
    }
    if (sbad) atomicOr(s + 2, 1ull);
  }
}

__global__ void k_center_final(const unsigned long long* s, double* c) {
  if (threadIdx.x == 0) {
    const double lo = key_value(s[0]), hi = key_value(s[1]);
    c[0] = (s[2] || s[0] == ~0ull || !(lo <= hi)) ? 0.0 : 0.5 * (lo + hi);
  }
}

// *center = device address of the midrange of `in` (kernels enqueued on st),
// or nullptr when centering does not apply (no scratch, constant mode).
int image_center(const hc_array* in, int mode, void* scratch, hipStream_t st,
                 const double** center) {
  *center = nullptr;
  if (!scratch || mode == HC_MODE_CONSTANT) return HC_OK;
  const int64_t n = numel(in);
  if (n == 0) return HC_OK;
  unsigned long long* s =
      reinterpret_cast<unsigned long long*>(reinterpret_cast<char*>(scratch) + 32);
  double* c = reinterpret_cast<double*>(reinterpret_cast<char*>(scratch) + 56);
  const unsigned grid = (unsigned)std::max<int64_t>(1, std::min<int64_t>(1024, (n + 4095) / 4096));
  HC_LAUNCH("k_center_reset", st, k_center_reset, dim3(1), dim3(64), 0, s);
  if (in->dtype == HC_FLOAT32)
    HC_LAUNCH("k_center_reduce", st, k_center_reduce<float>, dim3(grid), dim3(256), 0, in->data, in->dtype, n, s);
  else if (in->dtype == HC_FLOAT64)
    HC_LAUNCH("k_center_reduce", st, k_center_reduce<double>, dim3(grid), dim3(256), 0, in->data, in->dtype, n, s);
  else
    HC_LAUNCH("k_center_reduce", st, k_center_reduce<void>, dim3(grid), dim3(256), 0, in->data, in->dtype, n, s);
  HC_LAUNCH("k_center_final", st, k_center_final, dim3(1), dim3(64), 0, s, c);
  HC_HIP(hipGetLastError());
  *center = c;
  return HC_OK;
}

// Runs the non-final passes of the Hessian schedule from `img` into the
// buffers buf[k] (one per output pair).  After it, buf[k] holds output k's
// chains along every axis but the last.
template <typename W>
int hessian_prefix(const hc_array* img, W* const* buf, const Pair* P, int nP,
                   const AxisTaps<W>* tp, int mode, double cval, int negate, hipStream_t st,
                   const double* center) {
  const int nd = img->ndim;
  // pass over axis 0 from the image: distinct chains
  int node_of[6];
  int nnode = 0;
  int node_first[6], node_second[6], node_buf[6];
  for (int k = 0; k < nP; ++k) {
    const int f = chain_first(P[k], 0), s = chain_second(P[k], 0);
    int n = -1;
    for (int m = 0; m < nnode; ++m)
      if (node_first[m] == f && node_second[m] == s) n = m;
    if (n < 0) {
      n = nnode++;
      node_first[n] = f;
      node_second[n] = s;
      node_buf[n] = k;  // stored in its first child's buffer
    }
    node_of[k] = n;
  }
  if (nnode > 3) return fail(HC_ERR_VALUE, "internal: too many chains");
  ChainJob job0;
  std::memset(&job0, 0, sizeof(job0));
  job0.in = img->data;
  job0.in_dt = img->dtype;
  job0.nout = nnode;
  for (int m = 0; m < nnode; ++m)
    job0.o[m] = ChainOut{buf[node_buf[m]], node_first[m], node_second[m]};
  int rc = run_chain_pass<W>(&job0, 1, img, 0, tp[0], mode, cval, negate, st, center);
  if (rc != HC_OK) return rc;
  if (nd == 3) {
    // pass over axis 1: one job per axis-0 node
    ChainJob jobs[3];
    std::memset(jobs, 0, sizeof(jobs));
    for (int m = 0; m < nnode; ++m) {
      jobs[m].in = buf[node_buf[m]];
      jobs[m].in_dt = sizeof(W) == 4 ? HC_FLOAT32 : HC_FLOAT64;
      for (int k = 0; k < nP; ++k) {
        if (node_of[k] != m) continue;
        if (jobs[m].nout >= 3) return fail(HC_ERR_VALUE, "internal: too many chains");
        jobs[m].o[jobs[m].nout++] =
            ChainOut{buf[k], chain_first(P[k], 1), chain_second(P[k], 1)};
      }
    }
    rc = run_chain_pass<W>(jobs, nnode, img, 1, tp[1], mode, cval, 0, st);
    if (rc != HC_OK) return rc;
  }
  return HC_OK;
}

template <typename W>
int hessian_impl(const hc_array* in, const hc_array* H, int nH, const hc_taps* ts,
                 const hc_taps* td, int mode, double cval, int order_xy, void* scratch,
                 hipStream_t st) {
  const int nd = in->ndim;
  AxisTaps<W> tp[3];
  for (int x = 0; x < nd; ++x) {
    int rc = axis_taps<W>(&ts[x], &td[x], &tp[x]);
    if (rc != HC_OK) return rc;
  }
  int rc = plan_fits<W>(in, tp);
  if (rc != HC_OK) return rc;
  Pair P[6];
  const int nP = hessian_pairs(nd, order_xy, P);
  if (nP != nH) return fail(HC_ERR_VALUE, "wrong number of Hessian outputs");
  W* buf[6];
  for (int k = 0; k < nH; ++k) buf[k] = reinterpret_cast<W*>(H[k].data);
  if (numel(in) == 0) return HC_OK;
  const double* center = nullptr;
  rc = image_center(in, mode, scratch, st, &center);
  if (rc != HC_OK) return rc;
  rc = hessian_prefix<W>(in, buf, P, nP, tp, mode, cval, 0, st, center);
  if (rc != HC_OK) return rc;
  FinalArgs<W> a;
  std::memset(&a, 0, sizeof(a));
  rc = prepare_final<W>(&a, in, tp[nd - 1], mode, cval);
  if (rc != HC_OK) return rc;
  a.nin = nH;
  for (int k = 0; k < nH; ++k) {
    a.in[k] = buf[k];
    a.out[k] = buf[k];
    a.first[k] = chain_first(P[k], nd - 1);
    a.second[k] = chain_second(P[k], nd - 1);
  }
  a.epi = EPI_STORE;
  return launch_final<W>(a, st);
}

template <typename W>
int frangi_impl(const hc_array* in, const hc_array* out, const hc_taps* ts, const hc_taps* td,
                int nsig, double alpha, double beta, double gamma, bool gamma_known,
                int black_ridges, int mode,
                double cval, const hc_array* work, int nwork, void* scratch,
                const double* center_in, hipStream_t st) {
  const int nd = in->ndim;
  const int nH = nd == 3 ? 6 : 3;
  if (nwork < nH) return fail(HC_ERR_VALUE, "not enough work arrays");
  for (int k = 0; k < nH; ++k) {
    if (!same_shape(&work[k], in) || !is_c_contiguous(&work[k]))
      return fail(HC_ERR_VALUE, "work arrays must be C-contiguous with the image's shape");
    if (work[k].dtype != out->dtype) return fail(HC_ERR_VALUE, "work dtype mismatch");
  }
  std::vector<AxisTaps<W>> tps((size_t)nsig * nd);
  for (int s = 0; s < nsig; ++s)
    for (int x = 0; x < nd; ++x) {
      int rc = axis_taps<W>(&ts[s * nd + x], &td[s * nd + x], &tps[(size_t)s * nd + x]);
      if (rc != HC_OK) return rc;
    }
  for (int s = 0; s < nsig; ++s) {
    int rc = plan_fits<W>(in, &tps[(size_t)s * nd]);
    if (rc != HC_OK) return rc;
  }
  const int64_t n = numel(in);
  if (n == 0 || nsig == 0) return HC_OK;
  if (!scratch) return fail(HC_ERR_VALUE, "scratch is null");
  unsigned long long* rg = reinterpret_cast<unsigned long long*>(scratch);
  double* gsq_dev = reinterpret_cast<double*>(reinterpret_cast<char*>(scratch) + 16);
  const double alpha_sq = 2 * alpha * alpha, beta_sq = 2 * beta * beta;
  const double gamma_sq = gamma_known ? 2 * gamma * gamma : 0.0;
  Pair P[6];
  hessian_pairs(nd, 0, P);
  W* buf[6];
  for (int k = 0; k < nH; ++k) buf[k] = reinterpret_cast<W*>(work[k].data);
  W* fo = reinterpret_cast<W*>(out->data);
  const int negate = black_ridges ? 0 : 1;
  const double* center = nullptr;
  if (center_in) {
    if (mode != HC_MODE_CONSTANT) center = center_in;
  } else {
    const int rc = image_center(in, mode, scratch, st, &center);
    if (rc != HC_OK) return rc;
  }
  for (int s = 0; s < nsig; ++s) {
    const AxisTaps<W>* tp = &tps[(size_t)s * nd];
    int rc = hessian_prefix<W>(in, buf, P, nH, tp, mode, cval, negate, st, center);
    if (rc != HC_OK) return rc;
    FinalArgs<W> a;
    std::memset(&a, 0, sizeof(a));
    rc = prepare_final<W>(&a, in, tp[nd - 1], mode, cval);
    if (rc != HC_OK) return rc;
    a.nin = nH;
    for (int k = 0; k < nH; ++k) {
      a.in[k] = buf[k];
      a.first[k] = chain_first(P[k], nd - 1);
      a.second[k] = chain_second(P[k], nd - 1);
    }
    a.epi = EPI_FRANGI;
    a.fout = fo;
    a.alpha_sq = alpha_sq;
    a.beta_sq = beta_sq;
    a.first_sigma = s == 0;
    const bool store = (s == 0) && !gamma_known;
    a.store_eig = store;
    if (store) {
      for (int e = 0; e < nd; ++e) a.eig[e] = buf[e];
      a.rg_max = rg;
      HC_LAUNCH("k_frangi_reset", st, k_frangi_reset, dim3(1), dim3(64), 0, rg);
    }
    if (gamma_known) a.gamma_sq = gamma_sq;
    else if (s > 0) a.gamma_sq_dev = gsq_dev;
    rc = launch_final<W>(a, st);
    if (rc != HC_OK) return rc;
    if (store) {
      HC_LAUNCH("k_frangi_gamma", st, k_frangi_gamma<W>, dim3(1), dim3(64), 0, rg, gsq_dev);
      const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
      if (nd == 3)
        HC_LAUNCH("k_frangi_from_eig", st, (k_frangi_from_eig<W, 3>), dim3((unsigned)blocks), dim3(256), 0, buf[0], buf[1], buf[2], fo, n, alpha_sq, beta_sq, gsq_dev);
      else
        HC_LAUNCH("k_frangi_from_eig", st, (k_frangi_from_eig<W, 2>), dim3((unsigned)blocks), dim3(256), 0, buf[0], buf[1], buf[1], fo, n, alpha_sq, beta_sq, gsq_dev);
      HC_HIP(hipGetLastError());
    }
  }
  return HC_OK;
}

template <typename W>
int frangi_eig_impl(const hc_array* eig, int nd, const hc_array* out, double alpha, double beta,
                    double gamma, bool known, int first_sigma, void* scratch,
                    hipStream_t st) {
  const int64_t n = numel(out);
  if (n == 0) return HC_OK;
  const double alpha_sq = 2 * alpha * alpha, beta_sq = 2 * beta * beta;
  const W* e0 = (const W*)eig[0].data;
  const W* e1 = (const W*)eig[1].data;
  const W* e2 = (const W*)eig[nd == 3 ? 2 : 1].data;
  W* o = (W*)out->data;
  unsigned long long* rg = reinterpret_cast<unsigned long long*>(scratch);
  double* gsq_dev = reinterpret_cast<double*>(reinterpret_cast<char*>(scratch) + 16);
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
  if (!known && first_sigma) {
    HC_LAUNCH("k_frangi_reset", st, k_frangi_reset, dim3(1), dim3(64), 0, rg);
    if (nd == 3)
      HC_LAUNCH("k_rg_max", st, (k_rg_max<W, 3>), dim3((unsigned)blocks), dim3(256), 0, e0, e1, e2, n, (void*)rg);
    else
      HC_LAUNCH("k_rg_max", st, (k_rg_max<W, 2>), dim3((unsigned)blocks), dim3(256), 0, e0, e1, e2, n, (void*)rg);
    HC_LAUNCH("k_frangi_gamma", st, k_frangi_gamma<W>, dim3(1), dim3(64), 0, rg, gsq_dev);
  }
  const double gsq = known ? 2 * gamma * gamma : 0.0;
  const double* gdev = known ? nullptr : gsq_dev;
  if (nd == 3)
    HC_LAUNCH("k_frangi_eig_max", st, (k_frangi_eig_max<W, 3>), dim3((unsigned)blocks), dim3(256), 0, e0, e1, e2, o, n, alpha_sq, beta_sq, gsq, gdev, first_sigma);
  else
    HC_LAUNCH("k_frangi_eig_max", st, (k_frangi_eig_max<W, 2>), dim3((unsigned)blocks), dim3(256), 0, e0, e1, e2, o, n, alpha_sq, beta_sq, gsq, gdev, first_sigma);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

// Passes along different axes commute for the index-remapping modes and for
// zero padding; with mode='constant' and cval != 0 they do not (the padding
// value is not scaled by the other axes' kernels), so the fused schedule,
// which reorders the reference's passes, must not run.
int check_commuting(int mode, double cval) {
  if (mode == HC_MODE_CONSTANT && cval != 0.0)
    return fail(HC_ERR_UNSUPPORTED,
                "mode='constant' with cval != 0 needs the reference pass order");
  return HC_OK;
}

int check_float_image(const hc_array* in) {
  if (!in) return fail(HC_ERR_VALUE, "null array");
  if (in->ndim != 2 && in->ndim != 3)
    return fail(HC_ERR_VALUE, "the fused Hessian handles 2-D and 3-D images");
  if (!is_c_contiguous(in)) return fail(HC_ERR_VALUE, "input must be C-contiguous");
  for (int d = 0; d < in->ndim; ++d)
    if (in->shape[d] > (int64_t)1 << 30) return fail(HC_ERR_VALUE, "axis too long");
  return HC_OK;
}

}  // namespace
}  // namespace hc

using namespace hc;

extern "C" {

int hc_max_resident_line(int w_f64) {
  // strided passes bound it: 2 lines x (N + K) x 17 columns in LDS and at
  // most 256 x kChMaxItems eight-output items per work-group
  const int es = w_f64 ? 8 : 4;
  const int lds_bound = kMaxLds / (2 * kChS * es) - 136;
  const int item_bound = 256 * kChMaxItems / kChC * kChP;
  return lds_bound < item_bound ? lds_bound : item_bound;
}

int hc_hessian_gaussian(const hc_array* in, const hc_array* H, int nH, const hc_taps* taps_s,
                        const hc_taps* taps_d, int mode, double cval, int order_xy,
                        void* scratch, void* stream) {
  int rc = check_float_image(in);
  if (rc != HC_OK) return rc;
  if (in->dtype != HC_FLOAT32 && in->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "hessian input must be float32 or float64");
  if (order_xy && in->ndim != 2) return fail(HC_ERR_VALUE, "Order 'xy' is only supported for 2D images.");
  if (mode < HC_MODE_REFLECT || mode > HC_MODE_CONSTANT)
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  rc = check_commuting(mode, cval);
  if (rc != HC_OK) return rc;
  if (!H || !taps_s || !taps_d) return fail(HC_ERR_VALUE, "null argument");
  for (int k = 0; k < nH; ++k) {
    if (!same_shape(&H[k], in) || !is_c_contiguous(&H[k]) || H[k].dtype != in->dtype)
      return fail(HC_ERR_VALUE, "H arrays must be C-contiguous with the image's shape and dtype");
  }
  const bool f64 = in->dtype == HC_FLOAT64;
  for (int x = 0; x < in->ndim; ++x)
    if ((taps_s[x].w_f64 != 0) != f64 || (taps_d[x].w_f64 != 0) != f64)
      return fail(HC_ERR_VALUE, "weights dtype must match the image dtype");
  hipStream_t st = (hipStream_t)stream;
  return f64 ? hessian_impl<double>(in, H, nH, taps_s, taps_d, mode, cval, order_xy, scratch, st)
             : hessian_impl<float>(in, H, nH, taps_s, taps_d, mode, cval, order_xy, scratch, st);
}

int hc_sym_eigvals(const hc_array* S, int nS, const hc_array* eig, int sort, int abs_sort,
                   void* stream) {
  if (!S || !eig) return fail(HC_ERR_VALUE, "null argument");
  int ne = 0;  // matrix order: nS = ne (ne + 1) / 2
  while (ne < 8 && ne * (ne + 1) / 2 < nS) ++ne;
  if (nS < 1 || ne * (ne + 1) / 2 != nS)
    return fail(HC_ERR_VALUE, "element count must be n(n+1)/2 for a matrix order 1..8");
  const int dt = S[0].dtype;
  if (dt != HC_FLOAT32 && dt != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "expected real-valued floating point matrices");
  for (int k = 0; k < nS; ++k)
    if (!same_shape(&S[k], &S[0]) || !is_c_contiguous(&S[k]) || S[k].dtype != dt)
      return fail(HC_ERR_VALUE, "element arrays must be C-contiguous, same shape and dtype");
  for (int k = 0; k < ne; ++k)
    if (!same_shape(&eig[k], &S[0]) || !is_c_contiguous(&eig[k]) || eig[k].dtype != dt)
      return fail(HC_ERR_VALUE, "eigenvalue arrays must match the element arrays");
  const int64_t n = numel(&S[0]);
  if (n == 0) return HC_OK;
  hipStream_t st = (hipStream_t)stream;
  // other orders: the iterative solver of the n-dimensional branch
  if (nS != 3 && nS != 6) return sym_eigvals_nd(S, nS, eig, sort, abs_sort, st);
  const int64_t blocks = std::min<int64_t>((n + 255) / 256, 256 * 64);
  const void* s[6];
  for (int k = 0; k < 6; ++k) s[k] = S[k < nS ? k : 0].data;
  void* e[3];
  for (int k = 0; k < 3; ++k) e[k] = eig[k < ne ? k : 0].data;
  if (dt == HC_FLOAT64)
    HC_LAUNCH("k_sym_eig", st, k_sym_eig<double>, dim3((unsigned)blocks), dim3(256), 0, (const double*)s[0], (const double*)s[1], (const double*)s[2], (const double*)s[3], (const double*)s[4], (const double*)s[5], (double*)e[0], (double*)e[1], (double*)e[2], n, nS, sort, abs_sort);
  else
    HC_LAUNCH("k_sym_eig", st, k_sym_eig<float>, dim3((unsigned)blocks), dim3(256), 0, (const float*)s[0], (const float*)s[1], (const float*)s[2], (const float*)s[3], (const float*)s[4], (const float*)s[5], (float*)e[0], (float*)e[1], (float*)e[2], n, nS, sort, abs_sort);
  HC_HIP(hipGetLastError());
  return HC_OK;
}

int hc_frangi_from_eigvals(const hc_array* eig, int ndim, const hc_array* out, double alpha,
                           double beta, double gamma, int has_gamma, int first_sigma,
                           void* scratch, void* stream) {
  if (!eig || !out || !scratch) return fail(HC_ERR_VALUE, "null argument");
  if (ndim != 2 && ndim != 3) return fail(HC_ERR_VALUE, "frangi handles 2-D and 3-D images");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "frangi output must be float32 or float64");
  if (!is_c_contiguous(out)) return fail(HC_ERR_VALUE, "output must be C-contiguous");
  for (int k = 0; k < ndim; ++k)
    if (!same_shape(&eig[k], out) || !is_c_contiguous(&eig[k]) || eig[k].dtype != out->dtype)
      return fail(HC_ERR_VALUE, "eigenvalue arrays must match the output");
  hipStream_t st = (hipStream_t)stream;
  return out->dtype == HC_FLOAT64
             ? frangi_eig_impl<double>(eig, ndim, out, alpha, beta, gamma, has_gamma != 0,
                                       first_sigma, scratch, st)
             : frangi_eig_impl<float>(eig, ndim, out, alpha, beta, gamma, has_gamma != 0,
                                      first_sigma, scratch, st);
}

int hc_frangi(const hc_array* in, const hc_array* out, const hc_taps* taps_s,
              const hc_taps* taps_d, int nsig, double alpha, double beta, double gamma,
              int has_gamma, int black_ridges, int mode, double cval, const hc_array* work,
              int nwork, void* scratch, const double* center, void* stream) {
  int rc = check_float_image(in);
  if (rc != HC_OK) return rc;
  if (!out || !same_shape(in, out) || !is_c_contiguous(out))
    return fail(HC_ERR_VALUE, "output must be C-contiguous with the image's shape");
  if (out->dtype != HC_FLOAT32 && out->dtype != HC_FLOAT64)
    return fail(HC_ERR_VALUE, "frangi output must be float32 or float64");
  if (mode < HC_MODE_REFLECT || mode > HC_MODE_CONSTANT)
    return fail(HC_ERR_VALUE, "boundary mode not supported");
  if (nsig < 0 || (nsig > 0 && (!taps_s || !taps_d))) return fail(HC_ERR_VALUE, "bad sigmas");
  rc = check_commuting(mode, cval);
  if (rc != HC_OK) return rc;
  if (!work) return fail(HC_ERR_VALUE, "null work arrays");
  const bool f64 = out->dtype == HC_FLOAT64;
  for (int k = 0; k < nsig * in->ndim; ++k)
    if ((taps_s[k].w_f64 != 0) != f64 || (taps_d[k].w_f64 != 0) != f64)
      return fail(HC_ERR_VALUE, "weights dtype must match the work dtype");
  hipStream_t st = (hipStream_t)stream;
  return f64 ? frangi_impl<double>(in, out, taps_s, taps_d, nsig, alpha, beta, gamma,
                                   has_gamma != 0, black_ridges, mode, cval, work, nwork,
                                   scratch, center, st)
             : frangi_impl<float>(in, out, taps_s, taps_d, nsig, alpha, beta, gamma,
                                  has_gamma != 0, black_ridges, mode, cval, work, nwork,
                                  scratch, center, st);
}

}  // extern "C"
