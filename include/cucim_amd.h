/*
 * cucim_amd -- C ABI of the MI355X (gfx950) separable-filter / Hessian-ridge
 * library.  Every entry point is stream-ordered, never throws, returns an
 * hc_status code and leaves a message for hc_last_error() on failure.
 *
 * Arrays are described by hc_array (device pointer, element type, shape,
 * element strides: the information a DLPack tensor carries).  The caller owns
 * every buffer, including outputs and work arrays; the library allocates no
 * device memory.  Weights are passed from the host as float64 values already
 * rounded to the weights dtype, with a device copy for long kernels.
 *
 * The reference exposes this path only as Python functions (CuPy arrays in,
 * CuPy arrays out); each entry point below names the reference function it
 * replaces, paths relative to python/cucim/src/cucim/skimage/.
 */
#ifndef CUCIM_AMD_H
#define CUCIM_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HC_MAX_NDIM 8

/* element types */
typedef enum hc_dtype {
  HC_BOOL = 0,
  HC_INT8 = 1,
  HC_UINT8 = 2,
  HC_INT16 = 3,
  HC_UINT16 = 4,
  HC_INT32 = 5,
  HC_UINT32 = 6,
  HC_INT64 = 7,
  HC_UINT64 = 8,
  HC_FLOAT16 = 9,
  HC_FLOAT32 = 10,
  HC_FLOAT64 = 11,
  HC_BFLOAT16 = 12  /* surrogate network activations only */
} hc_dtype;

/* boundary modes, _vendored/_ndimage_util.py:186-308 ('wrap' == 'grid-wrap',
 * 'grid-mirror' == 'reflect', 'grid-constant' == 'constant') */
typedef enum hc_mode {
  HC_MODE_REFLECT = 0,
  HC_MODE_MIRROR = 1,
  HC_MODE_NEAREST = 2,
  HC_MODE_WRAP = 3,
  HC_MODE_CONSTANT = 4
} hc_mode;

/* line-kernel choice, the algorithm= keyword of
 * _vendored/_ndimage_filters.py:135-221 */
typedef enum hc_algorithm {
  HC_ALGO_AUTO = 0,
  HC_ALGO_TILED = 1,  /* 'shared_memory': LDS halo tiles */
  HC_ALGO_GATHER = 2  /* 'elementwise': one work-item per output */
} hc_algorithm;

typedef enum hc_status {
  HC_OK = 0,
  HC_ERR_VALUE = 1,     /* bad argument (ValueError / RuntimeError upstream) */
  HC_ERR_RESOURCE = 2,  /* LDS budget exceeded (ResourceLimitError upstream) */
  HC_ERR_HIP = 3,       /* HIP runtime error */
  HC_ERR_UNSUPPORTED = 4
} hc_status;

typedef struct hc_array {
  void* data;                     /* device pointer */
  int32_t dtype;                  /* hc_dtype */
  int32_t ndim;
  int64_t shape[HC_MAX_NDIM];
  int64_t strides[HC_MAX_NDIM];   /* element strides */
} hc_array;

typedef struct hc_taps {
  const double* w;   /* host weights, K values, rounded to the weights dtype */
  int64_t K;
  int32_t anchor;    /* K // 2 + origin: tap k reads position i + k - anchor */
  int32_t w_f64;     /* weights / accumulation dtype: 0 float32, 1 float64 */
  const void* w_dev; /* device copy (weights dtype); needed when
                        K > hc_max_inline_taps(w_f64), else may be NULL */
} hc_taps;

const char* hc_last_error(void);
int hc_version(void);
int hc_max_inline_taps(int w_f64);
/* largest line the fused Hessian / Frangi passes keep resident in LDS */
int hc_max_resident_line(int w_f64);

/* Launch tracing (no counterpart in the reference's skimage path, whose only
 * tracing is NVTX in the I/O library).  While enabled, every kernel launch is
 * bracketed by a hipEvent pair on its stream.  hc_timing_collect() waits for
 * the recorded launches, writes up to max_records elapsed times (ms) and the
 * kernel names ('\n'-separated) and clears the record. */
int hc_timing_enable(int on);
int hc_timing_collect(char* names, int names_len, double* ms, int max_records);

/* Tile-shape knobs (no counterpart in the reference, whose shared-memory
 * kernel fixes its tiles in filters/_separable_filtering.py:17-91).  A value
 * of 0 restores the built-in heuristic.  Process-wide, read at each
 * launch; the batched 2-D driver's autotuner sets HC_TUNE_SEP2D_TY-like
 * choices per plan (cucim_amd/batch.py). */
typedef enum hc_tune_knob {
  HC_TUNE_CONTIG_SEG = 0,   /* contiguous-axis correlate: outputs per wave
                               segment, 256..4096, multiple of 256 */
  HC_TUNE_SEP2D_TY = 1,     /* fused 2-D separable: output rows per tile,
                               8 (float32 register-blocked kernel only), 16,
                               32 or 64 */
  HC_TUNE_HESSIAN_VALU = 2, /* 1: fused Hessian passes on the vector ALUs
                               only, tap by tap as the reference computes
                               (default 0: float32 passes on the matrix
                               cores where they apply) */
  HC_TUNE_EDGE_ROWS = 3,    /* fused 2-D edge magnitude (float input of the
                               output's type): output rows per wave, 8, 16,
                               32, 64 or 128 */
  HC_TUNE_MFMA_WAVES = 4,   /* matrix-core Hessian passes, lines of 129..256:
                               waves per work-group, 4 or 8 */
  HC_TUNE_NKNOBS = 8
} hc_tune_knob;
int hc_tune_set(int knob, int value);
int hc_tune_get(int knob);

/* One image of a batched 2-D pipeline (no counterpart in the reference,
 * which filters one image per call).  item0 is filled by hc_batch2d_plan. */
typedef struct hc_tile2d {
  const void* in;   /* input image, C-contiguous */
  void* out;        /* output image, C-contiguous */
  int64_t H, W;     /* rows, columns */
  int64_t item0;    /* index of the image's first work item */
} hc_tile2d;

/* Host-side plan of a batched 2-D pipeline over n images: fills
 * tiles[i].item0 and, if item_tile is not NULL, item_tile[0 .. total) with
 * the image of each work item (cap entries available).  K: gaussian kernel
 * length, ty: output rows per work item (16 or 32).  Returns the number of
 * work items, or -1 (message in hc_last_error()). */
int64_t hc_batch2d_plan(hc_tile2d* tiles, int n, int K, int ty, int32_t* item_tile,
                        int64_t cap);

/* rgb2gray -> gaussian -> edge magnitude of many uneven uint8 RGB images in
 * one persistent launch, each output equal to
 *   filters.sobel(filters.gaussian(color.rgb2gray(rgb), sigma, mode=gauss_mode,
 *                                  cval=gauss_cval), mode=edge_mode)
 * (color/colorconv.py:1037-1095, _shared/filters.py:21-166,
 * filters/edges.py:148-241; the edge filter's weights given as correlation
 * weights, edge already divided by sqrt(2)).  tiles_dev / item_tile_dev:
 * device copies of the planned table (inputs (H, W, 3) uint8, outputs
 * (H, W) float32).  Gaussian taps float32, K <= 33.  edge_mode 'wrap' is
 * HC_ERR_UNSUPPORTED (run the three calls instead). */
int hc_batch2d_gray_gauss_sobel(const hc_tile2d* tiles_dev, int n,
                                const int32_t* item_tile_dev, int64_t n_items,
                                const hc_taps* gauss, int gauss_mode, double gauss_cval,
                                const double* edge3, const double* smooth3, int edge_mode,
                                int ty, void* stream);

/* correlate1d / convolve1d core: _vendored/_ndimage_filters.py:271-344
 * (_correlate_or_convolve1d), the shared-memory kernel of
 * filters/_separable_filtering.py:977-1070 and the elementwise kernel of
 * _vendored/_ndimage_filters_core.py:247-408.  Convolution is expressed by
 * the caller as correlation with reversed weights (_ndimage_filters.py:236-243).
 * out must be C-contiguous with in's shape and must not overlap in. */
int hc_correlate1d(const hc_array* in, const hc_array* out, const hc_taps* taps,
                   int axis, int mode, double cval, int algorithm, void* stream);

/* n-D correlate with a full weights array: _vendored/_ndimage_filters.py:224-257
 * (_correlate_or_convolve) + _ndimage_filters_core.py:247-408.  w_dev holds
 * prod(wshape) weights (weights dtype) in C order; offsets[d] = wshape[d]//2
 * + origin[d]; modes per axis. */
int hc_correlate_nd(const hc_array* in, const hc_array* out, const void* w_dev,
                    const int64_t* wshape, const int32_t* offsets,
                    const int32_t* modes, int w_f64, double cval, void* stream);

/* Two 1-D passes of a 2-D array in one launch (axis 0, then axis 1), the
 * intermediate rounded to out's dtype exactly as _run_1d_filters
 * (_vendored/_ndimage_filters_core.py:124-163) stores it.  Serves
 * gaussian_filter / uniform_filter / sobel / prewitt on 2-D inputs. */
int hc_separable2d(const hc_array* in, const hc_array* out, const hc_taps* taps0,
                   const hc_taps* taps1, int mode0, int mode1, double cval,
                   void* stream);

/* Hessian by Gaussian derivatives for 2-D / 3-D float32/float64 images:
 * feature/corner.py:161-254 (_hessian_matrix_with_gaussian).
 * taps_s[d] / taps_d[d]: order-0 / order-1 kernels of axis d.  H: the
 * ndim*(ndim+1)/2 outputs (C-contiguous, image dtype) in
 * combinations_with_replacement order; order_xy != 0 reverses the axes
 * (2-D only).  No work memory is needed.  The passes run per axis, which
 * reorders the reference's passes: exact for every mode except 'constant'
 * with cval != 0, which returns HC_ERR_UNSUPPORTED (run the 1-D passes in
 * the reference order with hc_correlate1d instead).  scratch (optional, >= 64
 * bytes of device memory): when given and mode is not 'constant', the first
 * pass subtracts the image's midrange (reduced on the device), which leaves
 * every Hessian element unchanged in exact arithmetic and shrinks float32
 * rounding; NULL runs on the image as it is. */
int hc_hessian_gaussian(const hc_array* in, const hc_array* H, int nH,
                        const hc_taps* taps_s, const hc_taps* taps_d, int mode,
                        double cval, int order_xy, void* scratch, void* stream);

/* Symmetric eigenvalues per voxel: feature/corner.py:699-755
 * (_symmetric_compute_eigenvalues).  S: nS = n(n+1)/2 element arrays in
 * combinations_with_replacement order (n = 1..8), eig: n outputs.  n = 2, 3:
 * the closed forms of :473-696; other n: the eigvalsh branch (:744-755,
 * _symmetric_image :758-783), by cyclic Jacobi rotations in float64.
 * sort: 0 ascending, 1 descending; abs_sort orders by |value| first. */
int hc_sym_eigvals(const hc_array* S, int nS, const hc_array* eig, int sort,
                   int abs_sort, void* stream);

/* Frangi vesselness: filters/ridges.py:322-476.  taps_s/taps_d hold
 * nsig*ndim kernels (sigma-major).  has_gamma == 0 means gamma=None (gamma
 * from the first sigma, computed on the device; no host synchronisation);
 * otherwise gamma is used as given (any value, 2*gamma^2 as the reference).
 * Except in 'constant' mode the first pass subtracts the image's midrange,
 * as hc_hessian_gaussian does with a scratch (bytes 32..63), or, when center
 * (device, one double) is not NULL, that value instead (a sub-volume then
 * reproduces the values of the whole volume it was cut from).  work: 6
 * (3-D) or 3 (2-D) arrays of the image's shape in the work dtype;
 * scratch: >= 64 bytes of device memory.  out: image shape, work dtype.
 * mode='constant' with cval != 0 returns HC_ERR_UNSUPPORTED (see
 * hc_hessian_gaussian). */
int hc_frangi(const hc_array* in, const hc_array* out, const hc_taps* taps_s,
              const hc_taps* taps_d, int nsig, double alpha, double beta,
              double gamma, int has_gamma, int black_ridges, int mode, double cval,
              const hc_array* work, int nwork, void* scratch, const double* center,
              void* stream);

/* Frangi vesselness of one sigma from eigenvalues already sorted ascending
 * by |value| (filters/ridges.py:428-475 after the eigenvalue step): eig holds
 * ndim arrays; first_sigma != 0 writes out, else out = max(out, v).  With
 * has_gamma == 0, the first sigma reduces max(sum lambda^2) into scratch
 * (>= 64 bytes) and later sigmas read gamma from there.  Serves images whose
 * lines do not fit the fused hc_frangi schedule. */
int hc_frangi_from_eigvals(const hc_array* eig, int ndim, const hc_array* out,
                           double alpha, double beta, double gamma, int has_gamma,
                           int first_sigma, void* scratch, void* stream);

/* Sato tubeness of one sigma folded into the running maximum over sigmas:
 * filters/ridges.py:148-225 after its hessian_matrix call.  H: the
 * ndim*(ndim+1)/2 Hessian elements (combinations_with_replacement order),
 * out: running maximum (first_sigma != 0 starts it from 0). */
int hc_sato_fold(const hc_array* H, int nH, const hc_array* out, double sigma,
                 int first_sigma, void* stream);

/* Meijering neuriteness of one sigma folded into the running maximum:
 * filters/ridges.py:50-145 after its hessian_matrix call (normalised by
 * the image maximum of the sigma's response, reduced on the device; no host
 * synchronisation).  tmp: an array like out; scratch: >= 8 bytes of device
 * memory. */
int hc_meijering_fold(const hc_array* H, int nH, const hc_array* out, const hc_array* tmp,
                      double alpha, int first_sigma, void* scratch, void* stream);

/* a[a <= threshold] = value (float32 / float64): the last step of the hybrid
 * Hessian filter, filters/ridges.py:479-565. */
int hc_fill_le(const hc_array* a, double threshold, double value, void* stream);

/* ---- 3-D convolutional surrogate of the Frangi stack (no counterpart in the
 * reference; it stands in for filters/ridges.py:322-476 and its output is
 * compared with the plain kernels').  Valid 3x3x3 convolutions on channels-last
 * bfloat16 activations, bf16 weights, float32 accumulation on the matrix
 * cores.  act: 0 none, 1 GELU (tanh form), 2 clamp to [0, 1]. */

/* bytes of the packed (matrix-core operand order) weights of a layer */
int hc_cnn3d_packed_size(int ci, int co);
/* pack float32 weights w[27][ci][co] (tap = (dz*3 + dy)*3 + dx) */
int hc_cnn3d_pack(const float* w_dev, void* packed_dev, int ci, int co, void* stream);
/* out (D-2, H-2, W-2, co) = act(conv(in (D, H, W, ci)) + bias); out bfloat16,
 * or float32 for the single-channel last layer; pre_dev (float32, out's
 * shape) receives the pre-activation when not NULL (training).  Compiled
 * layer shapes: 1->24 GELU, 24->24 GELU, 24->1 none / clamp. */
int hc_cnn3d_conv(const hc_array* in, const hc_array* out, const void* packed,
                  const float* bias_dev, int act, float* pre_dev, void* stream);
/* out (D+2p, H+2p, W+2p[, 1]) bfloat16 = boundary-extended float32 in */
int hc_cnn3d_pad(const hc_array* in, const hc_array* out, int pad, int mode, double cval,
                 void* stream);
/* training (float32): g *= gelu'(pre); input / weight / bias gradients of a
 * layer (w and dw: [27][ci][co]; dw, db accumulate); masked-free MSE
 * gradient (loss accumulates the mean squared error); Adam step */
int hc_cnn3d_gelu_bwd(float* g, const float* pre, int64_t n, void* stream);
int hc_cnn3d_dgrad(const float* dy, const float* w, float* dx, int D, int H, int W, int ci,
                   int co, void* stream);
int hc_cnn3d_wgrad(const void* x_bf16, const float* dy, float* dw, float* db, int D, int H,
                   int W, int ci, int co, void* stream);
int hc_cnn3d_mse_grad(const float* y, const float* t, float* g, int64_t n, float* loss,
                      void* stream);
int hc_cnn3d_adam(float* p, const float* g, float* m, float* v, int64_t n, double lr,
                  double b1, double b2, double eps, int step, void* stream);

/* Sampled plain-Frangi check of the surrogate: for each of nb lines, the
 * plain 3-D Frangi (filters/ridges.py:322-476 with a fixed gamma) evaluated
 * exactly on up to 64 voxels along axis 0 (lines[4i..4i+3] = x0, y0, za,
 * zb), placed by the caller where no Gaussian pass reaches the volume's
 * border; per sigma kc holds the composite kernels SS, SD, DD (full
 * convolutions of the pass kernels, L taps each, anchor A = 2 * anchor) and
 * kinfo (offset, L, A).  plain_out: nb x 64 values; err_out: per line
 * sum (sur - plain)^2, sum plain^2 and the number of samples.  A line with
 * zb - za outside [0, 64], or whose windows would reach outside the volume,
 * is not evaluated (0 samples, NaN plain values). */
int hc_frangi_check(const hc_array* vol, const hc_array* sur, const int32_t* lines_dev, int nb,
                    const float* kc_dev, const int32_t* kinfo_dev, const int32_t* kinfo_host,
                    int nsig, int black_ridges, double alpha, double beta, double gamma,
                    float* plain_out, float* err_out, void* stream);

/* rgb2gray: color/colorconv.py:1037-1095 (+ _prepare_colorarray :160-180).
 * in: (..., 3) C-contiguous, any real dtype; out: (...) float32/float64. */
int hc_rgb2gray(const hc_array* in, const hc_array* out, void* stream);

/* img_as_float32/64 style conversion (util/dtype.py:219-377): unsigned ints
 * scale by 1/imax, signed ints scale and clip at -1, bool/float cast.
 * in and out C-contiguous, same shape. */
int hc_to_float(const hc_array* in, const hc_array* out, void* stream);

/* Finite-difference gradient along one axis: _shared/_gradient.py. */
int hc_gradient(const hc_array* in, const hc_array* out, int axis, void* stream);

/* Edge magnitude sqrt((S_1 E_0 I)^2 + (S_0 E_1 I)^2) of a 2-D image,
 * filters/edges.py:148-241 (_generic_edge_filter with axis=None) in one
 * launch.  edge3/smooth3 are CORRELATION weights (the convolution weights
 * reversed; edge already divided by sqrt(2)); integer input is converted as
 * img_as_float does; constant mode pads with 0 (cval is not forwarded
 * upstream).  axis_mask must be 3; work is unused.  Other ranks are composed
 * from hc_correlate1d + hc_square_acc + hc_sqrt_inplace by the caller. */
int hc_edge_magnitude(const hc_array* in, const hc_array* out,
                      const double* edge3, const double* smooth3, int axis_mask,
                      int mode, const hc_array* work, void* stream);

/* out = out * binary_erosion(mask, full structure, border 0)
 * (filters/edges.py:75-85).  mask: bool/uint8 C-contiguous of out's shape. */
int hc_mask_eroded(const hc_array* out, const hc_array* mask, void* stream);

/* Elementwise NaN-propagating maximum a = max(a, b) (cp.maximum). */
int hc_maximum_inplace(const hc_array* a, const hc_array* b, void* stream);

/* acc = v*v (first != 0) or acc += v*v, each op rounded once (float32 /
 * float64) or wrapped (integer dtypes), C-contiguous: the `output *= output;
 * tmp *= tmp; output += tmp` of _vendored/_ndimage_filters.py:1012-1097. */
int hc_square_acc(const hc_array* acc, const hc_array* v, int first, void* stream);

/* a = sqrt(a), C-contiguous; integer arrays as sqrt(a, a, casting='unsafe')
 * (float for 8/16-bit, double for wider, truncated; NaN -> 0). */
int hc_sqrt_inplace(const hc_array* a, void* stream);

/* Sparse n-D correlation: filters/_sparse.py:94-118 (correlate_sparse) with
 * _correlate_sparse :55-91.  kshape: the kernel's shape (ndim values);
 * tap_idx (ntaps x ndim, host) / tap_w (ntaps, host, already rounded to the
 * output dtype): the nonzero taps in C order, led by the corner (0, ..., 0)
 * (value 0 when the corner tap is zero, as the reference prepends it).
 * valid != 0: no padding, out shape = in shape - kshape + 1; otherwise the
 * image is padded by kshape // 2 per side in `mode` (HC_MODE_*, constant
 * pads with 0; even kernel sizes are rejected) and out has in's shape.
 * out (float32/float64, C-contiguous) = sum_k w_k * P[o + idx_k], each
 * product and add rounded once in out's dtype (w = +-1: no multiply).  in:
 * any real dtype and strides, converted to out's dtype. */
int hc_correlate_sparse(const hc_array* in, const hc_array* out, const int32_t* kshape,
                        const int32_t* tap_idx, const double* tap_w, int ntaps, int valid,
                        int mode, void* stream);

/* a /= d in place, d rounded to a's dtype first and each quotient rounded
 * once (float32/float64, C-contiguous): the `out /= math.sqrt(2)` of
 * filters/edges.py:614-660 (roberts). */
int hc_div_scalar(const hc_array* a, double d, void* stream);

/* Elementwise out = a - b (HC_BIN_SUB), a * b (HC_BIN_MUL) or a + b
 * (HC_BIN_ADD), each result rounded once in the arrays' dtype (float32 /
 * float64), or wrapped modulo 2^bits for integer dtypes (equal shapes,
 * dtypes, C-contiguous; out may alias a or b).  The `out -= gaussian(...)` of
 * filters/_gaussian.py:154-172, the `der0 * der1` products of
 * feature/corner.py:137-148 and the `output += tmp` of
 * _vendored/_ndimage_filters.py:806-890 (generic_laplace). */
enum { HC_BIN_SUB = 0, HC_BIN_MUL = 1, HC_BIN_ADD = 2 };
int hc_binary(const hc_array* a, const hc_array* b, const hc_array* out, int op, void* stream);

/* Strided copy with dtype cast (cast<> semantics of
 * _ndimage_filters_core.py:227-244): out[...] = cast(in[...]). */
int hc_copy_cast(const hc_array* in, const hc_array* out, void* stream);

/* ---- entry points that take sigmas (the kernels are built inside) ---- */

/* The 1-D Gaussian correlation kernel of _vendored/_ndimage_filters.py:
 * 689-717 (_gaussian_kernel1d) in float64: 2 * radius + 1 values into out.
 * Host-only (no device work).  Evaluated as numpy evaluates it (pairwise
 * sum, fused multiply-adds in the polynomial), except that exp is the C
 * library's: values can differ from numpy's in the last bit. */
int hc_gaussian_kernel1d(double sigma, int order, int64_t radius, double* out);

/* gaussian_filter over every axis: _vendored/_ndimage_filters.py:580-686.
 * sigma, order, mode: ndim values each; radius: NULL or ndim values (< 0:
 * int(truncate * sigma + 0.5)); axes with sigma < 1e-15 or radius 0 are
 * skipped; no axis left copies in to out.  Weights in promote(in, float32).
 * out: C-contiguous, in's shape, any dtype (intermediates in out's dtype, as
 * the reference's pass scheduler keeps them).  work: an array like out,
 * needed when two or more passes run (except the fused 2-D float32 case);
 * otherwise may be NULL.  Kernels longer than hc_max_inline_taps return
 * HC_ERR_RESOURCE (run them with hc_correlate1d and a device copy of the
 * taps). */
int hc_gaussian_nd(const hc_array* in, const hc_array* out, const double* sigma,
                   const int* order, const int64_t* radius, const int* mode, double cval,
                   double truncate, const hc_array* work, void* stream);

/* hessian_matrix(use_gaussian_derivatives=True) from sigmas (ndim values):
 * feature/corner.py:161-254, the kernels of :218-232 built in the library
 * (sigma / sqrt(2), truncate 8 if every sigma > 1 else 100, rounded to the
 * image dtype, trimmed to the common nonzero support), then
 * hc_hessian_gaussian (same arrays, modes and errors). */
int hc_hessian_matrix(const hc_array* in, const hc_array* H, int nH, const double* sigma,
                      int mode, double cval, int order_xy, void* scratch, void* stream);

/* frangi from sigmas: filters/ridges.py:322-476 with the per-sigma kernels
 * built as hc_hessian_matrix builds them, then hc_frangi (same arguments
 * otherwise). */
int hc_frangi_sigmas(const hc_array* in, const hc_array* out, const double* sigmas, int nsig,
                     double alpha, double beta, double gamma, int has_gamma, int black_ridges,
                     int mode, double cval, const hc_array* work, int nwork, void* scratch,
                     void* stream);

/* ---- after the filter: image range, histogram, thresholds, labels ---- */

/* Minimum and maximum of a C-contiguous array (any real dtype; NaN
 * propagates): the image.min() / image.max() of exposure/exposure.py:51-63
 * and filters/thresholding.py.  result_dev receives two values of the key
 * type: double for float inputs, int64 for signed integers, uint64 for
 * unsigned ones and bool.  work_dev: >= 16 KiB of device memory. */
int hc_minmax(const hc_array* in, void* result_dev, void* work_dev, int64_t work_bytes,
              void* stream);

/* hc_minmax without a work argument: no counterpart in the reference; for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * serving loops).  The partial results go to a static 16 KiB device array of
 * the library (one per GPU, held until the process ends; its address is
 * looked up once per device on the first call).  Allocates nothing.  Calls
 * that overlap on one GPU race on that array. */
int hc_minmax_noalloc(const hc_array* in, void* result_dev, void* stream);

/* Histogram counts (uint64, nbins, zeroed first) of a C-contiguous array:
 * exposure/exposure.py:66-105, 308-345 (_histogram).  Integer images
 * (edges_dev NULL): bin = value - lo for values in [lo, lo + nbins), the
 * bincount of the reference's _bincount_histogram.  Float images: edges_dev
 * holds nbins + 1 increasing float64 edges; bin i takes edges[i] <= x <
 * edges[i + 1], the last bin closed, values outside and NaN dropped
 * (numpy.histogram's rule, which cupy.histogram follows). */
int hc_histogram(const hc_array* in, int64_t nbins, const double* edges_dev, int64_t lo,
                 uint64_t* counts_dev, void* stream);

/* One step of threshold_li's float iteration (filters/thresholding.py:
 * 760-779): over the finite values v = x - vmin (rounded in x's dtype, the
 * reference's `image -= image_min`), out4_dev = {sum v > t, count v > t,
 * sum v <= t, count v <= t} in float64.  work_dev: >= 32 KiB. */
int hc_li_means(const hc_array* in, double vmin, double t, double* out4_dev, void* work_dev,
                int64_t work_bytes, void* stream);

/* int32 elements of the work buffer hc_label needs for n voxels. */
int64_t hc_label_work_elems(int64_t n);

/* Connected-component labels: measure/_label.py:21-124 (label in greyscale
 * mode, _label_kernels.py).  Voxels whose integer key differs from
 * `background` are foreground; face/edge/corner neighbours within
 * `connectivity` (generate_binary_structure(ndim, connectivity)) with equal
 * keys are connected.  Keys: integer values (uint64 as their bit pattern),
 * floats truncated toward zero (the reference's astype(intp)).  out: int32,
 * the input's shape, both C-contiguous; labels 1..count numbered in raster
 * order of each component's first voxel, background 0.  work_dev:
 * hc_label_work_elems(n) int32; count_dev (optional): the component count. */
int hc_label(const hc_array* in, int64_t background, int connectivity, const hc_array* out,
             int32_t* work_dev, int64_t work_elems, int32_t* count_dev, void* stream);

/* ---- before the filters: whole-slide image tiles ---- */

/* An open TIFF / BigTIFF (pyramid) file: the reference's CuImage over the
 * cuslide plugin (cpp/include/cucim/cuimage.h:168, cpp/src/cuimage.cpp,
 * cpp/plugins/cucim.kit.cuslide/src/cuslide/tiff/tiff.cpp, ifd.cpp).
 * Levels: IFD 0 and every following tiled IFD smaller than the level before
 * it.  Codecs: none, LZW, deflate, baseline JPEG (TIFF JPEGTables
 * supported); 8-bit interleaved samples; horizontal predictor. */
typedef struct hc_wsi hc_wsi;
int hc_wsi_open(const char* path, hc_wsi** out);
void hc_wsi_close(hc_wsi* h);
int hc_wsi_level_count(const hc_wsi* h);
/* info[8]: width, height, tile width, tile height (strip: width x rows per
 * strip), output samples per pixel (JPEG: 3 or 1), compression, photometric,
 * tiled flag. */
int hc_wsi_level_info(const hc_wsi* h, int level, int64_t* info);
/* res[3]: x resolution, y resolution, resolution unit (TIFF tags 282, 283,
 * 296 of IFD 0); background: the fill value of pixels outside the image
 * (255 for Aperio / Philips files, else 0, ifd.cpp:605); desc: the image
 * description of IFD 0; vendor: "aperio", "philips" or "generic".  Any
 * pointer may be NULL. */
int hc_wsi_properties(const hc_wsi* h, double* res, int* background, char* desc, int desc_len,
                      char* vendor, int vendor_len);
/* read_region (cuimage.h:168, ifd.cpp:576-1260): the region [x, x + w) x
 * [y, y + h) of a level (location may be negative or reach past the edge;
 * outside pixels take the background value) as h x w x channels uint8 into
 * host memory, tiles decoded on num_workers host threads (0 or 1: the
 * calling thread). */
int hc_wsi_read_region(const hc_wsi* h, int level, int64_t x, int64_t y, int64_t w, int64_t hgt,
                       uint8_t* out, int num_workers);
/* Associated images of Aperio files (thumbnail, label, macro: untiled IFDs
 * named by their position and subfile type, tiff.cpp:670-700;
 * CuImage::associated_images / associated_image, cuimage.h).  info[4]: width,
 * height, output samples per pixel, compression.  read: the whole image as
 * height x width x samples uint8 into host memory (strips decoded on
 * num_workers host threads). */
int hc_wsi_associated_count(const hc_wsi* h);
int hc_wsi_associated_info(const hc_wsi* h, int k, char* name, int name_len, int64_t* info);
int hc_wsi_associated_read(const hc_wsi* h, int k, uint8_t* out, int num_workers);
/* The same region into device memory: JPEG tiles are uploaded compressed
 * and decoded on the GPU (entropy decoding one work-group per tile from LDS,
 * integer inverse DCT, fancy chroma upsampling and YCbCr -> RGB in device
 * kernels, bit-exact to the host path and to libjpeg-turbo); lossless tiles
 * are decoded on host threads and uploaded.  Returns after the region is
 * complete on the stream. */
int hc_wsi_read_region_device(const hc_wsi* h, int level, int64_t x, int64_t y, int64_t w,
                              int64_t hgt, uint8_t* out_dev, int num_workers, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* CUCIM_AMD_H */
