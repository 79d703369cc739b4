// Eigenvalues of small symmetric matrices, one per voxel (gfx950): the
// n-dimensional branch of feature/corner.py:699-755
// (_symmetric_compute_eigenvalues), which stacks the n(n+1)/2 element images
// into (..., n, n) matrices (_symmetric_image, :758-783) and calls eigvalsh.
//
// Each work-item loads its voxel's matrix into registers (double, whatever
// the element dtype), diagonalises it with the cyclic Jacobi method (plane
// rotations chosen by Rutishauser's stable formulas; sweeps over all (p, q)
// pairs until the off-diagonal mass is below 1e-32 of the matrix's squared
// Frobenius norm, at most kJacobiSweeps sweeps), and writes the eigenvalues
// ascending as eigvalsh returns them, then optionally re-ordered by
// magnitude (a stable sort, as argsort of |lambda| over a stable sort) and/or
// reversed for descending order.  Matrices of order 1..8 (the library's
// largest rank); the 2x2 and 3x3 cases have their own closed-form kernel.
#include <algorithm>
#include <cmath>

#include "hc_common.h"

namespace hc {

constexpr int kJacobiSweeps = 60;
constexpr int kEigMaxN = 8;

struct EigNdArgs {
  const void* S[kEigMaxN * (kEigMaxN + 1) / 2];  // combinations_with_replacement order
  void* E[kEigMaxN];                             // eigenvalue images
  int64_t n;                                     // voxels
};

template <int N>
__device__ __forceinline__ void jacobi_eigvals(double (&A)[N][N]) {
  for (int sweep = 0; sweep < kJacobiSweeps; ++sweep) {
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int p = 0; p < N; ++p) {
      tot += A[p][p] * A[p][p];
#pragma unroll
      for (int q = p + 1; q < N; ++q) off += A[p][q] * A[p][q];
    }
    tot += 2.0 * off;
    // converged, exactly diagonal, or not finite (NaN / Inf entries)
    if (!(off > 1e-32 * tot)) break;
#pragma unroll
    for (int p = 0; p < N - 1; ++p) {
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = A[p][q];
        if (apq == 0.0) continue;
        const double theta = (A[q][q] - A[p][p]) / (2.0 * apq);
        // t = tan of the rotation angle, the root of t^2 + 2 theta t - 1 = 0
        // of smaller magnitude
        double t;
        if (fabs(theta) > 1e150) {
          t = 0.5 / theta;
        } else {
          t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c, tau = s / (1.0 + c);
        A[p][p] -= t * apq;
        A[q][q] += t * apq;
        A[p][q] = A[q][p] = 0.0;
#pragma unroll
        for (int r = 0; r < N; ++r) {
          if (r == p || r == q) continue;
          const double g = A[r][p], h = A[r][q];
          const double gp = g - s * (h + g * tau), hq = h + s * (g - h * tau);
          A[r][p] = A[p][r] = gp;
          A[r][q] = A[q][r] = hq;
        }
      }
    }
  }
}

template <int N, typename T>
__global__ __launch_bounds__(256) void k_sym_eig_nd(const EigNdArgs a, int descending,
                                                    int abs_sort) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < a.n;
       i += (int64_t)gridDim.x * 256) {
    double A[N][N];
    int k = 0;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = r; c < N; ++c) {
        const double v = (double)reinterpret_cast<const T*>(a.S[k++])[i];
        A[r][c] = v;
        A[c][r] = v;
      }
    jacobi_eigvals<N>(A);
    double ev[N];
#pragma unroll
    for (int r = 0; r < N; ++r) ev[r] = A[r][r];
    // ascending (insertion sort; NaN compares false and stays in place)
#pragma unroll
    for (int r = 1; r < N; ++r)
#pragma unroll
      for (int j = r; j > 0; --j)
        if (ev[j] < ev[j - 1]) {
          const double tmp = ev[j];
          ev[j] = ev[j - 1];
          ev[j - 1] = tmp;
        }
    if (abs_sort) {  // stable: equal magnitudes keep the ascending order
#pragma unroll
      for (int r = 1; r < N; ++r)
#pragma unroll
        for (int j = r; j > 0; --j)
          if (fabs(ev[j]) < fabs(ev[j - 1])) {
            const double tmp = ev[j];
            ev[j] = ev[j - 1];
            ev[j - 1] = tmp;
          }
    }
#pragma unroll
    for (int r = 0; r < N; ++r)
      reinterpret_cast<T*>(a.E[r])[i] = (T)ev[descending ? N - 1 - r : r];
  }
}

template <typename T>
int launch_eig_nd(const EigNdArgs& a, int N, int descending, int abs_sort, hipStream_t st) {
  const dim3 g((unsigned)std::max<int64_t>(1, std::min<int64_t>((a.n + 255) / 256, 256 * 64)));
  switch (N) {
    case 1: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<1, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 2: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<2, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 3: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<3, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 4: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<4, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 5: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<5, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 6: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<6, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 7: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<7, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    case 8: HC_LAUNCH("k_sym_eig_nd", st, (k_sym_eig_nd<8, T>), g, dim3(256), 0, a, descending, abs_sort); break;
    default: return fail(HC_ERR_VALUE, "matrix order must be 1..8");
  }
  HC_HIP(hipGetLastError());
  return HC_OK;
}

// n(n+1)/2 element images -> n eigenvalue images (checked by the caller:
// same shape, dtype float32/float64, C-contiguous).
int sym_eigvals_nd(const hc_array* S, int nS, const hc_array* eig, int sort, int abs_sort,
                   hipStream_t st) {
  int N = 0;
  while (N <= kEigMaxN && N * (N + 1) / 2 < nS) ++N;
  if (N < 1 || N > kEigMaxN || N * (N + 1) / 2 != nS)
    return fail(HC_ERR_UNSUPPORTED, "element count is not n(n+1)/2 for a matrix order 1..8");
  EigNdArgs a;
  for (int k = 0; k < nS; ++k) a.S[k] = S[k].data;
  for (int k = 0; k < N; ++k) a.E[k] = eig[k].data;
  a.n = numel(&S[0]);
  if (a.n == 0) return HC_OK;
  return S[0].dtype == HC_FLOAT64 ? launch_eig_nd<double>(a, N, sort, abs_sort, st)
                                  : launch_eig_nd<float>(a, N, sort, abs_sort, st);
}

}  // namespace hc
