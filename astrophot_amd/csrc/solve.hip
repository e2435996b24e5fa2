// Damped Levenberg-Marquardt step, float64 (fit/lm.py:359-371):
//   A(L)_ii = H_ii + L (1 + H_ii),  A(L)_ij = H_ij / (1 + L),  h = A(L)^-1 g.
// A(L) is symmetric positive definite for L > 0 and H = J^T W J, so the solve
// is a Cholesky factorisation in LDS on one workgroup.  If a pivot is not
// positive (H numerically indefinite) the same workgroup falls back to
// Gaussian elimination with partial pivoting, the algorithm behind the
// reference's torch.linalg.solve; status reports which path ran.
#include <hip/hip_runtime.h>

#include "capi_internal.h"

namespace apx {

namespace {

constexpr int kMaxSolve = 64;

__global__ __launch_bounds__(256) void lm_solve_kernel(const double* hess, const double* grad,
                                                       int n, double L, double* h,
                                                       int32_t* status) {
  __shared__ double A[kMaxSolve][kMaxSolve + 1];
  __shared__ double b[kMaxSolve];
  __shared__ int fail;
  const int t = threadIdx.x;
  auto load = [&]() {
    for (int e = t; e < n * n; e += 256) {
      const int i = e / n, j = e % n;
      const double v = hess[e];
      A[i][j] = (i == j) ? v + L * (1.0 + v) : v / (1.0 + L);
    }
    for (int i = t; i < n; i += 256) b[i] = grad[i];
  };
  load();
  if (t == 0) fail = 0;
  __syncthreads();
  // right-looking Cholesky, lower triangle
  for (int k = 0; k < n; ++k) {
    if (t == 0) {
      const double piv = A[k][k];
      if (!(piv > 0.0) || !isfinite(piv)) fail = k + 1;
      A[k][k] = sqrt(piv);
    }
    __syncthreads();
    if (fail) break;
    for (int i = k + 1 + t; i < n; i += 256) A[i][k] /= A[k][k];
    __syncthreads();
    const int m = n - k - 1;
    for (int e = t; e < m * m; e += 256) {
      const int i = k + 1 + e / m, j = k + 1 + e % m;
      if (j <= i) A[i][j] -= A[i][k] * A[j][k];
    }
    __syncthreads();
  }
  if (!fail) {
    if (t == 0) {
      for (int i = 0; i < n; ++i) {  // L y = b
        double s = b[i];
        for (int j = 0; j < i; ++j) s -= A[i][j] * b[j];
        b[i] = s / A[i][i];
      }
      for (int i = n - 1; i >= 0; --i) {  // L^T x = y
        double s = b[i];
        for (int j = i + 1; j < n; ++j) s -= A[j][i] * b[j];
        b[i] = s / A[i][i];
      }
      status[0] = 0;
    }
    __syncthreads();
    for (int i = t; i < n; i += 256) h[i] = b[i];
    return;
  }
  // fallback: LU with partial pivoting on a fresh copy
  __syncthreads();
  load();
  __syncthreads();
  if (t == 0) {
    int sing = 0;
    for (int k = 0; k < n; ++k) {
      int pr = k;
      double best = fabs(A[k][k]);
      for (int i = k + 1; i < n; ++i)
        if (fabs(A[i][k]) > best) {
          best = fabs(A[i][k]);
          pr = i;
        }
      if (pr != k) {
        for (int j = 0; j < n; ++j) {
          const double tmp = A[k][j];
          A[k][j] = A[pr][j];
          A[pr][j] = tmp;
        }
        const double tb = b[k];
        b[k] = b[pr];
        b[pr] = tb;
      }
      if (A[k][k] == 0.0) sing = 1;
      for (int i = k + 1; i < n; ++i) {
        const double f = A[i][k] / A[k][k];
        for (int j = k; j < n; ++j) A[i][j] -= f * A[k][j];
        b[i] -= f * b[k];
      }
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = b[i];
      for (int j = i + 1; j < n; ++j) s -= A[i][j] * b[j];
      b[i] = s / A[i][i];
    }
    status[0] = sing ? -fail : fail;
  }
  __syncthreads();
  for (int i = t; i < n; i += 256) h[i] = b[i];
}

}  // namespace

int lm_solve(const double* hess, const double* grad, int32_t np, double L, double* h,
             int32_t* status, hipStream_t stream) {
  if (!hess || !grad || !h || !status) return set_error(APX_E_INVALID, "lm_solve: null pointer");
  if (np < 1 || np > kMaxSolve) return set_error(APX_E_LIMIT, "lm_solve: np must be 1..64");
  hipLaunchKernelGGL(lm_solve_kernel, dim3(1), dim3(256), 0, stream, hess, grad, np, L, h, status);
  return check_launch("lm_solve");
}

}  // namespace apx
