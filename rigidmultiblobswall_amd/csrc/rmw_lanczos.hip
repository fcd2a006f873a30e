// Lanczos square root: c = scale * sqrt+(T) e_1 for the symmetric
// tridiagonal T built by the Lanczos recurrence (FP64, on the device).
//
// Reference: stochastic_forcing/stochastic_forcing.py:208-227 forms the
// (i+1) x (i+1) tridiagonal h every iteration, takes numpy.linalg.eigh,
// replaces negative eigenvalues by 0 and evaluates
//   noise = V^T Q (|z| factor sqrt(lambda) * Q^T e_1).
// Here the eigen-decomposition is the implicit QL method with Wilkinson
// shifts (Bowdler, Martin, Reinsch & Wilkinson, Numer. Math. 11 (1968) 293,
// the algorithm behind EISPACK tql2), run by one wavefront per right-hand
// side: the scalar recurrence is evaluated redundantly by all 64 lanes
// (uniform), each plane rotation is applied to the eigenvector columns
// with the lanes spread over rows.  The kernel returns the coefficient
// vector c (noise = V c), not the noise itself.
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cmath>

#include "rmw_internal.h"

namespace rmw {

constexpr int kMaxLanczos = 1024;

// H of right-hand side r starts at H + r * h_rhs_stride; Lanczos step i wrote
// alpha_i = H[i * ldh + i] and beta_i = H[i * ldh + i + 1].
__global__ __launch_bounds__(64) void tridiag_sqrt_e1_kernel(const double* __restrict__ H, int64_t h_rhs_stride,
                                                             int64_t ldh, int m, const double* __restrict__ scale,
                                                             double* __restrict__ c_out, int64_t c_stride,
                                                             double* __restrict__ zscratch, int* __restrict__ status)
{
  __shared__ double d[kMaxLanczos];
  __shared__ double e[kMaxLanczos];
  const int r = blockIdx.x;
  const int lane = threadIdx.x;
  const double* Hr = H + r * h_rhs_stride;
  double* Z = zscratch + (int64_t)r * m * m;  // column-major eigenvectors
  for (int i = lane; i < m; i += 64) {
    d[i] = Hr[(int64_t)i * ldh + i];
    e[i] = (i + 1 < m) ? Hr[(int64_t)i * ldh + i + 1] : 0.0;
  }
  for (int64_t idx = lane; idx < (int64_t)m * m; idx += 64) Z[idx] = (idx % (m + 1) == 0) ? 1.0 : 0.0;
  __syncthreads();
  int fail = 0;
  for (int l = 0; l < m; ++l) {
    for (int iter = 0;; ++iter) {
      int mm = l;
      for (; mm < m - 1; ++mm) {
        const double dd = fabs(d[mm]) + fabs(d[mm + 1]);
        if (fabs(e[mm]) <= DBL_EPSILON * dd) break;
      }
      if (mm == l) break;
      if (iter == 60) { fail = 1; break; }
      double g = (d[l + 1] - d[l]) / (2.0 * e[l]);
      double rr = hypot(g, 1.0);
      g = d[mm] - d[l] + e[l] / (g + copysign(rr, g));
      double s = 1.0, c = 1.0, p = 0.0;
      bool deflated = false;
      for (int i = mm - 1; i >= l; --i) {
        const double f = s * e[i], b = c * e[i];
        rr = hypot(f, g);
        __syncthreads();
        if (lane == 0) e[i + 1] = rr;
        __syncthreads();
        if (rr == 0.0) {
          if (lane == 0) { d[i + 1] -= p; e[mm] = 0.0; }
          __syncthreads();
          deflated = true;
          break;
        }
        s = f / rr;
        c = g / rr;
        g = d[i + 1] - p;
        rr = (d[i] - g) * s + 2.0 * c * b;
        p = s * rr;
        __syncthreads();
        if (lane == 0) d[i + 1] = g + p;
        __syncthreads();
        g = c * rr - b;
        double* zi = Z + (int64_t)i * m;
        double* zi1 = Z + (int64_t)(i + 1) * m;
        for (int row = lane; row < m; row += 64) {
          const double a0 = zi[row], a1 = zi1[row];
          zi1[row] = s * a0 + c * a1;
          zi[row] = c * a0 - s * a1;
        }
      }
      if (deflated) continue;
      __syncthreads();
      if (lane == 0) { d[l] -= p; e[l] = g; e[mm] = 0.0; }
      __syncthreads();
    }
    if (fail) break;
  }
  __syncthreads();
  const double sc = scale[r];
  for (int j = lane; j < m; j += 64) {
    double acc = 0.0;
    for (int k = 0; k < m; ++k) {
      const double lam = d[k];
      const double sq = lam > 0.0 ? sqrt(lam) : 0.0;
      acc += Z[(int64_t)k * m + j] * sq * Z[(int64_t)k * m];
    }
    c_out[r * c_stride + j] = sc * acc;
  }
  if (lane == 0 && status) status[r] = fail;
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_lanczos_scratch_doubles(int m, int k, size_t* doubles)
{
  return guarded([&] {
    RMW_REQUIRE(doubles && m >= 1 && m <= kMaxLanczos && k >= 1, "bad sizes");
    *doubles = (size_t)k * m * m;
  });
}

// c_out[r * c_stride + j] = scale[r] * (sqrt+(T_r) e_1)_j, j < m, for the k
// tridiagonals T_r read from the Lanczos coefficient arrays H (see kernel).
extern "C" int rmw_lanczos_sqrt_coeffs(const double* H, int64_t h_rhs_stride, int64_t ldh, int m, int k,
                                       const double* scale, double* c_out, int64_t c_stride, double* scratch,
                                       size_t scratch_doubles, int* status, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(m >= 1 && m <= kMaxLanczos && k >= 1, "bad sizes");
    RMW_REQUIRE(ldh >= m + 1 && c_stride >= m, "bad strides");
    RMW_REQUIRE(scratch && scratch_doubles >= (size_t)k * m * m, "scratch too small");
    hipLaunchKernelGGL(tridiag_sqrt_e1_kernel, dim3(k), dim3(64), 0, (hipStream_t)stream, H, h_rhs_stride, ldh, m,
                       scale, c_out, c_stride, scratch, status);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
