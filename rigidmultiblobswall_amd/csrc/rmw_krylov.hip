// FP64 Krylov-space vector kernels for GMRES (Arnoldi with classical
// Gram-Schmidt and one re-orthogonalisation, CGS2) and Lanczos.
//
// Vectors are float64, length n; a basis V is stored as nvec rows of
// stride ldv.  Every reduction is two-pass with a fixed order (block
// partials, then one thread per output summing the partials), so results
// are bit-reproducible run to run.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "rmw_internal.h"

namespace rmw {

constexpr int kKBlock = 256;
constexpr int kMaxDotBlocks = 512;
constexpr int kMaxVec = 1024;

// partial[j * G + g] = sum over the g-th slice of V_j . w  (long vectors:
// each block walks a slice of (n + G - 1) / G elements once per vector)
__global__ __launch_bounds__(kKBlock) void dot_many_partial(const double* __restrict__ V, int64_t ldv, int nvec,
                                                            const double* __restrict__ w, int64_t n,
                                                            double* __restrict__ partial)
{
  __shared__ double red[kKBlock / 64];
  const int G = gridDim.x;
  const int64_t slice = (n + G - 1) / G;
  const int64_t b0 = slice * blockIdx.x;
  const int64_t b1 = b0 + slice < n ? b0 + slice : n;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  for (int j = 0; j < nvec; ++j) {
    const double* v = V + (int64_t)j * ldv;
    double acc = 0.0;
    for (int64_t i = b0 + threadIdx.x; i < b1; i += blockDim.x) acc += v[i] * w[i];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) red[wv] = acc;
    __syncthreads();
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int k = 0; k < (int)(blockDim.x >> 6); ++k) s += red[k];
      partial[(int64_t)j * G + blockIdx.x] = s;
    }
    __syncthreads();
  }
}

// Same partials when every block owns one slice of kKBlock * E elements:
// the slice of w stays in registers across the vectors, each vector's
// wave sums go to LDS and one barrier per group of 64 vectors publishes
// them (fixed summation order, as above).
constexpr int kDotGroup = 64;

template <int E>
__global__ __launch_bounds__(kKBlock) void dot_many_partial_reg(const double* __restrict__ V, int64_t ldv, int nvec,
                                                                const double* __restrict__ w, int64_t n,
                                                                double* __restrict__ partial)
{
  __shared__ double red[kDotGroup][kKBlock / 64];
  const int G = gridDim.x;
  const int64_t base = (int64_t)blockIdx.x * (kKBlock * E) + threadIdx.x;
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  double wr[E];
#pragma unroll
  for (int e = 0; e < E; ++e) {
    const int64_t i = base + (int64_t)e * kKBlock;
    wr[e] = i < n ? w[i] : 0.0;
  }
  for (int j0 = 0; j0 < nvec; j0 += kDotGroup) {
    const int jn = nvec - j0 < kDotGroup ? nvec - j0 : kDotGroup;
    for (int jj = 0; jj < jn; ++jj) {
      const double* v = V + (int64_t)(j0 + jj) * ldv;
      double acc = 0.0;
#pragma unroll
      for (int e = 0; e < E; ++e) {
        const int64_t i = base + (int64_t)e * kKBlock;
        if (i < n) acc = fma(v[i], wr[e], acc);
      }
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
      if (lane == 0) red[jj][wv] = acc;
    }
    __syncthreads();
    if ((int)threadIdx.x < jn) {
      double s = 0.0;
#pragma unroll
      for (int k = 0; k < kKBlock / 64; ++k) s += red[threadIdx.x][k];
      partial[(int64_t)(j0 + threadIdx.x) * G + blockIdx.x] = s;
    }
    __syncthreads();
  }
}

// out[j] = sum_g partial[j * G + g] (sqrt when sqrt_out; + out[j] when
// accumulate): one wavefront per output, lanes stride the partials, then a
// fixed shuffle tree.
__global__ __launch_bounds__(64) void dot_many_final(const double* __restrict__ partial, int G, int nvec,
                                                     double* __restrict__ out, int accumulate, int sqrt_out)
{
  const int j = blockIdx.x;
  if (j >= nvec) return;
  const double* p = partial + (int64_t)j * G;
  double s = 0.0;
  for (int g = threadIdx.x; g < G; g += 64) s += p[g];
  for (int o = 32; o > 0; o >>= 1) s += __shfl_down(s, o, 64);
  if (threadIdx.x == 0) {
    if (sqrt_out) s = sqrt(s);
    out[j] = accumulate ? out[j] + s : s;
  }
}

// w -= sum_j coef[j] V_j     (coef on the device)
__global__ __launch_bounds__(kKBlock) void sub_combination(const double* __restrict__ V, int64_t ldv, int nvec,
                                                           const double* __restrict__ coef,
                                                           double* __restrict__ w, int64_t n)
{
  __shared__ double c[kMaxVec];
  for (int j = threadIdx.x; j < nvec; j += blockDim.x) c[j] = coef[j];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = w[i];
    for (int j = 0; j < nvec; ++j) s -= c[j] * V[(int64_t)j * ldv + i];
    w[i] = s;
  }
}

// y = sum_j coef[j] V_j (+ y when accumulate)
__global__ __launch_bounds__(kKBlock) void combination(const double* __restrict__ V, int64_t ldv, int nvec,
                                                       const double* __restrict__ coef, double* __restrict__ y,
                                                       int64_t n, int accumulate)
{
  __shared__ double c[kMaxVec];
  for (int j = threadIdx.x; j < nvec; j += blockDim.x) c[j] = coef[j];
  __syncthreads();
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    double s = accumulate ? y[i] : 0.0;
    for (int j = 0; j < nvec; ++j) s += c[j] * V[(int64_t)j * ldv + i];
    y[i] = s;
  }
}

// y = alpha x + beta y, alpha optionally 1/(*dev_div) (0 -> unit e_0)
__global__ void axpby_kernel(double alpha, const double* __restrict__ x, double beta, double* __restrict__ y,
                             int64_t n, const double* __restrict__ dev_div)
{
  double al = alpha;
  if (dev_div) {
    const double d = *dev_div;
    al = d > 0.0 ? alpha / d : 0.0;
  }
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const double xv = (dev_div && !(*dev_div > 0.0)) ? (i == 0 ? 1.0 : 0.0) : al * x[i];
    y[i] = beta == 0.0 ? xv : xv + beta * y[i];
  }
}

// h[0..j] = h1 + h2 (CGS2 coefficients)
__global__ void add_vec(const double* __restrict__ a, double* __restrict__ b, int m)
{
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j < m) b[j] += a[j];
}

// Dot-product launch: elements per thread E of the register-slice kernel
// (0: the slice-walking kernel) and the number of blocks G <= kMaxDotBlocks.
struct DotLaunch {
  int E, G;
};

static DotLaunch dot_launch(int64_t n)
{
  const DeviceProps& dp = device_props();
  // aim for about two blocks per CU with the fewest elements per thread
  const int64_t want = std::min<int64_t>(kMaxDotBlocks, 2 * (int64_t)dp.cu_count);
  for (int E : {1, 2, 4, 8}) {
    const int64_t g = (n + (int64_t)kKBlock * E - 1) / ((int64_t)kKBlock * E);
    if (g <= want || (E == 8 && g <= kMaxDotBlocks)) return {E, (int)std::max<int64_t>(1, g)};
  }
  int64_t g = (n + 4 * kKBlock - 1) / (4 * kKBlock);
  return {0, (int)std::max<int64_t>(1, std::min<int64_t>(g, kMaxDotBlocks))};
}

static void launch_dot_partial(const DotLaunch& L, hipStream_t st, const double* V, int64_t ldv, int nvec,
                               const double* w, int64_t n, double* partial)
{
  switch (L.E) {
    case 1: hipLaunchKernelGGL(dot_many_partial_reg<1>, dim3(L.G), dim3(kKBlock), 0, st, V, ldv, nvec, w, n, partial); break;
    case 2: hipLaunchKernelGGL(dot_many_partial_reg<2>, dim3(L.G), dim3(kKBlock), 0, st, V, ldv, nvec, w, n, partial); break;
    case 4: hipLaunchKernelGGL(dot_many_partial_reg<4>, dim3(L.G), dim3(kKBlock), 0, st, V, ldv, nvec, w, n, partial); break;
    case 8: hipLaunchKernelGGL(dot_many_partial_reg<8>, dim3(L.G), dim3(kKBlock), 0, st, V, ldv, nvec, w, n, partial); break;
    default: hipLaunchKernelGGL(dot_many_partial, dim3(L.G), dim3(kKBlock), 0, st, V, ldv, nvec, w, n, partial);
  }
}

static void launch_dot_final(hipStream_t st, const double* partial, int G, int nvec, double* out, int accumulate,
                             int sqrt_out)
{
  hipLaunchKernelGGL(dot_many_final, dim3(nvec), dim3(64), 0, st, partial, G, nvec, out, accumulate, sqrt_out);
}

static int stream_blocks(int64_t n)
{
  const DeviceProps& dp = device_props();
  int64_t g = (n + kKBlock - 1) / kKBlock;
  return (int)std::max<int64_t>(1, std::min<int64_t>(g, 8 * (int64_t)dp.cu_count));
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_krylov_workspace_size(int64_t n, int max_vectors, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes && n >= 0 && max_vectors >= 1 && max_vectors <= kMaxVec, "bad arguments");
    *bytes = align_up((size_t)kMaxDotBlocks * max_vectors * sizeof(double)) + align_up(kMaxVec * sizeof(double));
  });
}

// out[j] = V_j . w for j < nvec (device out); with sqrt_out, out[j] = sqrt(...)
extern "C" int rmw_dot_many(const double* V, int64_t ldv, int nvec, const double* w, int64_t n, double* out,
                            int sqrt_out, void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(nvec >= 1 && nvec <= kMaxVec && n >= 0, "bad sizes");
    const DotLaunch L = dot_launch(n);
    RMW_REQUIRE(workspace_bytes >= (size_t)L.G * nvec * sizeof(double), "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    double* partial = (double*)workspace;
    launch_dot_partial(L, st, V, ldv, nvec, w, n, partial);
    launch_dot_final(st, partial, L.G, nvec, out, 0, sqrt_out);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

// w -= V[0..nvec) coef   (coef on the device)
extern "C" int rmw_sub_combination(const double* V, int64_t ldv, int nvec, const double* coef, double* w,
                                   int64_t n, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(nvec >= 1 && nvec <= kMaxVec, "bad nvec");
    hipLaunchKernelGGL(sub_combination, dim3(stream_blocks(n)), dim3(kKBlock), 0, (hipStream_t)stream, V, ldv, nvec,
                       coef, w, n);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

// y = V[0..nvec) coef (+ y if accumulate)
extern "C" int rmw_combination(const double* V, int64_t ldv, int nvec, const double* coef, double* y, int64_t n,
                               int accumulate, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(nvec >= 1 && nvec <= kMaxVec, "bad nvec");
    hipLaunchKernelGGL(combination, dim3(stream_blocks(n)), dim3(kKBlock), 0, (hipStream_t)stream, V, ldv, nvec,
                       coef, y, n, accumulate);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

// y = alpha x + beta y; when div is non-null alpha is divided by *div
// (device scalar; a zero divisor writes the unit vector e_0 instead).
extern "C" int rmw_axpby(double alpha, const double* x, double beta, double* y, int64_t n, const double* div,
                         void* stream)
{
  return guarded([&] {
    hipLaunchKernelGGL(axpby_kernel, dim3(stream_blocks(n)), dim3(kKBlock), 0, (hipStream_t)stream, alpha, x, beta,
                       y, n, div);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

// Arnoldi step with CGS2: given basis V[0..j] (j+1 vectors, stride ldv) and
// w, orthogonalise w against the basis twice, store the j+1 coefficients in
// h[0..j], ||w|| in h[j+1] and w/||w|| in V[j+1] (e_0 if ||w|| = 0).
// h and hscratch are device arrays of >= j+2 doubles.
extern "C" int rmw_arnoldi_cgs2(double* V, int64_t ldv, int j, double* w, int64_t n, double* h, double* hscratch,
                                void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    const int m = j + 1;
    RMW_REQUIRE(m >= 1 && m + 1 <= kMaxVec, "bad Arnoldi index");
    const DotLaunch L = dot_launch(n);
    RMW_REQUIRE(workspace_bytes >= (size_t)L.G * (m + 1) * sizeof(double), "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    double* partial = (double*)workspace;
    const int sb = stream_blocks(n);
    // pass 1
    launch_dot_partial(L, st, V, ldv, m, w, n, partial);
    launch_dot_final(st, partial, L.G, m, h, 0, 0);
    hipLaunchKernelGGL(sub_combination, dim3(sb), dim3(kKBlock), 0, st, V, ldv, m, h, w, n);
    // pass 2 (re-orthogonalisation)
    launch_dot_partial(L, st, V, ldv, m, w, n, partial);
    launch_dot_final(st, partial, L.G, m, hscratch, 0, 0);
    hipLaunchKernelGGL(sub_combination, dim3(sb), dim3(kKBlock), 0, st, V, ldv, m, hscratch, w, n);
    hipLaunchKernelGGL(add_vec, dim3((m + 63) / 64), dim3(64), 0, st, hscratch, h, m);
    // norm and normalisation
    launch_dot_partial(L, st, w, (int64_t)0, 1, w, n, partial);
    launch_dot_final(st, partial, L.G, 1, h + m, 0, 1);
    hipLaunchKernelGGL(axpby_kernel, dim3(sb), dim3(kKBlock), 0, st, 1.0, w, 0.0, V + (int64_t)m * ldv, n,
                       (const double*)(h + m));
    RMW_CHECK_HIP(hipGetLastError());
  });
}
