// Blob-blob soft repulsion: F_i = sum_{j != i} f0(r_ij) (r_j - r_i) over all
// blob pairs (gfx950).
//
// Reference: multi_bodies/forces_numba.py:12-56 (blob_blob_force_numba),
// the different-radii form :75-122 (a = (a_i + a_j) / 2) and the
// neighbour-list form :141-201 (pairs with r <= 2a + 30b only), the same
// force as forces.cpp:11-42 and forces_pycuda.py:24-110, selected by
// set_blob_blob_forces (multi_bodies/multi_bodies_functions.py:249-279).
// The force is minus the gradient of U(r) = eps (1 + (2a - r)/b) for
// r < 2a and eps exp(-(r - 2a)/b) beyond:
//   f0 = -(eps/b) exp(-(r - 2a)/b) / r   (r > 2a),   -(eps/b) / r   (r <= 2a),
// dr = r_j - r_i taken to its minimum image in each periodic direction.
//
// Two launches: the pair kernel (256 target blobs per workgroup, one per
// lane; sources streamed through LDS as FP64 structure-of-arrays; each work
// item one chunk of sources; FP64 partial force per (chunk, blob)) and a
// fixed-order chunk reduce.  A blob and its coincident partners (r = 0)
// contribute nothing, as in the reference (f0 finite, dr = 0), so the self
// pair needs no test.
//
// Precision: the geometry (differences, r, the exponent) and every sum are
// FP64 in both modes.  RMW_PREC_F64 evaluates sqrt, the quotient and exp in
// FP64; the other modes take 1/r from the hardware FP64 reciprocal square
// root refined by one Newton step (relative error ~1e-13) and the
// exponential from the hardware FP32 exp2 of the FP64 exponent rounded
// once, which leaves a relative error of ~1e-7 t e^-t per pair with
// t = (r - 2a)/b.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <string>

#include "rmw_accum.h"
#include "rmw_internal.h"

namespace rmw {

constexpr int kFBlock = 256;

struct ForceParams {
  const double* r;
  const double* rad;      // per-blob radius, or null: one radius a
  double a;
  int64_t n;
  double eb;              // eps / b
  double invb;            // 1 / b
  double t2;              // log2(e) / b: exponent scale of the FP32 path
  double cutoff;          // pairs with r > cutoff contribute nothing (+inf: none)
  int periodic;           // bit d set when L_d > 0
  double L[3];
  double* partial;        // [chunk][n][3]
  int64_t n_chunks, chunk_len, total_items;
};

// f0 of a pair at squared distance r2 with 2a = aa2 (0 for r2 = 0).
template <bool F64>
__device__ __forceinline__ double pair_f0(const ForceParams& p, double r2, double aa2)
{
  double r, inv, e;
  if constexpr (F64) {
    r = sqrt(r2);
    inv = 1.0 / r;
    e = exp(-(r - aa2) * p.invb);
  } else {
    const double y0 = __builtin_amdgcn_rsq(r2);
    inv = y0 * fma(-0.5 * r2 * y0, y0, 1.5);
    r = r2 * inv;
    e = (double)__builtin_amdgcn_exp2f((float)((aa2 - r) * p.t2));
  }
  const double f0 = -p.eb * inv * (r > aa2 ? e : 1.0);
  return (r2 > 0.0 && !(r > p.cutoff)) ? f0 : 0.0;
}

template <bool F64>
__global__ __launch_bounds__(kFBlock) void force_kernel(ForceParams p)
{
  __shared__ double s_x[kFBlock], s_y[kFBlock], s_z[kFBlock], s_a[kFBlock];
  int pbox[3];
  for (int d = 0; d < 3; ++d) pbox[d] = (p.periodic >> d) & 1;
  const int il = threadIdx.x;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const int64_t tt = item / p.n_chunks;
    const int64_t ch = item - tt * p.n_chunks;
    const int64_t i = tt * kFBlock + il;
    const bool valid = i < p.n;
    const int64_t gi = valid ? i : 0;
    const double xi = p.r[3 * gi], yi = p.r[3 * gi + 1], zi = p.r[3 * gi + 2];
    const double ai = p.rad ? p.rad[gi] : p.a;
    double fx = 0.0, fy = 0.0, fz = 0.0;

    const int64_t s_begin = ch * p.chunk_len;
    const int64_t s_end = s_begin + p.chunk_len < p.n ? s_begin + p.chunk_len : p.n;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += kFBlock) {
      const int cnt = (int)(s_end - s0 < kFBlock ? s_end - s0 : kFBlock);
      __syncthreads();
      if (il < cnt) {
        const int64_t j = s0 + il;
        s_x[il] = p.r[3 * j];
        s_y[il] = p.r[3 * j + 1];
        s_z[il] = p.r[3 * j + 2];
        s_a[il] = p.rad ? p.rad[j] : p.a;
      }
      __syncthreads();
      if (!p.periodic) {
#pragma unroll 4
        for (int j = 0; j < cnt; ++j) {
          const double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
          const double f0 = pair_f0<F64>(p, fma(dx, dx, fma(dy, dy, dz * dz)), ai + s_a[j]);
          fx = fma(f0, dx, fx);
          fy = fma(f0, dy, fy);
          fz = fma(f0, dz, fz);
        }
      } else {
        for (int j = 0; j < cnt; ++j) {
          double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
          if (pbox[0]) dx -= image_count(dx, p.L[0]) * p.L[0];
          if (pbox[1]) dy -= image_count(dy, p.L[1]) * p.L[1];
          if (pbox[2]) dz -= image_count(dz, p.L[2]) * p.L[2];
          const double f0 = pair_f0<F64>(p, fma(dx, dx, fma(dy, dy, dz * dz)), ai + s_a[j]);
          fx = fma(f0, dx, fx);
          fy = fma(f0, dy, fy);
          fz = fma(f0, dz, fz);
        }
      }
    }
    if (valid) {
      double* P = p.partial + (ch * p.n + i) * 3;
      P[0] = fx; P[1] = fy; P[2] = fz;
    }
  }
}

__global__ __launch_bounds__(256) void force_reduce_kernel(const double* __restrict__ partial, int64_t n_chunks,
                                                           int64_t n, double* __restrict__ out)
{
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= n * 3) return;
  double s0 = 0.0, s1 = 0.0;
  int64_t ch = 0;
  for (; ch + 2 <= n_chunks; ch += 2) {
    s0 += partial[ch * n * 3 + tid];
    s1 += partial[(ch + 1) * n * 3 + tid];
  }
  if (ch < n_chunks) s0 += partial[ch * n * 3 + tid];
  out[tid] = s0 + s1;
}

struct ForcePlan {
  int64_t chunk_len = 0, n_chunks = 0, n_tt = 0, total_items = 0;
  size_t bytes = 0;
};

static ForcePlan force_plan(int prec, int64_t n, const rmw_tuning* tuning)
{
  RMW_REQUIRE(prec == RMW_PREC_F32_F64ACC || prec == RMW_PREC_F32_KAHAN || prec == RMW_PREC_F64 ||
                  prec == RMW_PREC_F32_BLOCKED,
              "unknown precision mode");
  RMW_REQUIRE(n >= 0, "n must be >= 0");
  ForcePlan pl;
  int64_t chunk = tuning && tuning->chunk_len > 0 ? tuning->chunk_len : 0;
  if (chunk <= 0) {
    const double c = (double)n * (double)n / (16.0 * std::max(device_props().cu_count, 1) * kFBlock);
    chunk = std::max<int64_t>(kFBlock, (int64_t)std::ceil(c / 64.0) * 64);
  }
  chunk = std::max<int64_t>(64, (chunk + 63) / 64 * 64);
  if (n > 0) chunk = std::min<int64_t>(chunk, (n + 63) / 64 * 64);
  pl.chunk_len = chunk;
  pl.n_chunks = n > 0 ? (n + chunk - 1) / chunk : 0;
  pl.n_tt = (n + kFBlock - 1) / kFBlock;
  pl.total_items = pl.n_tt * pl.n_chunks;
  pl.bytes = align_up((size_t)pl.n_chunks * n * 3 * sizeof(double));
  return pl;
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_blob_forces_workspace_size(int prec, int64_t n, const rmw_tuning* tuning, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr, "bytes is null");
    *bytes = force_plan(prec, n, tuning).bytes;
  });
}

extern "C" int rmw_blob_forces(int prec, const double* r_vectors, const double* radius_blobs, double* force,
                               int64_t n, double repulsion_strength, double debye_length, double blob_radius,
                               double cutoff, const double* L, const rmw_tuning* tuning, void* workspace,
                               size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    const ForcePlan pl = force_plan(prec, n, tuning);
    RMW_REQUIRE(workspace_bytes >= pl.bytes, "workspace too small: need " + std::to_string(pl.bytes) + " bytes");
    RMW_REQUIRE(debye_length > 0.0 && std::isfinite(debye_length), "debye_length must be positive");
    RMW_REQUIRE(std::isfinite(repulsion_strength), "repulsion_strength must be finite");
    RMW_REQUIRE(radius_blobs != nullptr || (blob_radius >= 0.0 && std::isfinite(blob_radius)),
                "blob_radius must be >= 0 when no per-blob radii are given");
    RMW_REQUIRE(cutoff >= 0.0, "cutoff must be >= 0 (0: no cutoff)");
    if (n == 0) return;
    RMW_REQUIRE(r_vectors && force && workspace, "r_vectors, force and workspace are required");
    ForceParams q{};
    q.r = r_vectors;
    q.rad = radius_blobs;
    q.a = radius_blobs ? 0.0 : blob_radius;
    q.n = n;
    q.eb = repulsion_strength / debye_length;
    q.invb = 1.0 / debye_length;
    q.t2 = M_LOG2E / debye_length;
    q.cutoff = cutoff > 0.0 ? cutoff : INFINITY;
    for (int d = 0; d < 3; ++d) {
      const double Ld = L ? L[d] : 0.0;
      q.L[d] = Ld > 0.0 ? Ld : 0.0;
      if (Ld > 0.0) q.periodic |= 1 << d;
    }
    q.partial = (double*)workspace;
    q.n_chunks = pl.n_chunks;
    q.chunk_len = pl.chunk_len;
    q.total_items = pl.total_items;
    int grid = (int)std::min<int64_t>(pl.total_items, (int64_t)INT32_MAX);
    if (tuning && tuning->max_wgs > 0) grid = std::min(grid, tuning->max_wgs);
    hipStream_t st = (hipStream_t)stream;
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st);
    if (prec == RMW_PREC_F64)
      hipLaunchKernelGGL((force_kernel<true>), dim3(grid), dim3(kFBlock), 0, st, q);
    else
      hipLaunchKernelGGL((force_kernel<false>), dim3(grid), dim3(kFBlock), 0, st, q);
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);
    hipLaunchKernelGGL(force_reduce_kernel, dim3((unsigned)((n * 3 + 255) / 256)), dim3(256), 0, st, q.partial,
                       pl.n_chunks, n, force);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
