// Blobs of different radii: u = B_t M(x, y) B_s f from source blobs to
// target blobs, with or without the wall (gfx950).
//
// Reference: mobility_numba.py:1479-1658 (mobility_trans_times_force_
// source_target_numba) behind the wrappers mobility.py:551-591 (wall:
// heights shifted to max(z, radius), damping B = z/radius below the
// radius, shift_heights_different_radius :87-99 and damping_matrix_B_
// different_radius :102-119) and :593-615 (unbounded, no wrapper), as
// selected by the 'radii*' implementations of set_mobility_vector_prod
// (multi_bodies/multi_bodies.py:266-287).
//
// Two launches: the pair kernel (256 targets per workgroup in registers,
// sources streamed through LDS as structure-of-arrays, each work item one
// chunk of sources, FP64 partial sums per (chunk, target)) and a reduce
// (fixed-order chunk sum, target damping).  The self pair needs no special
// case (it is the inner RPY regime) and sources and targets may differ.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "rmw_accum.h"
#include "rmw_internal.h"
#include "rmw_pair_terms.h"

namespace rmw {

constexpr int kRBlock = 256;

struct RadiiParams {
  const double* src;
  const double* trg;
  const double* f;
  const double* rs;
  const double* rt;
  int64_t ns, nt;
  double norm;            // 1 / (8 pi eta)
  int wall;
  int periodic;           // bit d set when L_d > 0
  double L[3];
  double* partial;        // [chunk][nt][3]
  int64_t n_chunks, chunk_len, total_items;
};

// Height and damping of a blob of radius rad: z_eff = max(z, rad) and
// B = z / rad below the radius (wall only).
__device__ __forceinline__ void wall_shift(double z, double rad, int wall, double& ze, double& B)
{
  ze = z;
  B = 1.0;
  if (wall && !(z > rad)) {
    B = z < rad ? z / rad : 1.0;
    ze = rad;
  }
}

template <int ACC, typename T>
__global__ __launch_bounds__(kRBlock) void radii_kernel(RadiiParams p)
{
  __shared__ double s_x[kRBlock], s_y[kRBlock], s_z[kRBlock];
  __shared__ T s_b[kRBlock], s_ib[kRBlock], s_zr[kRBlock];
  __shared__ T s_f[3][kRBlock];
  int pbox[3];
  for (int d = 0; d < 3; ++d) pbox[d] = (p.periodic >> d) & 1;
  const int il = threadIdx.x;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const int64_t tt = item / p.n_chunks;
    const int64_t ch = item - tt * p.n_chunks;
    const int64_t i = tt * kRBlock + il;
    const bool valid = i < p.nt;
    const int64_t gi = valid ? i : 0;
    const double xi = p.trg[3 * gi], yi = p.trg[3 * gi + 1];
    const double a = p.rt[gi];
    double zi, Bi;
    wall_shift(p.trg[3 * gi + 2], a, p.wall, zi, Bi);
    const T ar = (T)a, inva = (T)(1.0 / a), zir = (T)zi;
    Acc<ACC> u[3];

    const int64_t s_begin = ch * p.chunk_len;
    const int64_t s_end = s_begin + p.chunk_len < p.ns ? s_begin + p.chunk_len : p.ns;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += kRBlock) {
      const int cnt = (int)(s_end - s0 < kRBlock ? s_end - s0 : kRBlock);
      __syncthreads();
      if (il < cnt) {
        const int64_t j = s0 + il;
        const double b = p.rs[j];
        double zj, Bj;
        wall_shift(p.src[3 * j + 2], b, p.wall, zj, Bj);
        s_x[il] = p.src[3 * j];
        s_y[il] = p.src[3 * j + 1];
        s_z[il] = zj;
        s_zr[il] = (T)zj;
        s_b[il] = (T)b;
        s_ib[il] = (T)(1.0 / b);
        const double c = p.norm * Bj;
        s_f[0][il] = (T)(c * p.f[3 * j]);
        s_f[1][il] = (T)(c * p.f[3 * j + 1]);
        s_f[2][il] = (T)(c * p.f[3 * j + 2]);
      }
      __syncthreads();
      for (int j = 0; j < cnt; ++j) {
        double ddx = xi - s_x[j], ddy = yi - s_y[j], ddz = zi - s_z[j];
        const T zj = s_zr[j];
        const T Rz = zir + zj;
        if (!p.periodic) {
          const T dx = (T)ddx, dy = (T)ddy, dz = (T)ddz;
          const CoefRadii<T> c = coef_radii<T>(dx, dy, dz, Rz, zir, zj, ar, s_b[j], inva, s_ib[j], p.wall != 0);
          const Vec3<T> v = apply_radii<T>(dx, dy, dz, Rz, c, s_f[0][j], s_f[1][j], s_f[2][j]);
          u[0].add(v.x);
          u[1].add(v.y);
          u[2].add(v.z);
          if ((j & (kSub - 1)) == kSub - 1) {
            u[0].flush(); u[1].flush(); u[2].flush();
          }
        } else {
          // minimum image of the FP64 difference, then the neighbouring boxes
          if (pbox[0]) ddx -= image_count(ddx, p.L[0]) * p.L[0];
          if (pbox[1]) ddy -= image_count(ddy, p.L[1]) * p.L[1];
          if (pbox[2]) ddz -= image_count(ddz, p.L[2]) * p.L[2];
          for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
            for (int by = -pbox[1]; by <= pbox[1]; ++by)
              for (int bz = -pbox[2]; bz <= pbox[2]; ++bz) {
                const T dx = (T)(ddx + bx * p.L[0]), dy = (T)(ddy + by * p.L[1]), dz = (T)(ddz + bz * p.L[2]);
                const CoefRadii<T> c =
                    coef_radii<T>(dx, dy, dz, Rz, zir, zj, ar, s_b[j], inva, s_ib[j], p.wall != 0);
                const Vec3<T> v = apply_radii<T>(dx, dy, dz, Rz, c, s_f[0][j], s_f[1][j], s_f[2][j]);
                u[0].add(v.x);
                u[1].add(v.y);
                u[2].add(v.z);
              }
          u[0].flush(); u[1].flush(); u[2].flush();
        }
      }
      u[0].flush(); u[1].flush(); u[2].flush();
    }
    if (valid) {
      double* P = p.partial + (ch * p.nt + i) * 3;
      P[0] = u[0].get(); P[1] = u[1].get(); P[2] = u[2].get();
    }
  }
}

__global__ __launch_bounds__(256) void radii_reduce_kernel(const double* __restrict__ partial, int64_t n_chunks,
                                                           int64_t nt, const double* __restrict__ trg,
                                                           const double* __restrict__ rt, int wall,
                                                           double* __restrict__ out)
{
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (tid >= nt * 3) return;
  const int64_t i = tid / 3;
  double s0 = 0.0, s1 = 0.0;
  int64_t ch = 0;
  for (; ch + 2 <= n_chunks; ch += 2) {
    s0 += partial[ch * nt * 3 + tid];
    s1 += partial[(ch + 1) * nt * 3 + tid];
  }
  if (ch < n_chunks) s0 += partial[ch * nt * 3 + tid];
  double ze, B;
  wall_shift(trg[3 * i + 2], rt[i], wall, ze, B);
  out[tid] = B * (s0 + s1);
}

__global__ void radii_zero_kernel(double* __restrict__ out, int64_t n)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = 0.0;
}

struct RadiiPlan {
  int64_t chunk_len = 0, n_chunks = 0, n_tt = 0, total_items = 0;
  size_t bytes = 0;
};

static RadiiPlan radii_plan(int prec, int64_t ns, int64_t nt, const rmw_tuning* tuning)
{
  RMW_REQUIRE(prec == RMW_PREC_F32_F64ACC || prec == RMW_PREC_F32_KAHAN || prec == RMW_PREC_F64 ||
                  prec == RMW_PREC_F32_BLOCKED,
              "unknown precision mode");
  RMW_REQUIRE(ns >= 0 && nt >= 0, "sizes must be >= 0");
  RadiiPlan pl;
  int64_t chunk = tuning && tuning->chunk_len > 0 ? tuning->chunk_len : 0;
  if (chunk <= 0) {
    const double c = (double)ns * (double)nt / (16.0 * std::max(device_props().cu_count, 1) * kRBlock);
    chunk = std::max<int64_t>(kRBlock, (int64_t)std::ceil(c / 64.0) * 64);
  }
  chunk = std::max<int64_t>(64, (chunk + 63) / 64 * 64);
  if (ns > 0) chunk = std::min<int64_t>(chunk, (ns + 63) / 64 * 64);
  pl.chunk_len = chunk;
  pl.n_chunks = ns > 0 ? (ns + chunk - 1) / chunk : 0;
  pl.n_tt = (nt + kRBlock - 1) / kRBlock;
  pl.total_items = pl.n_tt * pl.n_chunks;
  pl.bytes = align_up((size_t)pl.n_chunks * nt * 3 * sizeof(double));
  return pl;
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_radii_workspace_size(int prec, int64_t ns, int64_t nt, const rmw_tuning* tuning, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr, "bytes is null");
    *bytes = radii_plan(prec, ns, nt, tuning).bytes;
  });
}

extern "C" int rmw_radii_apply(int prec, int wall, const double* source, const double* target, const double* force,
                               const double* radius_source, const double* radius_target, double* velocity,
                               int64_t ns, int64_t nt, double eta, const double* L, const rmw_tuning* tuning,
                               void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    const RadiiPlan pl = radii_plan(prec, ns, nt, tuning);
    RMW_REQUIRE(workspace_bytes >= pl.bytes, "workspace too small: need " + std::to_string(pl.bytes) + " bytes");
    RMW_REQUIRE(eta > 0.0 && std::isfinite(eta), "eta must be positive");
    if (nt == 0) return;
    RMW_REQUIRE(target && radius_target && velocity, "target, radius_target and velocity are required");
    hipStream_t st = (hipStream_t)stream;
    if (ns == 0) {
      // a fill kernel rather than a memset, so the call can be graph-captured
      hipLaunchKernelGGL(radii_zero_kernel, dim3((unsigned)((nt * 3 + 255) / 256)), dim3(256), 0, st, velocity,
                         nt * 3);
      RMW_CHECK_HIP(hipGetLastError());
      return;
    }
    RMW_REQUIRE(source && force && radius_source && workspace, "source, force, radius_source and workspace are required");
    RadiiParams q{};
    q.src = source; q.trg = target; q.f = force; q.rs = radius_source; q.rt = radius_target;
    q.ns = ns; q.nt = nt;
    q.norm = 1.0 / (8.0 * M_PI * eta);
    q.wall = wall != 0;
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
    const dim3 g(grid), b(kRBlock);
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st, 1);
    if (prec == RMW_PREC_F32_F64ACC) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_F64ACC, float>), g, b, 0, st, q);
    else if (prec == RMW_PREC_F32_BLOCKED) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_BLOCKED, float>), g, b, 0, st, q);
    else if (prec == RMW_PREC_F32_KAHAN) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_KAHAN, float>), g, b, 0, st, q);
    else hipLaunchKernelGGL((radii_kernel<RMW_PREC_F64, double>), g, b, 0, st, q);
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);
    hipLaunchKernelGGL(radii_reduce_kernel, dim3((unsigned)((nt * 3 + 255) / 256)), dim3(256), 0, st, q.partial,
                       pl.n_chunks, nt, target, radius_target, q.wall, velocity);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
