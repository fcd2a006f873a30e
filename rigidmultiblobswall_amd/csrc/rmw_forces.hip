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
// Two launches: a pair kernel and a fixed-order reduce.  The force is
// antisymmetric (f0 depends on |dr| and (a_i + a_j) only, and the minimum
// image of -dr is exactly minus that of dr), so the default pair kernel is
// symmetric: each unordered pair of 256-blob tiles is visited once, each
// wave owning 64 target blobs and sweeping 64-source blocks with a rotating
// schedule (at step s lane l takes source (l + s) mod 64) while the
// source-side sum of each source travels with it one lane per step (DPP
// rotation); FP64 partial sums per work item (target side) and per tile
// pair (source side).  The ordered kernel (every target against a chunk of
// sources, one FP64 partial per (chunk, blob)) serves small systems and
// workspaces beyond the symmetric kernel's automatic limit.  Sums carry
// g = e/r; the reduce applies -eps/b.  A blob and its coincident partners
// contribute nothing: r^2 is clamped to 1e-50 (the reference's
// max(r, 1e-25)) and dr = 0.
//
// Precision: the geometry (differences, r, the exponent) and every sum are
// FP64 in both modes.  RMW_PREC_F64 takes 1/r from the FP64 reciprocal
// square root of the pair terms (rsq: <= 1.4e-16 relative) and exp in FP64;
// the other modes take 1/r from the hardware FP64 reciprocal square
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
#include "rmw_pair_terms.h"
#include "rmw_tile_pairs.h"

namespace rmw {

constexpr int kFBlock = 256;
// Largest workspace the automatic choice gives the symmetric kernel.
constexpr size_t kFMutAutoBytes = size_t(2) << 30;

struct ForceParams {
  const double* r;
  const double* rad;      // per-blob radius, or null: one radius a
  double a;
  int64_t n;
  double invb;            // 1 / b
  double t2;              // log2(e) / b: exponent scale of the FP32 path
  double cutoff;          // pairs with r > cutoff contribute nothing
  int periodic;           // bit d set when L_d > 0
  double L[3];
  double* partial;        // ordered: [chunk][n][3]; symmetric: target-side slots
  double* part_j;         // symmetric: source-side slots [tri(S, T)][kFBlock][3]
  int64_t n_chunks, chunk_len, total_items;
  int64_t n_t, chunk_tiles;  // symmetric: tiles, source tiles per item
};

// Per-blob radius term kept in LDS / registers: a (FP64 mode) or a t2.
template <bool F64>
__device__ __forceinline__ double radius_term(const ForceParams& p, int64_t j)
{
  const double a = p.rad ? p.rad[j] : p.a;
  return F64 ? a : a * p.t2;
}

// g = f0 / (-eps/b) of a pair at squared distance r2, with A = a_i + a_j
// (FP64 mode) or (a_i + a_j) t2.  FP32 path: 1/r = y0 (1 + (1 - r2 y0^2)/2)
// from the hardware estimate y0, and e = 2^min(t, 0) with 2^0 = 1 exactly
// for the overlap branch.
template <bool F64, bool CUT>
__device__ __forceinline__ double pair_g(const ForceParams& p, double r2, double A)
{
  r2 = fmax(r2, 1e-50);
  double r, inv, e;
  if constexpr (F64) {
    inv = rsq(r2);
    r = r2 * inv;
    e = r > A ? exp(-(r - A) * p.invb) : 1.0;
  } else {
    const double y0 = __builtin_amdgcn_rsq(r2);
    const double h = fma(-r2 * y0, y0, 1.0);
    inv = fma(y0 * h, 0.5, y0);
    r = r2 * inv;
    const float t = (float)fma(-r, p.t2, A);   // -(r - 2a)/b in units of ln 2
    e = (double)__builtin_amdgcn_exp2f(fminf(t, 0.0f));
  }
  const double g = inv * e;
  if constexpr (CUT) return r > p.cutoff ? 0.0 : g;
  return g;
}

// Minimum image of a difference (multi_bodies/forces_numba.py:40-42).
__device__ __forceinline__ void min_image(const ForceParams& p, double& dx, double& dy, double& dz)
{
  if (p.periodic & 1) dx -= image_count(dx, p.L[0]) * p.L[0];
  if (p.periodic & 2) dy -= image_count(dy, p.L[1]) * p.L[1];
  if (p.periodic & 4) dz -= image_count(dz, p.L[2]) * p.L[2];
}

// Ordered pair kernel: work item = (target tile, source chunk).
template <bool F64, bool CUT>
__global__ __launch_bounds__(kFBlock) void force_kernel(ForceParams p)
{
  __shared__ double s_x[kFBlock], s_y[kFBlock], s_z[kFBlock], s_a[kFBlock];
  const int il = threadIdx.x;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const int64_t tt = item / p.n_chunks;
    const int64_t ch = item - tt * p.n_chunks;
    const int64_t i = tt * kFBlock + il;
    const bool valid = i < p.n;
    const int64_t gi = valid ? i : 0;
    const double xi = p.r[3 * gi], yi = p.r[3 * gi + 1], zi = p.r[3 * gi + 2];
    const double ai = radius_term<F64>(p, gi);
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
        s_a[il] = radius_term<F64>(p, j);
      }
      __syncthreads();
      if (!p.periodic) {
#pragma unroll 4
        for (int j = 0; j < cnt; ++j) {
          const double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
          const double g = pair_g<F64, CUT>(p, fma(dx, dx, fma(dy, dy, dz * dz)), ai + s_a[j]);
          fx = fma(g, dx, fx);
          fy = fma(g, dy, fy);
          fz = fma(g, dz, fz);
        }
      } else {
        for (int j = 0; j < cnt; ++j) {
          double dx = s_x[j] - xi, dy = s_y[j] - yi, dz = s_z[j] - zi;
          min_image(p, dx, dy, dz);
          const double g = pair_g<F64, CUT>(p, fma(dx, dx, fma(dy, dy, dz * dz)), ai + s_a[j]);
          fx = fma(g, dx, fx);
          fy = fma(g, dy, fy);
          fz = fma(g, dz, fz);
        }
      }
    }
    if (valid) {
      double* P = p.partial + (ch * p.n + i) * 3;
      P[0] = fx; P[1] = fy; P[2] = fz;
    }
  }
}

// One step of the symmetric sweep: the pair (target in this lane, source
// at LDS slot js); g is zeroed when !on.
template <bool F64, bool CUT, bool PER, bool MASKED>
__device__ __forceinline__ void mutual_step(const ForceParams& p, const double* s_x, const double* s_y,
                                            const double* s_z, const double* s_a, int js, bool on, double xi,
                                            double yi, double zi, double ai, double& fx, double& fy, double& fz,
                                            double& jx, double& jy, double& jz)
{
  double dx = s_x[js] - xi, dy = s_y[js] - yi, dz = s_z[js] - zi;
  if constexpr (PER) min_image(p, dx, dy, dz);
  double g = pair_g<F64, CUT>(p, fma(dx, dx, fma(dy, dy, dz * dz)), ai + s_a[js]);
  if constexpr (MASKED) g = on ? g : 0.0;
  fx = fma(g, dx, fx);
  fy = fma(g, dy, fy);
  fz = fma(g, dz, fz);
  jx = fma(-g, dx, rol1(jx));
  jy = fma(-g, dy, rol1(jy));
  jz = fma(-g, dz, rol1(jz));
}

// Symmetric pair kernel: work item = (target tile T, run of source tiles
// S >= T); on the diagonal tile only sources above the target are taken.
// Each 64-source block sits twice in LDS ([b][0..63] and [b][64..127]) so
// that lane l reads source (l + s) mod 64 at slot l + s without wrapping.
template <bool F64, bool CUT, bool PER>
__global__ __launch_bounds__(kFBlock) void force_mutual_kernel(ForceParams p)
{
  constexpr int NB = kFBlock / 64;
  __shared__ double s_x[NB][128], s_y[NB][128], s_z[NB][128], s_a[NB][128];
  __shared__ double s_jd[NB][3][64];   // per-wave source-side sums of one block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int il = threadIdx.x;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const int64_t T = pair_target_tile(item, p.n_t, p.chunk_tiles);
    const int64_t S0 = T + (item - items_before(T, p.n_t, p.chunk_tiles)) * p.chunk_tiles;
    const int64_t S1 = S0 + p.chunk_tiles < p.n_t ? S0 + p.chunk_tiles : p.n_t;
    const int64_t i = T * kFBlock + il;
    const bool valid = i < p.n;
    const bool t_full = (T + 1) * kFBlock <= p.n;
    const int64_t gi = valid ? i : 0;
    const double xi = p.r[3 * gi], yi = p.r[3 * gi + 1], zi = p.r[3 * gi + 2];
    const double ai = radius_term<F64>(p, gi);
    double fx = 0.0, fy = 0.0, fz = 0.0;

    for (int64_t S = S0; S < S1; ++S) {
      const int64_t s0 = S * kFBlock;
      const int cnt = (int)(p.n - s0 < kFBlock ? p.n - s0 : kFBlock);
      __syncthreads();  // the previous tile has been consumed
      {
        double x = 0.0, y = 0.0, z = 0.0, a = 0.0;   // finite stand-ins beyond n: masked below
        if (il < cnt) {
          const int64_t j = s0 + il;
          x = p.r[3 * j]; y = p.r[3 * j + 1]; z = p.r[3 * j + 2];
          a = radius_term<F64>(p, j);
        }
        const int bb = il >> 6, q = il & 63;
        s_x[bb][q] = x; s_y[bb][q] = y; s_z[bb][q] = z; s_a[bb][q] = a;
        s_x[bb][q + 64] = x; s_y[bb][q + 64] = y; s_z[bb][q + 64] = z; s_a[bb][q + 64] = a;
      }
      __syncthreads();
      const bool masked = S == T || !t_full || cnt < kFBlock;
      for (int b = 0; b < NB && 64 * b < cnt; ++b) {
        double jx = 0.0, jy = 0.0, jz = 0.0;
        const double *bx = s_x[b] + lane, *by = s_y[b] + lane, *bz = s_z[b] + lane, *ba = s_a[b] + lane;
        if (!masked) {
#pragma unroll 8
          for (int st = 0; st < 64; ++st)
            mutual_step<F64, CUT, PER, false>(p, bx, by, bz, ba, st, true, xi, yi, zi, ai, fx, fy, fz, jx, jy, jz);
        } else {
          const bool diag = S == T;
#pragma unroll 4
          for (int st = 0; st < 64; ++st) {
            const int jl = 64 * b + ((lane + st) & 63);
            const bool on = valid && jl < cnt && (!diag || jl > il);
            mutual_step<F64, CUT, PER, true>(p, bx, by, bz, ba, st, on, xi, yi, zi, ai, fx, fy, fz, jx, jy, jz);
          }
        }
        // lane l now holds the sum of source (l + 63) mod 64 of the block
        const int jw = (lane + 63) & 63;
        s_jd[wv][0][jw] = jx;
        s_jd[wv][1][jw] = jy;
        s_jd[wv][2][jw] = jz;
        __syncthreads();
        if (il < 64 && 64 * b + il < cnt) {
          double* P = p.part_j + (tri_index(S, T) * kFBlock + 64 * b + il) * 3;
#pragma unroll
          for (int c = 0; c < 3; ++c) P[c] = ((s_jd[0][c][il] + s_jd[1][c][il]) + s_jd[2][c][il]) + s_jd[3][c][il];
        }
        __syncthreads();  // s_jd is free for the next block
      }
    }
    if (valid) {
      double* P = p.partial + (item * kFBlock + il) * 3;
      P[0] = fx; P[1] = fy; P[2] = fz;
    }
  }
}

// Ordered reduce: F = -(eps/b) * (fixed-order sum over chunks).
__global__ __launch_bounds__(256) void force_reduce_kernel(const double* __restrict__ partial, int64_t n_chunks,
                                                           int64_t n, double scale, double* __restrict__ out)
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
  out[tid] = scale * (s0 + s1);
}

// Symmetric reduce for blob i of tile X: the target-side slots of the items
// of tile X and the source-side slots (X, T) for T = 0..X, split over
// kFRedGroups threads per output, whose sums are added in a fixed order.
constexpr int kFRedGroups = 8, kFRedOut = 256 / kFRedGroups;
__global__ __launch_bounds__(256) void force_mutual_reduce_kernel(const double* __restrict__ part_i,
                                                                  const double* __restrict__ part_j, int64_t n,
                                                                  int64_t n_t, int64_t chunk_tiles, double scale,
                                                                  double* __restrict__ out)
{
  __shared__ double part[kFRedGroups][kFRedOut];
  const int g = threadIdx.x / kFRedOut, ol = threadIdx.x - g * kFRedOut;
  const int64_t tid = (int64_t)blockIdx.x * kFRedOut + ol;
  const bool valid = tid < n * 3;
  double s = 0.0;
  if (valid) {
    const int64_t i = tid / 3;
    const int c = (int)(tid - 3 * i);
    const int64_t X = i / kFBlock, l = i - X * kFBlock;
    const int64_t it0 = items_before(X, n_t, chunk_tiles), it1 = items_before(X + 1, n_t, chunk_tiles);
    const int64_t step = (int64_t)kFBlock * 3;
    s = group_strided_sum<kFRedGroups>(part_i + (it0 * kFBlock + l) * 3 + c, it1 - it0, step, g) +
        group_strided_sum<kFRedGroups>(part_j + (tri_index(X, 0) * kFBlock + l) * 3 + c, X + 1, step, g);
  }
  part[g][ol] = s;
  __syncthreads();
  if (g == 0 && valid) {
    double t = part[0][ol];
#pragma unroll
    for (int q = 1; q < kFRedGroups; ++q) t += part[q][ol];
    out[tid] = scale * t;
  }
}

struct ForcePlan {
  bool mutual = false;
  int64_t chunk_len = 0, n_chunks = 0, total_items = 0;
  int64_t n_t = 0, chunk_tiles = 0;
  size_t off_pj = 0, bytes = 0;
};

static ForcePlan force_plan(int prec, int64_t n, const rmw_tuning* tuning)
{
  RMW_REQUIRE(prec == RMW_PREC_F32_F64ACC || prec == RMW_PREC_F32_KAHAN || prec == RMW_PREC_F64 ||
                  prec == RMW_PREC_F32_BLOCKED,
              "unknown precision mode");
  RMW_REQUIRE(n >= 0, "n must be >= 0");
  ForcePlan pl;
  const int sym = tuning ? tuning->symmetric : 0;
  const int cu = std::max(device_props().cu_count, 1);
  pl.n_t = (n + kFBlock - 1) / kFBlock;
  const int64_t tile_pairs = pl.n_t * (pl.n_t + 1) / 2;
  pl.mutual = n > 0 && sym >= 0 && (sym > 0 || tile_pairs >= cu);
  if (pl.mutual) {
    if (tuning && tuning->chunk_len > 0)
      pl.chunk_tiles = std::max<int64_t>(1, (tuning->chunk_len + kFBlock - 1) / kFBlock);
    else
      pl.chunk_tiles = std::max<int64_t>(1, tile_pairs / (16 * (int64_t)cu));
    pl.total_items = items_before(pl.n_t, pl.n_t, pl.chunk_tiles);
    pl.off_pj = align_up((size_t)pl.total_items * kFBlock * 3 * sizeof(double));
    pl.bytes = pl.off_pj + align_up((size_t)tile_pairs * kFBlock * 3 * sizeof(double));
    if (sym == 0 && pl.bytes > kFMutAutoBytes) pl.mutual = false;
  }
  if (!pl.mutual) {
    int64_t chunk = tuning && tuning->chunk_len > 0 ? tuning->chunk_len : 0;
    if (chunk <= 0) {
      const double c = (double)n * (double)n / (16.0 * cu * kFBlock);
      chunk = std::max<int64_t>(kFBlock, (int64_t)std::ceil(c / 64.0) * 64);
    }
    chunk = std::max<int64_t>(64, (chunk + 63) / 64 * 64);
    if (n > 0) chunk = std::min<int64_t>(chunk, (n + 63) / 64 * 64);
    pl.chunk_len = chunk;
    pl.n_chunks = n > 0 ? (n + chunk - 1) / chunk : 0;
    pl.total_items = pl.n_t * pl.n_chunks;
    pl.off_pj = 0;
    pl.bytes = align_up((size_t)pl.n_chunks * n * 3 * sizeof(double));
  }
  return pl;
}

template <bool F64, bool CUT>
static void launch_force(bool mutual, const ForceParams& q, int grid, hipStream_t st)
{
  if (!mutual) hipLaunchKernelGGL((force_kernel<F64, CUT>), dim3(grid), dim3(kFBlock), 0, st, q);
  else if (q.periodic) hipLaunchKernelGGL((force_mutual_kernel<F64, CUT, true>), dim3(grid), dim3(kFBlock), 0, st, q);
  else hipLaunchKernelGGL((force_mutual_kernel<F64, CUT, false>), dim3(grid), dim3(kFBlock), 0, st, q);
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

extern "C" int rmw_blob_forces_describe(int prec, int64_t n, const rmw_tuning* tuning, rmw_plan_info* info)
{
  return guarded([&] {
    RMW_REQUIRE(info != nullptr, "info is null");
    const ForcePlan pl = force_plan(prec, n, tuning);
    info->symmetric = pl.mutual ? 1 : 0;
    info->rhs_group = 1;
    info->chunk_len = pl.mutual ? pl.chunk_tiles * kFBlock : pl.chunk_len;
    info->work_items = pl.total_items;
    info->workspace_bytes = (int64_t)pl.bytes;
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
    q.invb = 1.0 / debye_length;
    q.t2 = M_LOG2E / debye_length;
    q.cutoff = cutoff;
    for (int d = 0; d < 3; ++d) {
      const double Ld = L ? L[d] : 0.0;
      q.L[d] = Ld > 0.0 ? Ld : 0.0;
      if (Ld > 0.0) q.periodic |= 1 << d;
    }
    q.partial = (double*)workspace;
    q.part_j = (double*)((char*)workspace + pl.off_pj);
    q.n_chunks = pl.n_chunks;
    q.chunk_len = pl.chunk_len;
    q.total_items = pl.total_items;
    q.n_t = pl.n_t;
    q.chunk_tiles = pl.chunk_tiles;
    int grid = (int)std::min<int64_t>(pl.total_items, (int64_t)INT32_MAX);
    if (tuning && tuning->max_wgs > 0) grid = std::min(grid, tuning->max_wgs);
    hipStream_t st = (hipStream_t)stream;
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st, 2);
    const bool cut = cutoff > 0.0;
    if (prec == RMW_PREC_F64) {
      if (cut) launch_force<true, true>(pl.mutual, q, grid, st);
      else launch_force<true, false>(pl.mutual, q, grid, st);
    } else {
      if (cut) launch_force<false, true>(pl.mutual, q, grid, st);
      else launch_force<false, false>(pl.mutual, q, grid, st);
    }
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);
    const double scale = -repulsion_strength / debye_length;
    const dim3 rg((unsigned)((n * 3 + 255) / 256)), rb(256);
    if (pl.mutual)
      hipLaunchKernelGGL(force_mutual_reduce_kernel, dim3((unsigned)((n * 3 + kFRedOut - 1) / kFRedOut)), rb, 0, st,
                         q.partial, q.part_j, n, pl.n_t, pl.chunk_tiles, scale, force);
    else
      hipLaunchKernelGGL(force_reduce_kernel, rg, rb, 0, st, q.partial, pl.n_chunks, n, scale, force);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
