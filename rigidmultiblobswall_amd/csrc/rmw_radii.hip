// Blobs of different radii: u = B_t M(x, y) B_s f from source blobs to
// target blobs, above a wall, unbounded or below a free surface (gfx950).
//
// Reference: mobility_numba.py:1479-1658 (mobility_trans_times_force_
// source_target_numba) behind the wrappers mobility.py:551-591 (wall:
// heights shifted to max(z, radius), damping B = z/radius below the
// radius, shift_heights_different_radius :87-99 and damping_matrix_B_
// different_radius :102-119) and :593-615 (unbounded, no wrapper);
// mobility_numba.py:1941-2093 (free surface: the RPY of d plus the RPY of
// the image vector R = (d_x, d_y, z_i + z_j) times D = diag(1, 1, -1), no
// height shift or damping) behind mobility.py:1409-1429; as selected by
// the 'radii*' implementations of set_mobility_vector_prod
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
#include <type_traits>

#include "rmw_accum.h"
#include "rmw_internal.h"
#include "rmw_pair_terms.h"
#include "rmw_tile_pairs.h"

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
  int bnd;                // kRadiiUnbounded, kRadiiWall or kRadiiFreeSurface
  int periodic;           // bit d set when L_d > 0
  double L[3];
  double* partial;        // [chunk][nt][3]
  int64_t n_chunks, chunk_len, total_items;
};

// Height and damping of a blob of radius rad: z_eff = max(z, rad) and
// B = z / rad below the radius (wall only; the free surface and the
// unbounded kernel use the heights as they are).
__device__ __forceinline__ void wall_shift(double z, double rad, int bnd, double& ze, double& B)
{
  ze = z;
  B = 1.0;
  if (bnd == kRadiiWall && !(z > rad)) {
    B = z < rad ? z / rad : 1.0;
    ze = rad;
  }
}

template <int ACC, typename T, int BND>
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
    wall_shift(p.trg[3 * gi + 2], a, BND, zi, Bi);
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
        wall_shift(p.src[3 * j + 2], b, BND, zj, Bj);
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
          const CoefRadii<T> c = coef_radii<T>(dx, dy, dz, Rz, zir, zj, ar, s_b[j], inva, s_ib[j], BND);
          const Vec3<T> v = apply_radii<T>(dx, dy, dz, Rz, c, s_f[0][j], s_f[1][j], s_f[2][j]);
          u[0].add(v.x);
          u[1].add(v.y);
          u[2].add(v.z);
          if ((j & (kSub - 1)) == kSub - 1) {
            u[0].flush(); u[1].flush(); u[2].flush();
          }
        } else {
          // minimum image of the FP64 difference, then the neighbouring boxes;
          // R_z = z_i + z_j is not shifted, so with L_z > 0 the image term is
          // taken once per z box, as the reference kernels do
          if (pbox[0]) ddx -= image_count(ddx, p.L[0]) * p.L[0];
          if (pbox[1]) ddy -= image_count(ddy, p.L[1]) * p.L[1];
          if (pbox[2]) ddz -= image_count(ddz, p.L[2]) * p.L[2];
          for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
            for (int by = -pbox[1]; by <= pbox[1]; ++by)
              for (int bz = -pbox[2]; bz <= pbox[2]; ++bz) {
                const T dx = (T)(ddx + bx * p.L[0]), dy = (T)(ddy + by * p.L[1]), dz = (T)(ddz + bz * p.L[2]);
                const CoefRadii<T> c =
                    coef_radii<T>(dx, dy, dz, Rz, zir, zj, ar, s_b[j], inva, s_ib[j], BND);
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
                                                           const double* __restrict__ rt, int bnd,
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
  wall_shift(trg[3 * i + 2], rt[i], bnd, ze, B);
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

// ---------------------------------------------------------------------------
// Sources = targets (mobility.py:1369-1374): the tensor of a pair is its own
// transpose under exchange (M_ji = M_ij^T: the far, overlap and engulfed RPY
// coefficients, the wall's al, be, de and the free surface's image
// coefficients are symmetric in (a_i, a_j), while exchanging the blobs
// swaps g1 and g2, which the free surface does not have), so each unordered
// pair is evaluated once and applied to both blobs.
// ---------------------------------------------------------------------------

// M_ij^T v (the coefficients as computed for the ordered pair (i, j)).
template <typename T>
__device__ __forceinline__ Vec3<T> apply_radii_t(T dx, T dy, T dz, T Rz, const CoefRadii<T>& c, T vx, T vy, T vz)
{
  const T p = dx * vx + dy * vy;
  const T dv = p + dz * vz;
  const T Rv = p + Rz * vz;               // R . v
  const T h = c.be * Rv + c.g2 * vz;      // coefficient of -D R
  const T q = c.C2 * dv;
  Vec3<T> u;
  u.x = (c.C1 - c.al) * vx + (q - h) * dx;
  u.y = (c.C1 - c.al) * vy + (q - h) * dy;
  u.z = (c.C1 + c.al) * vz + q * dz + h * Rz + c.g1 * Rv + c.de * vz;
  return u;
}

struct RadiiMutParams {
  const double* r;
  const double* f;
  const double* rad;
  int64_t n;
  double norm;            // 1 / (8 pi eta)
  int bnd;                // kRadiiUnbounded, kRadiiWall or kRadiiFreeSurface
  double* part_i;         // [item][kRBlock][3] target-side sums
  double* part_j;         // [tri_index(S, T)][kRBlock][3] source-side sums
  int64_t n_t, chunk_tiles, total_items;
};

// Work item = (target tile T, run of source tiles S >= T); on the diagonal
// tile only sources above the target are taken, plus the self pair on the
// target side (it is the inner RPY regime, no special case).  Target-side sums use the
// precision's accumulator; source-side sums rotate one lane per step
// (DPP) so that lane l ends with source (l + 63) mod 64 of the block, in
// FP32 sub-sums of kSub steps folded into FP64 in the blocked mode.
template <int ACC, typename T, int BND>
__global__ __launch_bounds__(kRBlock) void radii_mutual_kernel(RadiiMutParams p)
{
  constexpr bool kFoldJ = ACC == RMW_PREC_F32_BLOCKED;
  using JT = typename std::conditional<kFoldJ, float, double>::type;
  __shared__ double s_x[kRBlock], s_y[kRBlock], s_z[kRBlock];
  __shared__ T s_zr[kRBlock], s_b[kRBlock], s_ib[kRBlock];
  __shared__ T s_f[3][kRBlock];
  __shared__ double s_jd[kRBlock / 64][3][64];   // per-wave source-side sums of one block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int il = threadIdx.x;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const int64_t Tt = pair_target_tile(item, p.n_t, p.chunk_tiles);
    const int64_t S0 = Tt + (item - items_before(Tt, p.n_t, p.chunk_tiles)) * p.chunk_tiles;
    const int64_t S1 = S0 + p.chunk_tiles < p.n_t ? S0 + p.chunk_tiles : p.n_t;
    const int64_t i = Tt * kRBlock + il;
    const bool valid = i < p.n;
    const int64_t gi = valid ? i : 0;
    const double xi = p.r[3 * gi], yi = p.r[3 * gi + 1];
    const double a = p.rad[gi];
    double zi, Bi;
    wall_shift(p.r[3 * gi + 2], a, BND, zi, Bi);
    const T ar = (T)a, inva = (T)(1.0 / a), zir = (T)zi;
    T fi[3];
    {
      const double c = valid ? p.norm * Bi : 0.0;
      fi[0] = (T)(c * p.f[3 * gi]); fi[1] = (T)(c * p.f[3 * gi + 1]); fi[2] = (T)(c * p.f[3 * gi + 2]);
    }
    Acc<ACC> u[3];

    for (int64_t S = S0; S < S1; ++S) {
      const int64_t s0 = S * kRBlock;
      const int cnt = (int)(p.n - s0 < kRBlock ? p.n - s0 : kRBlock);
      __syncthreads();  // the previous tile has been consumed
      if (il < cnt) {
        const int64_t j = s0 + il;
        const double b = p.rad[j];
        double zj, Bj;
        wall_shift(p.r[3 * j + 2], b, BND, zj, Bj);
        s_x[il] = p.r[3 * j]; s_y[il] = p.r[3 * j + 1]; s_z[il] = zj;
        s_zr[il] = (T)zj; s_b[il] = (T)b; s_ib[il] = (T)(1.0 / b);
        const double c = p.norm * Bj;
        s_f[0][il] = (T)(c * p.f[3 * j]); s_f[1][il] = (T)(c * p.f[3 * j + 1]); s_f[2][il] = (T)(c * p.f[3 * j + 2]);
      } else {
        // finite stand-ins beyond n (masked below)
        s_x[il] = 0.0; s_y[il] = 0.0; s_z[il] = 1.0; s_zr[il] = T(1); s_b[il] = T(1); s_ib[il] = T(1);
        s_f[0][il] = T(0); s_f[1][il] = T(0); s_f[2][il] = T(0);
      }
      __syncthreads();
      const bool diag = S == Tt;
      // full off-diagonal tile pair: every lane's pair is taken on both
      // sides, so the step loop runs without its selects
      const bool full = !diag && cnt == kRBlock && (Tt + 1) * kRBlock <= p.n;
      for (int b = 0; b < kRBlock / 64 && 64 * b < cnt; ++b) {
        JT ja[3] = {JT(0), JT(0), JT(0)};
        auto steps = [&](auto masked_c) {
        constexpr bool MASKED = decltype(masked_c)::value;
        for (int st = 0; st < 64; ++st) {
          const int jl = 64 * b + ((lane + st) & 63);
          // the self pair (the RPY self mobility and its image) once, on the target side
          const bool on = !MASKED || (valid && jl < cnt && (!diag || jl >= il));
          const bool on_j = !MASKED || (on && (!diag || jl != il));
          T dx, dy, dz;
          if constexpr (sizeof(T) == 4) {
            dx = (T)(xi - s_x[jl]); dy = (T)(yi - s_y[jl]); dz = (T)(zi - s_z[jl]);
          } else {
            dx = xi - s_x[jl]; dy = yi - s_y[jl]; dz = zi - s_z[jl];
          }
          const T zj = s_zr[jl];
          const T Rz = zir + zj;
          const CoefRadii<T> c = coef_radii<T>(dx, dy, dz, Rz, zir, zj, ar, s_b[jl], inva, s_ib[jl], BND);
          const Vec3<T> vi = apply_radii<T>(dx, dy, dz, Rz, c, s_f[0][jl], s_f[1][jl], s_f[2][jl]);
          const Vec3<T> vj = apply_radii_t<T>(dx, dy, dz, Rz, c, fi[0], fi[1], fi[2]);
          if constexpr (MASKED) {
            u[0].add(on ? vi.x : T(0));
            u[1].add(on ? vi.y : T(0));
            u[2].add(on ? vi.z : T(0));
            ja[0] = rol1(ja[0]) + (on_j ? (JT)vj.x : JT(0));
            ja[1] = rol1(ja[1]) + (on_j ? (JT)vj.y : JT(0));
            ja[2] = rol1(ja[2]) + (on_j ? (JT)vj.z : JT(0));
          } else {
            u[0].add(vi.x);
            u[1].add(vi.y);
            u[2].add(vi.z);
            ja[0] = rol1(ja[0]) + (JT)vj.x;
            ja[1] = rol1(ja[1]) + (JT)vj.y;
            ja[2] = rol1(ja[2]) + (JT)vj.z;
          }
          if ((st & (kSub - 1)) == kSub - 1) {
            u[0].flush(); u[1].flush(); u[2].flush();
            if constexpr (kFoldJ) {
              // lane l holds the sub-sum of source (l + st) mod 64 of the block
              const int jw = (lane + st) & 63;
#pragma unroll
              for (int k = 0; k < 3; ++k) {
                const double prev = st < kSub ? 0.0 : s_jd[wv][k][jw];
                s_jd[wv][k][jw] = prev + (double)ja[k];
                ja[k] = JT(0);
              }
            }
          }
        }
        };
        if (full) steps(std::false_type{}); else steps(std::true_type{});
        if constexpr (!kFoldJ) {
          const int jw = (lane + 63) & 63;
#pragma unroll
          for (int k = 0; k < 3; ++k) s_jd[wv][k][jw] = (double)ja[k];
        }
        __syncthreads();  // every wave's sums of this block are in LDS
        if (il < 64 && 64 * b + il < cnt) {
          double* P = p.part_j + (tri_index(S, Tt) * kRBlock + 64 * b + il) * 3;
#pragma unroll
          for (int k = 0; k < 3; ++k) P[k] = ((s_jd[0][k][il] + s_jd[1][k][il]) + s_jd[2][k][il]) + s_jd[3][k][il];
        }
        __syncthreads();  // s_jd is free for the next block
      }
      u[0].flush(); u[1].flush(); u[2].flush();
    }
    if (valid) {
      double* P = p.part_i + (item * kRBlock + il) * 3;
      P[0] = u[0].get(); P[1] = u[1].get(); P[2] = u[2].get();
    }
  }
}

// Blob i of tile X: the target-side sums of the items of tile X and the
// source-side sums (X, T) for T = 0..X, split over kRRedGroups threads per
// output whose sums are added in a fixed order; then the damping.
constexpr int kRRedGroups = 8, kRRedOut = 256 / kRRedGroups;
__global__ __launch_bounds__(256) void radii_mutual_reduce_kernel(const double* __restrict__ part_i,
                                                                  const double* __restrict__ part_j,
                                                                  const double* __restrict__ r,
                                                                  const double* __restrict__ rad, int bnd, int64_t n,
                                                                  int64_t n_t, int64_t chunk_tiles,
                                                                  double* __restrict__ out)
{
  __shared__ double part[kRRedGroups][kRRedOut];
  const int g = threadIdx.x / kRRedOut, ol = threadIdx.x - g * kRRedOut;
  const int64_t tid = (int64_t)blockIdx.x * kRRedOut + ol;
  const bool valid = tid < n * 3;
  double s = 0.0;
  if (valid) {
    const int64_t i = tid / 3;
    const int c = (int)(tid - 3 * i);
    const int64_t X = i / kRBlock, l = i - X * kRBlock;
    const int64_t it0 = items_before(X, n_t, chunk_tiles), it1 = items_before(X + 1, n_t, chunk_tiles);
    const int64_t step = (int64_t)kRBlock * 3;
    s = group_strided_sum<kRRedGroups>(part_i + (it0 * kRBlock + l) * 3 + c, it1 - it0, step, g) +
        group_strided_sum<kRRedGroups>(part_j + (tri_index(X, 0) * kRBlock + l) * 3 + c, X + 1, step, g);
  }
  part[g][ol] = s;
  __syncthreads();
  if (g == 0 && valid) {
    double t = part[0][ol];
#pragma unroll
    for (int q = 1; q < kRRedGroups; ++q) t += part[q][ol];
    const int64_t i = tid / 3;
    double ze, B;
    wall_shift(r[3 * i + 2], rad[i], bnd, ze, B);
    out[tid] = B * t;
  }
}

constexpr size_t kRMutAutoBytes = size_t(2) << 30;

struct RadiiSelfPlan {
  bool mutual = false;
  int64_t n_t = 0, chunk_tiles = 0, total_items = 0;
  size_t off_pj = 0, bytes = 0;
  RadiiPlan ordered;
};

// Symmetric kernel when the tile pairs fill the device (or tuning->symmetric
// > 0), the partial sums fit kRMutAutoBytes and there is no periodic box;
// otherwise the ordered kernel with sources = targets.
static RadiiSelfPlan radii_self_plan(int prec, int64_t n, bool periodic, const rmw_tuning* tuning)
{
  RadiiSelfPlan pl;
  pl.ordered = radii_plan(prec, n, n, tuning);
  const int sym = tuning ? tuning->symmetric : 0;
  const int cu = std::max(device_props().cu_count, 1);
  pl.n_t = (n + kRBlock - 1) / kRBlock;
  const int64_t tile_pairs = pl.n_t * (pl.n_t + 1) / 2;
  RMW_REQUIRE(!(sym > 0 && periodic), "the symmetric radii kernel has no periodic form");
  pl.mutual = n > 0 && !periodic && sym >= 0 && (sym > 0 || tile_pairs >= cu);
  if (pl.mutual) {
    if (tuning && tuning->chunk_len > 0)
      pl.chunk_tiles = std::max<int64_t>(1, (tuning->chunk_len + kRBlock - 1) / kRBlock);
    else
      pl.chunk_tiles = std::max<int64_t>(1, tile_pairs / (16 * (int64_t)cu));
    pl.total_items = items_before(pl.n_t, pl.n_t, pl.chunk_tiles);
    pl.off_pj = align_up((size_t)pl.total_items * kRBlock * 3 * sizeof(double));
    pl.bytes = pl.off_pj + align_up((size_t)tile_pairs * kRBlock * 3 * sizeof(double));
    if (sym == 0 && pl.bytes > kRMutAutoBytes) pl.mutual = false;
  }
  if (!pl.mutual) pl.bytes = pl.ordered.bytes;
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

// One instantiation per precision and boundary (the boundary's terms are
// resolved at compile time).
template <int BND>
static void launch_radii(int prec, dim3 g, dim3 b, hipStream_t st, const RadiiParams& q)
{
  if (prec == RMW_PREC_F32_F64ACC) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_F64ACC, float, BND>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_BLOCKED) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_BLOCKED, float, BND>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_KAHAN) hipLaunchKernelGGL((radii_kernel<RMW_PREC_F32_KAHAN, float, BND>), g, b, 0, st, q);
  else hipLaunchKernelGGL((radii_kernel<RMW_PREC_F64, double, BND>), g, b, 0, st, q);
}

template <int BND>
static void launch_radii_mutual(int prec, dim3 g, dim3 b, hipStream_t st, const RadiiMutParams& q)
{
  if (prec == RMW_PREC_F32_F64ACC) hipLaunchKernelGGL((radii_mutual_kernel<RMW_PREC_F32_F64ACC, float, BND>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_BLOCKED) hipLaunchKernelGGL((radii_mutual_kernel<RMW_PREC_F32_BLOCKED, float, BND>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_KAHAN) hipLaunchKernelGGL((radii_mutual_kernel<RMW_PREC_F32_KAHAN, float, BND>), g, b, 0, st, q);
  else hipLaunchKernelGGL((radii_mutual_kernel<RMW_PREC_F64, double, BND>), g, b, 0, st, q);
}

// The ordered pair kernel and its reduce, source -> target.
static void radii_ordered(const RadiiPlan& pl, int prec, int bnd, const double* source, const double* target,
                          const double* force, const double* radius_source, const double* radius_target,
                          double* velocity, int64_t ns, int64_t nt, double eta, const double* L,
                          const rmw_tuning* tuning, void* workspace, hipStream_t st)
{
  RadiiParams q{};
  q.src = source; q.trg = target; q.f = force; q.rs = radius_source; q.rt = radius_target;
  q.ns = ns; q.nt = nt;
  q.norm = 1.0 / (8.0 * M_PI * eta);
  q.bnd = bnd;
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
  if (bnd == kRadiiWall) launch_radii<kRadiiWall>(prec, g, b, st, q);
  else if (bnd == kRadiiFreeSurface) launch_radii<kRadiiFreeSurface>(prec, g, b, st, q);
  else launch_radii<kRadiiUnbounded>(prec, g, b, st, q);
  RMW_CHECK_HIP(hipGetLastError());
  prof_end(rec, prof, st);
  hipLaunchKernelGGL(radii_reduce_kernel, dim3((unsigned)((nt * 3 + 255) / 256)), dim3(256), 0, st, q.partial,
                     pl.n_chunks, nt, target, radius_target, q.bnd, velocity);
  RMW_CHECK_HIP(hipGetLastError());
}

static bool any_periodic(const double* L)
{
  return L && (L[0] > 0.0 || L[1] > 0.0 || L[2] > 0.0);
}

static void check_boundary(int boundary)
{
  RMW_REQUIRE(boundary == kRadiiUnbounded || boundary == kRadiiWall || boundary == kRadiiFreeSurface,
              "boundary must be 0 (unbounded), 1 (wall) or 2 (free surface)");
}

extern "C" int rmw_radii_apply(int prec, int boundary, const double* source, const double* target, const double* force,
                               const double* radius_source, const double* radius_target, double* velocity,
                               int64_t ns, int64_t nt, double eta, const double* L, const rmw_tuning* tuning,
                               void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    const RadiiPlan pl = radii_plan(prec, ns, nt, tuning);
    check_boundary(boundary);
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
    radii_ordered(pl, prec, boundary, source, target, force, radius_source, radius_target, velocity, ns, nt, eta, L,
                  tuning, workspace, st);
  });
}

extern "C" int rmw_radii_self_workspace_size(int prec, int64_t n, const double* L, const rmw_tuning* tuning,
                                             size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr, "bytes is null");
    *bytes = radii_self_plan(prec, n, any_periodic(L), tuning).bytes;
  });
}

extern "C" int rmw_radii_self_describe(int prec, int64_t n, const double* L, const rmw_tuning* tuning,
                                       int* symmetric, int64_t* work_items, size_t* bytes)
{
  return guarded([&] {
    const RadiiSelfPlan pl = radii_self_plan(prec, n, any_periodic(L), tuning);
    if (symmetric) *symmetric = pl.mutual ? 1 : 0;
    if (work_items) *work_items = pl.mutual ? pl.total_items : pl.ordered.total_items;
    if (bytes) *bytes = pl.bytes;
  });
}

extern "C" int rmw_radii_self_apply(int prec, int boundary, const double* r, const double* force, const double* radius,
                                    double* velocity, int64_t n, double eta, const double* L,
                                    const rmw_tuning* tuning, void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    const RadiiSelfPlan pl = radii_self_plan(prec, n, any_periodic(L), tuning);
    check_boundary(boundary);
    RMW_REQUIRE(workspace_bytes >= pl.bytes, "workspace too small: need " + std::to_string(pl.bytes) + " bytes");
    RMW_REQUIRE(eta > 0.0 && std::isfinite(eta), "eta must be positive");
    if (n == 0) return;
    RMW_REQUIRE(r && force && radius && velocity && workspace, "r, force, radius, velocity and workspace are required");
    hipStream_t st = (hipStream_t)stream;
    if (!pl.mutual) {
      radii_ordered(pl.ordered, prec, boundary, r, r, force, radius, radius, velocity, n, n, eta, L, tuning, workspace,
                    st);
      return;
    }
    RadiiMutParams q{};
    q.r = r; q.f = force; q.rad = radius; q.n = n;
    q.norm = 1.0 / (8.0 * M_PI * eta);
    q.bnd = boundary;
    q.part_i = (double*)workspace;
    q.part_j = (double*)((char*)workspace + pl.off_pj);
    q.n_t = pl.n_t; q.chunk_tiles = pl.chunk_tiles; q.total_items = pl.total_items;
    int grid = (int)std::min<int64_t>(pl.total_items, (int64_t)INT32_MAX);
    if (tuning && tuning->max_wgs > 0) grid = std::min(grid, tuning->max_wgs);
    const dim3 g(grid), b(kRBlock);
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st, 1);
    if (boundary == kRadiiWall) launch_radii_mutual<kRadiiWall>(prec, g, b, st, q);
    else if (boundary == kRadiiFreeSurface) launch_radii_mutual<kRadiiFreeSurface>(prec, g, b, st, q);
    else launch_radii_mutual<kRadiiUnbounded>(prec, g, b, st, q);
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);
    hipLaunchKernelGGL(radii_mutual_reduce_kernel, dim3((unsigned)((n * 3 + kRRedOut - 1) / kRRedOut)), dim3(256), 0,
                       st, q.part_i, q.part_j, r, radius, q.bnd, n, pl.n_t, pl.chunk_tiles, velocity);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
