// Wall-corrected RPY blob mobility products on gfx950 (MI355X).
//
// One call evaluates any subset of the grand mobility blocks
// [[UF, UT], [WF, WT]] for k right-hand sides on one suspension, or on a
// ragged batch of independent suspensions, in two launches:
//
//   1. pair   : all-pairs sums; each workgroup owns 256 targets (one per
//               lane, in registers) and one chunk of sources that it
//               streams through LDS in tiles of 256, reading positions and
//               sources straight from the caller's float64 arrays: z_eff =
//               max(z, a), damping B = min(z/a, 1) and 1/(8 pi eta a^p) are
//               applied as a tile is staged.  The RPY and wall image terms
//               are fused in one pass and the geometry is shared by all
//               blocks and right-hand sides.  Partial sums (FP64) go to the
//               workspace, one slot per (chunk, target).
//   2. reduce : fixed-order FP64 sum over chunks, FP64 self term, B and
//               the remaining 1/a factor, written as (k, n, 3) float64.
//
// Reference: mobility/mobility_numba.py:123-287 (UF), :547-686 (UT),
// :937-1073 (WF), :1188-1328 (WT); wrappers mobility/mobility.py:52-84,
// 1132-1342; pseudo-periodic images mobility_numba.py:149-198.
//
// FP32 modes.  Subtracting two FP32 coordinates of size X loses ~6e-8 X of
// absolute accuracy, which for near pairs (|d| ~ a, X up to ~500 a) would be
// a 3e-5 relative error.  The difference d = r_i - r_j (and its periodic
// shifts) is therefore formed in FP64 and rounded once to FP32, so d
// carries FP32 relative accuracy in d itself; everything after that is
// FP32.  (On gfx950 an FP64 add issues at the rate of an unpacked FP32 add.)
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "rmw_internal.h"
#include "rmw_pair_terms.h"
#include "rmw_tile_pairs.h"
#include "rmw_accum.h"

namespace rmw {

// ---------------------------------------------------------------------------
// Errors and device properties
// ---------------------------------------------------------------------------
static thread_local std::string g_last_error;
void set_last_error(const std::string& msg) { g_last_error = msg; }

const DeviceProps& device_props()
{
  static std::mutex mu;
  static DeviceProps cache[64];
  int dev = 0;
  RMW_CHECK_HIP(hipGetDevice(&dev));
  RMW_REQUIRE(dev >= 0 && dev < 64, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(mu);
  DeviceProps& p = cache[dev];
  if (p.device < 0) {
    hipDeviceProp_t prop;
    RMW_CHECK_HIP(hipGetDeviceProperties(&prop, dev));
    p.cu_count = prop.multiProcessorCount;
    p.clock_khz = prop.clockRate;
    p.warp = prop.warpSize;
    p.arch = prop.gcnArchName;
    p.device = dev;
  }
  return p;
}

constexpr int kBlock = 256;  // targets per workgroup = sources per LDS tile
enum : int { kUF = 1, kUT = 2, kWF = 4, kWT = 8 };

// One independent suspension of the batch and its share of the work list.
struct SysInfo {
  int64_t b0;         // first blob (global index)
  int64_t n;          // blobs
  int64_t n_tt;       // target tiles of kBlock
  int64_t n_chunks;   // source chunks
  int64_t chunk_len;  // sources per chunk (multiple of kBlock)
  int64_t item0;      // first work item
  int64_t part0;      // first partial (in doubles)
};

struct PairParams {
  const double* r;        // (ntot, 3) positions
  const double* f;        // K right-hand sides rhs_stride apart, or null
  const double* t;
  int64_t rhs_stride;
  double a, inva, c1, c2; // c1 = 1/(8 pi eta a), c2 = c1 / a
  double* partial;
  const SysInfo* table;   // nullptr -> single
  SysInfo single;
  int nsys;
  int64_t ntot;
  int64_t total_items;
  int periodic;           // bit d set when L_d > 0
  double L[3];            // periodic lengths / a
  int domain;             // kDomWall, kDomNoWall, kDomInPlane, kDomFreeSurface
};

__device__ inline SysInfo find_sys_item(const PairParams& p, int64_t item)
{
  if (!p.table) return p.single;
  int lo = 0, hi = p.nsys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (p.table[mid].item0 <= item) lo = mid; else hi = mid - 1;
  }
  return p.table[lo];
}

// ---------------------------------------------------------------------------
// Blob staging: scaled position (x/a, y/a, z_eff/a) with z_eff = max(z, a)
// (shift_heights, mobility.py:52-64) and damping B = min(z/a, 1)
// (damping_matrix_B, mobility.py:67-84) when a wall bounds the fluid; the
// no-wall and free-surface products take the heights as they are
// (mobility.py:1119-1129, 1390-1406).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_blob(const double* __restrict__ r, int64_t g, double a, double inva,
                                          double& X, double& Y, double& Z, double& B, int dom)
{
  const double x = r[3 * g], y = r[3 * g + 1], z = r[3 * g + 2];
  const bool wall = dom_has_wall(dom);
  const double ze = wall && z <= a ? a : z;
  X = x * inva;
  Y = y * inva;
  Z = ze * inva;
  B = wall && z < a ? z * inva : 1.0;
}

// Source record of the FP32 kernels: FP64 position for the difference, FP32
// height for the wall terms (two 16-byte LDS reads).
struct __align__(16) SrcF32 {
  double x, y, z;
  float zf, pad;
};

// ---------------------------------------------------------------------------
// 3. pair kernels
// ---------------------------------------------------------------------------
// Per-target accumulator policies: rmw_accum.h.

template <int K, int ACC>
__device__ __forceinline__ void flush_all(Acc<ACC> (&u)[K][3], Acc<ACC> (&w)[K][3])
{
#pragma unroll
  for (int k = 0; k < K; ++k)
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      u[k][c].flush();
      w[k][c].flush();
    }
}

template <int MASK> struct BlockTraits {
  static constexpr bool HU = (MASK & (kUF | kUT)) != 0;  // writes velocity
  static constexpr bool HW = (MASK & (kWF | kWT)) != 0;  // writes omega
  static constexpr bool NF = (MASK & (kUF | kWF)) != 0;  // reads force
  static constexpr bool NT = (MASK & (kUT | kWT)) != 0;  // reads torque
  static constexpr int OC = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
};

template <typename T> struct V4;
template <> struct V4<float> { using type = float4; };
template <> struct V4<double> { using type = double4; };

// One (target, source) pair for all blocks of MASK and all K right-hand
// sides.  `skip` drops the pair (the self pair of the central box).
template <int MASK, int K, int ACC, typename T, int DOM>
__device__ __forceinline__ void pair_eval(T dx, T dy, T dz, T zi, T zj,
                                          const typename V4<T>::type* __restrict__ sf,
                                          const typename V4<T>::type* __restrict__ st, int j,
                                          bool skip, Acc<ACC> (&u)[K][3], Acc<ACC> (&w)[K][3])
{
  using BT = BlockTraits<MASK>;
  Geom<T> g = make_geom(dx, dy, dz, zi, zj);
  CoefUF<T> cuf;
  CoefRot<T> cut, cwf;
  CoefWT<T> cwt;
  if constexpr ((MASK & kUF) != 0) cuf = coef_uf(g, zi, zj, DOM);
  if constexpr ((MASK & kUT) != 0) cut = coef_rot(g, zi, DOM);
  if constexpr ((MASK & kWF) != 0) cwf = coef_rot(g, zj, DOM);
  if constexpr ((MASK & kWT) != 0) cwt = coef_wt(g, DOM);
#pragma unroll
  for (int k = 0; k < K; ++k) {
    Vec3<T> uu{T(0), T(0), T(0)}, ww{T(0), T(0), T(0)};
    typename V4<T>::type fv, tv;
    if constexpr (BT::NF) fv = sf[k * kBlock + j];
    if constexpr (BT::NT) tv = st[k * kBlock + j];
    if constexpr ((MASK & kUF) != 0) {
      Vec3<T> c = apply_uf(g, cuf, fv.x, fv.y, fv.z);
      uu.x += c.x; uu.y += c.y; uu.z += c.z;
    }
    if constexpr ((MASK & kUT) != 0) {
      Vec3<T> c = apply_ut(g, cut, tv.x, tv.y, tv.z);
      uu.x += c.x; uu.y += c.y; uu.z += c.z;
    }
    if constexpr ((MASK & kWF) != 0) {
      Vec3<T> c = apply_wf(g, cwf, fv.x, fv.y, fv.z);
      ww.x += c.x; ww.y += c.y; ww.z += c.z;
    }
    if constexpr ((MASK & kWT) != 0) {
      Vec3<T> c = apply_wt(g, cwt, tv.x, tv.y, tv.z);
      ww.x += c.x; ww.y += c.y; ww.z += c.z;
    }
    if constexpr (BT::HU) {
      u[k][0].add(skip ? T(0) : uu.x);
      u[k][1].add(skip ? T(0) : uu.y);
      u[k][2].add(skip ? T(0) : uu.z);
    }
    if constexpr (BT::HW) {
      w[k][0].add(skip ? T(0) : ww.x);
      w[k][1].add(skip ? T(0) : ww.y);
      w[k][2].add(skip ? T(0) : ww.z);
    }
  }
}


// Partial sums of one work item: slot (chunk, rhs, target) of the system.
template <int MASK, int K, typename A>
__device__ __forceinline__ void store_partials(const PairParams& p, const SysInfo& S, int64_t ch, int64_t il,
                                               A (&u)[K][3], A (&w)[K][3])
{
  using BT = BlockTraits<MASK>;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double* P = p.partial + S.part0 + ((ch * K + k) * S.n + il) * BT::OC;
    int o = 0;
    if constexpr (BT::HU) {
      P[0] = u[k][0].get(); P[1] = u[k][1].get(); P[2] = u[k][2].get();
      o = 3;
    }
    if constexpr (BT::HW) {
      P[o] = w[k][0].get(); P[o + 1] = w[k][1].get(); P[o + 2] = w[k][2].get();
    }
  }
}

template <typename V> __device__ V make_v4(double x, double y, double z);
template <> __device__ __forceinline__ float4 make_v4<float4>(double x, double y, double z)
{
  return make_float4((float)x, (float)y, (float)z, 0.0f);
}
template <> __device__ __forceinline__ double4 make_v4<double4>(double x, double y, double z)
{
  return make_double4(x, y, z, 0.0);
}

// Stage the sources of blob gj for K right-hand sides into LDS slot `slot`:
// B f / (8 pi eta a) and B t / (8 pi eta a^2) as float4 or double4.
template <int MASK, int K, typename V, int DOM>
__device__ __forceinline__ void stage_sources(const PairParams& p, int64_t gj, double B, int slot, V* s_f, V* s_t)
{
  using BT = BlockTraits<MASK>;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    // in plane: only the x-y components act (P v)
    const double pz = DOM == kDomInPlane ? 0.0 : 1.0;
    if constexpr (BT::NF) {
      const double* v = p.f + k * p.rhs_stride + 3 * gj;
      const double c = p.c1 * B;
      s_f[k * kBlock + slot] = make_v4<V>(c * v[0], c * v[1], pz * c * v[2]);
    }
    if constexpr (BT::NT) {
      const double* v = p.t + k * p.rhs_stride + 3 * gj;
      const double c = p.c2 * B;
      s_t[k * kBlock + slot] = make_v4<V>(c * v[0], c * v[1], pz * c * v[2]);
    }
  }
}

// Minimum resident waves per SIMD asked of the register allocator for the
// FP32 pair kernels (1 = no constraint).
#ifndef RMW_PAIR_WAVES
#define RMW_PAIR_WAVES 1
#endif
template <int MASK, int K>
constexpr int pair_min_waves()
{
  return (K == 1 && (MASK == kUF || MASK == kUT || MASK == kWF || MASK == kWT)) ? RMW_PAIR_WAVES : 1;
}

template <int MASK, int K, int ACC, int DOM = kDomWall>
__global__ __launch_bounds__(kBlock, (pair_min_waves<MASK, K>())) void pair_kernel_f32(PairParams p)
{
  using BT = BlockTraits<MASK>;
  __shared__ SrcF32 s_pos[kBlock];
  __shared__ float4 s_f[BT::NF ? K * kBlock : 1];
  __shared__ float4 s_t[BT::NT ? K * kBlock : 1];
  int pbox[3];
  for (int d = 0; d < 3; ++d) pbox[d] = (p.periodic >> d) & 1;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const SysInfo S = find_sys_item(p, item);
    const int64_t local = item - S.item0;
    const int64_t tt = local / S.n_chunks;
    const int64_t ch = local - tt * S.n_chunks;
    const int64_t t0 = tt * kBlock;
    const int64_t il = t0 + threadIdx.x;
    const bool valid = il < S.n;
    double Xi, Yi, Zi, Bi;
    load_blob(p.r, S.b0 + (valid ? il : 0), p.a, p.inva, Xi, Yi, Zi, Bi, DOM);
    const float zi = (float)Zi;

    Acc<ACC> u[K][3], w[K][3];

    const int64_t s_begin = ch * S.chunk_len;
    const int64_t s_end = s_begin + S.chunk_len < S.n ? s_begin + S.chunk_len : S.n;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += kBlock) {
      const int cnt = (int)(s_end - s0 < kBlock ? s_end - s0 : kBlock);
      __syncthreads();  // the previous tile has been consumed
      if ((int)threadIdx.x < cnt) {
        const int64_t gj = S.b0 + s0 + threadIdx.x;
        double X, Y, Z, B;
        load_blob(p.r, gj, p.a, p.inva, X, Y, Z, B, DOM);
        s_pos[threadIdx.x] = SrcF32{X, Y, Z, (float)Z, 0.0f};
        stage_sources<MASK, K, float4, DOM>(p, gj, B, threadIdx.x, s_f, s_t);
      }
      __syncthreads();
      // Tiles that contain this workgroup's own targets need the self mask.
      const bool diag = (s0 < t0 + kBlock) && (t0 < s0 + cnt);
      const int self_j = (int)(il - s0);  // local source index of the target itself
      if (!p.periodic) {
        auto eval = [&](int j, bool masked) {
          const SrcF32 sp = s_pos[j];
          const float dx = (float)(Xi - sp.x);
          const float dy = (float)(Yi - sp.y);
          const float dz = (float)(Zi - sp.z);
          pair_eval<MASK, K, ACC, float, DOM>(dx, dy, dz, zi, sp.zf, s_f, s_t, j, masked && j == self_j, u, w);
        };
        int j = 0;
        if (diag) {
          for (; j + kSub <= cnt; j += kSub) {
#pragma unroll 2
            for (int jj = 0; jj < kSub; ++jj) eval(j + jj, true);
            flush_all<K, ACC>(u, w);
          }
          for (; j < cnt; ++j) eval(j, true);
        } else {
          for (; j + kSub <= cnt; j += kSub) {
#pragma unroll 2
            for (int jj = 0; jj < kSub; ++jj) eval(j + jj, false);
            flush_all<K, ACC>(u, w);
          }
          for (; j < cnt; ++j) eval(j, false);
        }
        flush_all<K, ACC>(u, w);
      } else {
        for (int j = 0; j < cnt; ++j) {
          const SrcF32 sp = s_pos[j];
          // minimum image of the FP64 difference, then every box shift in FP64
          double ddx = Xi - sp.x, ddy = Yi - sp.y, ddz = Zi - sp.z;
          if (pbox[0]) ddx -= image_count(ddx, p.L[0]) * p.L[0];
          if (pbox[1]) ddy -= image_count(ddy, p.L[1]) * p.L[1];
          if (pbox[2]) ddz -= image_count(ddz, p.L[2]) * p.L[2];
          for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
            for (int by = -pbox[1]; by <= pbox[1]; ++by)
              for (int bz = -pbox[2]; bz <= pbox[2]; ++bz) {
                const float dx = (float)(ddx + bx * p.L[0]);
                const float dy = (float)(ddy + by * p.L[1]);
                const float dz = (float)(ddz + bz * p.L[2]);
                const bool central = (bx == 0) && (by == 0) && (bz == 0);
                pair_eval<MASK, K, ACC, float, DOM>(dx, dy, dz, zi, sp.zf, s_f, s_t, j,
                                                    diag && central && (j == self_j), u, w);
              }
          flush_all<K, ACC>(u, w);
        }
      }
    }
    if (valid) store_partials<MASK, K>(p, S, ch, il, u, w);
  }
}

template <int MASK, int K, int DOM = kDomWall>
__global__ __launch_bounds__(kBlock) void pair_kernel_f64(PairParams p)
{
  using BT = BlockTraits<MASK>;
  __shared__ double4 s_pos[kBlock];
  __shared__ double4 s_f[BT::NF ? K * kBlock : 1];
  __shared__ double4 s_t[BT::NT ? K * kBlock : 1];
  int pbox[3];
  for (int d = 0; d < 3; ++d) pbox[d] = (p.periodic >> d) & 1;

  for (int64_t item = blockIdx.x; item < p.total_items; item += gridDim.x) {
    const SysInfo S = find_sys_item(p, item);
    const int64_t local = item - S.item0;
    const int64_t tt = local / S.n_chunks;
    const int64_t ch = local - tt * S.n_chunks;
    const int64_t t0 = tt * kBlock;
    const int64_t il = t0 + threadIdx.x;
    const bool valid = il < S.n;
    double Xi, Yi, Zi, Bi;
    load_blob(p.r, S.b0 + (valid ? il : 0), p.a, p.inva, Xi, Yi, Zi, Bi, DOM);

    Acc<RMW_PREC_F64> u[K][3], w[K][3];
    const int64_t s_begin = ch * S.chunk_len;
    const int64_t s_end = s_begin + S.chunk_len < S.n ? s_begin + S.chunk_len : S.n;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += kBlock) {
      const int cnt = (int)(s_end - s0 < kBlock ? s_end - s0 : kBlock);
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        const int64_t gj = S.b0 + s0 + threadIdx.x;
        double X, Y, Z, B;
        load_blob(p.r, gj, p.a, p.inva, X, Y, Z, B, DOM);
        s_pos[threadIdx.x] = make_double4(X, Y, Z, 0.0);
        stage_sources<MASK, K, double4, DOM>(p, gj, B, threadIdx.x, s_f, s_t);
      }
      __syncthreads();
      const bool diag = (s0 < t0 + kBlock) && (t0 < s0 + cnt);
      const int self_j = (int)(il - s0);
      for (int j = 0; j < cnt; ++j) {
        const double4 pj = s_pos[j];
        double rx = Xi - pj.x, ry = Yi - pj.y, rz = Zi - pj.z;
        if (!p.periodic) {
          pair_eval<MASK, K, RMW_PREC_F64, double, DOM>(rx, ry, rz, Zi, pj.z, s_f, s_t, j, diag && j == self_j, u,
                                                        w);
        } else {
          if (pbox[0]) rx -= image_count(rx, p.L[0]) * p.L[0];
          if (pbox[1]) ry -= image_count(ry, p.L[1]) * p.L[1];
          if (pbox[2]) rz -= image_count(rz, p.L[2]) * p.L[2];
          for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
            for (int by = -pbox[1]; by <= pbox[1]; ++by)
              for (int bz = -pbox[2]; bz <= pbox[2]; ++bz) {
                const bool central = (bx == 0) && (by == 0) && (bz == 0);
                pair_eval<MASK, K, RMW_PREC_F64, double, DOM>(rx + bx * p.L[0], ry + by * p.L[1],
                                                              rz + bz * p.L[2], Zi, pj.z, s_f, s_t, j,
                                                              diag && central && (j == self_j), u, w);
              }
        }
      }
    }
    if (valid) store_partials<MASK, K>(p, S, ch, il, u, w);
  }
}

// ---------------------------------------------------------------------------
// 3b. symmetric ("mutual") kernels.  The wall mobility is symmetric:
//     UF_ji = UF_ij^T, WT_ji = WT_ij^T, UT_ji = WF_ij^T, WF_ji = UT_ij^T, so
//     each unordered pair is evaluated once and serves both blobs.  Work
//     items are (target tile T, run of source tiles S >= T) of one system.
//     Each wave owns 64 targets (registers) and sweeps a source tile in
//     blocks of 64: at step s lane l takes source (l + s) mod 64 of the
//     block, so the 64 lanes touch 64 different sources (structure-of-arrays
//     LDS, conflict-free).  The source-side sum of a source travels with it:
//     every step the accumulator register moves one lane down (DPP
//     wave_rol:1, fused into the add), so after the 64 steps each source's
//     sum over the wave's targets sits in one lane.  After each block the
//     four waves' sums are added in a fixed order.  On the diagonal tile
//     (S == T) only sources above the target are taken.  Partial sums: one
//     slot per item (target side) and one per tile pair (S, T) (source side),
//     reduced in a fixed order.
// ---------------------------------------------------------------------------
// Largest workspace the automatic choice gives the symmetric kernels.
constexpr size_t kMutAutoBytes = size_t(2) << 30;

struct MutSys {
  int64_t b0, n, n_t;   // first blob, blobs, tiles of kBlock
  int64_t item0;        // first work item
  int64_t pi0, pj0;     // first target-side / source-side slot
};

struct MutParams {
  const double* r;
  const double* f;          // K right-hand sides rhs_stride apart, or null
  const double* t;
  int64_t rhs_stride;
  double a, inva, c1, c2;
  double* part_i;           // [slot][K][kBlock][OC]
  double* part_j;
  const MutSys* table;      // nullptr -> single
  MutSys single;
  int nsys;
  int64_t n_items, chunk_blocks;   // work item: a run of chunk_blocks 64-source blocks
  int periodic;             // bit d set when L_d > 0 (the PER instantiations)
  double L[3];              // periodic lengths / a
};

__device__ __forceinline__ MutSys mut_sys_of_item(const MutParams& p, int64_t item)
{
  if (!p.table) return p.single;
  int lo = 0, hi = p.nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (p.table[mid].item0 <= item) lo = mid; else hi = mid - 1;
  }
  return p.table[lo];
}

// Symmetric pair kernel.  R = float: FP32 pair terms (ACC: the target-side
// accumulator; source side: FP32 sums over 16 steps folded into FP64 in the
// blocked mode, an FP64 travelling accumulator otherwise); R = double: FP64.
// PER: pseudo-periodic images (mobility_numba.py:149-198).  The minimum
// image is odd in d and the box set {-1, 0, 1} is symmetric, so the images
// of (j, i) are the negated images of (i, j) and M_ji = M_ij^T holds image
// by image; each unordered pair takes all its images on both sides, sums
// are folded every step, and the self pair's images outside the central box
// are added by the reduce.
template <int MASK, int K, int ACC, typename R, int DOM = kDomWall, bool PER = false>
__global__ __launch_bounds__(kBlock) void mutual_kernel(MutParams p)
{
  using BT = BlockTraits<MASK>;
  constexpr bool UF = (MASK & kUF) != 0, UT = (MASK & kUT) != 0, WF = (MASK & kWF) != 0, WT = (MASK & kWT) != 0;
  constexpr int OC = BT::OC;
  constexpr bool kFoldJ = ACC == RMW_PREC_F32_BLOCKED;
  // several right-hand sides: form each pair's blocks once as 3x3 matrices
  constexpr bool kMat = K >= 2;
  using JT = typename std::conditional<kFoldJ, float, double>::type;
  // single-output kernels of one right-hand side store each 64-source
  // block twice (entry 128 b + m, m < 128, holds source 64 b + m mod 64), so
  // the rotated read of step s, entry 128 b + lane + s, needs no wrap
  // (1.5 % faster UF product at 32768 blobs; not for the two-output kernel)
  constexpr bool kDbl = !kMat && !PER && OC == 3;
  constexpr int kLds = kDbl ? 2 * kBlock : kBlock;
  __shared__ double s_x[kLds], s_y[kLds], s_z[kLds];
  __shared__ R s_zr[kLds];
  __shared__ R s_f[BT::NF ? K : 1][3][kLds];
  __shared__ R s_t[BT::NT ? K : 1][3][kLds];
  __shared__ double s_jd[kBlock / 64][K][OC][64];  // per-wave source-side sums of one block
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int il = threadIdx.x;
  // fold the FP32 sums every kSub sources (every source with the images)
  constexpr int kFold = PER ? 0 : kSub - 1;
  int pbox[3];
  for (int d = 0; d < 3; ++d) pbox[d] = PER ? (p.periodic >> d) & 1 : 0;

  for (int64_t item = blockIdx.x; item < p.n_items; item += gridDim.x) {
    const MutSys Sy = mut_sys_of_item(p, item);
    const int64_t local = item - Sy.item0;
    const int64_t T = pair_target_tile_blocks(local, Sy.n_t, p.chunk_blocks);
    const int64_t blk0 = 4 * T + (local - items_before_blocks(T, Sy.n_t, p.chunk_blocks)) * p.chunk_blocks;
    const int64_t blk1 = blk0 + p.chunk_blocks < 4 * Sy.n_t ? blk0 + p.chunk_blocks : 4 * Sy.n_t;
    const int64_t S0 = blk0 >> 2, S1 = (blk1 + 3) >> 2;
    const int64_t i = T * kBlock + il;  // within the system
    const bool valid = i < Sy.n;
    const int64_t gi = Sy.b0 + (valid ? i : 0);
    double Xi, Yi, Zi, Bi;
    load_blob(p.r, gi, p.a, p.inva, Xi, Yi, Zi, Bi, DOM);
    constexpr double pz = DOM == kDomInPlane ? 0.0 : 1.0;   // in plane: P v
    const R zi = (R)Zi;
    R fi[BT::NF ? K : 1][3], ti[BT::NT ? K : 1][3];
#pragma unroll
    for (int k = 0; k < K; ++k) {
      if constexpr (BT::NF) {
        const double* v = p.f + k * p.rhs_stride + 3 * gi;
        const double c = valid ? p.c1 * Bi : 0.0;
        fi[k][0] = (R)(c * v[0]); fi[k][1] = (R)(c * v[1]); fi[k][2] = (R)(pz * c * v[2]);
      }
      if constexpr (BT::NT) {
        const double* v = p.t + k * p.rhs_stride + 3 * gi;
        const double c = valid ? p.c2 * Bi : 0.0;
        ti[k][0] = (R)(c * v[0]); ti[k][1] = (R)(c * v[1]); ti[k][2] = (R)(pz * c * v[2]);
      }
    }
    Acc<ACC> u[K][3], w[K][3];

    for (int64_t S = S0; S < S1; ++S) {
      const int64_t s0 = S * kBlock;
      const int cnt = (int)(Sy.n - s0 < kBlock ? Sy.n - s0 : kBlock);
      // the 64-source blocks of this tile the item covers: only their
      // sources are staged (an item of one block reads a quarter of the tile)
      const int bb0 = S == S0 ? (int)(blk0 - 4 * S) : 0;
      const int bb1 = blk1 - 4 * S < kBlock / 64 ? (int)(blk1 - 4 * S) : kBlock / 64;
      __syncthreads();  // the previous tile has been consumed
      if ((il >> 6) >= bb0 && (il >> 6) < bb1) {
        // the source record of this thread (finite zeros beyond the system:
        // pairs not taken are masked through their geometry, make_geom_on,
        // and still multiply these values by zero coefficients)
        double X = 0.0, Y = 0.0, Z = 0.0;
        R fv[BT::NF ? K : 1][3], tv[BT::NT ? K : 1][3];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int c = 0; c < 3; ++c) {
            fv[k][c] = R(0);
            tv[k][c] = R(0);
          }
        if (il < cnt) {
          const int64_t gj = Sy.b0 + s0 + il;
          double B;
          load_blob(p.r, gj, p.a, p.inva, X, Y, Z, B, DOM);
#pragma unroll
          for (int k = 0; k < K; ++k) {
            if constexpr (BT::NF) {
              const double* v = p.f + k * p.rhs_stride + 3 * gj;
              const double c = p.c1 * B;
              fv[k][0] = (R)(c * v[0]); fv[k][1] = (R)(c * v[1]); fv[k][2] = (R)(pz * c * v[2]);
            }
            if constexpr (BT::NT) {
              const double* v = p.t + k * p.rhs_stride + 3 * gj;
              const double c = p.c2 * B;
              tv[k][0] = (R)(c * v[0]); tv[k][1] = (R)(c * v[1]); tv[k][2] = (R)(pz * c * v[2]);
            }
          }
        }
        const int e0 = kDbl ? 128 * (il >> 6) + (il & 63) : il;
#pragma unroll
        for (int e = e0; e <= e0 + (kDbl ? 64 : 0); e += 64) {
          s_x[e] = X; s_y[e] = Y; s_z[e] = Z; s_zr[e] = (R)Z;
#pragma unroll
          for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < 3; ++c) {
              if constexpr (BT::NF) s_f[k][c][e] = fv[k][c];
              if constexpr (BT::NT) s_t[k][c][e] = tv[k][c];
            }
        }
      }
      __syncthreads();
      const bool diag = S == T;
      // full off-diagonal tile pair: every lane's pair is taken, no masks.
      // Measured on MI355X with every kernel in one process, alternating:
      // the unmasked loop (8 fewer VALU per pair) makes the wall UF product
      // 17 % faster at 32768 blobs (fp32_blocked; 4 % in FP64) and UF + UT
      // 3 % faster; the two-output (grand) kernels keep the masked loop
      // only, as the second copy of their loop costs registers (K = 1:
      // 1 % slower; matrix form: 40 % slower).
      constexpr bool kUnmasked = DOM == kDomNoWall || DOM == kDomFreeSurface || (!kMat && OC == 3);
      const bool full = kUnmasked && !diag && cnt == kBlock && (T + 1) * kBlock <= Sy.n;
      for (int b = bb0; b < bb1 && 64 * b < cnt; ++b) {
        JT ja[K][OC];
#pragma unroll
        for (int k = 0; k < K; ++k)
#pragma unroll
          for (int c = 0; c < OC; ++c) ja[k][c] = JT(0);
        auto steps = [&](auto masked_c) {
        constexpr bool MASKED = decltype(masked_c)::value;
        for (int st = 0; st < 64; ++st) {
          const int jl = 64 * b + ((lane + st) & 63);   // the source, within the tile
          const int jr = kDbl ? 128 * b + lane + st : jl;   // its LDS entry
          const bool on = !MASKED || (valid && jl < cnt && (!diag || jl > il));
          const R zj = s_zr[jr];
          // matrix form (two waves per SIMD, too few to hide LDS latency):
          // the source's position, then its right-hand sides, read at the
          // top of the step in that order, so that the geometry waits for
          // the position only and the blocks' formation covers the rest
          // (2 % faster than reads where first used; FP32 pair terms only,
          // the FP64 kernels would lose a wave per SIMD to the registers)
          constexpr bool kEarly = kMat && !PER && sizeof(R) == 4;
          R fje[kEarly && BT::NF ? K : 1][3], tje[kEarly && BT::NT ? K : 1][3];
          double xje = 0.0, yje = 0.0, zje = 0.0;
          if constexpr (kEarly) {
            xje = s_x[jr]; yje = s_y[jr]; zje = s_z[jr];
            __builtin_amdgcn_sched_barrier(0);
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                if constexpr (BT::NF) fje[k][c] = s_f[k][c][jr];
                if constexpr (BT::NT) tje[k][c] = s_t[k][c][jr];
              }
            __builtin_amdgcn_sched_barrier(0);
          }
          // source-side sums move one lane per step (fused into the add; with
          // images, once per step before them)
          auto acc_j = [&](JT& a, JT x) {
            if constexpr (PER) a = a + x; else a = rol1(a) + x;
          };
          // one image d = (dx, dy, dz) of the pair, applied on both sides
          auto pair_apply = [&](R dx, R dy, R dz) {
          // a pair not taken contributes exact zeros (make_geom_on)
          const Geom<R> g = MASKED ? make_geom_on(dx, dy, dz, zi, zj, on) : make_geom(dx, dy, dz, zi, zj);
          CoefUF<R> cuf;
          CoefRot<R> cut, cwf;
          CoefWT<R> cwt;
          if constexpr (UF) cuf = coef_uf(g, zi, zj, DOM);
          if constexpr (UT || WF) { cut = coef_rot(g, zi, DOM); cwf = coef_rot(g, zj, DOM); }
          if constexpr (WT) cwt = coef_wt(g, DOM);
          if constexpr (kMat) {
            // the pair's blocks as 3x3 matrices (zero for pairs not taken)
            Mat3<R> Mu, Mc, Md, Mw;   // UF, UT, WF, WT of (i, j)
            if constexpr (UF) Mu = mat_uf(g, cuf);
            if constexpr (UT) Mc = mat_ut(g, cut);
            if constexpr (WF) Md = mat_wf(g, cwf);
            if constexpr (WT) Mw = mat_wt(g, cwt);
#pragma unroll
            for (int k = 0; k < K; ++k) {
              R fj[3], tj[3];
#pragma unroll
              for (int c = 0; c < 3; ++c) {
                if constexpr (BT::NF) fj[c] = kEarly ? fje[k][c] : s_f[k][c][jr];
                if constexpr (BT::NT) tj[c] = kEarly ? tje[k][c] : s_t[k][c][jr];
              }
#pragma unroll
              for (int r = 0; r < 3; ++r) {
                // target side: u_i += UF f_j + UT t_j, w_i += WF f_j + WT t_j
                // source side: u_j += UF^T f_i + WF^T t_i, w_j += UT^T f_i + WT^T t_i
                // (where the accumulator allows, its running sum seeds the
                // target-side chain: no separate add)
                constexpr bool kChain = Acc<ACC>::kChain;
                if constexpr (BT::HU) {
                  R x = R(0), y = R(0);
                  if constexpr (kChain) x = u[k][r].run();
                  if constexpr (UF) {
                    x = fma_(Mu.m[r][0], fj[0], fma_(Mu.m[r][1], fj[1], kChain ? fma_(Mu.m[r][2], fj[2], x) : Mu.m[r][2] * fj[2]));
                    y = fma_(Mu.m[0][r], fi[k][0], fma_(Mu.m[1][r], fi[k][1], Mu.m[2][r] * fi[k][2]));
                  }
                  if constexpr (UT)
                    x = fma_(Mc.m[r][0], tj[0], fma_(Mc.m[r][1], tj[1], r == 2 ? x : fma_(Mc.m[r][2], tj[2], x)));
                  if constexpr (WF)
                    y = fma_(Md.m[0][r], ti[k][0], fma_(Md.m[1][r], ti[k][1], r == 2 ? y : fma_(Md.m[2][r], ti[k][2], y)));
                  if constexpr (kChain) u[k][r].run() = x; else u[k][r].add(x);
                  acc_j(ja[k][r], (JT)y);
                }
                if constexpr (BT::HW) {
                  constexpr int o = BT::HU ? 3 : 0;
                  R x = R(0), y = R(0);
                  if constexpr (kChain) x = w[k][r].run();
                  if constexpr (WT) {
                    x = fma_(Mw.m[r][0], tj[0], fma_(Mw.m[r][1], tj[1], kChain ? fma_(Mw.m[r][2], tj[2], x) : Mw.m[r][2] * tj[2]));
                    y = fma_(Mw.m[0][r], ti[k][0], fma_(Mw.m[1][r], ti[k][1], Mw.m[2][r] * ti[k][2]));
                  }
                  if constexpr (WF)
                    x = fma_(Md.m[r][0], fj[0], fma_(Md.m[r][1], fj[1], r == 2 ? x : fma_(Md.m[r][2], fj[2], x)));
                  if constexpr (UT)
                    y = fma_(Mc.m[0][r], fi[k][0], fma_(Mc.m[1][r], fi[k][1], r == 2 ? y : fma_(Mc.m[2][r], fi[k][2], y)));
                  if constexpr (kChain) w[k][r].run() = x; else w[k][r].add(x);
                  acc_j(ja[k][o + r], (JT)y);
                }
              }
            }
          } else {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            // target side: u_i = UF f_j + UT t_j, w_i = WF f_j + WT t_j;
            // source side: u_j = UF^T f_i + WF^T t_i, w_j = UT^T f_i + WT^T t_i
            // (the first block present initialises each sum: no 0 + x)
            Vec3<R> ui, wi, uj, wj;
            auto put = [](Vec3<R>& s, const Vec3<R>& v, bool first) {
              if (first) { s = v; } else { s.x += v.x; s.y += v.y; s.z += v.z; }
            };
            if constexpr (UF) {
              put(ui, apply_uf(g, cuf, s_f[k][0][jr], s_f[k][1][jr], s_f[k][2][jr]), true);
              put(uj, apply_uf_t(g, cuf, fi[k][0], fi[k][1], fi[k][2]), true);
            }
            if constexpr (UT) {
              put(ui, apply_ut(g, cut, s_t[k][0][jr], s_t[k][1][jr], s_t[k][2][jr]), !UF);
              put(uj, apply_wf_t(g, cwf, ti[k][0], ti[k][1], ti[k][2]), !UF);
            }
            if constexpr (WF) {
              put(wi, apply_wf(g, cwf, s_f[k][0][jr], s_f[k][1][jr], s_f[k][2][jr]), true);
              put(wj, apply_ut_t(g, cut, fi[k][0], fi[k][1], fi[k][2]), true);
            }
            if constexpr (WT) {
              put(wi, apply_wt(g, cwt, s_t[k][0][jr], s_t[k][1][jr], s_t[k][2][jr]), !WF);
              put(wj, apply_wt_t(g, cwt, ti[k][0], ti[k][1], ti[k][2]), !WF);
            }
            if constexpr (BT::HU) {
              u[k][0].add(ui.x);
              u[k][1].add(ui.y);
              u[k][2].add(ui.z);
              acc_j(ja[k][0], (JT)uj.x);
              acc_j(ja[k][1], (JT)uj.y);
              acc_j(ja[k][2], (JT)uj.z);
            }
            if constexpr (BT::HW) {
              constexpr int o = BT::HU ? 3 : 0;
              w[k][0].add(wi.x);
              w[k][1].add(wi.y);
              w[k][2].add(wi.z);
              acc_j(ja[k][o], (JT)wj.x);
              acc_j(ja[k][o + 1], (JT)wj.y);
              acc_j(ja[k][o + 2], (JT)wj.z);
            }
          }
          }
          };
          if constexpr (kEarly) {
            if constexpr (sizeof(R) == 4)
              pair_apply((R)(Xi - xje), (R)(Yi - yje), (R)(Zi - zje));
            else
              pair_apply(Xi - xje, Yi - yje, Zi - zje);
          } else if constexpr (!PER) {
            if constexpr (sizeof(R) == 4)
              pair_apply((R)(Xi - s_x[jr]), (R)(Yi - s_y[jr]), (R)(Zi - s_z[jr]));
            else
              pair_apply(Xi - s_x[jr], Yi - s_y[jr], Zi - s_z[jr]);
          } else {
#pragma unroll
            for (int k = 0; k < K; ++k)
#pragma unroll
              for (int c = 0; c < OC; ++c) ja[k][c] = rol1(ja[k][c]);
            // minimum image of the FP64 difference, then the neighbouring boxes
            double ddx = Xi - s_x[jr], ddy = Yi - s_y[jr], ddz = Zi - s_z[jr];
            if (pbox[0]) ddx -= image_count(ddx, p.L[0]) * p.L[0];
            if (pbox[1]) ddy -= image_count(ddy, p.L[1]) * p.L[1];
            if (pbox[2]) ddz -= image_count(ddz, p.L[2]) * p.L[2];
            for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
              for (int by = -pbox[1]; by <= pbox[1]; ++by)
                for (int bz = -pbox[2]; bz <= pbox[2]; ++bz)
                  pair_apply((R)(ddx + bx * p.L[0]), (R)(ddy + by * p.L[1]), (R)(ddz + bz * p.L[2]));
          }
          if ((st & kFold) == kFold) {
            flush_all<K, ACC>(u, w);
            if constexpr (kFoldJ) {
              // lane l holds the sum of source (l + st) mod 64 of the block
              const int jw = (lane + st) & 63;
              // every read first, then the adds and writes: one LDS round
              // trip per fold (at the first fold of a block the reads return
              // stale values, a previous block's sums or uninitialised
              // memory, which the select drops)
              double prev[K][OC];
#pragma unroll
              for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < OC; ++c) prev[k][c] = s_jd[wv][k][c][jw];
              const bool first = st <= kFold;
#pragma unroll
              for (int k = 0; k < K; ++k)
#pragma unroll
                for (int c = 0; c < OC; ++c) {
                  s_jd[wv][k][c][jw] = (first ? 0.0 : prev[k][c]) + (double)ja[k][c];
                  ja[k][c] = JT(0);
                }
            }
          }
        }
        };
        if constexpr (!kUnmasked) {
          steps(std::true_type{});
        } else {
          if (full) steps(std::false_type{}); else steps(std::true_type{});
        }
        if constexpr (!kFoldJ) {
          const int jw = (lane + 63) & 63;
#pragma unroll
          for (int k = 0; k < K; ++k)
#pragma unroll
            for (int c = 0; c < OC; ++c) s_jd[wv][k][c][jw] = (double)ja[k][c];
        }
        __syncthreads();  // every wave's sums of this block are in LDS
        if (il < 64 && 64 * b + il < cnt) {
          const int64_t slot = Sy.pj0 + tri_index(S, T);
#pragma unroll
          for (int k = 0; k < K; ++k) {
            double* P = p.part_j + ((slot * K + k) * kBlock + 64 * b + il) * OC;
#pragma unroll
            for (int c = 0; c < OC; ++c)
              P[c] = ((s_jd[0][k][c][il] + s_jd[1][k][c][il]) + s_jd[2][k][c][il]) + s_jd[3][k][c][il];
          }
        }
        __syncthreads();  // s_jd is free for the next block
      }
      flush_all<K, ACC>(u, w);
    }
    if (valid) {
      const int64_t slot = Sy.pi0 + local;
#pragma unroll
      for (int k = 0; k < K; ++k) {
        double* P = p.part_i + ((slot * K + k) * kBlock + il) * OC;
        int o = 0;
        if constexpr (BT::HU) {
          P[0] = u[k][0].get(); P[1] = u[k][1].get(); P[2] = u[k][2].get();
          o = 3;
        }
        if constexpr (BT::HW) {
          P[o] = w[k][0].get(); P[o + 1] = w[k][1].get(); P[o + 2] = w[k][2].get();
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// 4. reduce: fixed-order chunk sum + FP64 self term + damping.
// ---------------------------------------------------------------------------
struct ReduceParams {
  const double* r;
  const double* f;
  const double* t;
  double* vel;
  double* omg;
  const double* partial;
  const SysInfo* table;
  SysInfo single;
  int nsys;
  int mask;
  int K;
  int64_t ntot;
  int64_t rhs_stride;  // doubles between RHS in f/t/vel/omg
  double a, inva, c1, c2;
  int domain;
};

__device__ inline SysInfo find_sys_blob(const ReduceParams& p, int64_t gi)
{
  if (!p.table) return p.single;
  int lo = 0, hi = p.nsys - 1;
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (p.table[mid].b0 <= gi) lo = mid; else hi = mid - 1;
  }
  return p.table[lo];
}

// Output component c (0..OC-1) of right-hand side k for blob gi: the FP64
// self term, damping B and the remaining 1/a of the rotational rows added
// to the pair sum.
__device__ __forceinline__ void finish_output(const double* __restrict__ r, const double* __restrict__ f,
                                              const double* __restrict__ t, double* __restrict__ vel,
                                              double* __restrict__ omg, int mask, int k, int64_t gi, int c,
                                              double sum, int64_t rhs_stride, double a, double inva, double c1,
                                              double c2, int dom, int periodic = 0, double L0 = 0.0,
                                              double L1 = 0.0, double L2 = 0.0)
{
  const bool HU = (mask & (kUF | kUT)) != 0;
  const bool HW = (mask & (kWF | kWT)) != 0;
  const bool wall = dom_has_wall(dom);
  const double z = r[3 * gi + 2];
  const double ze = wall && z <= a ? a : z;
  const double B = wall && z < a ? z * inva : 1.0;
  const double h = ze * inva;
  const double pz = dom == kDomInPlane ? 0.0 : 1.0;   // in plane: P v in, P u out
  const bool is_w = HW && (!HU || c >= 3);
  const int comp = c % 3;
  double fv[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
  if (f) {
    const double* v = f + k * rhs_stride + 3 * gi;
    fv[0] = c1 * B * v[0]; fv[1] = c1 * B * v[1]; fv[2] = pz * c1 * B * v[2];
  }
  if (t) {
    const double* v = t + k * rhs_stride + 3 * gi;
    tv[0] = c2 * B * v[0]; tv[1] = c2 * B * v[1]; tv[2] = pz * c2 * B * v[2];
  }
  double self = 0.0;
  // images of the blob itself outside the central box (pseudo-periodic
  // symmetric kernels): pairs at d = (b_x L_x, b_y L_y, b_z L_z), z_i = z_j
  for (int bx = -(periodic & 1); bx <= (periodic & 1); ++bx)
    for (int by = -((periodic >> 1) & 1); by <= ((periodic >> 1) & 1); ++by)
      for (int bz = -((periodic >> 2) & 1); bz <= ((periodic >> 2) & 1); ++bz) {
        if (bx == 0 && by == 0 && bz == 0) continue;
        const Geom<double> g = make_geom(bx * L0, by * L1, bz * L2, h, h);
        Vec3<double> v{0.0, 0.0, 0.0};
        if (!is_w) {
          if (mask & kUF) {
            const Vec3<double> q = apply_uf(g, coef_uf(g, h, h, dom), fv[0], fv[1], fv[2]);
            v.x += q.x; v.y += q.y; v.z += q.z;
          }
          if (mask & kUT) {
            const Vec3<double> q = apply_ut(g, coef_rot(g, h, dom), tv[0], tv[1], tv[2]);
            v.x += q.x; v.y += q.y; v.z += q.z;
          }
        } else {
          if (mask & kWF) {
            const Vec3<double> q = apply_wf(g, coef_rot(g, h, dom), fv[0], fv[1], fv[2]);
            v.x += q.x; v.y += q.y; v.z += q.z;
          }
          if (mask & kWT) {
            const Vec3<double> q = apply_wt(g, coef_wt(g, dom), tv[0], tv[1], tv[2]);
            v.x += q.x; v.y += q.y; v.z += q.z;
          }
        }
        self += comp == 0 ? v.x : (comp == 1 ? v.y : v.z);
      }
  if (!is_w) {
    if (mask & kUF) { Vec3<double> sv = self_uf_dom(dom, h, fv[0], fv[1], fv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    if (mask & kUT) { Vec3<double> sv = self_ut_dom(dom, h, tv[0], tv[1], tv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    vel[k * rhs_stride + 3 * gi + comp] = (comp == 2 ? pz : 1.0) * B * (sum + self);
  } else {
    if (mask & kWF) { Vec3<double> sv = self_wf_dom(dom, h, fv[0], fv[1], fv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    if (mask & kWT) { Vec3<double> sv = self_wt_dom(dom, h, tv[0], tv[1], tv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    omg[k * rhs_stride + 3 * gi + comp] = B * inva * (sum + self);
  }
}

__global__ __launch_bounds__(256) void reduce_kernel(ReduceParams p)
{
  // one thread per (rhs k, blob gi, output component c in [0, OC))
  const bool HU = (p.mask & (kUF | kUT)) != 0;
  const bool HW = (p.mask & (kWF | kWT)) != 0;
  const int OC = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
  const int64_t tid = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t per_k = p.ntot * OC;
  if (tid >= per_k * p.K) return;
  const int k = (int)(tid / per_k);
  const int64_t rem = tid - (int64_t)k * per_k;
  const int64_t gi = rem / OC;
  const int c = (int)(rem - gi * OC);
  const SysInfo S = find_sys_blob(p, gi);
  const int64_t il = gi - S.b0;
  const double* P = p.partial + S.part0 + ((int64_t)k * S.n + il) * OC + c;
  const int64_t cstride = (int64_t)p.K * S.n * OC;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
  int64_t ch = 0;
  for (; ch + 4 <= S.n_chunks; ch += 4) {
    s0 += P[(ch + 0) * cstride];
    s1 += P[(ch + 1) * cstride];
    s2 += P[(ch + 2) * cstride];
    s3 += P[(ch + 3) * cstride];
  }
  for (; ch < S.n_chunks; ++ch) s0 += P[ch * cstride];
  double sum = (s0 + s1) + (s2 + s3);

  finish_output(p.r, p.f, p.t, p.vel, p.omg, p.mask, k, gi, c, sum, p.rhs_stride, p.a, p.inva, p.c1, p.c2,
                p.domain);
}

// Reduce of the symmetric path: for blob gi of tile X of its system, the
// target-side slots of X's items, then the source-side slots (X, T) for
// T = 0..X, each with two running sums, in that order.
struct MutReduceParams {
  const double* r;
  const double* f;
  const double* t;
  double* vel;
  double* omg;
  const double* part_i;
  const double* part_j;
  const MutSys* table;
  MutSys single;
  int nsys;
  int mask;
  int K;
  int64_t ntot, chunk_blocks;
  int64_t rhs_stride;
  double a, inva, c1, c2;
  int domain;
  int periodic;             // bit d set when L_d > 0: add the self pair's outer images
  double L[3];              // periodic lengths / a
};

// Output o (component c of blob gi, RHS k) sums the target-side slots of the
// work items of its tile and the source-side slots (X, T), T <= X.  Eight
// groups of 32 threads share a block's 32 outputs: group g takes slots
// g, g + 8, ... of both lists (consecutive threads read consecutive
// outputs of one slot), then group 0 adds the eight partial sums in order.
constexpr int kRedGroups = 8, kRedOut = 256 / kRedGroups;

// PER: pseudo-periodic calls, which add the self pair's outer images; the
// image loop takes 168 VGPRs, so calls without images run an instantiation
// without it (70 VGPRs: seven waves per SIMD instead of three for this
// latency-bound loop).
template <bool PER>
__global__ __launch_bounds__(256) void mutual_reduce_kernel(MutReduceParams p)
{
  __shared__ double part[kRedGroups][kRedOut];
  const bool HU = (p.mask & (kUF | kUT)) != 0;
  const bool HW = (p.mask & (kWF | kWT)) != 0;
  const int OC = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
  const int g = threadIdx.x / kRedOut, ol = threadIdx.x - g * kRedOut;
  const int64_t tid = (int64_t)blockIdx.x * kRedOut + ol;
  const int64_t per_k = p.ntot * OC;
  const bool valid = tid < per_k * p.K;
  double sum = 0.0;
  int k = 0;
  int64_t gi = 0;
  int c = 0;
  if (valid) {
    k = (int)(tid / per_k);
    const int64_t rem = tid - (int64_t)k * per_k;
    gi = rem / OC;
    c = (int)(rem - gi * OC);
    MutSys Sy = p.single;
    if (p.table) {
      int lo = 0, hi = p.nsys - 1;
      while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (p.table[mid].b0 <= gi) lo = mid; else hi = mid - 1;
      }
      Sy = p.table[lo];
    }
    const int64_t il = gi - Sy.b0;
    const int64_t X = il / kBlock, l = il - X * kBlock;
    const int64_t it0 = items_before_blocks(X, Sy.n_t, p.chunk_blocks);
    const int64_t it1 = items_before_blocks(X + 1, Sy.n_t, p.chunk_blocks);
    const int64_t step = (int64_t)p.K * kBlock * OC;
    const double* Pi = p.part_i + (((Sy.pi0 + it0) * p.K + k) * kBlock + l) * OC + c;
    const double* Pj = p.part_j + (((Sy.pj0 + tri_index(X, 0)) * p.K + k) * kBlock + l) * OC + c;
    // four loads in flight per thread (the loop is latency-bound: one
    // dependent load per slot took ~2x longer at 32768 blobs)
    auto strided_sum = [&](const double* P, int64_t cnt) {
      double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
      int64_t m = g;
      for (; m + 3 * kRedGroups < cnt; m += 4 * kRedGroups) {
        a0 += P[m * step];
        a1 += P[(m + kRedGroups) * step];
        a2 += P[(m + 2 * kRedGroups) * step];
        a3 += P[(m + 3 * kRedGroups) * step];
      }
      for (; m < cnt; m += kRedGroups) a0 += P[m * step];
      return (a0 + a1) + (a2 + a3);
    };
    sum = strided_sum(Pi, it1 - it0) + strided_sum(Pj, X + 1);
  }
  part[g][ol] = sum;
  __syncthreads();
  if (g == 0 && valid) {
    double s = part[0][ol];
#pragma unroll
    for (int q = 1; q < kRedGroups; ++q) s += part[q][ol];
    finish_output(p.r, p.f, p.t, p.vel, p.omg, p.mask, k, gi, c, s, p.rhs_stride, p.a, p.inva, p.c1, p.c2,
                  p.domain, PER ? p.periodic : 0, p.L[0], p.L[1], p.L[2]);
  }
}

// ---------------------------------------------------------------------------
// 5. optional check of the outputs (RMW_FLAG_CHECK_FINITE).
// ---------------------------------------------------------------------------
__global__ void zero_count_kernel(unsigned int* count)
{
  if (threadIdx.x == 0) count[0] = 0u;
}

// count += number of inf / NaN entries among k rows of len doubles, rows
// `stride` apart.
__global__ __launch_bounds__(256) void count_nonfinite_kernel(const double* __restrict__ x, int64_t k, int64_t len,
                                                              int64_t stride, unsigned int* count)
{
  unsigned int bad = 0;
  const int64_t total = k * len;
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t row = idx / len;
    const double v = x[row * stride + (idx - row * len)];
    bad += isfinite(v) ? 0u : 1u;
  }
  if (bad) atomicAdd(count, bad);
}

// ---------------------------------------------------------------------------
// Host side: plan, workspace layout, dispatch.
// ---------------------------------------------------------------------------
// Vector with room for one element inside the object: the plan of a
// single suspension (every call but the ragged ones) uses no heap memory.
template <typename T>
class SmallVec {
 public:
  void push_back(const T& v)
  {
    if (n_ == 0 && heap_.empty()) {
      one_ = v;
    } else {
      if (heap_.empty()) heap_.push_back(one_);
      heap_.push_back(v);
    }
    ++n_;
  }
  size_t size() const { return n_; }
  bool empty() const { return n_ == 0; }
  void clear()
  {
    n_ = 0;
    heap_.clear();
  }
  const T* data() const { return heap_.empty() ? &one_ : heap_.data(); }
  const T& operator[](size_t i) const { return data()[i]; }

 private:
  T one_{};
  std::vector<T> heap_;
  size_t n_ = 0;
};

struct Plan {
  int domain = kDomWall;
  int mask = 0, prec = 0, kmax = 1;
  int slots = 0;             // resident pair-kernel workgroups on the device
  int64_t ntot = 0;
  SmallVec<SysInfo> sys;     // non-empty systems only
  int64_t total_items = 0;
  size_t partial_doubles = 0;
  size_t off_partial = 0, off_table = 0, total = 0;
  // symmetric path (UF or all four blocks, no periodic images)
  bool mut_ok = false;
  int kmax_mut = 1;
  int64_t mut_items = 0, chunk_blocks = 4;   // symmetric work item: chunk_blocks 64-source blocks
  SmallVec<MutSys> msys;     // non-empty systems only
  size_t off_pi = 0, off_pj = 0, off_mtable = 0;
  // RMW_FLAG_CHECK_FINITE: a counter of non-finite outputs after the rest
  bool check_finite = false;
  size_t off_status = 0;
};

static bool valid_mask(int mask) { return mask > 0 && mask < 16; }

// Right-hand sides per pair-kernel pass: 4, 2 or 1, at most the tuned cap.
static int group_k(int k, int cap)
{
  const int g = k >= 4 ? 4 : (k >= 2 ? 2 : 1);
  return cap > 0 && cap < g ? (cap >= 2 ? 2 : 1) : g;
}

static int rhs_cap(const rmw_tuning* tuning) { return tuning ? tuning->rhs_group : 0; }

constexpr int kChunkQuantum = 64;  // chunk lengths are multiples of this

template <int MASK, int K>
static const void* pair_ptr(int prec)
{
  if (prec == RMW_PREC_F32_F64ACC) return reinterpret_cast<const void*>(&pair_kernel_f32<MASK, K, RMW_PREC_F32_F64ACC>);
  if (prec == RMW_PREC_F32_BLOCKED) return reinterpret_cast<const void*>(&pair_kernel_f32<MASK, K, RMW_PREC_F32_BLOCKED>);
  if (prec == RMW_PREC_F32_KAHAN) return reinterpret_cast<const void*>(&pair_kernel_f32<MASK, K, RMW_PREC_F32_KAHAN>);
  return reinterpret_cast<const void*>(&pair_kernel_f64<MASK, K>);
}

template <int MASK>
static const void* pair_ptr_k(int prec, int K)
{
  return K == 4 ? pair_ptr<MASK, 4>(prec) : (K == 2 ? pair_ptr<MASK, 2>(prec) : pair_ptr<MASK, 1>(prec));
}

static const void* pair_kernel_ptr(int mask, int prec, int K)
{
  switch (mask) {
    case 1: return pair_ptr_k<1>(prec, K);
    case 2: return pair_ptr_k<2>(prec, K);
    case 3: return pair_ptr_k<3>(prec, K);
    case 4: return pair_ptr_k<4>(prec, K);
    case 5: return pair_ptr_k<5>(prec, K);
    case 8: return pair_ptr_k<8>(prec, K);
    case 10: return pair_ptr_k<10>(prec, K);
    case 12: return pair_ptr_k<12>(prec, K);
    default: return pair_ptr_k<15>(prec, K);
  }
}

// Workgroups of the (mask, prec, K) pair kernel resident on the whole
// device at once (register and LDS limited), cached per device.
static int resident_slots(int mask, int prec, int K)
{
  // fixed table (no heap use on the launch path): one entry per
  // (device, mask, prec, K) seen, first come first kept
  static std::mutex mu;
  static std::pair<int64_t, int> cache[256];
  static int used = 0;
  const DeviceProps& dp = device_props();
  const int64_t key = ((((int64_t)dp.device * 16 + mask) * 8 + prec) * 8) + K;
  {
    std::lock_guard<std::mutex> lk(mu);
    for (int i = 0; i < used; ++i)
      if (cache[i].first == key) return cache[i].second;
  }
  int per_cu = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, pair_kernel_ptr(mask, prec, K), kBlock, 0) !=
          hipSuccess ||
      per_cu < 1)
    per_cu = 1;
  const int slots = per_cu * std::max(dp.cu_count, 1);
  std::lock_guard<std::mutex> lk(mu);
  if (used < 256) cache[used++] = {key, slots};
  return slots;
}

// Built-in chunk length: ~16 work items per CU over the whole batch (small
// items keep the workgroups of a CU out of step and balance the persistent
// loop of ragged batches), at least one LDS tile.  The autotuner refines it.
static int64_t default_chunk(double pairs, int cu_count)
{
  const double c = pairs / (16.0 * std::max(cu_count, 1) * kBlock);
  return std::max<int64_t>(kBlock, (int64_t)std::ceil(c / kChunkQuantum) * kChunkQuantum);
}

// offsets: host array of nsys + 1 blob offsets (offsets[0] = 0).  periodic:
// the call has pseudo-periodic images (the symmetric wall kernels then take
// one right-hand side per pass, so their partial sums are sized for one).
static Plan make_plan(int mask, int prec, const int64_t* offsets, int nsys, int k,
                      const rmw_tuning* tuning, bool ragged, int domain = kDomWall, bool periodic = false)
{
  RMW_REQUIRE(valid_mask(mask), "mask must be a non-empty subset of UF|UT|WF|WT");
  RMW_REQUIRE(domain >= kDomWall && domain <= kDomFreeSurface, "unknown domain");
  RMW_REQUIRE(domain != kDomInPlane || (mask & ~(kUF | kUT)) == 0,
              "the in-plane mobility has the UF and UT blocks only");
  RMW_REQUIRE(domain != kDomFreeSurface || mask == kUF, "the free-surface mobility has the UF block only");
  RMW_REQUIRE(prec == RMW_PREC_F32_F64ACC || prec == RMW_PREC_F32_KAHAN || prec == RMW_PREC_F64 ||
                  prec == RMW_PREC_F32_BLOCKED,
              "unknown precision mode");
  RMW_REQUIRE(k >= 1, "k must be >= 1");
  Plan pl;
  pl.mask = mask;
  pl.prec = prec;
  pl.domain = domain;
  RMW_REQUIRE(!tuning || tuning->rhs_group == 0 || tuning->rhs_group == 1 || tuning->rhs_group == 2 ||
                  tuning->rhs_group == 4,
              "rhs_group must be 0, 1, 2 or 4");
  pl.kmax = group_k(k, domain == kDomWall ? rhs_cap(tuning) : 1);
  RMW_REQUIRE(offsets != nullptr && nsys >= 1 && offsets[0] == 0, "offsets must start at 0");
  for (int s = 0; s < nsys; ++s) RMW_REQUIRE(offsets[s + 1] >= offsets[s], "offsets must be non-decreasing");
  pl.ntot = offsets[nsys];

  double pairs = 0.0;
  int64_t nmax = 0;
  for (int s = 0; s < nsys; ++s) {
    double n = (double)(offsets[s + 1] - offsets[s]);
    pairs += n * n;
    nmax = std::max<int64_t>(nmax, offsets[s + 1] - offsets[s]);
  }
  pl.slots = resident_slots(mask, prec, pl.kmax);
  int64_t chunk = tuning && tuning->chunk_len > 0 ? tuning->chunk_len : 0;
  if (chunk <= 0) chunk = default_chunk(pairs, device_props().cu_count);
  chunk = std::max<int64_t>(kChunkQuantum, (chunk + kChunkQuantum - 1) / kChunkQuantum * kChunkQuantum);
  if (nmax > 0) chunk = std::min<int64_t>(chunk, (nmax + kChunkQuantum - 1) / kChunkQuantum * kChunkQuantum);

  const bool HU = (mask & (kUF | kUT)) != 0, HW = (mask & (kWF | kWT)) != 0;
  const int OC = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
  int64_t item = 0, part = 0;
  for (int s = 0; s < nsys; ++s) {
    SysInfo S;
    S.b0 = offsets[s];
    S.n = offsets[s + 1] - offsets[s];
    if (S.n == 0) continue;
    S.n_tt = (S.n + kBlock - 1) / kBlock;
    S.chunk_len = chunk;
    S.n_chunks = (S.n + chunk - 1) / chunk;
    S.item0 = item;
    S.part0 = part;
    item += S.n_tt * S.n_chunks;
    part += S.n_chunks * (int64_t)pl.kmax * S.n * OC;
    pl.sys.push_back(S);
  }
  pl.total_items = item;
  pl.partial_doubles = (size_t)part;

  // Symmetric layout, unless switched off: the UF block alone, UF + UT, the
  // single UT, WF and WT blocks (UT_ji = WF_ij^T, WT_ji = WT_ij^T: a pair's
  // source side of one of them is the transposed other) and all four
  // blocks, as far as the boundary serves them (the free surface: UF only;
  // in plane: UF, UT, UF + UT); several right-hand sides per pass for UF
  // and all four blocks above the wall, one otherwise
  const int sym = tuning ? tuning->symmetric : 0;
  const bool single_rot = mask == kUT || mask == kWF || mask == kWT;
  pl.mut_ok = (mask == kUF || ((mask == (kUF | kUT) || single_rot) && domain != kDomFreeSurface) ||
               (mask == (kUF | kUT | kWF | kWT) && (domain == kDomWall || domain == kDomNoWall))) &&
              pl.ntot > 0 && sym >= 0;
  RMW_REQUIRE(!(sym > 0 && !pl.mut_ok),
              "symmetric evaluation needs the UF block alone, UF + UT, one of UT, WF, WT, or all four blocks");
  int64_t tile_pairs = 0;
  for (int s = 0; s < nsys; ++s) {
    const int64_t nt = (offsets[s + 1] - offsets[s] + kBlock - 1) / kBlock;
    tile_pairs += nt * (nt + 1) / 2;
  }
  // automatic choice: only with enough tile pairs to fill the device
  if (pl.mut_ok && sym == 0) pl.mut_ok = tile_pairs >= device_props().cu_count;
  size_t mut_total = 0;
  if (pl.mut_ok) {
    // up to 4 right-hand sides per pass (each pair's blocks formed once as
    // 3x3 matrices and applied to all of them), fewer when the tuning caps it
    const int rc = rhs_cap(tuning);
    pl.kmax_mut = domain == kDomWall && mask != (kUF | kUT) && !single_rot && !periodic
                      ? group_k(k, rc > 0 ? std::min(rc, 4) : 4)
                      : 1;
    if (tuning && tuning->chunk_len > 0) {
      // 64 or 128 sources: a quarter or half of a tile pair per work item
      // (finer work for suspensions of few tiles); else whole source tiles
      const int64_t blocks = (tuning->chunk_len + 63) / 64;
      pl.chunk_blocks = blocks <= 2 ? blocks : (blocks + 3) / 4 * 4;
    } else {
      const DeviceProps& dp = device_props();
      pl.chunk_blocks = 4 * std::max<int64_t>(1, tile_pairs / (16 * (int64_t)std::max(dp.cu_count, 1)));
    }
    int64_t item = 0, pi = 0, pj = 0;
    for (int s = 0; s < nsys; ++s) {
      MutSys M;
      M.b0 = offsets[s];
      M.n = offsets[s + 1] - offsets[s];
      if (M.n == 0) continue;
      M.n_t = (M.n + kBlock - 1) / kBlock;
      M.item0 = item;
      M.pi0 = pi;
      M.pj0 = pj;
      const int64_t its = items_before_blocks(M.n_t, M.n_t, pl.chunk_blocks);
      item += its;
      pi += its;
      pj += M.n_t * (M.n_t + 1) / 2;
      pl.msys.push_back(M);
    }
    pl.mut_items = item;
    const int OCm = (mask & (kWF | kWT)) ? 6 : 3;
    if (rc <= 0 && pl.kmax_mut > 2 &&
        (size_t)(pi + pj) * pl.kmax_mut * kBlock * OCm * sizeof(double) > kMutAutoBytes)
      pl.kmax_mut = group_k(k, 2);   // partial sums of 4 RHS too large: 2 per pass
    size_t o = 0;
    pl.off_pi = o; o = align_up(o + (size_t)pi * pl.kmax_mut * kBlock * OCm * sizeof(double));
    pl.off_pj = o; o = align_up(o + (size_t)pj * pl.kmax_mut * kBlock * OCm * sizeof(double));
    if (pl.msys.size() > 1) { pl.off_mtable = o; o = align_up(o + pl.msys.size() * sizeof(MutSys)); }
    mut_total = o;
    // the source-side slots grow as N^2 / kBlock: the automatic choice
    // keeps them within kMutAutoBytes and uses the ordered kernel beyond
    if (sym == 0 && mut_total > kMutAutoBytes) {
      pl.mut_ok = false;
      pl.msys.clear();
      mut_total = 0;
    }
  }

  size_t o = 0;
  pl.off_partial = o; o = align_up(o + pl.partial_doubles * sizeof(double));
  if (ragged) { pl.off_table = o; o = align_up(o + pl.sys.size() * sizeof(SysInfo)); }
  pl.total = std::max(o, mut_total);
  pl.check_finite = tuning && (tuning->flags & RMW_FLAG_CHECK_FINITE) != 0;
  if (pl.check_finite) {
    pl.off_status = align_up(pl.total);
    pl.total = pl.off_status + RMW_STATUS_BYTES;
  }
  return pl;
}

// Pair-kernel launch for one (MASK, K) instantiation.
template <int MASK, int K>
static void launch_pair(int prec, const PairParams& pp, int grid, hipStream_t st)
{
  if (prec == RMW_PREC_F32_F64ACC)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, K, RMW_PREC_F32_F64ACC>), dim3(grid), dim3(kBlock), 0, st, pp);
  else if (prec == RMW_PREC_F32_BLOCKED)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, K, RMW_PREC_F32_BLOCKED>), dim3(grid), dim3(kBlock), 0, st, pp);
  else if (prec == RMW_PREC_F32_KAHAN)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, K, RMW_PREC_F32_KAHAN>), dim3(grid), dim3(kBlock), 0, st, pp);
  else
    hipLaunchKernelGGL((pair_kernel_f64<MASK, K>), dim3(grid), dim3(kBlock), 0, st, pp);
}

template <int MASK>
static void launch_pair_k(int prec, int K, const PairParams& pp, int grid, hipStream_t st)
{
  if (K == 4) launch_pair<MASK, 4>(prec, pp, grid, st);
  else if (K == 2) launch_pair<MASK, 2>(prec, pp, grid, st);
  else launch_pair<MASK, 1>(prec, pp, grid, st);
}

// Pair kernels of the no-wall, in-plane and free-surface products (one
// right-hand side per pass).
template <int MASK, int DOM>
static void launch_pair_dom(int prec, const PairParams& pp, int grid, hipStream_t st)
{
  const dim3 g(grid), b(kBlock);
  if (prec == RMW_PREC_F32_F64ACC)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, 1, RMW_PREC_F32_F64ACC, DOM>), g, b, 0, st, pp);
  else if (prec == RMW_PREC_F32_BLOCKED)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, 1, RMW_PREC_F32_BLOCKED, DOM>), g, b, 0, st, pp);
  else if (prec == RMW_PREC_F32_KAHAN)
    hipLaunchKernelGGL((pair_kernel_f32<MASK, 1, RMW_PREC_F32_KAHAN, DOM>), g, b, 0, st, pp);
  else
    hipLaunchKernelGGL((pair_kernel_f64<MASK, 1, DOM>), g, b, 0, st, pp);
}

static void dispatch_pair_domain(int domain, int mask, int prec, const PairParams& pp, int grid, hipStream_t st)
{
  if (domain == kDomFreeSurface) return launch_pair_dom<kUF, kDomFreeSurface>(prec, pp, grid, st);
  if (domain == kDomInPlane) {
    switch (mask) {
      case 1: return launch_pair_dom<1, kDomInPlane>(prec, pp, grid, st);
      case 2: return launch_pair_dom<2, kDomInPlane>(prec, pp, grid, st);
      default: return launch_pair_dom<3, kDomInPlane>(prec, pp, grid, st);
    }
  }
  switch (mask) {
    case 1: return launch_pair_dom<1, kDomNoWall>(prec, pp, grid, st);
    case 2: return launch_pair_dom<2, kDomNoWall>(prec, pp, grid, st);
    case 3: return launch_pair_dom<3, kDomNoWall>(prec, pp, grid, st);
    case 4: return launch_pair_dom<4, kDomNoWall>(prec, pp, grid, st);
    case 5: return launch_pair_dom<5, kDomNoWall>(prec, pp, grid, st);
    case 8: return launch_pair_dom<8, kDomNoWall>(prec, pp, grid, st);
    case 10: return launch_pair_dom<10, kDomNoWall>(prec, pp, grid, st);
    case 12: return launch_pair_dom<12, kDomNoWall>(prec, pp, grid, st);
    default: return launch_pair_dom<15, kDomNoWall>(prec, pp, grid, st);
  }
}

static void dispatch_pair(int mask, int prec, int K, const PairParams& pp, int grid, hipStream_t st)
{
  switch (mask) {
    case 1: launch_pair_k<1>(prec, K, pp, grid, st); break;
    case 2: launch_pair_k<2>(prec, K, pp, grid, st); break;
    case 3: launch_pair_k<3>(prec, K, pp, grid, st); break;
    case 4: launch_pair_k<4>(prec, K, pp, grid, st); break;
    case 5: launch_pair_k<5>(prec, K, pp, grid, st); break;
    case 8: launch_pair_k<8>(prec, K, pp, grid, st); break;
    case 10: launch_pair_k<10>(prec, K, pp, grid, st); break;
    case 12: launch_pair_k<12>(prec, K, pp, grid, st); break;
    default: launch_pair_k<15>(prec, K, pp, grid, st); break;
  }
}

// Masks with their own instantiation; the rest are rejected.
static int kernel_mask(int mask)
{
  switch (mask) {
    case 1: case 2: case 3: case 4: case 5: case 8: case 10: case 12: case 15: return mask;
    default: return 0;
  }
}

// Pinned host staging ring for small stream-ordered uploads (the ragged
// work table).  A slot is reused only after the copy that last read it has
// completed, so callers may free their host data right after the call.
struct StageSlot {
  void* host = nullptr;
  size_t cap = 0;
  hipEvent_t done = nullptr;
  bool pending = false;
};
static std::mutex g_stage_mu;
static StageSlot g_stage[8];
static int g_stage_next = 0;

static void stage_upload(void* dst, const void* src, size_t bytes, hipStream_t st)
{
  if (bytes == 0) return;
  std::lock_guard<std::mutex> lk(g_stage_mu);
  StageSlot& s = g_stage[g_stage_next];
  g_stage_next = (g_stage_next + 1) % 8;
  if (s.pending) {
    RMW_CHECK_HIP(hipEventSynchronize(s.done));
    s.pending = false;
  }
  if (s.cap < bytes) {
    if (s.host) RMW_CHECK_HIP(hipHostFree(s.host));
    s.host = nullptr;
    size_t cap = align_up(bytes, 4096);
    RMW_CHECK_HIP(hipHostMalloc(&s.host, cap, hipHostMallocDefault));
    s.cap = cap;
  }
  if (!s.done) RMW_CHECK_HIP(hipEventCreateWithFlags(&s.done, hipEventDisableTiming));
  std::memcpy(s.host, src, bytes);
  RMW_CHECK_HIP(hipMemcpyAsync(dst, s.host, bytes, hipMemcpyHostToDevice, st));
  RMW_CHECK_HIP(hipEventRecord(s.done, st));
  s.pending = true;
}

void stage_upload_public(void* dst, const void* src, size_t bytes, hipStream_t st)
{
  stage_upload(dst, src, bytes, st);
}

static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

template <int MASK, int DOM>
static void dispatch_mutual_dom(int prec, const MutParams& q, int grid, hipStream_t st)
{
  const dim3 g(grid), b(kBlock);
  if (prec == RMW_PREC_F32_BLOCKED)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_BLOCKED, float, DOM>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_F64ACC)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_F64ACC, float, DOM>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_KAHAN)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_KAHAN, float, DOM>), g, b, 0, st, q);
  else
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F64, double, DOM>), g, b, 0, st, q);
}

template <int MASK>
static void dispatch_mutual_m(int prec, int K, const MutParams& q, int grid, hipStream_t st)
{
  const dim3 g(grid), b(kBlock);
  if (prec == RMW_PREC_F32_BLOCKED) {
    if (K == 4) hipLaunchKernelGGL((mutual_kernel<MASK, 4, RMW_PREC_F32_BLOCKED, float>), g, b, 0, st, q);
    else if (K == 2) hipLaunchKernelGGL((mutual_kernel<MASK, 2, RMW_PREC_F32_BLOCKED, float>), g, b, 0, st, q);
    else hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_BLOCKED, float>), g, b, 0, st, q);
  } else if (prec == RMW_PREC_F32_F64ACC) {
    if (K == 4) hipLaunchKernelGGL((mutual_kernel<MASK, 4, RMW_PREC_F32_F64ACC, float>), g, b, 0, st, q);
    else if (K == 2) hipLaunchKernelGGL((mutual_kernel<MASK, 2, RMW_PREC_F32_F64ACC, float>), g, b, 0, st, q);
    else hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_F64ACC, float>), g, b, 0, st, q);
  } else if (prec == RMW_PREC_F32_KAHAN) {
    if (K == 4) hipLaunchKernelGGL((mutual_kernel<MASK, 4, RMW_PREC_F32_KAHAN, float>), g, b, 0, st, q);
    else if (K == 2) hipLaunchKernelGGL((mutual_kernel<MASK, 2, RMW_PREC_F32_KAHAN, float>), g, b, 0, st, q);
    else hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_KAHAN, float>), g, b, 0, st, q);
  } else {
    if (K == 4) hipLaunchKernelGGL((mutual_kernel<MASK, 4, RMW_PREC_F64, double>), g, b, 0, st, q);
    else if (K == 2) hipLaunchKernelGGL((mutual_kernel<MASK, 2, RMW_PREC_F64, double>), g, b, 0, st, q);
    else hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F64, double>), g, b, 0, st, q);
  }
}

// Pseudo-periodic images (wall domain, one right-hand side per pass).
template <int MASK>
static void dispatch_mutual_per_m(int prec, const MutParams& q, int grid, hipStream_t st)
{
  const dim3 g(grid), b(kBlock);
  if (prec == RMW_PREC_F32_BLOCKED)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_BLOCKED, float, kDomWall, true>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_F64ACC)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_F64ACC, float, kDomWall, true>), g, b, 0, st, q);
  else if (prec == RMW_PREC_F32_KAHAN)
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F32_KAHAN, float, kDomWall, true>), g, b, 0, st, q);
  else
    hipLaunchKernelGGL((mutual_kernel<MASK, 1, RMW_PREC_F64, double, kDomWall, true>), g, b, 0, st, q);
}

static void dispatch_mutual_periodic(int mask, int prec, const MutParams& q, int grid, hipStream_t st)
{
  if (mask == kUF) dispatch_mutual_per_m<kUF>(prec, q, grid, st);
  else if (mask == (kUF | kUT)) dispatch_mutual_per_m<kUF | kUT>(prec, q, grid, st);
  else if (mask == kUT) dispatch_mutual_per_m<kUT>(prec, q, grid, st);
  else if (mask == kWF) dispatch_mutual_per_m<kWF>(prec, q, grid, st);
  else if (mask == kWT) dispatch_mutual_per_m<kWT>(prec, q, grid, st);
  else dispatch_mutual_per_m<kUF | kUT | kWF | kWT>(prec, q, grid, st);
}

static void dispatch_mutual(int domain, int mask, int prec, int K, const MutParams& q, int grid, hipStream_t st)
{
  if (domain == kDomNoWall) {
    if (mask == kUF) dispatch_mutual_dom<kUF, kDomNoWall>(prec, q, grid, st);
    else if (mask == (kUF | kUT)) dispatch_mutual_dom<kUF | kUT, kDomNoWall>(prec, q, grid, st);
    else if (mask == kUT) dispatch_mutual_dom<kUT, kDomNoWall>(prec, q, grid, st);
    else if (mask == kWF) dispatch_mutual_dom<kWF, kDomNoWall>(prec, q, grid, st);
    else if (mask == kWT) dispatch_mutual_dom<kWT, kDomNoWall>(prec, q, grid, st);
    else dispatch_mutual_dom<kUF | kUT | kWF | kWT, kDomNoWall>(prec, q, grid, st);
  } else if (domain == kDomInPlane) {
    if (mask == kUF) dispatch_mutual_dom<kUF, kDomInPlane>(prec, q, grid, st);
    else if (mask == (kUF | kUT)) dispatch_mutual_dom<kUF | kUT, kDomInPlane>(prec, q, grid, st);
    else dispatch_mutual_dom<kUT, kDomInPlane>(prec, q, grid, st);
  } else if (domain == kDomFreeSurface) {
    dispatch_mutual_dom<kUF, kDomFreeSurface>(prec, q, grid, st);
  } else if (mask == kUF) {
    dispatch_mutual_m<kUF>(prec, K, q, grid, st);
  } else if (mask == (kUF | kUT)) {
    // u = UF f + UT t: target side UF, UT; source side UF^T f_i + WF^T t_i
    dispatch_mutual_dom<kUF | kUT, kDomWall>(prec, q, grid, st);
  } else if (mask == kUT) {
    // u = UT t: target side UT t_j, source side WF^T t_i
    dispatch_mutual_dom<kUT, kDomWall>(prec, q, grid, st);
  } else if (mask == kWF) {
    // w = WF f: target side WF f_j, source side UT^T f_i
    dispatch_mutual_dom<kWF, kDomWall>(prec, q, grid, st);
  } else if (mask == kWT) {
    dispatch_mutual_dom<kWT, kDomWall>(prec, q, grid, st);
  } else {
    dispatch_mutual_m<kUF | kUT | kWF | kWT>(prec, K, q, grid, st);
  }
}

// Set while the calling thread runs rmw_wall_rpy_apply_noalloc: its
// launches take no timing records (each record creates two HIP events).
static thread_local bool tl_no_records = false;

// Pair-kernel timing records (rmw_profile_*).
void prof_begin(ProfRec& rec, bool& on, hipStream_t st, int kind)
{
  rec.kind = kind;
  {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    on = g_prof_on && !tl_no_records;
  }
  if (on) {
    RMW_CHECK_HIP(hipEventCreate(&rec.a));
    RMW_CHECK_HIP(hipEventCreate(&rec.b));
    RMW_CHECK_HIP(hipEventRecord(rec.a, st));
  }
}

void prof_end(const ProfRec& rec, bool on, hipStream_t st)
{
  if (on) {
    RMW_CHECK_HIP(hipEventRecord(rec.b, st));
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof.push_back(rec);
  }
}

static void run_mutual(const Plan& pl, const double* r, const double* f, const double* t, double* vel,
                       double* omg, int k, int64_t stride, double a, double inva, double c1, double c2,
                       const rmw_tuning* tuning, char* base, hipStream_t st, int periodic, const double* Lp)
{
  const bool HU = (pl.mask & (kUF | kUT)) != 0, HW = (pl.mask & (kWF | kWT)) != 0;
  const bool NF = (pl.mask & (kUF | kWF)) != 0, NT = (pl.mask & (kUT | kWT)) != 0;
  MutSys* dtable = nullptr;
  if (pl.msys.size() > 1) {
    dtable = (MutSys*)(base + pl.off_mtable);
    stage_upload(dtable, pl.msys.data(), pl.msys.size() * sizeof(MutSys), st);
  }
  int grid = (int)std::min<int64_t>(pl.mut_items, (int64_t)INT32_MAX);
  if (tuning && tuning->max_wgs > 0) grid = std::min(grid, tuning->max_wgs);
  // periodic images: one right-hand side per pass
  const int kmax = periodic ? 1 : pl.kmax_mut;
  for (int k0 = 0, K = group_k(k, kmax); k0 < k; k0 += K) {
    K = group_k(k - k0, kmax);
    MutParams q{};
    q.r = r;
    q.f = NF ? f + (size_t)k0 * stride : nullptr;
    q.t = NT ? t + (size_t)k0 * stride : nullptr;
    q.rhs_stride = stride;
    q.a = a; q.inva = inva; q.c1 = c1; q.c2 = c2;
    q.part_i = (double*)(base + pl.off_pi);
    q.part_j = (double*)(base + pl.off_pj);
    q.table = dtable;
    q.single = pl.msys[0];
    q.nsys = (int)pl.msys.size();
    q.n_items = pl.mut_items;
    q.chunk_blocks = pl.chunk_blocks;
    q.periodic = periodic;
    for (int d = 0; d < 3; ++d) q.L[d] = Lp[d];
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st);
    if (periodic) dispatch_mutual_periodic(pl.mask, pl.prec, q, grid, st);
    else dispatch_mutual(pl.domain, pl.mask, pl.prec, K, q, grid, st);
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);

    MutReduceParams rp{};
    rp.r = r;
    rp.f = q.f; rp.t = q.t;
    rp.vel = HU ? vel + (size_t)k0 * stride : nullptr;
    rp.omg = HW ? omg + (size_t)k0 * stride : nullptr;
    rp.part_i = q.part_i;
    rp.part_j = q.part_j;
    rp.table = dtable;
    rp.single = q.single;
    rp.nsys = q.nsys;
    rp.mask = pl.mask;
    rp.K = K;
    rp.ntot = pl.ntot;
    rp.chunk_blocks = pl.chunk_blocks;
    rp.rhs_stride = stride;
    rp.a = a; rp.inva = inva; rp.c1 = c1; rp.c2 = c2;
    rp.domain = pl.domain;
    rp.periodic = periodic;
    for (int d = 0; d < 3; ++d) rp.L[d] = Lp[d];
    const int OC = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
    const int64_t nout = pl.ntot * K * OC;
    const dim3 rgrid((unsigned)((nout + kRedOut - 1) / kRedOut));
    if (periodic)
      hipLaunchKernelGGL(mutual_reduce_kernel<true>, rgrid, dim3(256), 0, st, rp);
    else
      hipLaunchKernelGGL(mutual_reduce_kernel<false>, rgrid, dim3(256), 0, st, rp);
    RMW_CHECK_HIP(hipGetLastError());
  }
}

// RMW_FLAG_CHECK_FINITE: count the non-finite output entries on the stream,
// wait for it and report them as RMW_E_NOTFINITE.
static void check_outputs(const Plan& pl, const double* vel, const double* omg, int k, int64_t stride, char* base,
                          hipStream_t st)
{
  if (!pl.check_finite) return;
  unsigned int* count = (unsigned int*)(base + pl.off_status);
  hipLaunchKernelGGL(zero_count_kernel, dim3(1), dim3(64), 0, st, count);
  RMW_CHECK_HIP(hipGetLastError());
  const int64_t len = 3 * pl.ntot;
  const int64_t grid = std::min<int64_t>(((int64_t)k * len + 255) / 256, 4096);
  for (const double* x : {vel, omg}) {
    if (!x || grid == 0) continue;
    hipLaunchKernelGGL(count_nonfinite_kernel, dim3((unsigned)grid), dim3(256), 0, st, x, (int64_t)k, len, stride,
                       count);
    RMW_CHECK_HIP(hipGetLastError());
  }
  RMW_CHECK_HIP(hipStreamSynchronize(st));
  unsigned int bad = 0;
  RMW_CHECK_HIP(hipMemcpy(&bad, count, sizeof(bad), hipMemcpyDeviceToHost));
  if (bad)
    throw Error(RMW_E_NOTFINITE, std::to_string(bad) + " output entries are inf or NaN (coincident distinct "
                                                       "blobs or non-finite inputs)");
}

static bool has_images(const double* L)
{
  return L && (L[0] > 0.0 || L[1] > 0.0 || L[2] > 0.0);
}

// Workspace bytes for a call whose lengths L the query does not see: the
// larger of the plans without and with pseudo-periodic images.
static size_t plan_bytes_any_images(int mask, int prec, const int64_t* offsets, int nsys, int k,
                                    const rmw_tuning* tuning, bool ragged, int domain = kDomWall)
{
  const Plan a = make_plan(mask, prec, offsets, nsys, k, tuning, ragged, domain, false);
  const Plan b = make_plan(mask, prec, offsets, nsys, k, tuning, ragged, domain, true);
  return std::max(a.total, b.total);
}

static void run_plan(const Plan& pl, const double* r, const double* f, const double* t, double* vel,
                     double* omg, int k, int64_t ld, double eta, double a, const double* L,
                     const rmw_tuning* tuning, void* ws, size_t ws_bytes, hipStream_t st, bool ragged)
{
  RMW_REQUIRE(ld == 0 || ld >= 3 * pl.ntot, "ld must be 0 or >= 3 n");
  RMW_REQUIRE(ws_bytes >= pl.total, "workspace too small: need " + std::to_string(pl.total) + " bytes");
  if (pl.ntot == 0) return;
  RMW_REQUIRE(r != nullptr && ws != nullptr, "r_vectors / workspace is null");
  RMW_REQUIRE(eta > 0.0 && a > 0.0 && std::isfinite(eta) && std::isfinite(a), "eta and a must be positive");
  const int kmask = kernel_mask(pl.mask);
  RMW_REQUIRE(kmask == pl.mask, "unsupported block combination (use UF, UT, WF, WT, UF|UT, UF|WF, "
                                "UT|WT, WF|WT or all four)");
  const bool NF = (kmask & (kUF | kWF)) != 0, NT = (kmask & (kUT | kWT)) != 0;
  const bool HU = (pl.mask & (kUF | kUT)) != 0, HW = (pl.mask & (kWF | kWT)) != 0;
  RMW_REQUIRE(!((pl.mask & (kUF | kWF)) && f == nullptr), "force is required by the mask");
  RMW_REQUIRE(!((pl.mask & (kUT | kWT)) && t == nullptr), "torque is required by the mask");
  RMW_REQUIRE(!(HU && vel == nullptr), "velocity output is required by the mask");
  RMW_REQUIRE(!(HW && omg == nullptr), "omega output is required by the mask");

  char* base = (char*)ws;
  double Lp[3] = {0, 0, 0};
  int periodic = 0;
  for (int d = 0; d < 3; ++d) {
    double Ld = L ? L[d] : 0.0;
    if (Ld > 0.0) {
      periodic |= 1 << d;
      Lp[d] = Ld / a;
    }
  }
  const double inva = 1.0 / a;
  const double c1 = 1.0 / (8.0 * M_PI * eta * a);
  const double c2 = c1 / a;

  const int64_t stride = ld > 0 ? ld : 3 * pl.ntot;
  if (pl.mut_ok && (periodic == 0 || pl.domain == kDomWall)) {
    run_mutual(pl, r, f, t, vel, omg, k, stride, a, inva, c1, c2, tuning, base, st, periodic, Lp);
    check_outputs(pl, HU ? vel : nullptr, HW ? omg : nullptr, k, stride, base, st);
    return;
  }

  SysInfo* dtable = nullptr;
  if (ragged) {
    dtable = (SysInfo*)(base + pl.off_table);
    stage_upload(dtable, pl.sys.data(), pl.sys.size() * sizeof(SysInfo), st);
  }

  int grid = (int)std::min<int64_t>(pl.total_items, (int64_t)INT32_MAX);
  if (ragged || (tuning && tuning->max_wgs > 0)) {
    const int cap = tuning && tuning->max_wgs > 0 ? tuning->max_wgs : std::max(pl.slots, 1);
    grid = (int)std::min<int64_t>(grid, cap);
  }

  const int kcap = pl.domain == kDomWall ? rhs_cap(tuning) : 1;
  for (int k0 = 0, K = group_k(k, kcap); k0 < k; k0 += K) {
    K = group_k(k - k0, kcap);
    PairParams q{};
    q.r = r;
    q.f = NF ? f + (size_t)k0 * stride : nullptr;
    q.t = NT ? t + (size_t)k0 * stride : nullptr;
    q.rhs_stride = stride;
    q.a = a; q.inva = inva; q.c1 = c1; q.c2 = c2;
    q.partial = (double*)(base + pl.off_partial);
    q.table = dtable;
    q.single = pl.sys.empty() ? SysInfo{} : pl.sys[0];
    q.nsys = (int)pl.sys.size();
    q.ntot = pl.ntot;
    q.total_items = pl.total_items;
    q.periodic = periodic;
    for (int d = 0; d < 3; ++d) q.L[d] = Lp[d];
    q.domain = pl.domain;
    // The partial layout was planned for kmax RHS; a smaller trailing group
    // reuses it with its own K (items only touch their own slots).
    ProfRec rec{};
    bool prof = false;
    prof_begin(rec, prof, st);
    if (pl.domain == kDomWall) dispatch_pair(kmask, pl.prec, K, q, grid, st);
    else dispatch_pair_domain(pl.domain, kmask, pl.prec, q, grid, st);
    RMW_CHECK_HIP(hipGetLastError());
    prof_end(rec, prof, st);

    ReduceParams rp{};
    rp.r = r;
    rp.f = q.f; rp.t = q.t;
    rp.vel = HU ? vel + (size_t)k0 * stride : nullptr;
    rp.omg = HW ? omg + (size_t)k0 * stride : nullptr;
    rp.partial = q.partial;
    rp.table = dtable;
    rp.single = q.single;
    rp.nsys = q.nsys;
    rp.mask = pl.mask;
    rp.K = K;
    rp.ntot = pl.ntot;
    rp.rhs_stride = stride;
    rp.a = a; rp.inva = inva; rp.c1 = c1; rp.c2 = c2;
    rp.domain = pl.domain;
    const int OCr = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
    const int64_t nthreads = pl.ntot * K * OCr;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, st, rp);
    RMW_CHECK_HIP(hipGetLastError());
  }
  check_outputs(pl, HU ? vel : nullptr, HW ? omg : nullptr, k, stride, base, st);
}

// ---------------------------------------------------------------------------
// No-allocation entry point: the partial sums live in a fixed device array
// instead of a caller workspace.  Each device holds its own copy of the
// array (for the life of the process); its address on the current device
// is looked up once and kept.  Nothing in it must start at zero: every
// partial-sum slot a reduce reads is written by the pair kernel of the same
// call first.
// ---------------------------------------------------------------------------
static __device__ __align__(256) unsigned char scratch[RMW_NOALLOC_SCRATCH_BYTES];

static void* scratch_address()
{
  static std::mutex mu;
  static void* ptrs[64] = {};
  int dev = 0;
  RMW_CHECK_HIP(hipGetDevice(&dev));
  RMW_REQUIRE(dev >= 0 && dev < 64, "device ordinal out of range");
  std::lock_guard<std::mutex> lk(mu);
  if (ptrs[dev] == nullptr) {
    void* p = nullptr;
    RMW_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    RMW_REQUIRE(p != nullptr, "no device address for the scratch array");
    ptrs[dev] = p;
  }
  return ptrs[dev];
}

// The plan of rmw_wall_rpy_apply for these sizes if its partial sums fit
// in `cap` bytes; otherwise (unless the tuning requires the symmetric
// kernel) the ordered kernel with source chunks as long as the fit needs,
// and fewer right-hand sides per pass if one chunk of k_group is still too
// large.  Throws RMW_E_LIMIT when nothing fits.
static Plan plan_within(int mask, int prec, int64_t n, int k, const rmw_tuning* tuning, size_t cap, bool periodic)
{
  const int64_t off[2] = {0, n};
  Plan pl = make_plan(mask, prec, off, 1, k, tuning, false, kDomWall, periodic);
  if (pl.total <= cap) return pl;
  const bool sym_required = tuning && tuning->symmetric > 0;
  if (!sym_required) {
    const int OC = 3 * (((mask & (kUF | kUT)) ? 1 : 0) + ((mask & (kWF | kWT)) ? 1 : 0));
    rmw_tuning t = tuning ? *tuning : rmw_tuning{};
    t.symmetric = -1;
    for (int g : {4, 2, 1}) {
      if (t.rhs_group != 0 && g > t.rhs_group) continue;
      const int kg = group_k(k, g);
      const double per_chunk = (double)kg * (double)n * OC * sizeof(double);
      const int64_t chunks = (int64_t)std::floor((double)(cap - 256) / per_chunk);
      if (chunks < 1) continue;
      const int64_t len = (n + chunks - 1) / chunks;
      t.chunk_len = std::max<int64_t>(t.chunk_len, (len + kChunkQuantum - 1) / kChunkQuantum * kChunkQuantum);
      t.rhs_group = kg;
      Plan p2 = make_plan(mask, prec, off, 1, k, &t, false, kDomWall, periodic);
      if (p2.total <= cap) return p2;
    }
  }
  throw Error(RMW_E_LIMIT, "rmw_wall_rpy_apply_noalloc: n = " + std::to_string(n) + ", k = " + std::to_string(k) +
                               " needs " + std::to_string(pl.total) + " bytes of partial sums, more than the " +
                               std::to_string((size_t)RMW_NOALLOC_SCRATCH_BYTES) + "-byte scratch array");
}

}  // namespace rmw

// ---------------------------------------------------------------------------
// C ABI
// ---------------------------------------------------------------------------
using namespace rmw;

extern "C" const char* rmw_last_error(void) { return g_last_error.c_str(); }

extern "C" const char* rmw_version(void) { return "rmw_hip 0.1 gfx950"; }

extern "C" int rmw_wall_rpy_workspace_size(int mask, int prec, int64_t n, int k, const rmw_tuning* tuning,
                                           size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr, "bytes is null");
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const int64_t off[2] = {0, n};
    *bytes = plan_bytes_any_images(mask, prec, off, 1, k, tuning, false);
  });
}

extern "C" int rmw_wall_rpy_describe(int mask, int prec, int64_t n, int k, const rmw_tuning* tuning,
                                     int periodic, rmw_plan_info* info)
{
  return guarded([&] {
    RMW_REQUIRE(info != nullptr, "info is null");
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const int64_t off[2] = {0, n};
    Plan pl = make_plan(mask, prec, off, 1, k, tuning, false, kDomWall, periodic != 0);
    const bool mut = pl.mut_ok && (!periodic || pl.domain == kDomWall);
    info->symmetric = mut ? 1 : 0;
    info->rhs_group = mut ? (periodic ? 1 : pl.kmax_mut) : pl.kmax;
    info->chunk_len = mut ? pl.chunk_blocks * 64 : (pl.sys.empty() ? 0 : pl.sys[0].chunk_len);
    info->work_items = mut ? pl.mut_items : pl.total_items;
    info->workspace_bytes = (int64_t)pl.total;
  });
}

extern "C" int rmw_wall_rpy_apply(int mask, int prec, const double* r_vectors, const double* force,
                                  const double* torque, double* velocity, double* omega, int64_t n, int k,
                                  int64_t ld, double eta, double a, const double* L, const rmw_tuning* tuning,
                                  void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const int64_t off[2] = {0, n};
    Plan pl = make_plan(mask, prec, off, 1, k, tuning, false, kDomWall, has_images(L));
    run_plan(pl, r_vectors, force, torque, velocity, omega, k, ld, eta, a, L, tuning, workspace,
             workspace_bytes, (hipStream_t)stream, false);
  });
}

extern "C" int rmw_wall_rpy_apply_noalloc(int mask, int prec, const double* r_vectors, const double* force,
                                          const double* torque, double* velocity, double* omega, int64_t n, int k,
                                          int64_t ld, double eta, double a, const double* L,
                                          const rmw_tuning* tuning, void* stream)
{
  struct NoRecords {
    NoRecords() { tl_no_records = true; }
    ~NoRecords() { tl_no_records = false; }
  } guard;
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    RMW_REQUIRE(!(tuning && (tuning->flags & RMW_FLAG_CHECK_FINITE)),
                "the finite check waits for the stream; use rmw_wall_rpy_apply for it");
    Plan pl = plan_within(mask, prec, n, k, tuning, RMW_NOALLOC_SCRATCH_BYTES, has_images(L));
    if (pl.ntot == 0) return;
    // the plan may differ from the caller's tuning (longer ordered chunks);
    // run_plan reads only max_wgs and rhs_group from it
    rmw_tuning t = tuning ? *tuning : rmw_tuning{};
    if (!pl.mut_ok) t.rhs_group = pl.kmax;
    run_plan(pl, r_vectors, force, torque, velocity, omega, k, ld, eta, a, L, &t, scratch_address(),
             RMW_NOALLOC_SCRATCH_BYTES, (hipStream_t)stream, false);
  });
}

extern "C" int rmw_mobility_workspace_size(int domain, int mask, int prec, int64_t n, int k,
                                           const rmw_tuning* tuning, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr, "bytes is null");
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const int64_t off[2] = {0, n};
    *bytes = plan_bytes_any_images(mask, prec, off, 1, k, tuning, false, domain);
  });
}

extern "C" int rmw_mobility_apply(int domain, int mask, int prec, const double* r_vectors, const double* force,
                                  const double* torque, double* velocity, double* omega, int64_t n, int k,
                                  int64_t ld, double eta, double a, const double* L, const rmw_tuning* tuning,
                                  void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const int64_t off[2] = {0, n};
    Plan pl = make_plan(mask, prec, off, 1, k, tuning, false, domain, has_images(L));
    run_plan(pl, r_vectors, force, torque, velocity, omega, k, ld, eta, a, L, tuning, workspace,
             workspace_bytes, (hipStream_t)stream, false);
  });
}

extern "C" int rmw_wall_rpy_ragged_workspace_size(int mask, int prec, const int64_t* offsets, int nsys, int k,
                                                  const rmw_tuning* tuning, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr && offsets != nullptr && nsys >= 1, "bad arguments");
    *bytes = plan_bytes_any_images(mask, prec, offsets, nsys, k, tuning, true);
  });
}

extern "C" int rmw_wall_rpy_ragged_apply(int mask, int prec, const double* r_vectors, const double* force,
                                         const double* torque, double* velocity, double* omega,
                                         const int64_t* offsets, int nsys, int k, int64_t ld, double eta,
                                         double a, const double* L, const rmw_tuning* tuning, void* workspace,
                                         size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(offsets != nullptr && nsys >= 1, "bad offsets");
    Plan pl = make_plan(mask, prec, offsets, nsys, k, tuning, true, kDomWall, has_images(L));
    run_plan(pl, r_vectors, force, torque, velocity, omega, k, ld, eta, a, L, tuning, workspace,
             workspace_bytes, (hipStream_t)stream, true);
  });
}

extern "C" int rmw_profile_enable(int on)
{
  return guarded([&] {
    std::lock_guard<std::mutex> lk(g_prof_mu);
    g_prof_on = on != 0;
  });
}

extern "C" int rmw_profile_collect(int64_t* launches, double* total_ms, double* mean_ms)
{
  return guarded([&] {
    std::vector<ProfRec> recs;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      recs.swap(g_prof);
    }
    double tot = 0.0;
    for (auto& r : recs) {
      RMW_CHECK_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      RMW_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      tot += ms;
      RMW_CHECK_HIP(hipEventDestroy(r.a));
      RMW_CHECK_HIP(hipEventDestroy(r.b));
    }
    if (launches) *launches = (int64_t)recs.size();
    if (total_ms) *total_ms = tot;
    if (mean_ms) *mean_ms = recs.empty() ? 0.0 : tot / recs.size();
  });
}

extern "C" int rmw_profile_collect_kinds(int64_t* launches, double* total_ms)
{
  return guarded([&] {
    RMW_REQUIRE(launches && total_ms, "launches and total_ms are required");
    std::vector<ProfRec> recs;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      recs.swap(g_prof);
    }
    for (int k = 0; k < kProfKinds; ++k) {
      launches[k] = 0;
      total_ms[k] = 0.0;
    }
    for (auto& r : recs) {
      RMW_CHECK_HIP(hipEventSynchronize(r.b));
      float ms = 0.f;
      RMW_CHECK_HIP(hipEventElapsedTime(&ms, r.a, r.b));
      launches[r.kind] += 1;
      total_ms[r.kind] += ms;
      RMW_CHECK_HIP(hipEventDestroy(r.a));
      RMW_CHECK_HIP(hipEventDestroy(r.b));
    }
  });
}

extern "C" int rmw_profile_pending(int64_t* launches)
{
  return guarded([&] {
    RMW_REQUIRE(launches != nullptr, "launches is null");
    std::lock_guard<std::mutex> lk(g_prof_mu);
    *launches = (int64_t)g_prof.size();
  });
}

extern "C" int rmw_profile_collect_launches(int32_t* kinds, double* ms, int64_t cap, int64_t* launches)
{
  return guarded([&] {
    RMW_REQUIRE(kinds && ms && launches && cap >= 0, "kinds, ms and launches are required");
    std::vector<ProfRec> recs;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      recs.swap(g_prof);
    }
    *launches = (int64_t)recs.size();
    for (size_t i = 0; i < recs.size(); ++i) {
      RMW_CHECK_HIP(hipEventSynchronize(recs[i].b));
      float t = 0.f;
      RMW_CHECK_HIP(hipEventElapsedTime(&t, recs[i].a, recs[i].b));
      if ((int64_t)i < cap) {
        kinds[i] = recs[i].kind;
        ms[i] = t;
      }
      RMW_CHECK_HIP(hipEventDestroy(recs[i].a));
      RMW_CHECK_HIP(hipEventDestroy(recs[i].b));
    }
  });
}

extern "C" int rmw_device_info(int* cu_count, int* clock_khz, int* wavefront, char* arch_name, int arch_len)
{
  return guarded([&] {
    const DeviceProps& p = device_props();
    if (cu_count) *cu_count = p.cu_count;
    if (clock_khz) *clock_khz = p.clock_khz;
    if (wavefront) *wavefront = p.warp;
    if (arch_name && arch_len > 0) {
      std::strncpy(arch_name, p.arch.c_str(), (size_t)arch_len - 1);
      arch_name[arch_len - 1] = '\0';
    }
  });
}
