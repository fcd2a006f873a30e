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
#include <vector>

#include "rmw_internal.h"
#include "rmw_pair_terms.h"

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
// (damping_matrix_B, mobility.py:67-84).
// ---------------------------------------------------------------------------
__device__ __forceinline__ void load_blob(const double* __restrict__ r, int64_t g, double a, double inva,
                                          double& X, double& Y, double& Z, double& B)
{
  const double x = r[3 * g], y = r[3 * g + 1], z = r[3 * g + 2];
  const double ze = z <= a ? a : z;
  X = x * inva;
  Y = y * inva;
  Z = ze * inva;
  B = z < a ? z * inva : 1.0;
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
// Per-target accumulator policies.
template <int ACC> struct Acc;
template <> struct Acc<RMW_PREC_F32_F64ACC> {
  double v = 0.0;
  __device__ __forceinline__ void add(float x) { v += (double)x; }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ double get() const { return v; }
};
template <> struct Acc<RMW_PREC_F32_BLOCKED> {
  double v = 0.0;
  float p = 0.0f;
  __device__ __forceinline__ void add(float x) { p += x; }
  __device__ __forceinline__ void flush()
  {
    v += (double)p;
    p = 0.0f;
  }
  __device__ __forceinline__ double get() const { return v + (double)p; }
};
template <> struct Acc<RMW_PREC_F32_KAHAN> {
  float s = 0.0f, c = 0.0f;
  __device__ __forceinline__ void add(float x)
  {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ double get() const { return (double)s - (double)c; }
};
template <> struct Acc<RMW_PREC_F64> {
  double v = 0.0;
  __device__ __forceinline__ void add(double x) { v += x; }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ double get() const { return v; }
};

constexpr int kSub = 16;  // sources per FP32 sub-sum in the blocked mode

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
template <int MASK, int K, int ACC, typename T>
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
  if constexpr ((MASK & kUF) != 0) cuf = coef_uf(g, zi, zj);
  if constexpr ((MASK & kUT) != 0) cut = coef_rot(g, zi);
  if constexpr ((MASK & kWF) != 0) cwf = coef_rot(g, zj);
  if constexpr ((MASK & kWT) != 0) cwt = coef_wt(g);
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

// Minimum-image shift count of mobility_numba.py:184-193: int(d/L + 0.5 sign(d)).
__device__ __forceinline__ double image_count(double d, double L)
{
  double s = (double)((d > 0.0) - (d < 0.0));
  return trunc(d / L + 0.5 * s);
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
template <int MASK, int K, typename V>
__device__ __forceinline__ void stage_sources(const PairParams& p, int64_t gj, double B, int slot, V* s_f, V* s_t)
{
  using BT = BlockTraits<MASK>;
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if constexpr (BT::NF) {
      const double* v = p.f + k * p.rhs_stride + 3 * gj;
      const double c = p.c1 * B;
      s_f[k * kBlock + slot] = make_v4<V>(c * v[0], c * v[1], c * v[2]);
    }
    if constexpr (BT::NT) {
      const double* v = p.t + k * p.rhs_stride + 3 * gj;
      const double c = p.c2 * B;
      s_t[k * kBlock + slot] = make_v4<V>(c * v[0], c * v[1], c * v[2]);
    }
  }
}

template <int MASK, int K, int ACC>
__global__ __launch_bounds__(kBlock) void pair_kernel_f32(PairParams p)
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
    load_blob(p.r, S.b0 + (valid ? il : 0), p.a, p.inva, Xi, Yi, Zi, Bi);
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
        load_blob(p.r, gj, p.a, p.inva, X, Y, Z, B);
        s_pos[threadIdx.x] = SrcF32{X, Y, Z, (float)Z, 0.0f};
        stage_sources<MASK, K, float4>(p, gj, B, threadIdx.x, s_f, s_t);
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
          pair_eval<MASK, K, ACC, float>(dx, dy, dz, zi, sp.zf, s_f, s_t, j, masked && j == self_j, u, w);
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
                pair_eval<MASK, K, ACC, float>(dx, dy, dz, zi, sp.zf, s_f, s_t, j,
                                               diag && central && (j == self_j), u, w);
              }
          flush_all<K, ACC>(u, w);
        }
      }
    }
    if (valid) store_partials<MASK, K>(p, S, ch, il, u, w);
  }
}

template <int MASK, int K>
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
    load_blob(p.r, S.b0 + (valid ? il : 0), p.a, p.inva, Xi, Yi, Zi, Bi);

    Acc<RMW_PREC_F64> u[K][3], w[K][3];
    const int64_t s_begin = ch * S.chunk_len;
    const int64_t s_end = s_begin + S.chunk_len < S.n ? s_begin + S.chunk_len : S.n;
    for (int64_t s0 = s_begin; s0 < s_end; s0 += kBlock) {
      const int cnt = (int)(s_end - s0 < kBlock ? s_end - s0 : kBlock);
      __syncthreads();
      if ((int)threadIdx.x < cnt) {
        const int64_t gj = S.b0 + s0 + threadIdx.x;
        double X, Y, Z, B;
        load_blob(p.r, gj, p.a, p.inva, X, Y, Z, B);
        s_pos[threadIdx.x] = make_double4(X, Y, Z, 0.0);
        stage_sources<MASK, K, double4>(p, gj, B, threadIdx.x, s_f, s_t);
      }
      __syncthreads();
      const bool diag = (s0 < t0 + kBlock) && (t0 < s0 + cnt);
      const int self_j = (int)(il - s0);
      for (int j = 0; j < cnt; ++j) {
        const double4 pj = s_pos[j];
        double rx = Xi - pj.x, ry = Yi - pj.y, rz = Zi - pj.z;
        if (!p.periodic) {
          pair_eval<MASK, K, RMW_PREC_F64, double>(rx, ry, rz, Zi, pj.z, s_f, s_t, j, diag && j == self_j, u, w);
        } else {
          if (pbox[0]) rx -= image_count(rx, p.L[0]) * p.L[0];
          if (pbox[1]) ry -= image_count(ry, p.L[1]) * p.L[1];
          if (pbox[2]) rz -= image_count(rz, p.L[2]) * p.L[2];
          for (int bx = -pbox[0]; bx <= pbox[0]; ++bx)
            for (int by = -pbox[1]; by <= pbox[1]; ++by)
              for (int bz = -pbox[2]; bz <= pbox[2]; ++bz) {
                const bool central = (bx == 0) && (by == 0) && (bz == 0);
                pair_eval<MASK, K, RMW_PREC_F64, double>(rx + bx * p.L[0], ry + by * p.L[1], rz + bz * p.L[2], Zi,
                                                         pj.z, s_f, s_t, j, diag && central && (j == self_j), u, w);
              }
        }
      }
    }
    if (valid) store_partials<MASK, K>(p, S, ch, il, u, w);
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

  const double z = p.r[3 * gi + 2];
  const double ze = z <= p.a ? p.a : z;
  const double B = z < p.a ? z * p.inva : 1.0;
  const double h = ze * p.inva;
  const bool is_w = HW && (!HU || c >= 3);
  const int comp = c % 3;
  double fv[3] = {0, 0, 0}, tv[3] = {0, 0, 0};
  if (p.f) {
    const double* v = p.f + k * p.rhs_stride + 3 * gi;
    fv[0] = p.c1 * B * v[0]; fv[1] = p.c1 * B * v[1]; fv[2] = p.c1 * B * v[2];
  }
  if (p.t) {
    const double* v = p.t + k * p.rhs_stride + 3 * gi;
    tv[0] = p.c2 * B * v[0]; tv[1] = p.c2 * B * v[1]; tv[2] = p.c2 * B * v[2];
  }
  double self = 0.0;
  if (!is_w) {
    if (p.mask & kUF) { Vec3<double> sv = self_uf(h, fv[0], fv[1], fv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    if (p.mask & kUT) { Vec3<double> sv = self_ut(h, tv[0], tv[1], tv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    p.vel[k * p.rhs_stride + 3 * gi + comp] = B * (sum + self);
  } else {
    if (p.mask & kWF) { Vec3<double> sv = self_wf(h, fv[0], fv[1], fv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    if (p.mask & kWT) { Vec3<double> sv = self_wt(h, tv[0], tv[1], tv[2]); self += comp == 0 ? sv.x : (comp == 1 ? sv.y : sv.z); }
    p.omg[k * p.rhs_stride + 3 * gi + comp] = B * p.inva * (sum + self);
  }
}

// ---------------------------------------------------------------------------
// Host side: plan, workspace layout, dispatch.
// ---------------------------------------------------------------------------
struct Plan {
  int mask = 0, prec = 0, kmax = 1;
  int64_t ntot = 0;
  std::vector<SysInfo> sys;  // non-empty systems only
  int64_t total_items = 0;
  size_t partial_doubles = 0;
  size_t off_partial = 0, off_table = 0, total = 0;
};

static bool valid_mask(int mask) { return mask > 0 && mask < 16; }

// Right-hand sides per pair-kernel pass: 4, 2 or 1, at most the tuned cap.
static int group_k(int k, int cap)
{
  const int g = k >= 4 ? 4 : (k >= 2 ? 2 : 1);
  return cap > 0 && cap < g ? (cap >= 2 ? 2 : 1) : g;
}

static int rhs_cap(const rmw_tuning* tuning) { return tuning ? tuning->rhs_group : 0; }

static Plan make_plan(int mask, int prec, const std::vector<int64_t>& offsets, int k,
                      const rmw_tuning* tuning, bool ragged)
{
  RMW_REQUIRE(valid_mask(mask), "mask must be a non-empty subset of UF|UT|WF|WT");
  RMW_REQUIRE(prec == RMW_PREC_F32_F64ACC || prec == RMW_PREC_F32_KAHAN || prec == RMW_PREC_F64 ||
                  prec == RMW_PREC_F32_BLOCKED,
              "unknown precision mode");
  RMW_REQUIRE(k >= 1, "k must be >= 1");
  Plan pl;
  pl.mask = mask;
  pl.prec = prec;
  RMW_REQUIRE(!tuning || tuning->rhs_group == 0 || tuning->rhs_group == 1 || tuning->rhs_group == 2 ||
                  tuning->rhs_group == 4,
              "rhs_group must be 0, 1, 2 or 4");
  pl.kmax = group_k(k, rhs_cap(tuning));
  const int nsys = (int)offsets.size() - 1;
  RMW_REQUIRE(nsys >= 1 && offsets[0] == 0, "offsets must start at 0");
  for (int s = 0; s < nsys; ++s) RMW_REQUIRE(offsets[s + 1] >= offsets[s], "offsets must be non-decreasing");
  pl.ntot = offsets[nsys];

  // chunk length: aim for ~8 workgroups per CU over the whole batch.
  double pairs = 0.0;
  int64_t nmax = 0;
  for (int s = 0; s < nsys; ++s) {
    double n = (double)(offsets[s + 1] - offsets[s]);
    pairs += n * n;
    nmax = std::max<int64_t>(nmax, offsets[s + 1] - offsets[s]);
  }
  int64_t chunk = tuning && tuning->chunk_len > 0 ? tuning->chunk_len : 0;
  if (chunk <= 0) {
    const DeviceProps& dp = device_props();
    double target_items = 8.0 * std::max(dp.cu_count, 1);
    double c = pairs / (target_items * kBlock);
    chunk = (int64_t)std::ceil(c / kBlock) * kBlock;
  }
  chunk = std::max<int64_t>(kBlock, (chunk + kBlock - 1) / kBlock * kBlock);
  if (nmax > 0) chunk = std::min<int64_t>(chunk, (nmax + kBlock - 1) / kBlock * kBlock);

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

  size_t o = 0;
  pl.off_partial = o; o = align_up(o + pl.partial_doubles * sizeof(double));
  if (ragged) { pl.off_table = o; o = align_up(o + pl.sys.size() * sizeof(SysInfo)); }
  pl.total = o;
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

// Profiling record.
struct ProfRec { hipEvent_t a, b; };
static std::mutex g_prof_mu;
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;

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

  SysInfo* dtable = nullptr;
  if (ragged) {
    dtable = (SysInfo*)(base + pl.off_table);
    stage_upload(dtable, pl.sys.data(), pl.sys.size() * sizeof(SysInfo), st);
  }

  int grid = (int)std::min<int64_t>(pl.total_items, (int64_t)INT32_MAX);
  if (ragged || (tuning && tuning->max_wgs > 0)) {
    const DeviceProps& dp = device_props();
    int cap = tuning && tuning->max_wgs > 0 ? tuning->max_wgs : dp.cu_count * 8;
    grid = (int)std::min<int64_t>(grid, cap);
  }

  const int64_t stride = ld > 0 ? ld : 3 * pl.ntot;
  for (int k0 = 0, K = group_k(k, rhs_cap(tuning)); k0 < k; k0 += K) {
    K = group_k(k - k0, rhs_cap(tuning));
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
    // The partial layout was planned for kmax RHS; a smaller trailing group
    // reuses it with its own K (items only touch their own slots).
    ProfRec rec{};
    bool prof = false;
    {
      std::lock_guard<std::mutex> lk(g_prof_mu);
      prof = g_prof_on;
    }
    if (prof) {
      RMW_CHECK_HIP(hipEventCreate(&rec.a));
      RMW_CHECK_HIP(hipEventCreate(&rec.b));
      RMW_CHECK_HIP(hipEventRecord(rec.a, st));
    }
    dispatch_pair(kmask, pl.prec, K, q, grid, st);
    RMW_CHECK_HIP(hipGetLastError());
    if (prof) {
      RMW_CHECK_HIP(hipEventRecord(rec.b, st));
      std::lock_guard<std::mutex> lk(g_prof_mu);
      g_prof.push_back(rec);
    }

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
    const int OCr = 3 * ((HU ? 1 : 0) + (HW ? 1 : 0));
    const int64_t nthreads = pl.ntot * K * OCr;
    hipLaunchKernelGGL(reduce_kernel, dim3((unsigned)((nthreads + 255) / 256)), dim3(256), 0, st, rp);
    RMW_CHECK_HIP(hipGetLastError());
  }
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
    Plan pl = make_plan(mask, prec, {0, n}, k, tuning, false);
    *bytes = pl.total;
  });
}

extern "C" int rmw_wall_rpy_apply(int mask, int prec, const double* r_vectors, const double* force,
                                  const double* torque, double* velocity, double* omega, int64_t n, int k,
                                  int64_t ld, double eta, double a, const double* L, const rmw_tuning* tuning,
                                  void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    Plan pl = make_plan(mask, prec, {0, n}, k, tuning, false);
    run_plan(pl, r_vectors, force, torque, velocity, omega, k, ld, eta, a, L, tuning, workspace,
             workspace_bytes, (hipStream_t)stream, false);
  });
}

extern "C" int rmw_wall_rpy_ragged_workspace_size(int mask, int prec, const int64_t* offsets, int nsys, int k,
                                                  const rmw_tuning* tuning, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr && offsets != nullptr && nsys >= 1, "bad arguments");
    std::vector<int64_t> off(offsets, offsets + nsys + 1);
    Plan pl = make_plan(mask, prec, off, k, tuning, true);
    *bytes = pl.total;
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
    std::vector<int64_t> off(offsets, offsets + nsys + 1);
    Plan pl = make_plan(mask, prec, off, k, tuning, true);
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
