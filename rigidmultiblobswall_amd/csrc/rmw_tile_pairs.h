// Tile-pair schedule of the symmetric pair kernels: each unordered pair of
// blob tiles (S >= T) is visited once; a work item is a target tile T and
// a run of C consecutive source tiles starting at S >= T.  Also the DPP
// lane rotation that carries source-side sums across a wave.
#pragma once

#include <cstdint>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RMW_TP __host__ __device__ __forceinline__
#else
#define RMW_TP inline   // host build of the schedule for the CPU tests
#endif

namespace rmw {

RMW_TP int64_t tri_index(int64_t S, int64_t T) { return S * (S + 1) / 2 + T; }

// sum_{m=1}^{M} ceil(m / C).  M and C are tile counts (< 2^31): the
// quotient is a 32-bit division (a 64-bit one expands to ~40 instructions,
// and the work-item decode below evaluates this in a binary search).
RMW_TP int64_t ceil_sum(int64_t M, int64_t C)
{
  const uint32_t m = (uint32_t)M, c = (uint32_t)C;
  const int64_t q = m / c, r = m - (uint32_t)q * c;
  return C * q * (q + 1) / 2 + r * (q + 1);
}

// Work items of target tiles 0..T-1 of a system (target tile t has
// ceil((n_t - t) / C) runs of C source tiles): closed form, no table.
RMW_TP int64_t items_before(int64_t T, int64_t n_t, int64_t C)
{
  return ceil_sum(n_t, C) - ceil_sum(n_t - T, C);
}

#ifdef __HIPCC__
// Lane l receives lane (l + 1) mod 64 (DPP wave_rol:1).
// (bound_ctrl set: every lane of the wave is active where this is used, so it
// changes nothing, and it lets the rotate fuse into the add that consumes it,
// v_add_f32_dpp, instead of a separate v_mov_b32_dpp)
__device__ __forceinline__ float rol1(float v)
{
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x134, 0xF, 0xF, true));
}
__device__ __forceinline__ double rol1(double v)
{
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), 0x134, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), 0x134, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

// Partial-sum reduces: G threads share one output; thread g adds the slots
// P[m * step] for m = g, g + G, g + 2G, ... below cnt with four loads in
// flight (the loop is latency-bound), in a fixed order.  The caller adds
// the G results in order g = 0..G-1.
template <int G>
__device__ __forceinline__ double group_strided_sum(const double* __restrict__ P, int64_t cnt, int64_t step, int g)
{
  double a0 = 0.0, a1 = 0.0, a2 = 0.0, a3 = 0.0;
  int64_t m = g;
  for (; m + 3 * G < cnt; m += 4 * G) {
    a0 += P[m * step];
    a1 += P[(m + G) * step];
    a2 += P[(m + 2 * G) * step];
    a3 += P[(m + 3 * G) * step];
  }
  for (; m < cnt; m += G) a0 += P[m * step];
  return (a0 + a1) + (a2 + a3);
}
#endif

// Work items counted in runs of C 64-source blocks (C = 1, 2 or a multiple
// of 4; four blocks to a tile): target tile t has ceil(4 (n_t - t) / C)
// runs of blocks from its own tile on.  sum_{m=1}^{M} ceil(4 m / C):
RMW_TP int64_t block_ceil_sum(int64_t M, int64_t C)
{
  return (C & 3) == 0 ? ceil_sum(M, C >> 2) : (C == 1 ? 4 : 2) * (M * (M + 1) >> 1);
}

RMW_TP int64_t items_before_blocks(int64_t T, int64_t n_t, int64_t C)
{
  return block_ceil_sum(n_t, C) - block_ceil_sum(n_t - T, C);
}

RMW_TP int64_t pair_target_tile_blocks(int64_t local, int64_t n_t, int64_t C)
{
  int64_t lo = 0, hi = n_t - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (items_before_blocks(mid, n_t, C) <= local) lo = mid; else hi = mid - 1;
  }
  return lo;
}

RMW_TP int64_t pair_target_tile(int64_t local, int64_t n_t, int64_t C)
{
  int64_t lo = 0, hi = n_t - 1;
  while (lo < hi) {
    const int64_t mid = (lo + hi + 1) >> 1;
    if (items_before(mid, n_t, C) <= local) lo = mid; else hi = mid - 1;
  }
  return lo;
}

}  // namespace rmw
