// Per-target accumulator policies of the pair kernels and the minimum-image
// shift count, shared by the kernel files.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/rmw_hip.h"

namespace rmw {

// Per-target accumulator policies.
template <int ACC> struct Acc;
template <> struct Acc<RMW_PREC_F32_F64ACC> {
  static constexpr bool kChain = false;
  double v = 0.0;
  __device__ __forceinline__ void add(float x) { v += (double)x; }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ double get() const { return v; }
};
// kChain: the running sum may seed a product's FMA chain (run()), which
// saves the separate add of each term.
template <> struct Acc<RMW_PREC_F32_BLOCKED> {
  static constexpr bool kChain = true;
  double v = 0.0;
  float p = 0.0f;
  __device__ __forceinline__ void add(float x) { p += x; }
  __device__ __forceinline__ float& run() { return p; }
  __device__ __forceinline__ void flush()
  {
    v += (double)p;
    p = 0.0f;
  }
  __device__ __forceinline__ double get() const { return v + (double)p; }
};
template <> struct Acc<RMW_PREC_F32_KAHAN> {
  static constexpr bool kChain = false;
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
  static constexpr bool kChain = true;
  double v = 0.0;
  __device__ __forceinline__ void add(double x) { v += x; }
  __device__ __forceinline__ double& run() { return v; }
  __device__ __forceinline__ void flush() {}
  __device__ __forceinline__ double get() const { return v; }
};

constexpr int kSub = 16;  // sources per FP32 sub-sum in the blocked mode

// Minimum-image shift count of mobility_numba.py:184-193: int(d/L + 0.5 sign(d)).
__device__ __forceinline__ double image_count(double d, double L)
{
  double s = (double)((d > 0.0) - (d < 0.0));
  return trunc(d / L + 0.5 * s);
}

}  // namespace rmw
