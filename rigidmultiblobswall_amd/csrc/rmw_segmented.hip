// Segmented FP64 Krylov kernels for ragged batches of independent
// suspensions stacked in one vector (Lanczos and GMRES over many replicas,
// every step one launch sequence for the whole batch).
//
// System s owns blobs [off[s], off[s+1]) of the batch.  A vector holds
// `nparts` parts (the forces and the torques of the grand mobility, say),
// part p starting at p * part_stride, each with `comp` values per blob, so
// system s owns elements p * part_stride + comp * b + c, off[s] <= b <
// off[s+1].  Work is cut into units of kUnit blobs; a unit that straddles
// systems handles each of them in turn.  A reduction writes one partial per
// (unit, system) pair at index unit + system -- unique, because consecutive
// units share at most their boundary system -- and a second kernel sums
// each system's partials in unit order, so results are bit-reproducible
// and independent of how the batch is launched.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>

#include "rmw_internal.h"

namespace rmw {

constexpr int kSegBlock = 256;
constexpr int kUnit = 128;  // blobs per unit
constexpr int kSegMaxVec = 1024;

struct Seg {
  const int64_t* off;
  int nsys;
  int comp;
  int nparts;
  int64_t nblobs;
  int64_t part_stride;
};

// Largest s in [0, nsys) with off[s] <= b: the (non-empty) system of blob b.
__device__ inline int seg_of(const Seg& S, int64_t b)
{
  int lo = 0, hi = S.nsys - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (S.off[mid] <= b) lo = mid; else hi = mid - 1;
  }
  return lo;
}

// Sum over the 256 threads; the result is valid in thread 0.
__device__ inline double block_sum(double v, double* red)
{
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  __syncthreads();  // red[] of the previous call has been read
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
  if (threadIdx.x == 0)
    for (int k = 0; k < kSegBlock / 64; ++k) s += red[k];
  return s;
}

// partial[(u + s) * nvec + j] = V_j|s,u . w|s,u  (u = this block's unit)
__global__ __launch_bounds__(kSegBlock) void seg_dot_partial(const double* __restrict__ V, int64_t ldv, int nvec,
                                                             const double* __restrict__ w, Seg S,
                                                             double* __restrict__ partial)
{
  __shared__ double red[kSegBlock / 64];
  const int64_t u = blockIdx.x;
  const int64_t b0 = u * kUnit;
  const int64_t b1 = b0 + kUnit < S.nblobs ? b0 + kUnit : S.nblobs;
  const int s0 = seg_of(S, b0), s1 = seg_of(S, b1 - 1);
  for (int s = s0; s <= s1; ++s) {
    const int64_t lo = max(b0, S.off[s]), hi = min(b1, S.off[s + 1]);
    const int64_t cnt = hi > lo ? (hi - lo) * S.comp : 0;
    for (int j = 0; j < nvec; ++j) {
      const double* v = V + (int64_t)j * ldv;
      double acc = 0.0;
      for (int p = 0; p < S.nparts; ++p) {
        const int64_t base = p * S.part_stride + lo * S.comp;
        for (int64_t e = threadIdx.x; e < cnt; e += kSegBlock) acc += v[base + e] * w[base + e];
      }
      const double tot = block_sum(acc, red);
      if (threadIdx.x == 0) partial[(u + s) * nvec + j] = tot;
    }
  }
}

// out[s * ldo + j] (+)= sum over the units of system s, in unit order.
__global__ void seg_dot_final(const double* __restrict__ partial, int nvec, Seg S, double* __restrict__ out,
                              int64_t ldo, int accumulate, int sqrt_out)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)S.nsys * nvec) return;
  const int s = (int)(t / nvec), j = (int)(t - (int64_t)s * nvec);
  const int64_t a = S.off[s], b = S.off[s + 1];
  double acc = 0.0;
  if (b > a)
    for (int64_t u = a / kUnit; u <= (b - 1) / kUnit; ++u) acc += partial[(u + s) * nvec + j];
  if (sqrt_out) acc = sqrt(acc);
  double* o = out + (int64_t)s * ldo + j;
  *o = accumulate ? *o + acc : acc;
}

// y|s = sgn * sum_j coef[s * ldc + j] V_j|s (+ y|s when accumulate)
__global__ __launch_bounds__(kSegBlock) void seg_combination(const double* __restrict__ V, int64_t ldv, int nvec,
                                                             const double* __restrict__ coef, int64_t ldc,
                                                             double* __restrict__ y, Seg S, int accumulate,
                                                             double sgn)
{
  __shared__ double c[kSegMaxVec];
  const int64_t u = blockIdx.x;
  const int64_t b0 = u * kUnit;
  const int64_t b1 = b0 + kUnit < S.nblobs ? b0 + kUnit : S.nblobs;
  const int s0 = seg_of(S, b0), s1 = seg_of(S, b1 - 1);
  for (int s = s0; s <= s1; ++s) {
    const int64_t lo = max(b0, S.off[s]), hi = min(b1, S.off[s + 1]);
    if (hi <= lo) continue;  // uniform over the block
    __syncthreads();
    for (int j = threadIdx.x; j < nvec; j += kSegBlock) c[j] = sgn * coef[(int64_t)s * ldc + j];
    __syncthreads();
    const int64_t cnt = (hi - lo) * S.comp;
    for (int p = 0; p < S.nparts; ++p) {
      const int64_t base = p * S.part_stride + lo * S.comp;
      for (int64_t e = threadIdx.x; e < cnt; e += kSegBlock) {
        const int64_t i = base + e;
        double acc = accumulate ? y[i] : 0.0;
        for (int j = 0; j < nvec; ++j) acc += c[j] * V[(int64_t)j * ldv + i];
        y[i] = acc;
      }
    }
  }
}

// y|s = x|s / div[s * ld_div]; a non-positive divisor writes the unit vector
// e_0 of the system (first value of its first blob, part 0).
__global__ __launch_bounds__(kSegBlock) void seg_scale(const double* __restrict__ x, double* __restrict__ y,
                                                       const double* __restrict__ div, int64_t ld_div, Seg S)
{
  const int64_t u = blockIdx.x;
  const int64_t b0 = u * kUnit;
  const int64_t b1 = b0 + kUnit < S.nblobs ? b0 + kUnit : S.nblobs;
  const int s0 = seg_of(S, b0), s1 = seg_of(S, b1 - 1);
  for (int s = s0; s <= s1; ++s) {
    const int64_t lo = max(b0, S.off[s]), hi = min(b1, S.off[s + 1]);
    if (hi <= lo) continue;
    const double d = div[(int64_t)s * ld_div];
    const bool unit = !(d > 0.0);
    const double inv = unit ? 0.0 : 1.0 / d;
    const int64_t cnt = (hi - lo) * S.comp;
    for (int p = 0; p < S.nparts; ++p) {
      const int64_t base = p * S.part_stride + lo * S.comp;
      for (int64_t e = threadIdx.x; e < cnt; e += kSegBlock) {
        const int64_t i = base + e;
        y[i] = unit ? ((p == 0 && lo == S.off[s] && e == 0) ? 1.0 : 0.0) : x[i] * inv;
      }
    }
  }
}

// h[s * ldh + j] += a[s * lda + j], j < m
__global__ void seg_add(const double* __restrict__ a, int64_t lda, double* __restrict__ h, int64_t ldh, int m,
                        int nsys)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nsys * m) return;
  const int s = (int)(t / m), j = (int)(t - (int64_t)s * m);
  h[(int64_t)s * ldh + j] += a[(int64_t)s * lda + j];
}

static Seg make_seg(const rmw_segments* g)
{
  RMW_REQUIRE(g != nullptr, "segments is null");
  RMW_REQUIRE(g->offsets != nullptr && g->nsys >= 1, "segments need device offsets and nsys >= 1");
  RMW_REQUIRE(g->comp >= 1 && g->nparts >= 1 && g->nblobs >= 0, "bad segment layout");
  RMW_REQUIRE(g->nparts == 1 || g->part_stride >= (int64_t)g->comp * g->nblobs,
              "part_stride must be >= comp * nblobs");
  Seg S;
  S.off = g->offsets;
  S.nsys = g->nsys;
  S.comp = g->comp;
  S.nparts = g->nparts;
  S.nblobs = g->nblobs;
  S.part_stride = g->part_stride;
  return S;
}

static int64_t seg_units(const Seg& S) { return (S.nblobs + kUnit - 1) / kUnit; }

static size_t seg_partial_doubles(const Seg& S, int nvec) { return (size_t)(seg_units(S) + S.nsys) * nvec; }

static void seg_dot(const double* V, int64_t ldv, int nvec, const double* w, const Seg& S, double* out, int64_t ldo,
                    int accumulate, int sqrt_out, double* partial, hipStream_t st)
{
  const int64_t nu = seg_units(S);
  if (nu > 0) hipLaunchKernelGGL(seg_dot_partial, dim3((unsigned)nu), dim3(kSegBlock), 0, st, V, ldv, nvec, w, S,
                                 partial);
  const int64_t nt = (int64_t)S.nsys * nvec;
  hipLaunchKernelGGL(seg_dot_final, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, partial, nvec, S, out, ldo,
                     accumulate, sqrt_out);
}

static void seg_comb(const double* V, int64_t ldv, int nvec, const double* coef, int64_t ldc, double* y,
                     const Seg& S, int accumulate, double sgn, hipStream_t st)
{
  const int64_t nu = seg_units(S);
  if (nu > 0) hipLaunchKernelGGL(seg_combination, dim3((unsigned)nu), dim3(kSegBlock), 0, st, V, ldv, nvec, coef,
                                 ldc, y, S, accumulate, sgn);
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_seg_workspace_size(const rmw_segments* seg, int max_vectors, size_t* bytes)
{
  return guarded([&] {
    const Seg S = make_seg(seg);
    RMW_REQUIRE(bytes && max_vectors >= 1 && max_vectors <= kSegMaxVec, "bad arguments");
    *bytes = align_up(seg_partial_doubles(S, max_vectors) * sizeof(double));
  });
}

extern "C" int rmw_seg_dot_many(const double* V, int64_t ldv, int nvec, const double* w, const rmw_segments* seg,
                                double* out, int64_t ldo, int sqrt_out, void* workspace, size_t workspace_bytes,
                                void* stream)
{
  return guarded([&] {
    const Seg S = make_seg(seg);
    RMW_REQUIRE(nvec >= 1 && nvec <= kSegMaxVec && ldo >= nvec, "bad sizes");
    RMW_REQUIRE(workspace_bytes >= seg_partial_doubles(S, nvec) * sizeof(double), "workspace too small");
    seg_dot(V, ldv, nvec, w, S, out, ldo, 0, sqrt_out, (double*)workspace, (hipStream_t)stream);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_seg_combination(const double* V, int64_t ldv, int nvec, const double* coef, int64_t ldc,
                                   double* y, const rmw_segments* seg, int accumulate, void* stream)
{
  return guarded([&] {
    const Seg S = make_seg(seg);
    RMW_REQUIRE(nvec >= 1 && nvec <= kSegMaxVec && ldc >= nvec, "bad sizes");
    seg_comb(V, ldv, nvec, coef, ldc, y, S, accumulate, 1.0, (hipStream_t)stream);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_seg_scale(const double* x, double* y, const double* div, int64_t ld_div, const rmw_segments* seg,
                             void* stream)
{
  return guarded([&] {
    const Seg S = make_seg(seg);
    RMW_REQUIRE(ld_div >= 1 || S.nsys == 1, "bad ld_div");
    const int64_t nu = seg_units(S);
    if (nu > 0) hipLaunchKernelGGL(seg_scale, dim3((unsigned)nu), dim3(kSegBlock), 0, (hipStream_t)stream, x, y,
                                   div, ld_div, S);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_seg_arnoldi_cgs2(double* V, int64_t ldv, int j, double* w, const rmw_segments* seg, double* H,
                                    int64_t h_sys_stride, double* hscratch, void* workspace, size_t workspace_bytes,
                                    void* stream)
{
  return guarded([&] {
    const Seg S = make_seg(seg);
    const int m = j + 1;
    RMW_REQUIRE(m >= 1 && m + 1 <= kSegMaxVec, "bad Arnoldi index");
    RMW_REQUIRE(h_sys_stride >= m + 1 || S.nsys == 1, "h_sys_stride too small");
    RMW_REQUIRE(workspace_bytes >= seg_partial_doubles(S, m) * sizeof(double), "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    double* partial = (double*)workspace;
    // pass 1: h = V^T w, w -= V h
    seg_dot(V, ldv, m, w, S, H, h_sys_stride, 0, 0, partial, st);
    seg_comb(V, ldv, m, H, h_sys_stride, w, S, 1, -1.0, st);
    // pass 2 (re-orthogonalisation): h2 = V^T w, w -= V h2, h += h2
    seg_dot(V, ldv, m, w, S, hscratch, m, 0, 0, partial, st);
    seg_comb(V, ldv, m, hscratch, m, w, S, 1, -1.0, st);
    const int64_t nt = (int64_t)S.nsys * m;
    hipLaunchKernelGGL(seg_add, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, st, hscratch, (int64_t)m, H,
                       h_sys_stride, m, S.nsys);
    // ||w|| per system and the next basis vector
    seg_dot(w, 0, 1, w, S, H + m, h_sys_stride, 0, 1, partial, st);
    const int64_t nu = seg_units(S);
    if (nu > 0) hipLaunchKernelGGL(seg_scale, dim3((unsigned)nu), dim3(kSegBlock), 0, st, w, V + (int64_t)m * ldv,
                                   H + m, h_sys_stride, S);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
