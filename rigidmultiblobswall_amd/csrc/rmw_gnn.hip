// Kernels of the learned GMRES warm start: a BF16 MFMA GEMM (FP32
// accumulation) from which every layer of the per-body graph network and
// its backward pass is built, the nearest-neighbour graph and features,
// and the elementwise pieces (SiLU and its derivative, message sums, loss,
// Adam).  No reference counterpart: the network only supplies the initial
// guess of the rigid-body GMRES solve.
//
// GEMM: C = op(A) op(B) (+ bias) with op(A) (M x K) and op(B) (K x N) in
// BF16 (stored as uint16), FP32 accumulation by v_mfma_f32_16x16x16bf16_1k.
// Workgroup tile 64 x 64 (four wavefronts, 32 x 32 each = 2 x 2 MFMA
// tiles), K staged through LDS 32 at a time.  MFMA operand layout (16x16x16,
// 64 lanes): lane l supplies A[l % 16][4 (l / 16) + 0..3] and
// B[4 (l / 16) + 0..3][l % 16] and receives D[4 (l / 16) + i][l % 16].
#include <hip/hip_runtime.h>

#include <climits>

#include <algorithm>
#include <cmath>
#include <cstdint>

#include "rmw_internal.h"

namespace rmw {

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(((uint32_t)h) << 16); }

// round to nearest even
__device__ __forceinline__ uint16_t f2bf(float f)
{
  uint32_t u = __float_as_uint(f);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  u += 0x7fffu + ((u >> 16) & 1u);
  return (uint16_t)(u >> 16);
}

constexpr int GM = 64, GN = 64, GK = 32, GKP = GK + 8;  // LDS pitch in bf16 (80 B rows)

struct GemmParams {
  const uint16_t* A; int64_t lda; int transA;
  const uint16_t* B; int64_t ldb; int transB;
  float* C; int64_t ldc; float beta;      // optional FP32 output, C = AB + bias + beta C
  uint16_t* H; int64_t ldh;               // optional BF16 output of act(AB + bias)
  const float* bias;                      // optional, length N
  int act;                                // 0 none, 1 SiLU
  int M, N, K;
  int kchunk;                             // K range per blockIdx.z (split K)
  float* partial;                         // split K: raw sums [z][M][N], no epilogue
};

__global__ __launch_bounds__(256) void gemm_bf16_kernel(GemmParams p)
{
  __shared__ __align__(16) uint16_t As[GM * GKP];
  __shared__ __align__(16) uint16_t Bs[GN * GKP];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int wm = wv >> 1, wn = wv & 1;
  const int m0 = blockIdx.y * GM, n0 = blockIdx.x * GN;
  f32x4 acc[2][2];
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) acc[i][j] = f32x4{0.f, 0.f, 0.f, 0.f};
  const int kb = blockIdx.z * p.kchunk;
  const int ke = kb + p.kchunk < p.K ? kb + p.kchunk : p.K;

  for (int k0 = kb; k0 < ke; k0 += GK) {
    // stage A tile (rows m, cols k) and B tile (rows n, cols k) as [row][k];
    // consecutive threads take consecutive addresses of the operand in memory
    for (int e = tid; e < GM * GK; e += 256) {
      int r, c;
      if (p.transA) { c = e / GM; r = e - c * GM; } else { r = e / GK; c = e - r * GK; }
      const int m = m0 + r, k = k0 + c;
      uint16_t va = 0;
      if (m < p.M && k < ke) va = p.transA ? p.A[(int64_t)k * p.lda + m] : p.A[(int64_t)m * p.lda + k];
      As[r * GKP + c] = va;
    }
    for (int e = tid; e < GN * GK; e += 256) {
      int r, c;
      if (!p.transB) { c = e / GN; r = e - c * GN; } else { r = e / GK; c = e - r * GK; }
      const int n = n0 + r, k = k0 + c;
      uint16_t vb = 0;
      if (n < p.N && k < ke) vb = p.transB ? p.B[(int64_t)n * p.ldb + k] : p.B[(int64_t)k * p.ldb + n];
      Bs[r * GKP + c] = vb;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < GK; kk += 16) {
      const int kq = kk + 4 * (lane >> 4);
      bf16x4 af[2], bfr[2];
      for (int i = 0; i < 2; ++i) {
        const int row = wm * 32 + i * 16 + (lane & 15);
        af[i] = *reinterpret_cast<const bf16x4*>(&As[row * GKP + kq]);
      }
      for (int j = 0; j < 2; ++j) {
        const int col = wn * 32 + j * 16 + (lane & 15);
        bfr[j] = *reinterpret_cast<const bf16x4*>(&Bs[col * GKP + kq]);
      }
      for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(af[i], bfr[j], acc[i][j], 0, 0, 0);
    }
    __syncthreads();
  }
  if (p.partial) {
    // split K: this slice's raw sums
    float* P = p.partial + (int64_t)blockIdx.z * p.M * p.N;
    for (int i = 0; i < 2; ++i)
      for (int j = 0; j < 2; ++j) {
        const int n = n0 + wn * 32 + j * 16 + (lane & 15);
        if (n >= p.N) continue;
        for (int q = 0; q < 4; ++q) {
          const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
          if (m < p.M) P[(int64_t)m * p.N + n] = acc[i][j][q];
        }
      }
    return;
  }
  // epilogue
  for (int i = 0; i < 2; ++i)
    for (int j = 0; j < 2; ++j) {
      const int n = n0 + wn * 32 + j * 16 + (lane & 15);
      if (n >= p.N) continue;
      const float b = p.bias ? p.bias[n] : 0.f;
      for (int q = 0; q < 4; ++q) {
        const int m = m0 + wm * 32 + i * 16 + 4 * (lane >> 4) + q;
        if (m >= p.M) continue;
        float z = acc[i][j][q] + b;
        if (p.C) {
          float* c = p.C + (int64_t)m * p.ldc + n;
          *c = p.beta != 0.f ? z + p.beta * *c : z;
        }
        if (p.H) {
          float h = z;
          if (p.act == 1) h = z / (1.f + __expf(-z));
          p.H[(int64_t)m * p.ldh + n] = f2bf(h);
        }
      }
    }
}

// ---------------------------------------------------------------------------
// elementwise pieces
// ---------------------------------------------------------------------------
__global__ void f32_to_bf16_kernel(const float* __restrict__ x, uint16_t* __restrict__ y, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    y[i] = f2bf(x[i]);
}

// dZ = dH * silu'(Z)  (bf16 out), silu'(z) = s (1 + z (1 - s)), s = sigmoid(z)
__global__ void silu_backward_kernel(const float* __restrict__ dH, const float* __restrict__ Z,
                                     uint16_t* __restrict__ dZ, int64_t n)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float z = Z[i];
    const float s = 1.f / (1.f + __expf(-z));
    dZ[i] = f2bf(dH[i] * s * (1.f + z * (1.f - s)));
  }
}

// split K: C = sum_z partial[z] (fixed order) + beta C
__global__ void gemm_splitk_reduce_kernel(const float* __restrict__ partial, int splits, int M, int N,
                                          float* __restrict__ C, int64_t ldc, float beta)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)M * N) return;
  const int m = (int)(t / N), n = (int)(t - (int64_t)m * N);
  float s = 0.f;
  for (int z = 0; z < splits; ++z) s += partial[(int64_t)z * M * N + t];
  float* c = C + (int64_t)m * ldc + n;
  *c = beta != 0.f ? s + beta * *c : s;
}

// column sums: out[j] (+)= sum_i X[i * ldx + j] over i < rows (bf16 input).
// One 1024-thread workgroup per 64 columns: thread t sums columns
// 4 (t mod 16) .. +3 over rows t/16, t/16 + 64, ...; then the 64 row phases
// are added in order.
constexpr int kColsBlock = 1024, kColsPhases = kColsBlock / 16;

__global__ __launch_bounds__(kColsBlock) void colsum_bf16_kernel(const uint16_t* __restrict__ X, int64_t ldx,
                                                                 int rows, int cols, float* __restrict__ out,
                                                                 float beta)
{
  __shared__ float part[kColsPhases][64];
  const int t = threadIdx.x, cg = t & 15, ph = t >> 4;
  const int j0 = blockIdx.x * 64 + 4 * cg;
  float s[4] = {0.f, 0.f, 0.f, 0.f};
  for (int i = ph; i < rows; i += kColsPhases) {
    const uint16_t* row = X + (int64_t)i * ldx;
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (j0 + q < cols) s[q] += bf2f(row[j0 + q]);
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) part[ph][4 * cg + q] = s[q];
  __syncthreads();
  if (t < 64) {
    const int j = blockIdx.x * 64 + t;
    float v = 0.f;
    for (int k = 0; k < kColsPhases; ++k) v += part[k][t];
    if (j < cols) out[j] = beta != 0.f ? v + beta * out[j] : v;
  }
}

// node input: Xc[b] = [Xn[b] (nin) | sum_{e<deg} M[b*deg+e] (nmsg) | 0 pad], bf16
__global__ void gather_messages_kernel(const float* __restrict__ Xn, int nin, const float* __restrict__ Msg,
                                       int nmsg, int deg, int nb, uint16_t* __restrict__ Xc, int ldc)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb * ldc) return;
  const int b = (int)(t / ldc), c = (int)(t - (int64_t)(t / ldc) * ldc);
  float v = 0.f;
  if (c < nin) {
    v = Xn[(int64_t)b * nin + c];
  } else if (c < nin + nmsg) {
    const int j = c - nin;
    for (int e = 0; e < deg; ++e) v += Msg[((int64_t)b * deg + e) * nmsg + j];
  }
  Xc[t] = f2bf(v);
}

// dMsg broadcast: dE[b*deg+e][j] = dXc[b][nin + j]   (bf16 out from fp32 in)
__global__ void scatter_messages_kernel(const float* __restrict__ dXc, int ldc, int nin, int nmsg, int deg, int nb,
                                        uint16_t* __restrict__ dE)
{
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)nb * deg * nmsg) return;
  const int64_t row = t / nmsg;
  const int j = (int)(t - row * nmsg);
  const int b = (int)(row / deg);
  dE[t] = f2bf(dXc[(int64_t)b * ldc + nin + j]);
}

// weighted MSE: loss_part[block] and dY = 2 w (Y - T) / (rows) as bf16
__global__ void mse_kernel(const float* __restrict__ Y, const float* __restrict__ T, const float* __restrict__ w,
                           int rows, int cols, uint16_t* __restrict__ dY, float* __restrict__ loss_part)
{
  __shared__ float red[256];
  float acc = 0.f;
  const int64_t n = (int64_t)rows * cols;
  const float inv = 1.f / (float)rows;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const int c = (int)(i % cols);
    const float d = Y[i] - T[i];
    const float wc = w ? w[c] : 1.f;
    acc += wc * d * d;
    dY[i] = f2bf(2.f * wc * d * inv);
  }
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_part[blockIdx.x] = red[0] * inv;
}

// Adam on FP32 master weights, BF16 copy refreshed
__global__ void adam_kernel(float* __restrict__ w, const float* __restrict__ g, float* __restrict__ m,
                            float* __restrict__ v, uint16_t* __restrict__ wb, int64_t n, float lr, float b1, float b2,
                            float eps, float bc1, float bc2, float wd)
{
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
    const float gi = g[i] + wd * w[i];
    const float mi = b1 * m[i] + (1.f - b1) * gi;
    const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    const float wi = w[i] - lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
    w[i] = wi;
    if (wb) wb[i] = f2bf(wi);
  }
}

// ---------------------------------------------------------------------------
// graph and features (FP64 geometry in, FP32 features out)
// ---------------------------------------------------------------------------
__device__ inline void quat_to_rot(const double* q, double R[9])
{
  const double s = q[0], p0 = q[1], p1 = q[2], p2 = q[3], dg = s * s - 0.5;
  R[0] = 2 * (p0 * p0 + dg); R[1] = 2 * (p0 * p1 - s * p2); R[2] = 2 * (p0 * p2 + s * p1);
  R[3] = 2 * (p1 * p0 + s * p2); R[4] = 2 * (p1 * p1 + dg); R[5] = 2 * (p1 * p2 - s * p0);
  R[6] = 2 * (p2 * p0 - s * p1); R[7] = 2 * (p2 * p1 + s * p0); R[8] = 2 * (p2 * p2 + dg);
}

// 8 nearest neighbours of every body by centre distance (ties by index)
// k nearest neighbours of each body (ascending distance, ties by index,
// the body itself excluded): one workgroup per body.  Each thread keeps a
// sorted list of its nearest candidates (bodies j = tid, tid + 256, ...)
// in registers -- a compare-and-swap insertion over compile-time slots --
// then deg rounds of a block-wide (distance, index) minimum over the list
// heads take the overall nearest in order.
constexpr int kKnnBlock = 256;
constexpr int kKnnMax = 16;

__device__ __forceinline__ bool knn_less(double da, int ia, double db, int ib)
{
  return da < db || (da == db && ia < ib);
}

__global__ __launch_bounds__(kKnnBlock) void knn_kernel(const double* __restrict__ c, int nb, int deg,
                                                        int* __restrict__ nbr)
{
  __shared__ double s_d[kKnnBlock / 64];
  __shared__ int s_i[kKnnBlock / 64];
  const int b = blockIdx.x;
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const double x = c[3 * b], y = c[3 * b + 1], z = c[3 * b + 2];
  double bd[kKnnMax];
  int bi[kKnnMax];
#pragma unroll
  for (int e = 0; e < kKnnMax; ++e) { bd[e] = INFINITY; bi[e] = INT_MAX; }
  for (int j = tid; j < nb; j += kKnnBlock) {
    if (j == b) continue;
    const double dx = c[3 * j] - x, dy = c[3 * j + 1] - y, dz = c[3 * j + 2] - z;
    double d = dx * dx + dy * dy + dz * dz;
    int id = j;
    // insertion: the candidate sinks to its slot, the slot's entry moves down
#pragma unroll
    for (int e = 0; e < kKnnMax; ++e) {
      if (e < deg && knn_less(d, id, bd[e], bi[e])) {
        const double td = bd[e]; const int ti = bi[e];
        bd[e] = d; bi[e] = id;
        d = td; id = ti;
      }
    }
  }
  for (int r = 0; r < deg; ++r) {
    double d = bd[0];
    int id = bi[0];
    for (int o = 32; o > 0; o >>= 1) {
      const double od = __shfl_xor(d, o, 64);
      const int oi = __shfl_xor(id, o, 64);
      if (knn_less(od, oi, d, id)) { d = od; id = oi; }
    }
    if (lane == 0) { s_d[wv] = d; s_i[wv] = id; }
    __syncthreads();
    d = s_d[0]; id = s_i[0];
#pragma unroll
    for (int w2 = 1; w2 < kKnnBlock / 64; ++w2)
      if (knn_less(s_d[w2], s_i[w2], d, id)) { d = s_d[w2]; id = s_i[w2]; }
    __syncthreads();   // s_d / s_i are rewritten next round
    if (tid == 0) nbr[(int64_t)b * deg + r] = id == INT_MAX ? b : id;
    if (bi[0] == id) {
      // the winner's list pops its head
#pragma unroll
      for (int e = 0; e < kKnnMax - 1; ++e) { bd[e] = bd[e + 1]; bi[e] = bi[e + 1]; }
      bd[kKnnMax - 1] = INFINITY; bi[kKnnMax - 1] = INT_MAX;
    }
  }
}

// Node features (10): z-hat in the body frame (3), centre height (1),
//   force and torque in the body frame (6), force scaled by fscale.
// Edge features (16) for neighbour n of body b: R_b^T (c_n - c_b) (3), |d|,
//   1/|d|, relative quaternion q_b^-1 q_n (4), height of n, R_b^T F_n, R_b^T T_n.
__global__ void features_kernel(const double* __restrict__ c, const double* __restrict__ q,
                                const double* __restrict__ ft, const int* __restrict__ nbr, int nb, int deg,
                                double fscale, float* __restrict__ Xn, float* __restrict__ Xe)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double R[9];
  quat_to_rot(q + 4 * b, R);
  // body frame of a lab vector v: R^T v
  auto tobody = [&](const double* v, double* o) {
    o[0] = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
    o[1] = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
    o[2] = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  };
  const double zhat[3] = {0.0, 0.0, 1.0};
  double t3[3];
  float* xn = Xn + (int64_t)b * 10;
  tobody(zhat, t3);
  xn[0] = (float)t3[0]; xn[1] = (float)t3[1]; xn[2] = (float)t3[2];
  xn[3] = (float)c[3 * b + 2];
  tobody(ft + 6 * b, t3);
  xn[4] = (float)(t3[0] / fscale); xn[5] = (float)(t3[1] / fscale); xn[6] = (float)(t3[2] / fscale);
  tobody(ft + 6 * b + 3, t3);
  xn[7] = (float)(t3[0] / fscale); xn[8] = (float)(t3[1] / fscale); xn[9] = (float)(t3[2] / fscale);
  const double* qb = q + 4 * b;
  for (int e = 0; e < deg; ++e) {
    const int n = nbr[b * deg + e];
    float* xe = Xe + ((int64_t)b * deg + e) * 16;
    const double dl[3] = {c[3 * n] - c[3 * b], c[3 * n + 1] - c[3 * b + 1], c[3 * n + 2] - c[3 * b + 2]};
    tobody(dl, t3);
    const double dn = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    xe[0] = (float)t3[0]; xe[1] = (float)t3[1]; xe[2] = (float)t3[2];
    xe[3] = (float)dn;
    xe[4] = (float)(dn > 0 ? 1.0 / dn : 0.0);
    // q_b^-1 q_n  (conjugate of b times n), sign fixed so that the scalar part is >= 0
    const double* qn = q + 4 * n;
    const double s1 = qb[0], x1 = -qb[1], y1 = -qb[2], z1 = -qb[3];
    const double s2 = qn[0], x2 = qn[1], y2 = qn[2], z2 = qn[3];
    double rs = s1 * s2 - x1 * x2 - y1 * y2 - z1 * z2;
    double rx = s1 * x2 + s2 * x1 + y1 * z2 - z1 * y2;
    double ry = s1 * y2 + s2 * y1 + z1 * x2 - x1 * z2;
    double rz = s1 * z2 + s2 * z1 + x1 * y2 - y1 * x2;
    if (rs < 0) { rs = -rs; rx = -rx; ry = -ry; rz = -rz; }
    xe[5] = (float)rs; xe[6] = (float)rx; xe[7] = (float)ry; xe[8] = (float)rz;
    xe[9] = (float)c[3 * n + 2];
    tobody(ft + 6 * n, t3);
    xe[10] = (float)(t3[0] / fscale); xe[11] = (float)(t3[1] / fscale); xe[12] = (float)(t3[2] / fscale);
    tobody(ft + 6 * n + 3, t3);
    xe[13] = (float)(t3[0] / fscale); xe[14] = (float)(t3[1] / fscale); xe[15] = (float)(t3[2] / fscale);
  }
}

// Features of the motion since a previous configuration (c0, q0, ft0), for
// the network that corrects the previous step's solution.
// Node features (22): the 10 of features_kernel, then in the body frame
//   the displacement R_b^T (c_b - c0_b) / dscale (3), the rotation vector of
//   q_b q0_b^-1 / rscale (3), and the changes of force and torque / fscale (6).
// Edge features (16): those of features_kernel with entries 5..8 (the
//   relative orientation) replaced by the change of the offset to the
//   neighbour since the previous configuration, R_b^T (d - d0) / dscale (3),
//   and the change of its length / dscale (1).
__global__ void motion_features_kernel(const double* __restrict__ c, const double* __restrict__ q,
                                       const double* __restrict__ ft, const double* __restrict__ c0,
                                       const double* __restrict__ q0, const double* __restrict__ ft0,
                                       const int* __restrict__ nbr, int nb, int deg, double fscale, double dscale,
                                       double rscale, float* __restrict__ Xn, float* __restrict__ Xe)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  double R[9];
  quat_to_rot(q + 4 * b, R);
  auto tobody = [&](const double* v, double* o) {
    o[0] = R[0] * v[0] + R[3] * v[1] + R[6] * v[2];
    o[1] = R[1] * v[0] + R[4] * v[1] + R[7] * v[2];
    o[2] = R[2] * v[0] + R[5] * v[1] + R[8] * v[2];
  };
  const double zhat[3] = {0.0, 0.0, 1.0};
  double t3[3];
  float* xn = Xn + (int64_t)b * 22;
  tobody(zhat, t3);
  xn[0] = (float)t3[0]; xn[1] = (float)t3[1]; xn[2] = (float)t3[2];
  xn[3] = (float)c[3 * b + 2];
  tobody(ft + 6 * b, t3);
  xn[4] = (float)(t3[0] / fscale); xn[5] = (float)(t3[1] / fscale); xn[6] = (float)(t3[2] / fscale);
  tobody(ft + 6 * b + 3, t3);
  xn[7] = (float)(t3[0] / fscale); xn[8] = (float)(t3[1] / fscale); xn[9] = (float)(t3[2] / fscale);
  const double dc[3] = {c[3 * b] - c0[3 * b], c[3 * b + 1] - c0[3 * b + 1], c[3 * b + 2] - c0[3 * b + 2]};
  tobody(dc, t3);
  xn[10] = (float)(t3[0] / dscale); xn[11] = (float)(t3[1] / dscale); xn[12] = (float)(t3[2] / dscale);
  // q q0^-1 = (s, p) * (s0, -p0): the rotation that carried q0 to q (lab frame)
  {
    const double s1 = q[4 * b], x1 = q[4 * b + 1], y1 = q[4 * b + 2], z1 = q[4 * b + 3];
    const double s2 = q0[4 * b], x2 = -q0[4 * b + 1], y2 = -q0[4 * b + 2], z2 = -q0[4 * b + 3];
    double rs = s1 * s2 - x1 * x2 - y1 * y2 - z1 * z2;
    double rx = s1 * x2 + s2 * x1 + y1 * z2 - z1 * y2;
    double ry = s1 * y2 + s2 * y1 + z1 * x2 - x1 * z2;
    double rz = s1 * z2 + s2 * z1 + x1 * y2 - y1 * x2;
    if (rs < 0) { rs = -rs; rx = -rx; ry = -ry; rz = -rz; }
    const double pn = sqrt(rx * rx + ry * ry + rz * rz);
    const double k = pn > 0.0 ? 2.0 * atan2(pn, rs) / pn : 2.0;
    const double th[3] = {k * rx, k * ry, k * rz};
    tobody(th, t3);
    xn[13] = (float)(t3[0] / rscale); xn[14] = (float)(t3[1] / rscale); xn[15] = (float)(t3[2] / rscale);
  }
  const double dF[3] = {ft[6 * b] - ft0[6 * b], ft[6 * b + 1] - ft0[6 * b + 1], ft[6 * b + 2] - ft0[6 * b + 2]};
  const double dT[3] = {ft[6 * b + 3] - ft0[6 * b + 3], ft[6 * b + 4] - ft0[6 * b + 4],
                        ft[6 * b + 5] - ft0[6 * b + 5]};
  tobody(dF, t3);
  xn[16] = (float)(t3[0] / fscale); xn[17] = (float)(t3[1] / fscale); xn[18] = (float)(t3[2] / fscale);
  tobody(dT, t3);
  xn[19] = (float)(t3[0] / fscale); xn[20] = (float)(t3[1] / fscale); xn[21] = (float)(t3[2] / fscale);
  for (int e = 0; e < deg; ++e) {
    const int n = nbr[b * deg + e];
    float* xe = Xe + ((int64_t)b * deg + e) * 16;
    const double dl[3] = {c[3 * n] - c[3 * b], c[3 * n + 1] - c[3 * b + 1], c[3 * n + 2] - c[3 * b + 2]};
    const double d0[3] = {c0[3 * n] - c0[3 * b], c0[3 * n + 1] - c0[3 * b + 1], c0[3 * n + 2] - c0[3 * b + 2]};
    tobody(dl, t3);
    const double dn = sqrt(dl[0] * dl[0] + dl[1] * dl[1] + dl[2] * dl[2]);
    const double dn0 = sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
    xe[0] = (float)t3[0]; xe[1] = (float)t3[1]; xe[2] = (float)t3[2];
    xe[3] = (float)dn;
    xe[4] = (float)(dn > 0 ? 1.0 / dn : 0.0);
    const double dd[3] = {dl[0] - d0[0], dl[1] - d0[1], dl[2] - d0[2]};
    tobody(dd, t3);
    xe[5] = (float)(t3[0] / dscale); xe[6] = (float)(t3[1] / dscale); xe[7] = (float)(t3[2] / dscale);
    xe[8] = (float)((dn - dn0) / dscale);
    xe[9] = (float)c[3 * n + 2];
    tobody(ft + 6 * n, t3);
    xe[10] = (float)(t3[0] / fscale); xe[11] = (float)(t3[1] / fscale); xe[12] = (float)(t3[2] / fscale);
    tobody(ft + 6 * n + 3, t3);
    xe[13] = (float)(t3[0] / fscale); xe[14] = (float)(t3[1] / fscale); xe[15] = (float)(t3[2] / fscale);
  }
}

// Body-frame targets from a solution: T[b] = (R_b^T (lambda_b - lpc_b) per blob,
// R_b^T (U_b - Upc_b)) / fscale, standardised by (mean, inv_std) per column.
// mode 0: solution -> target; mode 1: prediction -> x0 = pc + R_b (pred * std + mean) * fscale
__global__ void body_frame_kernel(const double* __restrict__ q, const int64_t* __restrict__ offs, int nb,
                                  int64_t nblobs, int nbb, double* __restrict__ sol, const double* __restrict__ pc,
                                  float* __restrict__ T, const float* __restrict__ mean, const float* __restrict__ sd,
                                  double fscale, int mode)
{
  const int b = blockIdx.x;
  if (b >= nb) return;
  double R[9];
  quat_to_rot(q + 4 * b, R);
  const int64_t i0 = offs[b];
  const int cols = 3 * nbb + 6;
  for (int c3 = threadIdx.x; c3 < nbb + 2; c3 += blockDim.x) {
    // c3 < nbb: blob c3 of this body; nbb: v; nbb + 1: w
    double* s;
    const double* pcv;
    if (c3 < nbb) { s = sol + 3 * (i0 + c3); pcv = pc + 3 * (i0 + c3); }
    else { s = sol + 3 * nblobs + 6 * (int64_t)b + 3 * (c3 - nbb); pcv = pc + 3 * nblobs + 6 * (int64_t)b + 3 * (c3 - nbb); }
    float* t = T + (int64_t)b * cols + 3 * c3;
    const float* mu = mean + 3 * c3;
    const float* is = sd + 3 * c3;
    if (mode == 0) {
      const double d[3] = {s[0] - pcv[0], s[1] - pcv[1], s[2] - pcv[2]};
      for (int k = 0; k < 3; ++k) {
        const double v = (R[0 + k] * d[0] + R[3 + k] * d[1] + R[6 + k] * d[2]) / fscale;
        t[k] = (float)((v - mu[k]) / is[k]);
      }
    } else {
      double v[3];
      for (int k = 0; k < 3; ++k) v[k] = ((double)t[k] * is[k] + mu[k]) * fscale;
      for (int k = 0; k < 3; ++k) s[k] = pcv[k] + R[3 * k] * v[0] + R[3 * k + 1] * v[1] + R[3 * k + 2] * v[2];
    }
  }
}

static int grid_for(int64_t n) { return (int)std::max<int64_t>(1, std::min<int64_t>((n + 255) / 256, 4096)); }

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_gemm_bf16(int transA, int transB, int M, int N, int K, const uint16_t* A, int64_t lda,
                             const uint16_t* B, int64_t ldb, float* C, int64_t ldc, float beta, uint16_t* H,
                             int64_t ldh, const float* bias, int act, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(M >= 0 && N >= 0 && K >= 0, "bad sizes");
    RMW_REQUIRE(C || H, "no output");
    RMW_REQUIRE(act == 0 || act == 1, "act must be 0 (none) or 1 (SiLU)");
    if (M == 0 || N == 0) return;
    GemmParams p{A, lda, transA, B, ldb, transB, C, ldc, beta, H, ldh, bias, act, M, N, K, K > 0 ? K : GK, nullptr};
    dim3 grid((N + GN - 1) / GN, (M + GM - 1) / GM);
    RMW_REQUIRE(grid.y <= 65535, "M too large");
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, (hipStream_t)stream, p);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gemm_bf16_splitk_workspace(int M, int N, int K, int splits, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(M >= 0 && N >= 0 && K >= 0 && splits >= 1 && bytes, "bad sizes");
    *bytes = (size_t)splits * M * N * sizeof(float);
  });
}

extern "C" int rmw_gemm_bf16_splitk(int transA, int transB, int M, int N, int K, const uint16_t* A, int64_t lda,
                                    const uint16_t* B, int64_t ldb, float* C, int64_t ldc, float beta, int splits,
                                    float* partial, size_t partial_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(M >= 0 && N >= 0 && K >= 0 && splits >= 1 && splits <= 65535, "bad sizes");
    RMW_REQUIRE(C != nullptr, "no output");
    if (M == 0 || N == 0) return;
    // K slices of whole GK steps
    int kchunk = (K + splits - 1) / splits;
    kchunk = (kchunk + GK - 1) / GK * GK;
    if (kchunk <= 0) kchunk = GK;
    const int used = K > 0 ? (K + kchunk - 1) / kchunk : 1;
    RMW_REQUIRE(partial && partial_bytes >= (size_t)used * M * N * sizeof(float), "split-K workspace too small");
    GemmParams p{A, lda, transA, B, ldb, transB, C, ldc, beta, nullptr, 0, nullptr, 0, M, N, K, kchunk, partial};
    dim3 grid((N + GN - 1) / GN, (M + GM - 1) / GM, used);
    RMW_REQUIRE(grid.y <= 65535, "M too large");
    hipStream_t st = (hipStream_t)stream;
    hipLaunchKernelGGL(gemm_bf16_kernel, grid, dim3(256), 0, st, p);
    const int64_t mn = (int64_t)M * N;
    hipLaunchKernelGGL(gemm_splitk_reduce_kernel, dim3((unsigned)((mn + 255) / 256)), dim3(256), 0, st, partial, used,
                       M, N, C, ldc, beta);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_f32_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream)
{
  return guarded([&] {
    hipLaunchKernelGGL(f32_to_bf16_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, x, y, n);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_silu_backward(const float* dH, const float* Z, uint16_t* dZ, int64_t n, void* stream)
{
  return guarded([&] {
    hipLaunchKernelGGL(silu_backward_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, dH, Z, dZ, n);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_colsum_bf16(const uint16_t* X, int64_t ldx, int rows, int cols, float* out, float beta,
                               void* stream)
{
  return guarded([&] {
    hipLaunchKernelGGL(colsum_bf16_kernel, dim3((cols + 63) / 64), dim3(kColsBlock), 0, (hipStream_t)stream, X, ldx,
                       rows, cols, out, beta);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_gather(const float* Xn, int nin, const float* Msg, int nmsg, int deg, int nb, uint16_t* Xc,
                              int ldc, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(ldc >= nin + nmsg, "ldc too small");
    const int64_t n = (int64_t)nb * ldc;
    hipLaunchKernelGGL(gather_messages_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       Xn, nin, Msg, nmsg, deg, nb, Xc, ldc);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_scatter(const float* dXc, int ldc, int nin, int nmsg, int deg, int nb, uint16_t* dE,
                               void* stream)
{
  return guarded([&] {
    const int64_t n = (int64_t)nb * deg * nmsg;
    hipLaunchKernelGGL(scatter_messages_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, (hipStream_t)stream,
                       dXc, ldc, nin, nmsg, deg, nb, dE);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_mse(const float* Y, const float* T, const float* w, int rows, int cols, uint16_t* dY,
                       float* loss_part, int nparts, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(nparts >= 1 && nparts <= 4096, "bad nparts");
    hipLaunchKernelGGL(mse_kernel, dim3(nparts), dim3(256), 0, (hipStream_t)stream, Y, T, w, rows, cols, dY,
                       loss_part);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr, float b1,
                        float b2, float eps, int step, float weight_decay, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(step >= 1, "step counts from 1");
    const float bc1 = 1.f - std::pow(b1, (float)step), bc2 = 1.f - std::pow(b2, (float)step);
    hipLaunchKernelGGL(adam_kernel, dim3(grid_for(n)), dim3(256), 0, (hipStream_t)stream, w, g, m, v, wb, n, lr, b1,
                       b2, eps, bc1, bc2, weight_decay);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_knn(const double* centres, int nb, int deg, int* nbr, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(deg >= 1 && deg <= 16 && nb > deg, "deg must be in [1, 16] and below the body count");
    hipLaunchKernelGGL(knn_kernel, dim3(nb), dim3(kKnnBlock), 0, (hipStream_t)stream, centres, nb, deg, nbr);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_features(const double* centres, const double* quats, const double* force_torque,
                                const int* nbr, int nb, int deg, double fscale, float* Xn, float* Xe, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(fscale > 0, "fscale must be positive");
    hipLaunchKernelGGL(features_kernel, dim3((nb + 127) / 128), dim3(128), 0, (hipStream_t)stream, centres, quats,
                       force_torque, nbr, nb, deg, fscale, Xn, Xe);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_motion_features(const double* centres, const double* quats, const double* force_torque,
                                       const double* centres0, const double* quats0, const double* force_torque0,
                                       const int* nbr, int nb, int deg, double fscale, double dscale, double rscale,
                                       float* Xn, float* Xe, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(fscale > 0 && dscale > 0 && rscale > 0, "fscale, dscale and rscale must be positive");
    RMW_REQUIRE(centres && quats && force_torque && centres0 && quats0 && force_torque0 && nbr && Xn && Xe,
                "null pointer");
    hipLaunchKernelGGL(motion_features_kernel, dim3((nb + 127) / 128), dim3(128), 0, (hipStream_t)stream, centres,
                       quats, force_torque, centres0, quats0, force_torque0, nbr, nb, deg, fscale, dscale, rscale, Xn,
                       Xe);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_gnn_body_frame(const double* quats, const int64_t* body_offsets_dev, int nb, int64_t nblobs,
                                  int blobs_per_body, double* sol, const double* pc, float* T, const float* mean,
                                  const float* sd, double fscale, int mode, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(mode == 0 || mode == 1, "mode must be 0 (targets) or 1 (initial guess)");
    hipLaunchKernelGGL(body_frame_kernel, dim3(nb), dim3(64), 0, (hipStream_t)stream, quats, body_offsets_dev, nb,
                       nblobs, blobs_per_body, sol, pc, T, mean, sd, fscale, mode);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
