// Rigid multiblob bodies on gfx950: blob positions from body poses, the
// geometric matrix K = [J, rot] and its transpose, the rigid saddle-point
// operator, the dense per-body wall mobility and the block-diagonal body
// preconditioner -- all FP64.
//
// Reference:
//   body/body.py:64-77 (blob positions), :81-113 (K), quaternion.py:42-52
//   multi_bodies/multi_bodies.py:328-370 (K U, K^T lambda), :424-477
//     (linear_operator_rigid), :752-907 (build_block_diagonal_preconditioner
//     and its apply), mobility/mobility.py:1018-1116 (dense wall mobility)
//
// Preconditioner build, one workgroup per body (n = 3 * blobs of the body):
//   M_b      dense wall mobility B M(z_eff) B of the body's blobs
//   L        Cholesky factor, in place                  (M_b = L L^T)
//   X        L^-1, in place (lower triangular inverse)
//   M_b^-1 = X^T X             written to global memory (row-major n x n)
//   XK = X K,  M_b^-1 K = X^T (X K),  K^T M_b^-1 K = (XK)^T (XK)
//   N_b = pinv(K^T M_b^-1 K)   (6x6, symmetric Jacobi eigen-solve, rcond 1e-15)
// Bodies with n <= kLdsMaxN keep the whole factorisation in LDS (160 KB per
// workgroup on gfx950); larger bodies run the same code on a global scratch.
// Apply, one workgroup per body (multi_bodies.py:836-849):
//   lt = M_b^-1 s,  Y = N_b (-F - K^T lt),  lambda = lt + (M_b^-1 K) Y
// (lambda equals M_b^-1 (s + K Y) of the reference up to rounding).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "rmw_internal.h"
#include "rmw_pair_terms.h"

namespace rmw {

constexpr int kLdsMaxN = 136;          // (n^2 + 8 n) doubles <= 160 KB
constexpr int kRigidBlock = 256;

// Pinned staging of small host arrays (defined in rmw_wall_rpy.hip).
void stage_upload_public(void* dst, const void* src, size_t bytes, hipStream_t st);

// ---------------------------------------------------------------------------
// Blob positions and K
// ---------------------------------------------------------------------------
__global__ void body_blobs_kernel(const double* __restrict__ centres, const double* __restrict__ quats,
                                  const double* __restrict__ ref, const int64_t* __restrict__ offs,
                                  int nbodies, int64_t nblobs, double* __restrict__ r,
                                  double* __restrict__ rrel)
{
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nblobs) return;
  int lo = 0, hi = nbodies - 1;  // body owning blob i
  while (lo < hi) {
    int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= i) lo = mid; else hi = mid - 1;
  }
  const double* q = quats + 4 * lo;
  const double s = q[0], p0 = q[1], p1 = q[2], p2 = q[3];
  const double dg = s * s - 0.5;
  // rotation matrix of quaternion.py:42-52
  const double R00 = 2 * (p0 * p0 + dg), R01 = 2 * (p0 * p1 - s * p2), R02 = 2 * (p0 * p2 + s * p1);
  const double R10 = 2 * (p1 * p0 + s * p2), R11 = 2 * (p1 * p1 + dg), R12 = 2 * (p1 * p2 - s * p0);
  const double R20 = 2 * (p2 * p0 - s * p1), R21 = 2 * (p2 * p1 + s * p0), R22 = 2 * (p2 * p2 + dg);
  const double x = ref[3 * i], y = ref[3 * i + 1], z = ref[3 * i + 2];
  const double ax = R00 * x + R01 * y + R02 * z;
  const double ay = R10 * x + R11 * y + R12 * z;
  const double az = R20 * x + R21 * y + R22 * z;
  rrel[3 * i] = ax; rrel[3 * i + 1] = ay; rrel[3 * i + 2] = az;
  r[3 * i] = ax + centres[3 * lo];
  r[3 * i + 1] = ay + centres[3 * lo + 1];
  r[3 * i + 2] = az + centres[3 * lo + 2];
}

// Second half of the rigid operator (multi_bodies.py:424-477) once
// m_lambda = M lambda is known: y_lambda = m_lambda - K U, y_U = -K^T lambda.
// One workgroup per body.
__global__ __launch_bounds__(kRigidBlock) void rigid_post_kernel(const double* __restrict__ rrel,
                                                                 const int64_t* __restrict__ offs,
                                                                 const double* __restrict__ x,
                                                                 const double* __restrict__ mlam,
                                                                 double* __restrict__ y, int64_t nblobs)
{
  const int b = blockIdx.x;
  const int64_t i0 = offs[b], i1 = offs[b + 1];
  const double* U = x + 3 * nblobs + 6 * (int64_t)b;
  const double vx = U[0], vy = U[1], vz = U[2], wx = U[3], wy = U[4], wz = U[5];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double rx = rrel[3 * i], ry = rrel[3 * i + 1], rz = rrel[3 * i + 2];
    // K U = v + w x r   (body.py:81-113: rot_i x = -(r_i x x))
    const double kx = vx + (wy * rz - wz * ry);
    const double ky = vy + (wz * rx - wx * rz);
    const double kz = vz + (wx * ry - wy * rx);
    const double lx = x[3 * i], ly = x[3 * i + 1], lz = x[3 * i + 2];
    y[3 * i] = mlam[3 * i] - kx;
    y[3 * i + 1] = mlam[3 * i + 1] - ky;
    y[3 * i + 2] = mlam[3 * i + 2] - kz;
    // K^T lambda = [sum lambda; sum r x lambda]
    acc[0] += lx; acc[1] += ly; acc[2] += lz;
    acc[3] += ry * lz - rz * ly;
    acc[4] += rz * lx - rx * lz;
    acc[5] += rx * ly - ry * lx;
  }
  __shared__ double red[6][kRigidBlock];
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = kRigidBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) y[3 * nblobs + 6 * (int64_t)b + threadIdx.x] = -red[threadIdx.x][0];
}

// Body force-torque from blob forces: F_b = sum f_i, T_b = sum r_i x f_i
// over the blobs of body b (force_torque_calculator_sort_by_bodies,
// multi_bodies_functions.py:411-445, T = rot^T f).  One workgroup per body,
// fixed-order tree reduction.
__global__ __launch_bounds__(kRigidBlock) void body_force_torque_kernel(const double* __restrict__ rrel,
                                                                        const int64_t* __restrict__ offs,
                                                                        const double* __restrict__ f,
                                                                        double* __restrict__ ft)
{
  const int b = blockIdx.x;
  const int64_t i0 = offs[b], i1 = offs[b + 1];
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const double rx = rrel[3 * i], ry = rrel[3 * i + 1], rz = rrel[3 * i + 2];
    const double fx = f[3 * i], fy = f[3 * i + 1], fz = f[3 * i + 2];
    acc[0] += fx; acc[1] += fy; acc[2] += fz;
    acc[3] += ry * fz - rz * fy;
    acc[4] += rz * fx - rx * fz;
    acc[5] += rx * fy - ry * fx;
  }
  __shared__ double red[6][kRigidBlock];
  for (int c = 0; c < 6; ++c) red[c][threadIdx.x] = acc[c];
  __syncthreads();
  for (int s = kRigidBlock / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s)
      for (int c = 0; c < 6; ++c) red[c][threadIdx.x] += red[c][threadIdx.x + s];
    __syncthreads();
  }
  if (threadIdx.x < 6) ft[6 * (int64_t)b + threadIdx.x] = red[threadIdx.x][0];
}

// Rigid-body position update of the deterministic schemes
// (quaternion_integrator_multi_bodies.py:75-186): with V = c_det U - c_prev
// U_prev, location += V_v dt and orientation = q(V_w dt) * orientation, q(phi)
// = (cos(|phi|/2), sin(|phi|/2) phi/|phi|) (quaternion.py:15-38, product
// :41-49; no renormalisation).
__global__ void bodies_advance_kernel(const double* __restrict__ centres, const double* __restrict__ quats,
                                      const double* __restrict__ U, const double* __restrict__ Uprev, int nbodies,
                                      double c_det, double c_prev, double dt, double* __restrict__ cnew,
                                      double* __restrict__ qnew)
{
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nbodies) return;
  double V[6];
  for (int c = 0; c < 6; ++c) {
    V[c] = c_det * U[6 * b + c];
    if (Uprev) V[c] = V[c] - c_prev * Uprev[6 * b + c];
  }
  for (int c = 0; c < 3; ++c) cnew[3 * b + c] = centres[3 * b + c] + V[c] * dt;
  const double ph0 = V[3] * dt, ph1 = V[4] * dt, ph2 = V[5] * dt;
  const double pn = sqrt(ph0 * ph0 + ph1 * ph1 + ph2 * ph2);
  const double s1 = cos(pn / 2.0);
  double p1[3] = {0.0, 0.0, 0.0};
  if (pn != 0.0) {
    const double k = sin(pn / 2.0) / pn;
    p1[0] = k * ph0; p1[1] = k * ph1; p1[2] = k * ph2;
  }
  const double s2 = quats[4 * b], p2[3] = {quats[4 * b + 1], quats[4 * b + 2], quats[4 * b + 3]};
  qnew[4 * b] = s1 * s2 - (p1[0] * p2[0] + p1[1] * p2[1] + p1[2] * p2[2]);
  qnew[4 * b + 1] = s1 * p2[0] + s2 * p1[0] + (p1[1] * p2[2] - p1[2] * p2[1]);
  qnew[4 * b + 2] = s1 * p2[1] + s2 * p1[1] + (p1[2] * p2[0] - p1[0] * p2[2]);
  qnew[4 * b + 3] = s1 * p2[2] + s2 * p1[2] + (p1[0] * p2[1] - p1[1] * p2[0]);
}

// Number of heights r[3i+2] below `level` (body.check_function,
// body/body.py:118-140): per-workgroup counts, then one workgroup sums them.
__global__ __launch_bounds__(256) void count_below_kernel(const double* __restrict__ r, int64_t n, double level,
                                                          int64_t* __restrict__ block_counts)
{
  __shared__ int s_cnt[4];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int below = i < n && r[3 * i + 2] < level;
  const unsigned long long m = __ballot(below);
  if ((threadIdx.x & 63) == 0) s_cnt[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (threadIdx.x == 0) block_counts[blockIdx.x] = (int64_t)s_cnt[0] + s_cnt[1] + s_cnt[2] + s_cnt[3];
}

__global__ __launch_bounds__(256) void count_sum_kernel(const int64_t* __restrict__ block_counts, int64_t nblocks,
                                                        int64_t* __restrict__ count)
{
  __shared__ int64_t s[256];
  int64_t a = 0;
  for (int64_t k = threadIdx.x; k < nblocks; k += 256) a += block_counts[k];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) s[threadIdx.x] += s[threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0) *count = s[0];
}

// ---------------------------------------------------------------------------
// Dense blob mobility entries (FP64), one 3x3 block: above the wall
// B M(z_eff) B (single_wall_fluid_mobility, mobility.py:1018-1116), or the
// unbounded RPY tensor (rotne_prager_tensor, mobility.py:967-1012) for
// dom = kDomNoWall.
// ---------------------------------------------------------------------------
__device__ inline void uf_block(int dom, const double* ri, const double* rj, bool self, double a, double inva,
                                double norm, double m[9])
{
  const bool wall = dom == kDomWall;
  const double zi_raw = ri[2], zj_raw = rj[2];
  const double zie = wall && zi_raw <= a ? a : zi_raw, zje = wall && zj_raw <= a ? a : zj_raw;
  const double Bi = wall && zi_raw < a ? zi_raw * inva : 1.0, Bj = wall && zj_raw < a ? zj_raw * inva : 1.0;
  const double zi = zie * inva, zj = zje * inva;
  for (int c = 0; c < 3; ++c) {
    double v[3] = {0.0, 0.0, 0.0};
    v[c] = 1.0;
    Vec3<double> u;
    if (self) {
      u = self_uf_dom(dom, zi, v[0], v[1], v[2]);
    } else {
      Geom<double> g = make_geom((ri[0] - rj[0]) * inva, (ri[1] - rj[1]) * inva, (zie - zje) * inva, zi, zj);
      u = apply_uf(g, coef_uf(g, zi, zj, dom), v[0], v[1], v[2]);
    }
    const double s = norm * Bi * Bj;
    m[0 * 3 + c] = s * u.x;
    m[1 * 3 + c] = s * u.y;
    m[2 * 3 + c] = s * u.z;
  }
}

// Dense 3N x 3N wall mobility of each body (or of one system = one body),
// row-major at out + moff[b].
__global__ void dense_mobility_kernel(int dom, const double* __restrict__ r, const int64_t* __restrict__ offs,
                                      const int64_t* __restrict__ moff, double a, double eta,
                                      double* __restrict__ out)
{
  const int b = blockIdx.y;
  const int64_t i0 = offs[b], nb = offs[b + 1] - i0;
  const int64_t n = 3 * nb;
  const double inva = 1.0 / a, norm = 1.0 / (8.0 * M_PI * eta * a);
  double* M = out + moff[b];
  for (int64_t idx = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; idx < nb * nb;
       idx += (int64_t)gridDim.x * blockDim.x) {
    const int64_t i = idx / nb, j = idx - (idx / nb) * nb;
    double m[9];
    uf_block(dom, r + 3 * (i0 + i), r + 3 * (i0 + j), i == j, a, inva, norm, m);
    for (int p = 0; p < 3; ++p)
      for (int q = 0; q < 3; ++q) M[(3 * i + p) * n + 3 * j + q] = m[p * 3 + q];
  }
}

// ---------------------------------------------------------------------------
// Symmetric 6x6 pseudo-inverse by cyclic Jacobi (one thread).
// ---------------------------------------------------------------------------
__device__ void pinv6(const double* A_in, double* P)
{
  double A[36], V[36];
  for (int i = 0; i < 36; ++i) { A[i] = A_in[i]; V[i] = (i % 7 == 0) ? 1.0 : 0.0; }
  for (int sweep = 0; sweep < 60; ++sweep) {
    double off = 0.0, diag = 0.0;
    for (int p = 0; p < 6; ++p)
      for (int q = 0; q < 6; ++q) {
        if (p == q) diag += A[p * 6 + q] * A[p * 6 + q];
        else off += A[p * 6 + q] * A[p * 6 + q];
      }
    if (off <= 1e-32 * diag || off == 0.0) break;
    for (int p = 0; p < 5; ++p)
      for (int q = p + 1; q < 6; ++q) {
        const double apq = A[p * 6 + q];
        if (apq == 0.0) continue;
        const double app = A[p * 6 + p], aqq = A[q * 6 + q];
        const double theta = (aqq - app) / (2.0 * apq);
        const double t = (theta >= 0 ? 1.0 : -1.0) / (fabs(theta) + sqrt(theta * theta + 1.0));
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
        for (int k = 0; k < 6; ++k) {  // A <- A J (columns p, q)
          const double akp = A[k * 6 + p], akq = A[k * 6 + q];
          A[k * 6 + p] = c * akp - s * akq;
          A[k * 6 + q] = s * akp + c * akq;
        }
        for (int k = 0; k < 6; ++k) {  // A <- J^T A (rows p, q)
          const double apk = A[p * 6 + k], aqk = A[q * 6 + k];
          A[p * 6 + k] = c * apk - s * aqk;
          A[q * 6 + k] = s * apk + c * aqk;
        }
        for (int k = 0; k < 6; ++k) {
          const double vkp = V[k * 6 + p], vkq = V[k * 6 + q];
          V[k * 6 + p] = c * vkp - s * vkq;
          V[k * 6 + q] = s * vkp + c * vkq;
        }
      }
  }
  double lmax = 0.0;
  for (int i = 0; i < 6; ++i) lmax = fmax(lmax, fabs(A[i * 7]));
  const double cut = 1e-15 * lmax;  // numpy.linalg.pinv default rcond
  for (int i = 0; i < 36; ++i) P[i] = 0.0;
  for (int e = 0; e < 6; ++e) {
    const double lam = A[e * 7];
    if (fabs(lam) <= cut) continue;
    const double il = 1.0 / lam;
    for (int i = 0; i < 6; ++i)
      for (int j = 0; j < 6; ++j) P[i * 6 + j] += V[i * 6 + e] * il * V[j * 6 + e];
  }
}

// ---------------------------------------------------------------------------
// Preconditioner build: one workgroup per body.
// ---------------------------------------------------------------------------
struct PcBuildParams {
  const double* r;        // blob positions (N, 3)
  const double* rrel;     // blob - centre
  const int64_t* offs;    // body blob offsets (B+1)
  const int64_t* moff;    // offsets of the per-body n x n blocks (B+1)
  double* minv;           // out: M_b^-1 blocks
  double* minvk;          // out: (3N, 6) M_b^-1 K
  double* nbody;          // out: (B, 6, 6)
  double* scratch;        // global scratch for bodies above kLdsMaxN (same layout as minv)
  double* scratch_vec;    // global scratch, 8 n doubles per body at 8 * 3 * offs[b]
  int* status;            // out: per body 0 ok, 1 not positive definite
  double a, eta;
  int use_lds;
  double* chol;           // out (optional): lower Cholesky factor L of M_b (M_b = L L^T), packed as minv
  double* chol_inv;       // out (optional): L^-1 (lower), packed as minv
  int dom;                // kDomWall or kDomNoWall: the blob mobility M_b is built for
};

template <bool LDS>
__global__ __launch_bounds__(kRigidBlock) void pc_build_kernel(PcBuildParams p)
{
  extern __shared__ double smem[];
  const int b = blockIdx.x;
  const int64_t i0 = p.offs[b], nb = p.offs[b + 1] - i0;
  const int n = (int)(3 * nb);
  if (LDS != (n <= kLdsMaxN)) return;  // the other instantiation owns this body
  double* A = LDS ? smem : p.scratch + p.moff[b];
  double* tmp = LDS ? smem + (size_t)n * n : p.scratch_vec + 8 * 3 * i0;
  double* XK = tmp + n;  // n x 6
  __shared__ double a6[36];
  __shared__ int bad;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int lane = tid & 63, wv = tid >> 6, nw = nt >> 6;
  const double inva = 1.0 / p.a, norm = 1.0 / (8.0 * M_PI * p.eta * p.a);
  if (tid == 0) bad = 0;

  // 1. assemble M_b
  for (int64_t idx = tid; idx < nb * nb; idx += nt) {
    const int64_t i = idx / nb, j = idx - (idx / nb) * nb;
    double m[9];
    uf_block(p.dom, p.r + 3 * (i0 + i), p.r + 3 * (i0 + j), i == j, p.a, inva, norm, m);
    for (int q = 0; q < 3; ++q)
      for (int s = 0; s < 3; ++s) A[(3 * i + q) * n + 3 * j + s] = m[q * 3 + s];
  }
  __syncthreads();

  // 2. Cholesky, lower triangle in place (row-major)
  for (int k = 0; k < n; ++k) {
    if (tid == 0) {
      double d = A[k * n + k];
      if (!(d > 0.0)) { bad = 1; d = 1.0; }
      A[k * n + k] = sqrt(d);
    }
    __syncthreads();
    const double dk = A[k * n + k];
    for (int i = k + 1 + tid; i < n; i += nt) A[i * n + k] /= dk;
    __syncthreads();
    for (int i = k + 1 + wv; i < n; i += nw) {
      const double lik = A[i * n + k];
      for (int j = k + 1 + lane; j <= i; j += 64) A[i * n + j] -= lik * A[j * n + k];
    }
    __syncthreads();
  }
  if (p.chol) {
    // L with its upper triangle zeroed (A above the diagonal still holds M_b)
    double* Lb = p.chol + p.moff[b];
    for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
      const int i = (int)(idx / n), j = (int)(idx - (int64_t)i * n);
      Lb[idx] = j <= i ? A[idx] : 0.0;
    }
  }

  // 3. X = L^-1 in place, columns from last to first
  for (int j = n - 1; j >= 0; --j) {
    if (tid == 0) A[j * n + j] = 1.0 / A[j * n + j];
    for (int i = j + 1 + tid; i < n; i += nt) tmp[i] = A[i * n + j];
    __syncthreads();
    const double xjj = A[j * n + j];
    for (int i = j + 1 + tid; i < n; i += nt) {
      double y = 0.0;
      for (int k = j + 1; k <= i; ++k) y += A[i * n + k] * tmp[k];
      A[i * n + j] = -y * xjj;
    }
    __syncthreads();
  }

  if (p.chol_inv) {
    double* Xb = p.chol_inv + p.moff[b];
    for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
      const int i = (int)(idx / n), j = (int)(idx - (int64_t)i * n);
      Xb[idx] = j <= i ? A[idx] : 0.0;
    }
  }

  // 4. M^-1 = X^T X  (X lower: (X^T X)_ij = sum_{k >= max(i,j)} X_ki X_kj)
  double* Minv = p.minv + p.moff[b];
  for (int64_t idx = tid; idx < (int64_t)n * n; idx += nt) {
    const int i = (int)(idx / n), j = (int)(idx - (int64_t)(idx / n) * n);
    double s = 0.0;
    for (int k = i > j ? i : j; k < n; ++k) s += A[k * n + i] * A[k * n + j];
    Minv[idx] = s;
  }

  // 5. XK = X K  with K rows [e_c | rot_c]; rot_i x = -(r_i x x)
  for (int idx = tid; idx < n * 6; idx += nt) {
    const int k = idx / 6, c = idx - (idx / 6) * 6;
    double s = 0.0;
    for (int i = 0; i <= k; ++i) {
      const int blob = i / 3, comp = i - (i / 3) * 3;
      const double* rr = p.rrel + 3 * (i0 + blob);
      double kic;
      if (c < 3) {
        kic = comp == c ? 1.0 : 0.0;
      } else {
        // K[3m+comp][3+d] = rot entries of body.py:92-99
        const int d = c - 3;
        const double rx = rr[0], ry = rr[1], rz = rr[2];
        const double rot[3][3] = {{0.0, rz, -ry}, {-rz, 0.0, rx}, {ry, -rx, 0.0}};
        kic = rot[comp][d];
      }
      s += A[k * n + i] * kic;
    }
    XK[idx] = s;
  }
  __syncthreads();
  // M^-1 K = X^T XK
  double* MK = p.minvk + 3 * i0 * 6;
  for (int idx = tid; idx < n * 6; idx += nt) {
    const int i = idx / 6, c = idx - (idx / 6) * 6;
    double s = 0.0;
    for (int k = i; k < n; ++k) s += A[k * n + i] * XK[k * 6 + c];
    MK[idx] = s;
  }
  // K^T M^-1 K = XK^T XK
  if (tid < 36) {
    const int r6 = tid / 6, c6 = tid - (tid / 6) * 6;
    double s = 0.0;
    for (int k = 0; k < n; ++k) s += XK[k * 6 + r6] * XK[k * 6 + c6];
    a6[tid] = s;
  }
  __syncthreads();
  if (tid == 0) {
    pinv6(a6, p.nbody + 36 * (int64_t)b);
    p.status[b] = bad;
  }
}

// Preconditioner build for bodies that fit LDS (n = 3 nb <= kLdsMaxN): M_b
// assembled in LDS, then inverted in registers by Gauss-Jordan elimination
// without pivoting (M_b is symmetric positive definite: every pivot is a
// Schur-complement diagonal, positive; a non-positive one flags the body).
// Thread (ty, tx) of the 16 x 16 grid holds elements (ty + 16 a, tx + 16 c),
// a, c < R, of the matrix; each elimination step broadcasts the pivot row
// and column through LDS (double-buffered, one barrier per step) and
// applies the rank-one update to the R x R register block:
//   A'_ij = [i != k][j != k] A_ij - c'_i r'_j / A_kk,
//   c'_i = A_ik (c'_k = -1),  r'_j = A_kj (r'_k = 1),
// which also writes the scaled pivot row, column and 1/A_kk in place.
// Then M^-1 K from LDS, K^T M^-1 K and its pseudo-inverse.
template <int R>
__global__ __launch_bounds__(kRigidBlock) void pc_build_gj_kernel(PcBuildParams p)
{
  extern __shared__ double smem[];   // A (n x n) | pivot rows/cols during the elimination, then M^-1 K (n x 6)
  const int b = blockIdx.x;
  const int64_t i0 = p.offs[b], nb = p.offs[b + 1] - i0;
  const int n = (int)(3 * nb);
  if (n > kLdsMaxN || (n + 15) / 16 > R) return;   // another launch owns this body
  double* A = smem;
  double* prow = smem + (size_t)n * n;               // [2][16 R]
  double* pcol = prow + 2 * 16 * R;                  // [2][16 R]
  double* MK = smem + (size_t)n * n;                 // n x 6, after the elimination
  __shared__ double a6[36];
  __shared__ int bad;
  const int tid = threadIdx.x, nt = blockDim.x;
  const int tx = tid & 15, ty = tid >> 4;
  const double inva = 1.0 / p.a, norm = 1.0 / (8.0 * M_PI * p.eta * p.a);
  if (tid == 0) bad = 0;

  // 1. assemble M_b
  for (int64_t idx = tid; idx < nb * nb; idx += nt) {
    const int64_t i = idx / nb, j = idx - (idx / nb) * nb;
    double m[9];
    uf_block(p.dom, p.r + 3 * (i0 + i), p.r + 3 * (i0 + j), i == j, p.a, inva, norm, m);
    for (int q = 0; q < 3; ++q)
      for (int s = 0; s < 3; ++s) A[(3 * i + q) * n + 3 * j + s] = m[q * 3 + s];
  }
  __syncthreads();
  double E[R][R];
#pragma unroll
  for (int ra = 0; ra < R; ++ra)
#pragma unroll
    for (int cb = 0; cb < R; ++cb) {
      const int i = ty + 16 * ra, j = tx + 16 * cb;
      E[ra][cb] = (i < n && j < n) ? A[i * n + j] : 0.0;
    }

  // 2. in-place Gauss-Jordan inverse
  for (int k = 0; k < n; ++k) {
    double* pr = prow + (k & 1) * 16 * R;
    double* pc = pcol + (k & 1) * 16 * R;
    const int ka = k >> 4, kt = k & 15;
    if (ty == kt) {
#pragma unroll
      for (int ra = 0; ra < R; ++ra)
        if (ra == ka) {
#pragma unroll
          for (int cb = 0; cb < R; ++cb) pr[tx + 16 * cb] = E[ra][cb];
        }
    }
    if (tx == kt) {
#pragma unroll
      for (int cb = 0; cb < R; ++cb)
        if (cb == ka) {
#pragma unroll
          for (int ra = 0; ra < R; ++ra) pc[ty + 16 * ra] = E[ra][cb];
        }
    }
    __syncthreads();
    double piv = pr[k];
    if (!(piv > 0.0)) {
      if (tid == 0) bad = 1;
      piv = 1.0;
    }
    const double ip = 1.0 / piv;
    double c[R], r[R];
#pragma unroll
    for (int ra = 0; ra < R; ++ra) {
      const int i = ty + 16 * ra;
      c[ra] = i == k ? -ip : (i < n ? pc[i] * ip : 0.0);
    }
#pragma unroll
    for (int cb = 0; cb < R; ++cb) {
      const int j = tx + 16 * cb;
      r[cb] = j == k ? 1.0 : (j < n ? pr[j] : 0.0);
    }
#pragma unroll
    for (int ra = 0; ra < R; ++ra) {
      const bool rowk = ty + 16 * ra == k;
#pragma unroll
      for (int cb = 0; cb < R; ++cb) {
        const bool keep = !rowk && tx + 16 * cb != k;
        E[ra][cb] = fma(-c[ra], r[cb], keep ? E[ra][cb] : 0.0);
      }
    }
  }
  __syncthreads();   // the pivot buffers are free; A is rewritten below

  // 3. M^-1 to LDS and to the output blocks
  double* Minv = p.minv + p.moff[b];
#pragma unroll
  for (int ra = 0; ra < R; ++ra)
#pragma unroll
    for (int cb = 0; cb < R; ++cb) {
      const int i = ty + 16 * ra, j = tx + 16 * cb;
      if (i < n && j < n) {
        A[i * n + j] = E[ra][cb];
        Minv[(int64_t)i * n + j] = E[ra][cb];
      }
    }
  __syncthreads();

  // 4. M^-1 K with K rows [e_c | rot_c] (body.py:92-99: rot_i x = -(r_i x x))
  double* MKg = p.minvk + 3 * i0 * 6;
  for (int idx = tid; idx < n * 6; idx += nt) {
    const int i = idx / 6, c = idx - (idx / 6) * 6;
    const double* row = A + (int64_t)i * n;
    double s = 0.0;
    if (c < 3) {
      for (int m = 0; m < (int)nb; ++m) s += row[3 * m + c];
    } else {
      const int d = c - 3;
      for (int m = 0; m < (int)nb; ++m) {
        const double* rr = p.rrel + 3 * (i0 + m);
        const double rx = rr[0], ry = rr[1], rz = rr[2];
        // column d of rot = [[0, rz, -ry], [-rz, 0, rx], [ry, -rx, 0]]
        const double k0 = d == 0 ? 0.0 : (d == 1 ? rz : -ry);
        const double k1 = d == 0 ? -rz : (d == 1 ? 0.0 : rx);
        const double k2 = d == 0 ? ry : (d == 1 ? -rx : 0.0);
        s += row[3 * m] * k0 + row[3 * m + 1] * k1 + row[3 * m + 2] * k2;
      }
    }
    MK[idx] = s;
    MKg[idx] = s;
  }
  __syncthreads();
  // K^T M^-1 K
  if (tid < 36) {
    const int r6 = tid / 6, c6 = tid - (tid / 6) * 6;
    double s = 0.0;
    if (r6 < 3) {
      for (int m = 0; m < (int)nb; ++m) s += MK[(3 * m + r6) * 6 + c6];
    } else {
      const int d = r6 - 3;
      for (int m = 0; m < (int)nb; ++m) {
        const double* rr = p.rrel + 3 * (i0 + m);
        const double rx = rr[0], ry = rr[1], rz = rr[2];
        const double k0 = d == 0 ? 0.0 : (d == 1 ? rz : -ry);
        const double k1 = d == 0 ? -rz : (d == 1 ? 0.0 : rx);
        const double k2 = d == 0 ? ry : (d == 1 ? -rx : 0.0);
        s += k0 * MK[(3 * m) * 6 + c6] + k1 * MK[(3 * m + 1) * 6 + c6] + k2 * MK[(3 * m + 2) * 6 + c6];
      }
    }
    a6[tid] = s;
  }
  __syncthreads();
  if (tid == 0) {
    // symmetrise the rounding of K^T M^-1 K before the pseudo-inverse
    for (int r6 = 0; r6 < 6; ++r6)
      for (int c6 = r6 + 1; c6 < 6; ++c6) {
        const double v = 0.5 * (a6[r6 * 6 + c6] + a6[c6 * 6 + r6]);
        a6[r6 * 6 + c6] = v;
        a6[c6 * 6 + r6] = v;
      }
    pinv6(a6, p.nbody + 36 * (int64_t)b);
    p.status[b] = bad;
  }
}

// ---------------------------------------------------------------------------
// Point field of the other bodies: the velocity at every blob induced by the
// net forces of all other bodies, each body a sphere of radius R at its
// centre -- the different-radii wall RPY (coef_radii) between the blob
// (radius a) and the sphere, heights shifted to max(z, radius) and damped
// below it as the radii product does:
//   u_i = B_i sum_{c != body(i)} M(r_i, R_c; a, R) B_c F_c.
// It is the far-field part of the off-diagonal blocks of the rigid
// operator acting on any iterate whose body force sums are F, so that
// P^-1 (b - [u; 0]) is a block-Jacobi step from P^-1 b without the product.
// ---------------------------------------------------------------------------
constexpr int kPfBlock = 256;

__global__ __launch_bounds__(kPfBlock) void point_field_kernel(const double* __restrict__ r, int64_t nblobs,
                                                               const int64_t* __restrict__ offs, int nbodies,
                                                               const double* __restrict__ centres,
                                                               const double* __restrict__ force, int64_t fstride,
                                                               double R, double a, double norm, int wall,
                                                               double* __restrict__ out)
{
  __shared__ double s_x[kPfBlock], s_y[kPfBlock], s_z[kPfBlock];
  __shared__ double s_f[3][kPfBlock];
  const int64_t i = (int64_t)blockIdx.x * kPfBlock + threadIdx.x;
  const bool valid = i < nblobs;
  const int64_t gi = valid ? i : 0;
  int lo = 0, hi = nbodies - 1;   // own body: the last b with offs[b] <= gi
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (offs[mid] <= gi) lo = mid; else hi = mid - 1;
  }
  const int own = lo;
  const double xi = r[3 * gi], yi = r[3 * gi + 1], z0 = r[3 * gi + 2];
  const double zi = (wall && !(z0 > a)) ? a : z0;
  const double Bi = (wall && z0 < a) ? z0 / a : 1.0;
  const double inva = 1.0 / a, invR = 1.0 / R;
  double ux = 0.0, uy = 0.0, uz = 0.0;
  for (int c0 = 0; c0 < nbodies; c0 += kPfBlock) {
    const int cnt = nbodies - c0 < kPfBlock ? nbodies - c0 : kPfBlock;
    __syncthreads();
    if ((int)threadIdx.x < cnt) {
      const int c = c0 + threadIdx.x;
      const double zc = centres[3 * c + 2];
      const double Bc = (wall && zc < R) ? zc / R : 1.0;
      s_x[threadIdx.x] = centres[3 * c];
      s_y[threadIdx.x] = centres[3 * c + 1];
      s_z[threadIdx.x] = (wall && !(zc > R)) ? R : zc;
      const double k = norm * Bc;
      s_f[0][threadIdx.x] = k * force[c * fstride];
      s_f[1][threadIdx.x] = k * force[c * fstride + 1];
      s_f[2][threadIdx.x] = k * force[c * fstride + 2];
    }
    __syncthreads();
    for (int k = 0; k < cnt; ++k) {
      if (c0 + k == own) continue;
      const double dx = xi - s_x[k], dy = yi - s_y[k], dz = zi - s_z[k];
      const double Rz = zi + s_z[k];
      const CoefRadii<double> co = coef_radii<double>(dx, dy, dz, Rz, zi, s_z[k], a, R, inva, invR,
                                                      wall != 0 ? kRadiiWall : kRadiiUnbounded);
      const Vec3<double> v = apply_radii<double>(dx, dy, dz, Rz, co, s_f[0][k], s_f[1][k], s_f[2][k]);
      ux += v.x; uy += v.y; uz += v.z;
    }
  }
  if (valid) {
    out[3 * i] = Bi * ux;
    out[3 * i + 1] = Bi * uy;
    out[3 * i + 2] = Bi * uz;
  }
}

// Preconditioner apply: y = P^-1 x, one workgroup per body.
__global__ __launch_bounds__(kRigidBlock) void pc_apply_kernel(const int64_t* __restrict__ offs,
                                                               const int64_t* __restrict__ moff,
                                                               const double* __restrict__ minv,
                                                               const double* __restrict__ minvk,
                                                               const double* __restrict__ nbody,
                                                               const double* __restrict__ rrel,
                                                               const double* __restrict__ x,
                                                               double* __restrict__ y, int64_t nblobs)
{
  extern __shared__ double sm[];  // s (n) | lt (n)
  const int b = blockIdx.x;
  const int64_t i0 = offs[b], nb = offs[b + 1] - i0;
  const int n = (int)(3 * nb);
  double* s = sm;
  double* lt = sm + n;
  __shared__ double t6[6], Y[6];
  __shared__ double red[6][kRigidBlock / 64];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int i = tid; i < n; i += blockDim.x) s[i] = x[3 * i0 + i];
  __syncthreads();
  // lt = M^-1 s: one wavefront per row, lanes across columns (coalesced)
  const double* M = minv + moff[b];
  for (int i = wv; i < n; i += nw) {
    double acc = 0.0;
    for (int j = lane; j < n; j += 64) acc += M[(int64_t)i * n + j] * s[j];
    for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
    if (lane == 0) lt[i] = acc;
  }
  __syncthreads();
  // t = -F - K^T lt
  double acc[6] = {0, 0, 0, 0, 0, 0};
  for (int m = tid; m < nb; m += blockDim.x) {
    const double* rr = rrel + 3 * (i0 + m);
    const double lx = lt[3 * m], ly = lt[3 * m + 1], lz = lt[3 * m + 2];
    acc[0] += lx; acc[1] += ly; acc[2] += lz;
    acc[3] += rr[1] * lz - rr[2] * ly;
    acc[4] += rr[2] * lx - rr[0] * lz;
    acc[5] += rr[0] * ly - rr[1] * lx;
  }
  for (int c = 0; c < 6; ++c) {
    double v = acc[c];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (lane == 0) red[c][wv] = v;
  }
  __syncthreads();
  if (tid < 6) {
    double v = 0.0;
    for (int w = 0; w < nw; ++w) v += red[tid][w];
    t6[tid] = -x[3 * nblobs + 6 * (int64_t)b + tid] - v;
  }
  __syncthreads();
  if (tid < 6) {
    const double* Nb = nbody + 36 * (int64_t)b;
    double v = 0.0;
    for (int c = 0; c < 6; ++c) v += Nb[tid * 6 + c] * t6[c];
    Y[tid] = v;
    y[3 * nblobs + 6 * (int64_t)b + tid] = v;
  }
  __syncthreads();
  const double* MK = minvk + 3 * i0 * 6;
  for (int i = tid; i < n; i += blockDim.x) {
    double v = lt[i];
    for (int c = 0; c < 6; ++c) v += MK[i * 6 + c] * Y[c];
    y[3 * i0 + i] = v;
  }
}

// Per-body block matvec y_b = B_b x_b (trans = 0) or B_b^T x_b (trans = 1)
// for k vectors (blockIdx.y), B_b row-major n_b x n_b packed as minv;
// lower != 0: B_b is lower triangular and only its lower part is read.
// trans = 0: one wavefront per row, lanes across the columns; trans = 1:
// one thread per output entry, reading a row of B per step -- both
// coalesced.
__global__ __launch_bounds__(kRigidBlock) void block_apply_kernel(const int64_t* __restrict__ offs,
                                                                  const int64_t* __restrict__ moff,
                                                                  const double* __restrict__ B, int trans, int lower,
                                                                  const double* __restrict__ x, double* __restrict__ y,
                                                                  int64_t ld)
{
  extern __shared__ double sx[];
  const int b = blockIdx.x;
  const int64_t i0 = offs[b], nb = offs[b + 1] - i0;
  const int n = (int)(3 * nb);
  const double* xb = x + (int64_t)blockIdx.y * ld + 3 * i0;
  double* yb = y + (int64_t)blockIdx.y * ld + 3 * i0;
  const double* M = B + moff[b];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6, nw = blockDim.x >> 6;
  for (int i = tid; i < n; i += blockDim.x) sx[i] = xb[i];
  __syncthreads();
  if (!trans) {
    for (int i = wv; i < n; i += nw) {
      const int jend = lower ? i + 1 : n;
      double acc = 0.0;
      for (int j = lane; j < jend; j += 64) acc += M[(int64_t)i * n + j] * sx[j];
      for (int o = 32; o > 0; o >>= 1) acc += __shfl_down(acc, o, 64);
      if (lane == 0) yb[i] = acc;
    }
  } else {
    for (int i = tid; i < n; i += blockDim.x) {
      double acc = 0.0;
      for (int j = lower ? i : 0; j < n; ++j) acc += M[(int64_t)j * n + i] * sx[j];
      yb[i] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
struct RigidLayout {
  int nbodies = 0;
  int64_t nblobs = 0;
  int max_n = 0;
  std::vector<int64_t> offs, moff;
  size_t off_offs = 0, off_moff = 0, off_status = 0, total = 0;
  size_t minv_doubles = 0;
  size_t scratch_doubles = 0;  // factorisation scratch for bodies above kLdsMaxN
};

static RigidLayout rigid_layout(const int64_t* body_offsets, int nbodies)
{
  RMW_REQUIRE(body_offsets != nullptr && nbodies >= 1, "body_offsets / nbodies invalid");
  RMW_REQUIRE(body_offsets[0] == 0, "body_offsets[0] must be 0");
  RigidLayout L;
  L.nbodies = nbodies;
  L.offs.assign(body_offsets, body_offsets + nbodies + 1);
  L.moff.resize(nbodies + 1);
  L.moff[0] = 0;
  bool big = false;
  for (int b = 0; b < nbodies; ++b) {
    const int64_t nb = L.offs[b + 1] - L.offs[b];
    RMW_REQUIRE(nb >= 1, "every body needs at least one blob");
    const int64_t n = 3 * nb;
    RMW_REQUIRE(n <= 30000, "body too large for a dense block");
    L.moff[b + 1] = L.moff[b] + n * n;
    L.max_n = std::max<int>(L.max_n, (int)n);
    if (n > kLdsMaxN) big = true;
  }
  L.nblobs = L.offs[nbodies];
  L.minv_doubles = (size_t)L.moff[nbodies];
  size_t o = 0;
  L.off_offs = o; o = align_up(o + (nbodies + 1) * sizeof(int64_t));
  L.off_moff = o; o = align_up(o + (nbodies + 1) * sizeof(int64_t));
  L.off_status = o; o = align_up(o + nbodies * sizeof(int));
  L.total = o;
  if (big) L.scratch_doubles = L.minv_doubles + (size_t)8 * 3 * L.nblobs;
  return L;
}

static void upload_offsets(const RigidLayout& L, char* ws, hipStream_t st)
{
  stage_upload_public(ws + L.off_offs, L.offs.data(), L.offs.size() * sizeof(int64_t), st);
  stage_upload_public(ws + L.off_moff, L.moff.data(), L.moff.size() * sizeof(int64_t), st);
}

// Every rigid entry point below reads the body offsets from the workspace;
// rmw_rigid_setup must have been called on that workspace (stream-ordered)
// with the same body_offsets.

template <int R>
static void launch_pc_gj_r(const PcBuildParams& p, int nbodies, size_t lds, hipStream_t st)
{
  RMW_CHECK_HIP(hipFuncSetAttribute((const void*)pc_build_gj_kernel<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds));
  hipLaunchKernelGGL(pc_build_gj_kernel<R>, dim3(nbodies), dim3(kRigidBlock), lds, st, p);
  RMW_CHECK_HIP(hipGetLastError());
}

static void launch_pc_gj(int R, const PcBuildParams& p, int nbodies, size_t lds, hipStream_t st)
{
  switch (R) {
    case 1: launch_pc_gj_r<1>(p, nbodies, lds, st); break;
    case 2: launch_pc_gj_r<2>(p, nbodies, lds, st); break;
    case 3: launch_pc_gj_r<3>(p, nbodies, lds, st); break;
    case 4: launch_pc_gj_r<4>(p, nbodies, lds, st); break;
    case 5: launch_pc_gj_r<5>(p, nbodies, lds, st); break;
    case 6: launch_pc_gj_r<6>(p, nbodies, lds, st); break;
    case 7: launch_pc_gj_r<7>(p, nbodies, lds, st); break;
    case 8: launch_pc_gj_r<8>(p, nbodies, lds, st); break;
    case 9: launch_pc_gj_r<9>(p, nbodies, lds, st); break;
    default: RMW_REQUIRE(false, "body too large for the LDS preconditioner build");
  }
}

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_rigid_workspace_size(const int64_t* body_offsets, int nbodies, size_t* bytes,
                                        size_t* minv_doubles, size_t* pc_scratch_doubles)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    if (bytes) *bytes = L.total;
    if (minv_doubles) *minv_doubles = L.minv_doubles;
    if (pc_scratch_doubles) *pc_scratch_doubles = L.scratch_doubles;
  });
}

extern "C" int rmw_rigid_setup(const int64_t* body_offsets, int nbodies, void* workspace,
                               size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace && workspace_bytes >= L.total, "workspace too small");
    upload_offsets(L, (char*)workspace, (hipStream_t)stream);
  });
}

extern "C" int rmw_body_blobs(const double* centres, const double* quats, const double* ref,
                              const int64_t* body_offsets, int nbodies, double* r_blobs, double* r_rel,
                              void* workspace, size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    const int grid = (int)((L.nblobs + 255) / 256);
    hipLaunchKernelGGL(body_blobs_kernel, dim3(grid), dim3(256), 0, st, centres, quats, ref,
                       (const int64_t*)(ws + L.off_offs), nbodies, L.nblobs, r_blobs, r_rel);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rigid_operator_post(const double* r_rel, const int64_t* body_offsets, int nbodies,
                                       const double* x, const double* m_lambda, double* y, void* workspace,
                                       size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(x && m_lambda && y && r_rel, "null pointer");
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    hipLaunchKernelGGL(rigid_post_kernel, dim3(nbodies), dim3(kRigidBlock), 0, st, r_rel,
                       (const int64_t*)(ws + L.off_offs), x, m_lambda, y, L.nblobs);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rigid_force_torque(const double* r_rel, const int64_t* body_offsets, int nbodies,
                                      const double* blob_force, double* force_torque, void* workspace,
                                      size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(r_rel && blob_force && force_torque, "null pointer");
    hipLaunchKernelGGL(body_force_torque_kernel, dim3(nbodies), dim3(kRigidBlock), 0, (hipStream_t)stream, r_rel,
                       (const int64_t*)((char*)workspace + L.off_offs), blob_force, force_torque);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_bodies_advance(const double* centres, const double* quats, const double* velocities,
                                  const double* prev_velocities, int nbodies, double c_det, double c_prev,
                                  double dt, double* centres_new, double* quats_new, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(nbodies >= 0, "nbodies must be >= 0");
    if (nbodies == 0) return;
    RMW_REQUIRE(centres && quats && velocities && centres_new && quats_new, "null pointer");
    hipLaunchKernelGGL(bodies_advance_kernel, dim3((nbodies + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       centres, quats, velocities, prev_velocities, nbodies, c_det, c_prev, dt, centres_new,
                       quats_new);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_count_below_workspace_size(int64_t n, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr && n >= 0, "bad arguments");
    *bytes = align_up((size_t)std::max<int64_t>((n + 255) / 256, 1) * sizeof(int64_t));
  });
}

extern "C" int rmw_count_below(const double* r, int64_t n, double level, int64_t* count, void* workspace,
                               size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0 && count != nullptr, "bad arguments");
    const int64_t nb = (n + 255) / 256;
    RMW_REQUIRE(workspace && workspace_bytes >= (size_t)std::max<int64_t>(nb, 1) * sizeof(int64_t),
                "workspace too small");
    hipStream_t st = (hipStream_t)stream;
    if (nb > 0) {
      RMW_REQUIRE(r != nullptr, "r is null");
      hipLaunchKernelGGL(count_below_kernel, dim3((unsigned)nb), dim3(256), 0, st, r, n, level, (int64_t*)workspace);
      RMW_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(count_sum_kernel, dim3(1), dim3(256), 0, st, (const int64_t*)workspace, nb, count);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

static void check_rigid_domain(int domain)
{
  RMW_REQUIRE(domain == RMW_DOMAIN_WALL || domain == RMW_DOMAIN_NO_WALL,
              "rigid bodies: domain must be RMW_DOMAIN_WALL or RMW_DOMAIN_NO_WALL");
}

extern "C" int rmw_dense_wall_mobility(const double* r_blobs, const int64_t* body_offsets, int nbodies,
                                       double eta, double a, double* out, void* workspace,
                                       size_t workspace_bytes, void* stream)
{
  return rmw_dense_mobility(RMW_DOMAIN_WALL, r_blobs, body_offsets, nbodies, eta, a, out, workspace,
                            workspace_bytes, stream);
}

extern "C" int rmw_dense_mobility(int domain, const double* r_blobs, const int64_t* body_offsets, int nbodies,
                                  double eta, double a, double* out, void* workspace, size_t workspace_bytes,
                                  void* stream)
{
  return guarded([&] {
    check_rigid_domain(domain);
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(eta > 0 && a > 0, "eta and a must be positive");
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    int64_t maxpairs = 0;
    for (int b = 0; b < nbodies; ++b) {
      int64_t nb = L.offs[b + 1] - L.offs[b];
      maxpairs = std::max(maxpairs, nb * nb);
    }
    const int gx = (int)std::min<int64_t>((maxpairs + 255) / 256, 4096);
    hipLaunchKernelGGL(dense_mobility_kernel, dim3(gx, nbodies), dim3(256), 0, st,
                       domain == RMW_DOMAIN_WALL ? (int)kDomWall : (int)kDomNoWall, r_blobs,
                       (const int64_t*)(ws + L.off_offs), (const int64_t*)(ws + L.off_moff), a, eta, out);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rigid_pc_build(const double* r_blobs, const double* r_rel, const int64_t* body_offsets,
                                  int nbodies, double eta, double a, double* minv, double* minvk,
                                  double* nbody, int* status, double* scratch, size_t scratch_doubles,
                                  void* workspace, size_t workspace_bytes, void* stream)
{
  return rmw_rigid_pc_build_ex(RMW_DOMAIN_WALL, r_blobs, r_rel, body_offsets, nbodies, eta, a, minv, minvk, nbody,
                               nullptr, nullptr, status, scratch, scratch_doubles, workspace, workspace_bytes, stream);
}

extern "C" int rmw_rigid_pc_build_ex(int domain, const double* r_blobs, const double* r_rel,
                                     const int64_t* body_offsets, int nbodies, double eta, double a, double* minv,
                                     double* minvk, double* nbody, double* chol, double* chol_inv, int* status,
                                     double* scratch, size_t scratch_doubles, void* workspace,
                                     size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    check_rigid_domain(domain);
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(eta > 0 && a > 0, "eta and a must be positive");
    RMW_REQUIRE(r_blobs && r_rel && minv && minvk && nbody, "null pointer");
    RMW_REQUIRE((chol == nullptr) == (chol_inv == nullptr), "chol and chol_inv go together");
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    PcBuildParams p{};
    p.r = r_blobs; p.rrel = r_rel;
    p.offs = (const int64_t*)(ws + L.off_offs);
    p.moff = (const int64_t*)(ws + L.off_moff);
    p.minv = minv; p.minvk = minvk; p.nbody = nbody;
    p.chol = chol; p.chol_inv = chol_inv;
    p.dom = domain == RMW_DOMAIN_WALL ? (int)kDomWall : (int)kDomNoWall;
    RMW_REQUIRE(scratch_doubles >= L.scratch_doubles && (L.scratch_doubles == 0 || scratch != nullptr),
                "factorisation scratch too small (bodies above 45 blobs need it)");
    p.scratch = L.scratch_doubles ? scratch : nullptr;
    p.scratch_vec = L.scratch_doubles ? scratch + L.minv_doubles : nullptr;
    p.status = status ? status : (int*)(ws + L.off_status);
    p.a = a; p.eta = eta;
    int max_small = 0;
    bool big = false;
    for (int b = 0; b < nbodies; ++b) {
      const int n = (int)(3 * (L.offs[b + 1] - L.offs[b]));
      if (n <= kLdsMaxN) max_small = std::max(max_small, n); else big = true;
    }
    if (max_small && chol == nullptr) {
      // Gauss-Jordan in registers, R x R elements per thread for the largest body
      const int R = std::max(1, (max_small + 15) / 16);
      const size_t lds = ((size_t)max_small * max_small + std::max<size_t>(6 * (size_t)max_small, 4 * 16 * (size_t)R)) *
                         sizeof(double);
      RMW_REQUIRE(lds <= 160 * 1024, "preconditioner build exceeds LDS");
      launch_pc_gj(R, p, nbodies, lds, st);
    } else if (max_small) {
      // the Cholesky factorisation in LDS (it also writes L and L^-1): M_b
      // (n^2), then n + 6 n of vectors
      const size_t lds = ((size_t)max_small * max_small + 8 * (size_t)max_small) * sizeof(double);
      RMW_REQUIRE(lds <= 160 * 1024, "preconditioner build exceeds LDS");
      RMW_CHECK_HIP(hipFuncSetAttribute((const void*)pc_build_kernel<true>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      hipLaunchKernelGGL(pc_build_kernel<true>, dim3(nbodies), dim3(kRigidBlock), lds, st, p);
      RMW_CHECK_HIP(hipGetLastError());
    }
    if (big) {
      hipLaunchKernelGGL(pc_build_kernel<false>, dim3(nbodies), dim3(kRigidBlock), 0, st, p);
      RMW_CHECK_HIP(hipGetLastError());
    }
  });
}

extern "C" int rmw_rigid_block_apply(const int64_t* body_offsets, int nbodies, const double* blocks, int trans,
                                     int lower, const double* x, double* y, int k, int64_t ld, void* workspace,
                                     size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(blocks && x && y, "null pointer");
    RMW_REQUIRE(x != y, "the block apply is out of place");
    RMW_REQUIRE(k >= 1 && k <= 65535, "k must be in [1, 65535]");
    RMW_REQUIRE(ld == 0 || ld >= 3 * L.nblobs, "ld must be 0 or >= 3 N");
    const size_t lds = (size_t)L.max_n * sizeof(double);
    RMW_REQUIRE(lds <= 160 * 1024, "body too large for the LDS block apply");
    if (lds > 64 * 1024)
      RMW_CHECK_HIP(hipFuncSetAttribute((const void*)block_apply_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                        (int)lds));
    char* ws = (char*)workspace;
    hipLaunchKernelGGL(block_apply_kernel, dim3(nbodies, k), dim3(kRigidBlock), lds, (hipStream_t)stream,
                       (const int64_t*)(ws + L.off_offs), (const int64_t*)(ws + L.off_moff), blocks, trans ? 1 : 0,
                       lower ? 1 : 0, x, y, ld > 0 ? ld : 3 * L.nblobs);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rigid_point_field(const double* r_blobs, const double* centres, const double* force,
                                     int64_t force_stride, const int64_t* body_offsets, int nbodies, double radius,
                                     double a, double eta, int wall, double* velocity, void* workspace,
                                     size_t workspace_bytes, void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(r_blobs && centres && force && velocity, "null pointer");
    RMW_REQUIRE(force_stride >= 3, "force_stride must be >= 3");
    RMW_REQUIRE(radius > 0.0 && a > 0.0 && eta > 0.0, "radius, a and eta must be positive");
    if (L.nblobs == 0) return;
    const char* ws = (const char*)workspace;
    const unsigned grid = (unsigned)((L.nblobs + kPfBlock - 1) / kPfBlock);
    hipLaunchKernelGGL(point_field_kernel, dim3(grid), dim3(kPfBlock), 0, (hipStream_t)stream, r_blobs, L.nblobs,
                       (const int64_t*)(ws + L.off_offs), nbodies, centres, force, force_stride, radius, a,
                       1.0 / (8.0 * M_PI * eta), wall, velocity);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rigid_pc_apply(const int64_t* body_offsets, int nbodies, const double* minv,
                                  const double* minvk, const double* nbody, const double* r_rel,
                                  const double* x, double* y, void* workspace, size_t workspace_bytes,
                                  void* stream)
{
  return guarded([&] {
    RigidLayout L = rigid_layout(body_offsets, nbodies);
    RMW_REQUIRE(workspace_bytes >= L.total, "workspace too small");
    RMW_REQUIRE(x != y, "pc apply is out of place");
    hipStream_t st = (hipStream_t)stream;
    char* ws = (char*)workspace;
    const size_t lds = 2 * (size_t)L.max_n * sizeof(double);
    RMW_REQUIRE(lds <= 160 * 1024, "body too large for the LDS preconditioner apply");
    if (lds > 64 * 1024)
      RMW_CHECK_HIP(hipFuncSetAttribute((const void*)pc_apply_kernel,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(pc_apply_kernel, dim3(nbodies), dim3(kRigidBlock), lds, st,
                       (const int64_t*)(ws + L.off_offs), (const int64_t*)(ws + L.off_moff), minv, minvk,
                       nbody, r_rel, x, y, L.nblobs);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
