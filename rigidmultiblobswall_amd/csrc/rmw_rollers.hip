// Per-roller (O(N)) parts of the rollers time step (gfx950): the one-blob
// forces, the uncorrelated (self-mobility) velocities, and the position
// update with the wall check.  The O(N^2) parts (mobility products, blob-blob
// forces) are the pair kernels of the other files.
//
// Reference:
//   one-blob force (gravity + wall repulsion)   multi_bodies/multi_bodies_functions.py:151-196
//                                               (blob_external_forces), bound by multi_bodies.py:1325-1328
//   uncorrelated velocity and torque            quaternion_integrator_rollers.py:985-1080
//   uncorrelated stochastic velocity            quaternion_integrator_rollers.py:1263-1313
//   position update, Adams-Bashforth, wall test quaternion_integrator_rollers.py:119-302
//   free-kinematics torque                      quaternion_integrator_rollers.py:1586-1597
#include <hip/hip_runtime.h>

#include <cmath>
#include <string>

#include "rmw_internal.h"

namespace rmw {

constexpr int kRollBlock = 256;

// f_z = -g m + (eps/b) exp(-(z - a)/b) above the radius, + eps/b at or
// below it; f_x = f_y = 0.
__global__ __launch_bounds__(kRollBlock) void one_blob_force_kernel(const double* __restrict__ r, int64_t n,
                                                                     double mg, double a, double eb, double invb,
                                                                     int accumulate, double* __restrict__ f)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double h = r[3 * i + 2];
  double fz = -mg;
  fz += h > a ? eb * exp(-(h - a) * invb) : eb;
  if (accumulate) {
    f[3 * i + 2] += fz;
  } else {
    f[3 * i] = 0.0;
    f[3 * i + 1] = 0.0;
    f[3 * i + 2] = fz;
  }
}

// Self-mobility coefficients of one roller at height z (Swan-Brady, with
// the effective height max(z/a, 1) and the damping of the reference).
struct SelfCoef {
  double tt_para, tt_perp, rt_para, damping, h;
};

__device__ __forceinline__ SelfCoef self_coef(double z, double eta, double a)
{
  SelfCoef c;
  const double h_over_a = z / a;
  c.h = h_over_a > 1.0 ? h_over_a : 1.0;
  c.damping = 1.0;
  if (h_over_a < 0.0) c.damping = 0.0;
  else if (h_over_a <= 1.0) c.damping = h_over_a;
  const double ftt = 1.0 / (6.0 * M_PI * eta * a);
  const double frt = 1.0 / (6.0 * M_PI * eta * a * a);
  const double h2 = c.h * c.h, h3 = h2 * c.h, h4 = h2 * h2, h5 = h4 * c.h;
  c.rt_para = frt * (3.0 / (32.0 * h4)) * c.damping;
  c.tt_perp = ftt * (1.0 - 9.0 / (8.0 * c.h) + 1.0 / (2.0 * h3) - 1.0 / (8.0 * h5)) * c.damping;
  c.tt_para = ftt * (1.0 - 9.0 / (16.0 * c.h) + 2.0 / (16.0 * h3) - 1.0 / (16.0 * h5)) * c.damping;
  return c;
}

// Deterministic velocity (and torque) without hydrodynamic interactions.
// free != 0: torque = omega 8 pi eta a^3; else the torque that makes the
// roller spin at omega given its own force.
__global__ __launch_bounds__(kRollBlock) void uncorrelated_det_kernel(const double* __restrict__ r,
                                                                      const double* __restrict__ f, int64_t n,
                                                                      double ox, double oy, double oz, int free_kin,
                                                                      double eta, double a,
                                                                      double* __restrict__ v, double* __restrict__ t)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SelfCoef c = self_coef(r[3 * i + 2], eta, a);
  const double fx = f[3 * i], fy = f[3 * i + 1], fz = f[3 * i + 2];
  double tx, ty, tz;
  if (free_kin) {
    const double s = 8.0 * M_PI * eta * a * a * a;
    tx = ox * s; ty = oy * s; tz = oz * s;
  } else {
    const double frr = 1.0 / (6.0 * M_PI * eta * a * a * a);
    const double h3 = c.h * c.h * c.h;
    const double rr_perp_inv = 1.0 / (frr * (3.0 / 4.0 - 3.0 / (32.0 * h3)) * c.damping);
    const double rr_para_inv = 1.0 / (frr * (3.0 / 4.0 - 15.0 / (64.0 * h3)) * c.damping);
    tx = rr_para_inv * (ox + c.rt_para * fy);
    ty = rr_para_inv * (oy - c.rt_para * fx);
    tz = rr_perp_inv * oz;
  }
  v[3 * i] = c.tt_para * fx + c.rt_para * ty;
  v[3 * i + 1] = c.tt_para * fy - c.rt_para * tx;
  v[3 * i + 2] = c.tt_perp * fz;
  if (t) {
    t[3 * i] = tx; t[3 * i + 1] = ty; t[3 * i + 2] = tz;
  }
}

// Stochastic velocity without hydrodynamic interactions:
// sqrt(2 kT/dt) sqrt(mu) z plus the drift kT d(mu_perp)/dh along z.
__global__ __launch_bounds__(kRollBlock) void uncorrelated_stoch_kernel(const double* __restrict__ r,
                                                                        const double* __restrict__ z, int64_t n,
                                                                        double eta, double a, double kT,
                                                                        double factor_disp,
                                                                        double* __restrict__ v)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const SelfCoef c = self_coef(r[3 * i + 2], eta, a);
  const double ftt = 1.0 / (6.0 * M_PI * eta * a);
  const double h2 = c.h * c.h, h4 = h2 * h2, h6 = h4 * h2;
  const double deriv = ftt * (9.0 / (8.0 * h2) - 3.0 / (2.0 * h4) + 5.0 / (8.0 * h6)) * c.damping;
  const double sp = factor_disp * sqrt(c.tt_para);
  v[3 * i] = sp * z[3 * i];
  v[3 * i + 1] = sp * z[3 * i + 1];
  v[3 * i + 2] = factor_disp * sqrt(c.tt_perp) * z[3 * i + 2] + kT * deriv;
}

// r_new = r_old + dt v with v = c_det det + c_prev prev + stoch (prev and
// stoch optional); per-workgroup counts of new heights below 0 and below a.
__global__ __launch_bounds__(kRollBlock) void advance_kernel(const double* __restrict__ r_old,
                                                             const double* __restrict__ det,
                                                             const double* __restrict__ prev,
                                                             const double* __restrict__ stoch, int64_t n,
                                                             double c_det, double c_prev, double dt, double a,
                                                             double* __restrict__ r_new,
                                                             int64_t* __restrict__ block_counts)
{
  __shared__ int s_neg[kRollBlock / 64], s_low[kRollBlock / 64];
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int neg = 0, low = 0;
  if (i < n) {
    for (int c = 0; c < 3; ++c) {
      const int64_t k = 3 * i + c;
      double vel = c_det * det[k];
      if (prev) vel = vel - c_prev * prev[k];
      if (stoch) vel = vel + stoch[k];
      r_new[k] = r_old[k] + dt * vel;
    }
    const double z = r_new[3 * i + 2];
    neg = z < 0.0;
    low = z < a;
  }
  // wave sums through the ballot masks, then one lane per workgroup
  const unsigned long long bn = __ballot(neg), bl = __ballot(low);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    s_neg[wv] = __popcll(bn);
    s_low[wv] = __popcll(bl);
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    int sn = 0, sl = 0;
    for (int w = 0; w < kRollBlock / 64; ++w) {
      sn += s_neg[w];
      sl += s_low[w];
    }
    block_counts[2 * blockIdx.x] = sn;
    block_counts[2 * blockIdx.x + 1] = sl;
  }
}

__global__ __launch_bounds__(256) void count_reduce_kernel(const int64_t* __restrict__ block_counts, int64_t nblocks,
                                                           int64_t* __restrict__ counts)
{
  __shared__ int64_t s[2][256];
  int64_t a = 0, b = 0;
  for (int64_t k = threadIdx.x; k < nblocks; k += 256) {
    a += block_counts[2 * k];
    b += block_counts[2 * k + 1];
  }
  s[0][threadIdx.x] = a;
  s[1][threadIdx.x] = b;
  __syncthreads();
  for (int w = 128; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w) {
      s[0][threadIdx.x] += s[0][threadIdx.x + w];
      s[1][threadIdx.x] += s[1][threadIdx.x + w];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    counts[0] = s[0][0];
    counts[1] = s[1][0];
  }
}

static unsigned grid_of(int64_t n) { return (unsigned)((n + kRollBlock - 1) / kRollBlock); }

}  // namespace rmw

using namespace rmw;

extern "C" int rmw_rollers_one_blob_forces(const double* r, int64_t n, double blob_mass, double g, double blob_radius,
                                           double repulsion_strength_wall, double debye_length_wall, int accumulate,
                                           double* force, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    RMW_REQUIRE(debye_length_wall > 0.0, "debye_length_wall must be positive");
    if (n == 0) return;
    RMW_REQUIRE(r && force, "r and force are required");
    hipLaunchKernelGGL(one_blob_force_kernel, dim3(grid_of(n)), dim3(kRollBlock), 0, (hipStream_t)stream, r, n,
                       g * blob_mass, blob_radius, repulsion_strength_wall / debye_length_wall,
                       1.0 / debye_length_wall, accumulate, force);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rollers_uncorrelated_velocity(const double* r, const double* force, int64_t n, const double* omega,
                                                 int free_kinematics, double eta, double a, double* velocity,
                                                 double* torque, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    RMW_REQUIRE(eta > 0.0 && a > 0.0, "eta and a must be positive");
    RMW_REQUIRE(omega != nullptr, "omega (3 host values) is required");
    if (n == 0) return;
    RMW_REQUIRE(r && force && velocity, "r, force and velocity are required");
    hipLaunchKernelGGL(uncorrelated_det_kernel, dim3(grid_of(n)), dim3(kRollBlock), 0, (hipStream_t)stream, r,
                       force, n, omega[0], omega[1], omega[2], free_kinematics, eta, a, velocity, torque);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rollers_uncorrelated_noise(const double* r, const double* z, int64_t n, double eta, double a,
                                              double kT, double dt, double* velocity, void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    RMW_REQUIRE(eta > 0.0 && a > 0.0 && dt > 0.0 && kT >= 0.0, "eta, a, dt must be positive and kT >= 0");
    if (n == 0) return;
    RMW_REQUIRE(r && z && velocity, "r, z and velocity are required");
    hipLaunchKernelGGL(uncorrelated_stoch_kernel, dim3(grid_of(n)), dim3(kRollBlock), 0, (hipStream_t)stream, r, z,
                       n, eta, a, kT, std::sqrt(2.0 * kT / dt), velocity);
    RMW_CHECK_HIP(hipGetLastError());
  });
}

extern "C" int rmw_rollers_advance_workspace_size(int64_t n, size_t* bytes)
{
  return guarded([&] {
    RMW_REQUIRE(bytes != nullptr && n >= 0, "bad arguments");
    *bytes = align_up((size_t)(grid_of(n) > 0 ? grid_of(n) : 1) * 2 * sizeof(int64_t));
  });
}

extern "C" int rmw_rollers_advance(const double* r_old, const double* det_velocity, const double* prev_velocity,
                                   const double* stoch_velocity, int64_t n, double c_det, double c_prev, double dt,
                                   double a, double* r_new, int64_t* counts, void* workspace, size_t workspace_bytes,
                                   void* stream)
{
  return guarded([&] {
    RMW_REQUIRE(n >= 0, "n must be >= 0");
    const size_t need = align_up((size_t)(grid_of(n) > 0 ? grid_of(n) : 1) * 2 * sizeof(int64_t));
    RMW_REQUIRE(workspace && workspace_bytes >= need, "workspace too small: need " + std::to_string(need) + " bytes");
    RMW_REQUIRE(counts != nullptr, "counts (2 device int64) is required");
    hipStream_t st = (hipStream_t)stream;
    const unsigned nb = grid_of(n);
    if (nb > 0) {
      RMW_REQUIRE(r_old && det_velocity && r_new, "r_old, det_velocity and r_new are required");
      hipLaunchKernelGGL(advance_kernel, dim3(nb), dim3(kRollBlock), 0, st, r_old, det_velocity, prev_velocity,
                         stoch_velocity, n, c_det, c_prev, dt, a, r_new, (int64_t*)workspace);
      RMW_CHECK_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(count_reduce_kernel, dim3(1), dim3(256), 0, st, (const int64_t*)workspace, (int64_t)nb,
                       counts);
    RMW_CHECK_HIP(hipGetLastError());
  });
}
