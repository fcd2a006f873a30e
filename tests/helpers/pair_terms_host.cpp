// Test helper: compiles the product's pair-term header for the host so the
// decomposed block forms can be compared with the oracle on the CPU.
// Exposes the dense 3x3 block (kernel level, units of a, normalised like
// the reference: 1/(8 pi eta a^p)) obtained by applying the block to the
// three unit vectors.
#include <cmath>
#include <cstdint>
#include "../../rigidmultiblobswall_amd/csrc/rmw_pair_terms.h"

using namespace rmw;

template <typename T>
static Vec3<T> apply_block(int block, const Geom<T>& g, T zi, T zj, T vx, T vy, T vz)
{
  switch (block) {
    case 0: return apply_uf(g, coef_uf(g, zi, zj), vx, vy, vz);
    case 1: return apply_ut(g, coef_rot(g, zi), vx, vy, vz);
    case 2: return apply_wf(g, coef_rot(g, zj), vx, vy, vz);
    case 4: return apply_uf_t(g, coef_uf(g, zi, zj), vx, vy, vz);  // UF_ij^T
    case 5: return apply_ut_t(g, coef_rot(g, zi), vx, vy, vz);     // UT_ij^T
    case 6: return apply_wf_t(g, coef_rot(g, zj), vx, vy, vz);     // WF_ij^T
    case 7: return apply_wt_t(g, coef_wt(g), vx, vy, vz);          // WT_ij^T
    default: return apply_wt(g, coef_wt(g), vx, vy, vz);
  }
}

template <typename T>
static Vec3<T> apply_self(int block, T h, T vx, T vy, T vz)
{
  switch (block) {
    case 0: return self_uf(h, vx, vy, vz);
    case 1: return self_ut(h, vx, vy, vz);
    case 2: return self_wf(h, vx, vy, vz);
    default: return self_wt(h, vx, vy, vz);
  }
}

// Blocks 8-11: the explicit matrices mat_uf, mat_ut, mat_wf, mat_wt.
template <typename T>
static Mat3<T> explicit_block(int block, const Geom<T>& g, T zi, T zj)
{
  switch (block) {
    case 8: return mat_uf(g, coef_uf(g, zi, zj));
    case 9: return mat_ut(g, coef_rot(g, zi));
    case 10: return mat_wf(g, coef_rot(g, zj));
    default: return mat_wt(g, coef_wt(g));
  }
}

template <typename T>
static int pair_matrix(int block, const double* ri, const double* rj, int self, double eta, double a, double* m9)
{
  if (block < 0 || block > 11) return -1;
  const int b4 = block & 3;
  const int p = b4 == 0 ? 1 : (b4 == 3 ? 3 : 2);
  const double norm = 1.0 / (8.0 * M_PI * eta * std::pow(a, p));
  T dx = T((ri[0] - rj[0]) / a), dy = T((ri[1] - rj[1]) / a), dz = T((ri[2] - rj[2]) / a);
  T zi = T(ri[2] / a), zj = T(rj[2] / a);
  Geom<T> g = make_geom(dx, dy, dz, zi, zj);
  if (block >= 8) {
    if (self) return -1;
    const Mat3<T> M = explicit_block<T>(block, g, zi, zj);
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m9[3 * r + c] = double(M.m[r][c]) * norm;
    return 0;
  }
  for (int c = 0; c < 3; ++c) {
    T v[3] = {T(0), T(0), T(0)};
    v[c] = T(1);
    Vec3<T> u = self ? apply_self<T>(block < 4 ? block : 0, zi, v[0], v[1], v[2])
                     : apply_block<T>(block, g, zi, zj, v[0], v[1], v[2]);
    m9[0 * 3 + c] = double(u.x) * norm;
    m9[1 * 3 + c] = double(u.y) * norm;
    m9[2 * 3 + c] = double(u.z) * norm;
  }
  return 0;
}

extern "C" int pt_pair_matrix_f64(int block, const double* ri, const double* rj, int self, double eta, double a, double* m9)
{
  return pair_matrix<double>(block, ri, rj, self, eta, a, m9);
}

extern "C" int pt_pair_matrix_f32(int block, const double* ri, const double* rj, int self, double eta, double a, double* m9)
{
  return pair_matrix<float>(block, ri, rj, self, eta, a, m9);
}

// Different radii: the (target radius a, source radius b) block, normalised
// by 1/(8 pi eta), positions and radii in absolute units.
template <typename T>
static int radii_matrix(const double* ri, const double* rj, double a, double b, double eta, int wall, double* m9)
{
  const double norm = 1.0 / (8.0 * M_PI * eta);
  T dx = T(ri[0] - rj[0]), dy = T(ri[1] - rj[1]), dz = T(ri[2] - rj[2]);
  T Rz = T(ri[2] + rj[2]);
  CoefRadii<T> c = coef_radii<T>(dx, dy, dz, Rz, T(ri[2]), T(rj[2]), T(a), T(b), T(1.0 / a), T(1.0 / b), wall != 0);
  for (int k = 0; k < 3; ++k) {
    T v[3] = {T(0), T(0), T(0)};
    v[k] = T(1);
    Vec3<T> u = apply_radii<T>(dx, dy, dz, Rz, c, v[0], v[1], v[2]);
    m9[0 * 3 + k] = double(u.x) * norm;
    m9[1 * 3 + k] = double(u.y) * norm;
    m9[2 * 3 + k] = double(u.z) * norm;
  }
  return 0;
}

extern "C" int pt_radii_matrix_f64(const double* ri, const double* rj, double a, double b, double eta, int wall,
                                   double* m9)
{
  return radii_matrix<double>(ri, rj, a, b, eta, wall, m9);
}

extern "C" int pt_radii_matrix_f32(const double* ri, const double* rj, double a, double b, double eta, int wall,
                                   double* m9)
{
  return radii_matrix<float>(ri, rj, a, b, eta, wall, m9);
}

// make_geom_on: every block of every boundary, dyadic and explicit forms,
// applied to a unit vector, for a pair that is (on = 1) or is not (on = 0)
// taken.  m_out: the four blocks and their transposes in every domain
// (8 block forms x 4 domains x 3 unit vectors x 3 components = 288 values),
// then 4 explicit blocks x 9 (36 values).
template <typename T>
static int masked_blocks(const double* ri, const double* rj, int on, double* m_out)
{
  T dx = T(ri[0] - rj[0]), dy = T(ri[1] - rj[1]), dz = T(ri[2] - rj[2]);
  T zi = T(ri[2]), zj = T(rj[2]);
  Geom<T> g = make_geom_on(dx, dy, dz, zi, zj, on != 0);
  int o = 0;
  for (int dom = 0; dom < 4; ++dom) {
    const CoefUF<T> cu = coef_uf(g, zi, zj, dom);
    const CoefRot<T> ci = coef_rot(g, zi, dom), cj = coef_rot(g, zj, dom);
    const CoefWT<T> cw = coef_wt(g, dom);
    for (int c = 0; c < 3; ++c) {
      T v[3] = {T(0), T(0), T(0)};
      v[c] = T(1);
      const Vec3<T> r[8] = {apply_uf(g, cu, v[0], v[1], v[2]),   apply_uf_t(g, cu, v[0], v[1], v[2]),
                            apply_ut(g, ci, v[0], v[1], v[2]),   apply_ut_t(g, ci, v[0], v[1], v[2]),
                            apply_wf(g, cj, v[0], v[1], v[2]),   apply_wf_t(g, cj, v[0], v[1], v[2]),
                            apply_wt(g, cw, v[0], v[1], v[2]),   apply_wt_t(g, cw, v[0], v[1], v[2])};
      for (int b = 0; b < 8; ++b) {
        m_out[o++] = double(r[b].x);
        m_out[o++] = double(r[b].y);
        m_out[o++] = double(r[b].z);
      }
    }
  }
  const CoefUF<T> cu = coef_uf(g, zi, zj);
  const Mat3<T> M[4] = {mat_uf(g, cu), mat_ut(g, coef_rot(g, zi)), mat_wf(g, coef_rot(g, zj)), mat_wt(g, coef_wt(g))};
  for (int b = 0; b < 4; ++b)
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < 3; ++c) m_out[o++] = double(M[b].m[r][c]);
  return o;
}

extern "C" int pt_masked_blocks_f32(const double* ri, const double* rj, int on, double* m_out)
{
  return masked_blocks<float>(ri, rj, on, m_out);
}

extern "C" int pt_masked_blocks_f64(const double* ri, const double* rj, int on, double* m_out)
{
  return masked_blocks<double>(ri, rj, on, m_out);
}

// ---------------------------------------------------------------------------
// Work-item schedule of the symmetric kernels (rmw_tile_pairs.h): count the
// (target tile, source block or tile) pairs that a decode of every item
// visits other than exactly once; 0 when each unordered pair with source
// >= target is covered once and nothing else.
#include <vector>
#include "../../rigidmultiblobswall_amd/csrc/rmw_tile_pairs.h"

extern "C" int64_t tp_block_cover_errors(int64_t n_t, int64_t C)
{
  const int64_t nb = 4 * n_t, items = items_before_blocks(n_t, n_t, C);
  std::vector<int> cov((size_t)(n_t * nb), 0);
  int64_t bad = 0;
  for (int64_t local = 0; local < items; ++local) {
    const int64_t T = pair_target_tile_blocks(local, n_t, C);
    const int64_t b0 = 4 * T + (local - items_before_blocks(T, n_t, C)) * C;
    const int64_t b1 = b0 + C < nb ? b0 + C : nb;
    if (T < 0 || T >= n_t || b0 < 4 * T || b0 >= nb) { ++bad; continue; }
    for (int64_t b = b0; b < b1; ++b) ++cov[(size_t)(T * nb + b)];
  }
  for (int64_t T = 0; T < n_t; ++T)
    for (int64_t b = 0; b < nb; ++b) bad += cov[(size_t)(T * nb + b)] != (b >= 4 * T ? 1 : 0);
  return bad;
}

extern "C" int64_t tp_tile_cover_errors(int64_t n_t, int64_t C)
{
  const int64_t items = items_before(n_t, n_t, C);
  std::vector<int> cov((size_t)(n_t * n_t), 0);
  int64_t bad = 0;
  for (int64_t it = 0; it < items; ++it) {
    const int64_t T = pair_target_tile(it, n_t, C);
    const int64_t S0 = T + (it - items_before(T, n_t, C)) * C;
    const int64_t S1 = S0 + C < n_t ? S0 + C : n_t;
    if (T < 0 || T >= n_t || S0 < T || S0 >= n_t) { ++bad; continue; }
    for (int64_t S = S0; S < S1; ++S) ++cov[(size_t)(T * n_t + S)];
  }
  for (int64_t T = 0; T < n_t; ++T)
    for (int64_t S = 0; S < n_t; ++S) bad += cov[(size_t)(T * n_t + S)] != (S >= T ? 1 : 0);
  return bad;
}

extern "C" int64_t tp_ceil_sum(int64_t M, int64_t C) { return ceil_sum(M, C); }
