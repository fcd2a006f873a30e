// Pair terms of the wall-corrected RPY blob mobility (Rotne-Prager-Yamakawa
// tensor plus the Swan-Brady single-wall image correction), written once as
// host/device templates over the real type T (float or double).
//
// Reference math: mobility/mobility_numba.py
//   UF  :200-276  (RPY :209-234, wall :241-276)
//   UT  :614-679  (wall in the transposed source/target role, :646-679)
//   WF  :1004-1065
//   WT  :1256-1321
//
// The reference forms the 3x3 block M_ij element by element and multiplies
// it by the source vector.  Here each block is decomposed into a few scalar
// coefficients times dyads built from d = r_i - r_j, the image vector
// R = (d_x, d_y, z_i + z_j) and z-hat, so that the product M_ij v costs
// ~16-20 operations per right-hand side and the geometry part is shared by
// k right-hand sides and by the four blocks of the grand mobility:
//
//   UF v = (A + f1) v + Bc (d.v) d + [f2 (e.v) + f3 v_z] e + [f4 (e.v) + f5 v_z] z
//   WF v = c (v x d) - [F1 (v x e) + (F2 v_z + F3 (e.v)) (z x e) + F4 (v x z)]
//   UT v = c (v x d) - [F1 (e' x v) + (F2 z + F3 e') ((z x e').v) + F4 (z x v)]
//   WT v = (A + G1) v + Bc (d.v) d + G2 (e.v) e + G3 (e.v) z + G4 ((z x e).v)(z x e)
//
// with e = R/|R| and e' = (-R_x, -R_y, R_z)/|R| (UT's image seen from the
// source).  All lengths are in units of the blob radius a; the 1/(8 pi eta
// a^p) normalisations are folded into the source vectors by the caller.
// Self pairs (i == j in the central box) are handled by the caller.
#pragma once

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define RMW_HD __host__ __device__ __forceinline__
#else
#include <cmath>
#define RMW_HD inline
#endif

namespace rmw {

template <typename T> struct Vec3 { T x, y, z; };

// Boundary of the fluid (the `domain` options of the reference):
//   wall          RPY + Swan-Brady image (mobility_numba.py:123-287, ...)
//   no wall       RPY alone (:12-120, 440-545, 832-936, 1077-1186)
//   in plane      the wall blocks projected on x-y, P M P with
//                 P = diag(1, 1, 0) (:291-437, 690-830); the caller
//                 projects sources and outputs
//   free surface  RPY(d) + RPY(R) D, D = diag(1, 1, -1) (:1770-1938); UF only
enum : int { kDomWall = 0, kDomNoWall = 1, kDomInPlane = 2, kDomFreeSurface = 3 };
RMW_HD bool dom_has_wall(int dom) { return dom == kDomWall || dom == kDomInPlane; }

// Reciprocal square root.  FP32 on the device uses the hardware v_rsq_f32
// (about 1 ulp).  FP64 on the device refines the hardware estimate v_rsq_f64
// (relative error eps <= 5.3e-8 = 2^-24.2, measured on gfx950 over 4M
// arguments in [1e-7, 1e7]) by one third-order step of Newton's iteration
// for y^-2 = x: with e = 1 - x y^2, x^-1/2 = y (1 - e)^-1/2
// = y (1 + e/2 + 3 e^2/8 + O(e^3)), so the error falls to ~(5/16) eps^3
// plus the rounding of the step (measured: <= 1.39e-16 relative, against
// 1.66e-16 for 1/sqrt(x) with both operations rounded) in 5 FP64 operations
// instead of the ~25 of a correctly rounded sqrt and division.
RMW_HD float rsq(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rsqf(x);
#else
  return 1.0f / std::sqrt(x);
#endif
}
RMW_HD double rsq(double x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  const double y = __builtin_amdgcn_rsq(x);
  const double e = __builtin_fma(-(x * y), y, 1.0);
  return __builtin_fma(y * e, __builtin_fma(0.375, e, 0.5), y);
#else
  return 1.0 / std::sqrt(x);
#endif
}

RMW_HD float rcp(float x)
{
#if defined(__HIP_DEVICE_COMPILE__)
  return __builtin_amdgcn_rcpf(x);
#else
  return 1.0f / x;
#endif
}
RMW_HD double rcp(double x) { return 1.0 / x; }

// Fused multiply-add a*b + c (single rounding on the device and the host).
RMW_HD float fma_(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
RMW_HD double fma_(double a, double b, double c) { return __builtin_fma(a, b, c); }

// ---------------------------------------------------------------------------
// Shared geometry of one (target, source) pair.
// ---------------------------------------------------------------------------
template <typename T> struct Geom {
  T dx, dy, dz;   // d = r_i - r_j (units of a)
  T Rz;           // z_i + z_j
  T r2, invr, r;  // |d|^2, 1/|d|, |d|
  T invR;         // 1/|R|
  T ez, ez2;      // R_z/|R| and its square
};

template <typename T>
RMW_HD Geom<T> make_geom(T dx, T dy, T dz, T zi, T zj)
{
  Geom<T> g;
  g.dx = dx; g.dy = dy; g.dz = dz;
  g.Rz = zi + zj;
  T p2 = dx * dx + dy * dy;
  g.r2 = p2 + dz * dz;
  g.invr = rsq(g.r2);
  g.r = g.r2 * g.invr;
  T R2 = p2 + g.Rz * g.Rz;
  g.invR = rsq(R2);
  g.ez = g.Rz * g.invR;
  g.ez2 = g.ez * g.ez;
  return g;
}

// make_geom of a pair that may not be taken (on = false): then 1/|d| =
// 1/|R| = 0 and |d| = 3 (the far-field branches), which makes every
// coefficient of every block and boundary exactly zero, so the pair adds +-0
// to both of its blobs without a select per output.  d and the heights must
// be finite; a taken pair gets the same values as make_geom.
template <typename T>
RMW_HD Geom<T> make_geom_on(T dx, T dy, T dz, T zi, T zj, bool on)
{
  Geom<T> g;
  g.dx = dx; g.dy = dy; g.dz = dz;
  g.Rz = zi + zj;
  T p2 = dx * dx + dy * dy;
  g.r2 = p2 + dz * dz;
  const T ir = rsq(g.r2);
  g.invr = on ? ir : T(0);
  g.r = on ? g.r2 * ir : T(3);
  T R2 = p2 + g.Rz * g.Rz;
  const T iR = rsq(R2);
  g.invR = on ? iR : T(0);
  g.ez = g.Rz * g.invR;
  g.ez2 = g.ez * g.ez;
  return g;
}

// ---------------------------------------------------------------------------
// UF: translation from force.   mobility_numba.py:209-276
// ---------------------------------------------------------------------------
template <typename T> struct CoefUF {
  T s;      // A + f1 (coefficient of v)
  T s5;     // s + f5 (coefficient of v_z in the z row)
  T Bc;     // coefficient of (d.v) d
  T g2;     // f2 / |R|^2  -> coefficient of (R.v) R
  T g3;     // f3 / |R|    -> coefficient of v_z R
  T g4;     // f4 / |R|    -> coefficient of (R.v) z
};

// RPY translational coefficients A I + B r r^T at distance r (units of a):
// far (r > 2) and overlapping branches as selects.
template <typename T>
RMW_HD void rpy_tt(T r, T invr, T& A, T& B)
{
  const T one = T(1), two = T(2);
  T invr2 = invr * invr;
  T A_far = (one + T(2.0 / 3.0) * invr2) * invr;
  T B_far = (one - two * invr2) * (invr2 * invr);
  T A_near = T(4.0 / 3.0) - T(4.0 / 3.0 * 0.28125) * r;
  T B_near = T(4.0 / 3.0 * 0.09375) * invr;
  bool far = r > two;
  A = far ? A_far : A_near;
  B = far ? B_far : B_near;
}

template <typename T>
RMW_HD CoefUF<T> coef_uf(const Geom<T>& g, T zi, T zj, int dom = kDomWall)
{
  const T one = T(1);
  CoefUF<T> c;
  T A;
  rpy_tt(g.r, g.invr, A, c.Bc);
  if (dom == kDomNoWall) {
    c.s = A; c.s5 = A; c.g2 = T(0); c.g3 = T(0); c.g4 = T(0);
    return c;
  }
  if (dom == kDomFreeSurface) {
    // + RPY(R) D v = A_R (v - 2 v_z z) + B_R R (R.v) - 2 B_R R_z v_z R
    T AR, BR;
    rpy_tt(one / g.invR, g.invR, AR, BR);
    c.s = A + AR; c.s5 = A - AR; c.g2 = BR; c.g3 = T(-2) * BR * g.Rz; c.g4 = T(0);
    return c;
  }

  // Swan-Brady wall correction (mobility_numba.py:254-264) regrouped as
  // polynomials in w = 1/|R|^2 with E = e_z^2 = R_z^2 w.  With
  // h_hat = z_j/R_z the h_hat factors combine into division-free monomials:
  //   h_hat (1-h_hat) e_z^2 = z_i z_j w,  h_hat e_z / |R| = z_j w,
  //   (1-h_hat) e_z^2 = z_i R_z w,        h_hat^2 e_z^2 = z_j^2 w,
  // giving
  //   f1 = -q [1 + w (2 zz + 2/3 - 2E) + w^2 (10E/3 - 2/3)]
  //   f2 = -q [1 + w (-6 zz - 2 + 10E) + w^2 (10/3 - 70E/3)]
  //   f3 = w [2 z_j + R_z w (-12 zz - 4 + 20E + w (40/3 - 140E/3))]
  //   f4 = w [2 z_j - (20/3) R_z w^2]
  //   f5 = -q w [4 z_j^2 + 4E + w (8/3 - 20E)]
  // with q = 1/|R| and zz = z_i z_j.
  T q = g.invR;
  T w = q * q;
  T E = g.ez2;
  T zz = zi * zj;
  T t1 = fma_(T(2), zz, fma_(T(-2), E, T(2.0 / 3.0)));
  T t2 = fma_(T(10.0 / 3.0), E, T(-2.0 / 3.0));
  T f1 = -q * fma_(w, fma_(w, t2, t1), one);
  T t3 = fma_(T(-6), zz, fma_(T(10), E, T(-2)));
  T t4 = fma_(T(-70.0 / 3.0), E, T(10.0 / 3.0));
  T qw = q * w;
  c.g2 = -qw * fma_(w, fma_(w, t4, t3), one);              // f2 w
  T Rzw = g.Rz * w;
  T t5 = fma_(T(-12), zz, fma_(T(20), E, T(-4)));
  T t6 = fma_(T(-140.0 / 3.0), E, T(40.0 / 3.0));
  T two_zj = zj + zj;
  c.g3 = qw * fma_(Rzw, fma_(w, t6, t5), two_zj);          // f3 q
  c.g4 = qw * fma_(T(-20.0 / 3.0) * Rzw, w, two_zj);       // f4 q
  T t7 = fma_(T(-20), E, T(8.0 / 3.0));
  T t8 = fma_(T(4) * zj, zj, T(4) * E);
  T f5 = -qw * fma_(w, t7, t8);
  c.s = A + f1;
  c.s5 = c.s + f5;
  return c;
}

template <typename T>
RMW_HD Vec3<T> apply_uf(const Geom<T>& g, const CoefUF<T>& c, T vx, T vy, T vz)
{
  T p = g.dx * vx + g.dy * vy;
  T dv = p + g.dz * vz;          // d.v
  T Rv = p + g.Rz * vz;          // R.v
  T ge = c.g2 * Rv + c.g3 * vz;  // coefficient of R
  T q = c.Bc * dv;               // coefficient of d
  T qe = q + ge;                 // x and y rows share d_x = R_x, d_y = R_y
  Vec3<T> u;
  u.x = c.s * vx + qe * g.dx;
  u.y = c.s * vy + qe * g.dy;
  u.z = c.s5 * vz + q * g.dz + ge * g.Rz + c.g4 * Rv;
  return u;
}

// Transposed block: UF_ij^T v, which is UF_ji v (the mobility is symmetric),
// so one evaluation of a pair serves both of its blobs.  The coefficients of
// v_z R and (R.v) z trade places.
template <typename T>
RMW_HD Vec3<T> apply_uf_t(const Geom<T>& g, const CoefUF<T>& c, T vx, T vy, T vz)
{
  T p = g.dx * vx + g.dy * vy;
  T dv = p + g.dz * vz;
  T Rv = p + g.Rz * vz;
  T ge = c.g2 * Rv + c.g4 * vz;
  T q = c.Bc * dv;
  T qe = q + ge;
  Vec3<T> u;
  u.x = c.s * vx + qe * g.dx;
  u.y = c.s * vy + qe * g.dy;
  u.z = c.s5 * vz + q * g.dz + ge * g.Rz + c.g3 * Rv;
  return u;
}

// ---------------------------------------------------------------------------
// Wall coefficients shared by UT and WF (mobility_numba.py:659-665 and
// :1046-1052), for a given h_hat.  Returned pre-divided by powers of |R| so
// that they multiply R rather than e.
// ---------------------------------------------------------------------------
template <typename T> struct CoefRot {
  T c;    // RPY antisymmetric coefficient (1/r^3 far, (1 - 3r/8)/2 near)
  T F1r;  // F1 / |R|
  T F2r;  // F2 / |R|
  T F3r2; // F3 / |R|^2
  T F4;   // F4
};

template <typename T>
RMW_HD CoefRot<T> coef_rot(const Geom<T>& g, T hnum, int dom = kDomWall)
{
  const T one = T(1), two = T(2);
  CoefRot<T> c;
  T invr3 = g.invr * g.invr * g.invr;
  T c_near = T(0.5) * (one - T(0.375) * g.r);
  c.c = (g.r >= two) ? invr3 : c_near;
  if (!dom_has_wall(dom)) {
    c.F1r = T(0); c.F2r = T(0); c.F3r2 = T(0); c.F4 = T(0);
    return c;
  }
  // h_hat = h/R_z enters only as h_hat e_z^2 = h R_z/|R|^2 and
  // h_hat e_z = h/|R| (no division needed).
  T invR = g.invR;
  T invR2 = invR * invR;
  T invR4 = invR2 * invR2;
  T ez2 = g.ez2;
  T F1 = invR2;
  T F2 = (T(6) * hnum * g.Rz + one - T(10) * ez2) * invR4 * two;
  T F3 = -(T(3) * hnum * invR * invR2 - T(5) * g.ez * invR4) * two;
  T F4 = -(hnum * invR * invR2 - g.ez * invR4) * two;
  c.F1r = F1 * invR;
  c.F2r = F2 * invR;
  c.F3r2 = F3 * invR2;
  c.F4 = F4;
  return c;
}

// WF v = c (v x d) - [F1 (v x e) + (F2 v_z + F3 (e.v)) (z x e) + F4 (v x z)]
template <typename T>
RMW_HD Vec3<T> apply_wf(const Geom<T>& g, const CoefRot<T>& c, T vx, T vy, T vz)
{
  T Rv = g.dx * vx + g.dy * vy + g.Rz * vz;
  T h = c.F2r * vz + c.F3r2 * Rv;           // (F2 v_z + F3 e.v) / |R|
  T zc = vx * g.dy - vy * g.dx;              // (v x d)_z = (v x R)_z
  Vec3<T> w;
  w.x = c.c * (vy * g.dz - vz * g.dy) - c.F1r * (vy * g.Rz - vz * g.dy) + h * g.dy - c.F4 * vy;
  w.y = c.c * (vz * g.dx - vx * g.dz) - c.F1r * (vz * g.dx - vx * g.Rz) - h * g.dx + c.F4 * vx;
  w.z = (c.c - c.F1r) * zc;
  return w;
}

// UT v = c (v x d) - [F1 (e' x v) + (F2 z + F3 e') ((z x e').v) + F4 (z x v)],
// e' = (-d_x, -d_y, R_z)/|R|, coefficients built with h_hat = z_i / R_z.
template <typename T>
RMW_HD Vec3<T> apply_ut(const Geom<T>& g, const CoefRot<T>& c, T vx, T vy, T vz)
{
  T q = g.dy * vx - g.dx * vy;               // |R| (z x e').v
  T m3 = c.F3r2 * q;                          // F3 ((z x e').v) / |R|
  Vec3<T> u;
  u.x = c.c * (vy * g.dz - vz * g.dy) + c.F1r * (g.dy * vz + g.Rz * vy) + m3 * g.dx + c.F4 * vy;
  u.y = c.c * (vz * g.dx - vx * g.dz) - c.F1r * (g.Rz * vx + g.dx * vz) + m3 * g.dy - c.F4 * vx;
  u.z = (c.c - c.F1r - c.F2r) * q - m3 * g.Rz;
  return u;
}

// Transposes of the rotational couplings, read off the component forms
// above.  Reciprocity gives UT_ji = WF_ij^T and WF_ji = UT_ij^T, so with
// the coefficients of the pair (i, j) these are the blocks of (j, i).
// WF_ij^T v, with p = d_x v_x + d_y v_y, q = d_y v_x - d_x v_y:
//   x = F3 d_x q / |R|^2 + (F1 R_z / |R| - c d_z + F4) v_y + (c - F1/|R|) d_y v_z
//   y = F3 d_y q / |R|^2 - (F1 R_z / |R| - c d_z + F4) v_x - (c - F1/|R|) d_x v_z
//   z = (F1/|R| + F2/|R| + F3 R_z/|R|^2 - c) q
template <typename T>
RMW_HD Vec3<T> apply_wf_t(const Geom<T>& g, const CoefRot<T>& c, T vx, T vy, T vz)
{
  T q = g.dy * vx - g.dx * vy;
  T m3 = c.F3r2 * q;
  T G = c.F1r * g.Rz - c.c * g.dz + c.F4;
  T cf = c.c - c.F1r;
  Vec3<T> w;
  w.x = m3 * g.dx + G * vy + cf * g.dy * vz;
  w.y = m3 * g.dy - G * vx - cf * g.dx * vz;
  w.z = (c.F1r + c.F2r + c.F3r2 * g.Rz - c.c) * q;
  return w;
}

// UT_ij^T v:
//   x =  F3 d_y p / |R|^2 - G v_y + K d_y v_z
//   y = -F3 d_x p / |R|^2 + G v_x - K d_x v_z
//   z = (F1/|R| - c) q
// with G = c d_z + F1 R_z / |R| + F4 and K = c - F1/|R| - F2/|R| - F3 R_z/|R|^2.
template <typename T>
RMW_HD Vec3<T> apply_ut_t(const Geom<T>& g, const CoefRot<T>& c, T vx, T vy, T vz)
{
  T p = g.dx * vx + g.dy * vy;
  T q = g.dy * vx - g.dx * vy;
  T m3 = c.F3r2 * p;
  T G = c.c * g.dz + c.F1r * g.Rz + c.F4;
  T K = c.c - c.F1r - c.F2r - c.F3r2 * g.Rz;
  Vec3<T> u;
  u.x = m3 * g.dy - G * vy + K * g.dy * vz;
  u.y = -m3 * g.dx + G * vx - K * g.dx * vz;
  u.z = (c.F1r - c.c) * q;
  return u;
}

// ---------------------------------------------------------------------------
// WT: rotation from torque.  mobility_numba.py:1256-1321
// ---------------------------------------------------------------------------
template <typename T> struct CoefWT {
  T s;    // A + G1
  T Bc;   // coefficient of (d.v) d
  T G2r;  // G2 / |R|^2 -> (R.v) R
  T G3r;  // G3 / |R|   -> (R.v) z
  T G4r;  // G4 / |R|^2 -> ((z x R).v)(z x R)
};

template <typename T>
RMW_HD CoefWT<T> coef_wt(const Geom<T>& g, int dom = kDomWall)
{
  const T one = T(1), two = T(2);
  CoefWT<T> c;
  T invr2 = g.invr * g.invr;
  T invr3 = invr2 * g.invr;
  T r3 = g.r2 * g.r;
  T A_far = T(-0.5) * invr3;
  T B_far = T(1.5) * invr2 * invr3;
  T A_near = one - T(0.84375) * g.r + T(0.078125) * r3;
  T B_near = T(0.28125) * g.invr - T(0.046875) * g.r;
  bool far = g.r >= two;
  T A = far ? A_far : A_near;
  c.Bc = far ? B_far : B_near;
  if (!dom_has_wall(dom)) {
    c.s = A; c.G2r = T(0); c.G3r = T(0); c.G4r = T(0);
    return c;
  }
  T invR = g.invR;
  T invR2 = invR * invR;
  T invR3 = invR2 * invR;
  T G1 = (one - T(6) * g.ez2) * invR3 * T(0.5);
  T G2 = T(-1.5) * invR3;
  T G3 = T(3) * invR3 * g.ez;
  T G4 = T(3) * invR3;
  c.s = A + G1;
  c.G2r = G2 * invR2;
  c.G3r = G3 * invR;
  c.G4r = G4 * invR2;
  return c;
}

template <typename T>
RMW_HD Vec3<T> apply_wt(const Geom<T>& g, const CoefWT<T>& c, T vx, T vy, T vz)
{
  T p = g.dx * vx + g.dy * vy;
  T dv = p + g.dz * vz;
  T Rv = p + g.Rz * vz;
  T zt = g.dx * vy - g.dy * vx;      // (z x R).v
  T q = c.Bc * dv;
  T ge = c.G2r * Rv;
  T gz = c.G4r * zt;
  Vec3<T> w;
  w.x = c.s * vx + (q + ge) * g.dx - gz * g.dy;
  w.y = c.s * vy + (q + ge) * g.dy + gz * g.dx;
  w.z = c.s * vz + q * g.dz + ge * g.Rz + c.G3r * Rv;
  return w;
}

// WT_ij^T v = WT_ji v: the (R.v) z term becomes v_z R.
template <typename T>
RMW_HD Vec3<T> apply_wt_t(const Geom<T>& g, const CoefWT<T>& c, T vx, T vy, T vz)
{
  T p = g.dx * vx + g.dy * vy;
  T dv = p + g.dz * vz;
  T Rv = p + g.Rz * vz;
  T zt = g.dx * vy - g.dy * vx;
  T q = c.Bc * dv;
  T ge = c.G2r * Rv + c.G3r * vz;
  T gz = c.G4r * zt;
  Vec3<T> w;
  w.x = c.s * vx + (q + ge) * g.dx - gz * g.dy;
  w.y = c.s * vy + (q + ge) * g.dy + gz * g.dx;
  w.z = c.s * vz + q * g.dz + ge * g.Rz;
  return w;
}

// ---------------------------------------------------------------------------
// Explicit 3x3 blocks, m[r][c] = e_r . B e_c, of the four dyadic forms above
// (read off apply_uf / apply_ut / apply_wf / apply_wt at the unit vectors).
// A kernel that applies one pair's blocks to several right-hand sides forms
// them once and then spends 9 FMAs per block, side and right-hand side
// (the transposed blocks are the same entries read by columns).
// ---------------------------------------------------------------------------
template <typename T> struct Mat3 { T m[3][3]; };

template <typename T>
RMW_HD Mat3<T> mat_uf(const Geom<T>& g, const CoefUF<T>& c)
{
  Mat3<T> M;
  const T P = c.Bc + c.g2;
  const T h = fma_(c.Bc, g.dz, c.g2 * g.Rz);
  const T a3 = h + c.g3, b3 = h + c.g4;
  const T Pdx = P * g.dx;
  M.m[0][0] = fma_(Pdx, g.dx, c.s);
  M.m[0][1] = Pdx * g.dy;
  M.m[0][2] = a3 * g.dx;
  M.m[1][0] = M.m[0][1];
  M.m[1][1] = fma_(P * g.dy, g.dy, c.s);
  M.m[1][2] = a3 * g.dy;
  M.m[2][0] = b3 * g.dx;
  M.m[2][1] = b3 * g.dy;
  M.m[2][2] = fma_(c.Bc * g.dz, g.dz, fma_(fma_(c.g2, g.Rz, c.g3 + c.g4), g.Rz, c.s5));
  return M;
}

// UT (coefficients of coef_rot(g, z_i)); m[2][2] = 0.
template <typename T>
RMW_HD Mat3<T> mat_ut(const Geom<T>& g, const CoefRot<T>& c)
{
  Mat3<T> M;
  const T G = fma_(c.c, g.dz, fma_(c.F1r, g.Rz, c.F4));
  const T K = c.c - c.F1r - fma_(c.F3r2, g.Rz, c.F2r);
  const T cf = c.c - c.F1r;
  const T Fx = c.F3r2 * g.dx;
  M.m[0][0] = Fx * g.dy;
  M.m[0][1] = fma_(-Fx, g.dx, G);
  M.m[0][2] = -cf * g.dy;
  M.m[1][0] = fma_(c.F3r2 * g.dy, g.dy, -G);
  M.m[1][1] = -M.m[0][0];
  M.m[1][2] = cf * g.dx;
  M.m[2][0] = K * g.dy;
  M.m[2][1] = -K * g.dx;
  M.m[2][2] = T(0);
  return M;
}

// WF (coefficients of coef_rot(g, z_j)); m[2][2] = 0.
template <typename T>
RMW_HD Mat3<T> mat_wf(const Geom<T>& g, const CoefRot<T>& c)
{
  Mat3<T> M;
  const T G = fma_(c.c, g.dz, -fma_(c.F1r, g.Rz, c.F4));
  const T K = c.c - c.F1r - fma_(c.F3r2, g.Rz, c.F2r);
  const T cf = c.c - c.F1r;
  const T Fx = c.F3r2 * g.dx;
  M.m[0][0] = Fx * g.dy;
  M.m[0][1] = fma_(c.F3r2 * g.dy, g.dy, G);
  M.m[0][2] = -K * g.dy;
  M.m[1][0] = -fma_(Fx, g.dx, G);
  M.m[1][1] = -M.m[0][0];
  M.m[1][2] = K * g.dx;
  M.m[2][0] = cf * g.dy;
  M.m[2][1] = -cf * g.dx;
  M.m[2][2] = T(0);
  return M;
}

template <typename T>
RMW_HD Mat3<T> mat_wt(const Geom<T>& g, const CoefWT<T>& c)
{
  Mat3<T> M;
  const T P = c.Bc + c.G2r;
  const T h = fma_(c.Bc, g.dz, c.G2r * g.Rz);
  const T h3 = h + c.G3r;
  const T dxy = g.dx * g.dy;
  M.m[0][0] = fma_(P * g.dx, g.dx, fma_(c.G4r * g.dy, g.dy, c.s));
  M.m[0][1] = (P - c.G4r) * dxy;
  M.m[0][2] = h * g.dx;
  M.m[1][0] = M.m[0][1];
  M.m[1][1] = fma_(P * g.dy, g.dy, fma_(c.G4r * g.dx, g.dx, c.s));
  M.m[1][2] = h * g.dy;
  M.m[2][0] = h3 * g.dx;
  M.m[2][1] = h3 * g.dy;
  M.m[2][2] = fma_(c.Bc * g.dz, g.dz, fma_(fma_(c.G2r, g.Rz, c.G3r), g.Rz, c.s));
  return M;
}

// ---------------------------------------------------------------------------
// Blobs of different radii (target radius a, source radius b), lengths in
// absolute units, normalisation 1/(8 pi eta) folded in by the caller.
// mobility_numba.py:1479-1658.  The self pair (d = 0) needs no special
// case: it falls in the inner RPY regime.
//
//   RPY:  C1 I + C2 d d^T with, for r = |d| and s = a^2 + b^2,
//         r > a + b:     C1 = (1 + s/(3r^2))/r,  C2 = (1 - s/r^2)/r^3
//         r > |b - a|:   C1 = 4/(3ab) [(a + b)/2 - ((b-a)^2 + 3r^2)^2/(32 r^3)]
//                        C2 = 4/(3ab) (3/32) ((b-a)^2 - r^2)^2 / r^5
//         otherwise:     C1 = 4/(3 max(a, b)),   C2 = 0
//   wall: -S D v with D = diag(1, 1, -1), R = (d_x, d_y, x3 + y3) (x3, y3 the
//         target and source heights), rho = |R| and
//         S = alpha I + beta R R^T + g1 R z^T + g2 z R^T + delta z z^T,
//   the image Stokeslet, its doublets weighted by ca = 2a^2/3, cb = 2b^2/3
//   and the quadrupole weighted by cab = 2a^2 b^2/3 collected per dyad.
//   free surface (mobility_numba.py:1941-2093): + (C1' I + C2' R R^T) D v,
//         the RPY coefficients of the three regimes evaluated at rho, i.e.
//         S = -(C1' I + C2' R R^T) in the wall's form.
// ---------------------------------------------------------------------------
enum : int { kRadiiUnbounded = 0, kRadiiWall = 1, kRadiiFreeSurface = 2 };

template <typename T> struct CoefRadii {
  T C1, C2;                 // RPY
  T al, be, g1, g2, de;     // wall (or image) dyads
};

// RPY coefficients C1, C2 of radii a, b at squared distance r2.
template <typename T>
RMW_HD void rpy_radii(T r2, T a, T b, T inva, T invb, T& C1, T& C2)
{
  const T invr = rsq(r2);
  const T r = r2 > T(0) ? r2 * invr : T(0);
  const T w = invr * invr;
  const T s = a * a + b * b;
  const T C1_far = fma_(s * w, T(1.0 / 3.0), T(1)) * invr;
  const T C2_far = fma_(-s, w, T(1)) * w * invr;
  const T dab = b - a;
  const T dab2 = dab * dab;
  const T k1 = T(4.0 / 3.0) * inva * invb;
  const T t = fma_(T(3), r2, dab2);
  const T invr3 = w * invr;
  const T C1_mid = k1 * (T(0.5) * (a + b) - T(1.0 / 32.0) * t * t * invr3);
  const T u = dab2 - r2;
  const T C2_mid = k1 * T(3.0 / 32.0) * u * u * invr3 * w;
  const T C1_in = T(4.0 / 3.0) * (inva < invb ? inva : invb);
  const bool far = r > a + b;
  const bool mid = r > (dab < T(0) ? -dab : dab);
  C1 = far ? C1_far : (mid ? C1_mid : C1_in);
  C2 = far ? C2_far : (mid ? C2_mid : T(0));
}

// bnd: kRadiiUnbounded, kRadiiWall or kRadiiFreeSurface.
template <typename T>
RMW_HD CoefRadii<T> coef_radii(T dx, T dy, T dz, T Rz, T x3, T y3, T a, T b, T inva, T invb, int bnd)
{
  CoefRadii<T> c;
  const T r2 = dx * dx + dy * dy + dz * dz;
  rpy_radii(r2, a, b, inva, invb, c.C1, c.C2);
  const T s = a * a + b * b;
  if (bnd == kRadiiWall) {
    const T rho2 = dx * dx + dy * dy + Rz * Rz;
    const T q = rsq(rho2);
    const T qw = q * q;
    const T q3 = q * qw, q5 = q3 * qw, q7 = q5 * qw, q9 = q7 * qw;
    const T ca = T(2.0 / 3.0) * a * a, cb = T(2.0 / 3.0) * b * b, cab = ca * b * b;
    const T Rz2 = Rz * Rz;
    const T xy = x3 * y3;
    const T m = ca * y3 + cb * x3;
    const T P = fma_(s * qw, T(1.0 / 3.0), T(1)) * q;
    const T Q = fma_(-s, qw, T(1)) * q3;
    c.al = P + T(2) * xy * q3 - T(3) * m * Rz * q5 + cab * (T(5) * Rz2 * q7 - q5);
    c.be = Q - T(6) * xy * q5 + T(15) * m * Rz * q7 + cab * (T(5) * q7 - T(35) * Rz2 * q9);
    const T e = T(3) * (ca - cb) * q5;
    const T z10 = T(10) * cab * Rz * q7;
    c.g1 = x3 * (e - T(2) * q3) + z10;
    c.g2 = y3 * (-e - T(2) * q3) + z10;
    c.de = -T(2) * q - (ca + cb) * q3 - T(2) * cab * q5;
  } else if (bnd == kRadiiFreeSurface) {
    T I1, I2;
    rpy_radii(dx * dx + dy * dy + Rz * Rz, a, b, inva, invb, I1, I2);
    c.al = -I1;
    c.be = -I2;
    c.g1 = c.g2 = c.de = T(0);
  } else {
    c.al = c.be = c.g1 = c.g2 = c.de = T(0);
  }
  return c;
}

template <typename T>
RMW_HD Vec3<T> apply_radii(T dx, T dy, T dz, T Rz, const CoefRadii<T>& c, T vx, T vy, T vz)
{
  const T p = dx * vx + dy * vy;
  const T dv = p + dz * vz;
  const T Rv = p - Rz * vz;               // R . D v
  const T h = c.be * Rv - c.g1 * vz;      // coefficient of R
  const T q = c.C2 * dv;
  Vec3<T> u;
  u.x = (c.C1 - c.al) * vx + (q - h) * dx;
  u.y = (c.C1 - c.al) * vy + (q - h) * dy;
  u.z = (c.C1 + c.al) * vz + q * dz - h * Rz - (c.g2 * Rv - c.de * vz);
  return u;
}

// ---------------------------------------------------------------------------
// Self blocks (i == j, central box), mobility_numba.py:202-208 + 245-252,
// :653-657, :1040-1044, :1291-1297.  h = z_i / a (effective height).
// ---------------------------------------------------------------------------
template <typename T>
RMW_HD Vec3<T> self_uf(T h, T vx, T vy, T vz)
{
  T ih = T(1) / h;
  T ih3 = ih * ih * ih;
  T ih5 = ih3 * ih * ih;
  T pxy = T(4) / T(3) - (T(9) * ih - T(2) * ih3 + ih5) / T(12);
  T pz = T(4) / T(3) - (T(9) * ih - T(4) * ih3 + ih5) / T(6);
  return Vec3<T>{pxy * vx, pxy * vy, pz * vz};
}

// UT self: M_xy = +h^-4/8, M_yx = -h^-4/8.
template <typename T>
RMW_HD Vec3<T> self_ut(T h, T vx, T vy, T vz)
{
  T ih = T(1) / h;
  T c = T(0.125) * ih * ih * ih * ih;
  return Vec3<T>{c * vy, -c * vx, T(0)};
}

// WF self: M_xy = -h^-4/8, M_yx = +h^-4/8.
template <typename T>
RMW_HD Vec3<T> self_wf(T h, T vx, T vy, T vz)
{
  T ih = T(1) / h;
  T c = T(0.125) * ih * ih * ih * ih;
  return Vec3<T>{-c * vy, c * vx, T(0)};
}

template <typename T>
RMW_HD Vec3<T> self_wt(T h, T vx, T vy, T vz)
{
  T ih = T(1) / h;
  T ih3 = ih * ih * ih;
  T pxy = T(1) - T(0.3125) * ih3;
  T pz = T(1) - T(0.125) * ih3;
  return Vec3<T>{pxy * vx, pxy * vy, pz * vz};
}

// Self terms of every domain (same units as the wall ones): no wall, the
// bare RPY self mobility (4/3 I, 0, 0, I); free surface, 4/3 I plus the
// image RPY(R) D at R = (0, 0, 2h); in plane, the wall terms (the caller
// projects).
template <typename T>
RMW_HD Vec3<T> self_uf_dom(int dom, T h, T vx, T vy, T vz)
{
  const T s = T(4) / T(3);
  if (dom == kDomNoWall) return Vec3<T>{s * vx, s * vy, s * vz};
  if (dom == kDomFreeSurface) {
    const T R = h + h;
    T AR, BR;
    rpy_tt(R, T(1) / R, AR, BR);
    return Vec3<T>{(s + AR) * vx, (s + AR) * vy, (s - AR - R * R * BR) * vz};
  }
  return self_uf(h, vx, vy, vz);
}
template <typename T>
RMW_HD Vec3<T> self_ut_dom(int dom, T h, T vx, T vy, T vz)
{
  return dom_has_wall(dom) ? self_ut(h, vx, vy, vz) : Vec3<T>{T(0), T(0), T(0)};
}
template <typename T>
RMW_HD Vec3<T> self_wf_dom(int dom, T h, T vx, T vy, T vz)
{
  return dom_has_wall(dom) ? self_wf(h, vx, vy, vz) : Vec3<T>{T(0), T(0), T(0)};
}
template <typename T>
RMW_HD Vec3<T> self_wt_dom(int dom, T h, T vx, T vy, T vz)
{
  return dom_has_wall(dom) ? self_wt(h, vx, vy, vz) : Vec3<T>{vx, vy, vz};
}

}  // namespace rmw
