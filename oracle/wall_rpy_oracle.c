/*
 * FP64 CPU oracle for the wall-corrected RPY blob mobility products.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only the test suite,
 * the smoke check and the CPU-baseline leg of the benchmark load it.  The
 * product path (the HIP library) never calls or links it.
 *
 * It restates, in the same expression order, the four single-wall numba
 * kernels of the reference plus the unbounded RPY kernel:
 *
 *   UF  mobility/mobility_numba.py:123-287   single_wall_mobility_trans_times_force_numba
 *   UT  mobility/mobility_numba.py:547-686   single_wall_mobility_trans_times_torque_numba
 *   WF  mobility/mobility_numba.py:937-1073  single_wall_mobility_rot_times_force_numba
 *   WT  mobility/mobility_numba.py:1188-1328 single_wall_mobility_rot_times_torque_numba
 *   no-wall UF  mobility/mobility_numba.py:12-120
 *
 * and the Python wrappers that apply the wall-overlap regularisation
 * u = B M(z_eff) B f (mobility/mobility.py:52-84 shift_heights and
 * damping_matrix_B; :1132-1163, :1213-1232, :1273-1298, :1316-1342).
 *
 * Pseudo-periodic images follow mobility_numba.py:149-198: for L_d > 0 the
 * displacement is projected to the minimum image with truncation toward zero
 * and the 3 neighbouring boxes per periodic direction are summed; the self
 * term exists only in the central box.
 *
 * All loops are parallel over targets (OpenMP), serial over boxes and
 * sources, like the numba prange.  Every kernel takes a target range
 * [t0, t1) so that a bounded sample of rows can be timed.
 *
 * Built with -DORACLE_MUTANT it becomes a deliberately breakable copy: the
 * global `orc_mutation` selects one planted mistake (dropped term, wrong
 * sign, transposed index) so the tests can show that the pins catch it.
 */
#include <math.h>
#include <stdint.h>
#include <string.h>

#ifdef ORACLE_MUTANT
int orc_mutation = 0;
#define MUT(id, good, bad) ((orc_mutation == (id)) ? (bad) : (good))
#else
#define MUT(id, good, bad) (good)
#endif

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Periodic projection of one displacement component (mobility_numba.py:184-198). */
static inline double project_image(double d, double L, int box)
{
  if (L > 0.0) {
    double s = (double)((d > 0.0) - (d < 0.0));
    long shift = (long)(d / L + 0.5 * s);  /* int() truncates toward zero */
    d = d - (double)shift * L;
    d = d + (double)box * L;
  }
  return d;
}

typedef struct { double xx, xy, xz, yx, yy, yz, zx, zy, zz; } mat3;

/* ---------------------------------------------------------------------- */
/* Translation-translation (UF), mobility_numba.py:200-276                  */
/* (rx, ry, rz) in units of a; zi, zj the effective heights in units of a.  */
/* ---------------------------------------------------------------------- */
static inline mat3 uf_pair(double rx, double ry, double rz, double zi, double zj, int self)
{
  const double fourOverThree = 4.0 / 3.0;
  mat3 M;
  if (self) {
    M.xx = fourOverThree; M.xy = 0; M.xz = 0;
    M.yy = M.xx; M.yz = 0; M.zz = M.xx;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double invr = 1.0 / r;
    double invr2 = invr * invr;
    if (r > 2) {
      double c1 = 1.0 + 2.0 / (3.0 * r2);
      double c2 = (1.0 - MUT(1, 2.0, 3.0) * invr2) * invr2;
      M.xx = (c1 + c2 * rx * rx) * invr;
      M.xy = (c2 * rx * ry) * invr;
      M.xz = (c2 * rx * rz) * invr;
      M.yy = (c1 + c2 * ry * ry) * invr;
      M.yz = (c2 * ry * rz) * invr;
      M.zz = (c1 + c2 * rz * rz) * invr;
    } else {
      double c1 = fourOverThree * (1.0 - 0.28125 * r);
      double c2 = fourOverThree * 0.09375 * invr;
      M.xx = c1 + c2 * rx * rx;
      M.xy = c2 * rx * ry;
      M.xz = c2 * rx * rz;
      M.yy = c1 + c2 * ry * ry;
      M.yz = c2 * ry * rz;
      M.zz = c1 + c2 * rz * rz;
    }
  }
  M.yx = M.xy; M.zx = M.xz; M.zy = M.yz;

  /* Wall correction: image vector R = (rx, ry, zi + zj), h_j = zj. */
  double Rz = zi + zj;
  double hj = zj;
  if (self) {
    double invZi = 1.0 / hj;
    double invZi3 = invZi * invZi * invZi;
    double invZi5 = invZi3 * invZi * invZi;
    M.xx += -(9.0 * invZi - 2.0 * invZi3 + invZi5) / 12.0;
    M.yy += -(9.0 * invZi - 2.0 * invZi3 + invZi5) / 12.0;
    M.zz += -(9.0 * invZi - MUT(2, 4.0, 2.0) * invZi3 + invZi5) / 6.0;
  } else {
    double h_hat = hj / Rz;
    double invR = 1.0 / sqrt(rx * rx + ry * ry + Rz * Rz);
    double ex = rx * invR;
    double ey = ry * invR;
    double ez = Rz * invR;
    double invR3 = invR * invR * invR;
    double invR5 = invR3 * invR * invR;
    double ez2 = ez * ez;

    double fact1 = -(3.0 * (1.0 + 2.0 * h_hat * (1.0 - h_hat) * ez2) * invR
                     + 2.0 * (1.0 - 3.0 * ez2) * invR3
                     - 2.0 * (1.0 - 5.0 * ez2) * invR5) / 3.0;
    double fact2 = -(3.0 * (1.0 - 6.0 * h_hat * (1.0 - h_hat) * ez2) * invR
                     - 6.0 * (1.0 - 5.0 * ez2) * invR3
                     + 10.0 * (1.0 - 7.0 * ez2) * invR5) / 3.0;
    double fact3 = ez * (3.0 * h_hat * (1.0 - 6.0 * (1.0 - h_hat) * ez2) * invR
                         - 6.0 * (1.0 - 5.0 * ez2) * invR3
                         + 10.0 * (2.0 - 7.0 * ez2) * invR5) * 2.0 / 3.0;
    double fact4 = ez * (3.0 * h_hat * invR - 10.0 * invR5) * 2.0 / 3.0;
    double fact5 = -(3.0 * h_hat * h_hat * ez2 * invR + 3.0 * ez2 * invR3
                     + MUT(3, (2.0 - 15.0 * ez2), 0.0) * invR5) * 4.0 / 3.0;

    M.xx += fact1 + fact2 * ex * ex;
    M.xy += fact2 * ex * ey;
    M.xz += fact2 * ex * ez + fact3 * ex;
    M.yx += fact2 * ey * ex;
    M.yy += fact1 + fact2 * ey * ey;
    M.yz += fact2 * ey * ez + fact3 * ey;
    M.zx += fact2 * ez * ex + MUT(4, fact4 * ex, fact3 * ex);
    M.zy += fact2 * ez * ey + fact4 * ey;
    M.zz += fact1 + fact2 * ez * ez + fact3 * ez + fact4 * ez + fact5;
  }
  return M;
}

/* ---------------------------------------------------------------------- */
/* Translation-rotation (UT), mobility_numba.py:614-679                     */
/* Wall term in the transposed source/target role: R = (-rx,-ry, zi+zj),   */
/* h = zi.                                                                  */
/* ---------------------------------------------------------------------- */
static inline mat3 ut_pair(double rx, double ry, double rz, double zi, double zj, int self)
{
  mat3 M;
  if (self) {
    M.xx = 0; M.xy = 0; M.xz = 0; M.yy = 0; M.yz = 0; M.zz = 0;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double r3 = r2 * r;
    double invr3 = 1.0 / r3;
    if (r >= 2) {
      M.xx = 0;
      M.xy = rz * invr3;
      M.xz = -ry * invr3;
      M.yy = 0;
      M.yz = rx * invr3;
      M.zz = 0;
    } else {
      double c1 = 0.5 * (1.0 - 0.375 * r);
      M.xx = 0;
      M.xy = c1 * rz;
      M.xz = -c1 * ry;
      M.yy = 0;
      M.yz = c1 * rx;
      M.zz = 0;
    }
  }
  M.yx = -M.xy; M.zx = -M.xz; M.zy = -M.yz;

  double Rx = -rx, Ry = -ry;
  double Rz = zi + zj;
  double hj = zi;
  if (self) {
    double invZi = 1.0 / hj;
    double invZi4 = invZi * invZi * invZi * invZi;
    M.xy -= -invZi4 * 0.125;
    M.yx -= invZi4 * 0.125;
  } else {
    double h_hat = hj / Rz;
    double invR = 1.0 / sqrt(Rx * Rx + Ry * Ry + Rz * Rz);
    double invR2 = invR * invR;
    double invR4 = invR2 * invR2;
    double ex = Rx * invR;
    double ey = Ry * invR;
    double ez = Rz * invR;
    double fact1 = invR2;
    double fact2 = (6.0 * h_hat * ez * ez * invR2 + (1.0 - 10.0 * ez * ez) * invR4) * 2.0;
    double fact3 = -ez * (3.0 * h_hat * invR2 - 5.0 * invR4) * 2.0;
    double fact4 = -ez * (h_hat * invR2 - invR4) * 2.0;

    M.xx -= -fact3 * ex * ey;
    M.xy -= -fact1 * ez + fact3 * ex * ex - fact4;
    M.xz -= fact1 * ey;
    M.yx -= fact1 * ez - fact3 * ey * ey + fact4;
    M.yy -= fact3 * ex * ey;
    M.yz -= -fact1 * ex;
    M.zx -= -fact1 * ey - fact2 * ey - fact3 * ey * ez;
    M.zy -= fact1 * ex + fact2 * ex + fact3 * ex * ez;
  }
  return M;
}

/* ---------------------------------------------------------------------- */
/* Rotation-translation (WF), mobility_numba.py:1004-1065                   */
/* Wall term: R = (rx, ry, zi+zj), h = zj.                                  */
/* ---------------------------------------------------------------------- */
static inline mat3 wf_pair(double rx, double ry, double rz, double zi, double zj, int self)
{
  mat3 M;
  if (self) {
    M.xx = 0; M.xy = 0; M.xz = 0; M.yy = 0; M.yz = 0; M.zz = 0;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double r3 = r2 * r;
    double invr3 = 1.0 / r3;
    if (r >= 2.0) {
      M.xx = 0;
      M.xy = rz * invr3;
      M.xz = -ry * invr3;
      M.yy = 0;
      M.yz = rx * invr3;
      M.zz = 0;
    } else {
      double c1 = 0.5 * (1.0 - 0.375 * r);
      M.xx = 0;
      M.xy = c1 * rz;
      M.xz = -c1 * ry;
      M.yy = 0;
      M.yz = c1 * rx;
      M.zz = 0;
    }
  }
  M.yx = -M.xy; M.zx = -M.xz; M.zy = -M.yz;

  double Rz = zi + zj;
  double hj = zj;
  if (self) {
    double invZi = 1.0 / hj;
    double invZi4 = invZi * invZi * invZi * invZi;
    M.xy += -invZi4 * 0.125;
    M.yx += invZi4 * 0.125;
  } else {
    double h_hat = hj / Rz;
    double invR = 1.0 / sqrt(rx * rx + ry * ry + Rz * Rz);
    double invR2 = invR * invR;
    double invR4 = invR2 * invR2;
    double ex = rx * invR;
    double ey = ry * invR;
    double ez = Rz * invR;
    double fact1 = invR2;
    double fact2 = (6.0 * h_hat * ez * ez * invR2 + (1.0 - 10.0 * ez * ez) * invR4) * 2.0;
    double fact3 = -ez * (3.0 * h_hat * invR2 - 5.0 * invR4) * 2.0;
    double fact4 = -ez * (h_hat * invR2 - invR4) * 2.0;

    M.xx -= -fact3 * ex * ey;
    M.xy -= fact1 * ez - fact3 * ey * ey + fact4;
    M.xz -= -fact1 * ey - fact2 * ey - fact3 * ey * ez;
    M.yx -= -fact1 * ez + fact3 * ex * ex - fact4;
    M.yy -= fact3 * ex * ey;
    M.yz -= fact1 * ex + fact2 * ex + fact3 * ex * ez;
    M.zx -= fact1 * ey;
    M.zy -= MUT(5, -fact1 * ex, fact1 * ex);
  }
  return M;
}

/* ---------------------------------------------------------------------- */
/* Rotation-rotation (WT), mobility_numba.py:1256-1321                      */
/* Wall term: R = (rx, ry, zi+zj), h = zj.                                  */
/* ---------------------------------------------------------------------- */
static inline mat3 wt_pair(double rx, double ry, double rz, double zi, double zj, int self)
{
  mat3 M;
  if (self) {
    M.xx = 1.0; M.xy = 0; M.xz = 0; M.yy = M.xx; M.yz = 0; M.zz = M.xx;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double r3 = r2 * r;
    double invr = 1.0 / r;
    double invr2 = 1.0 / r2;
    double invr3 = 1.0 / r3;
    if (r >= 2) {
      double c1 = -0.5;
      double c2 = 1.5 * invr2;
      M.xx = (c1 + c2 * rx * rx) * invr3;
      M.xy = (c2 * rx * ry) * invr3;
      M.xz = (c2 * rx * rz) * invr3;
      M.yy = (c1 + c2 * ry * ry) * invr3;
      M.yz = (c2 * ry * rz) * invr3;
      M.zz = (c1 + c2 * rz * rz) * invr3;
    } else {
      double c1 = (1.0 - 0.84375 * r + 0.078125 * r3);
      double c2 = 0.28125 * invr - 0.046875 * r;
      M.xx = c1 + c2 * rx * rx;
      M.xy = c2 * rx * ry;
      M.xz = c2 * rx * rz;
      M.yy = c1 + c2 * ry * ry;
      M.yz = c2 * ry * rz;
      M.zz = c1 + c2 * rz * rz;
    }
  }
  M.yx = M.xy; M.zx = M.xz; M.zy = M.yz;

  double Rz = zi + zj;
  double hj = zj;
  if (self) {
    double invZi = 1.0 / hj;
    double invZi3 = invZi * invZi * invZi;
    M.xx += -invZi3 * 0.3125;
    M.yy += -invZi3 * 0.3125;
    M.zz += -invZi3 * 0.125;
  } else {
    double invR = 1.0 / sqrt(rx * rx + ry * ry + Rz * Rz);
    double invR3 = invR * invR * invR;
    double ex = rx * invR;
    double ey = ry * invR;
    double ez = Rz * invR;
    double fact1 = ((1.0 - 6.0 * ez * ez) * invR3) * 0.5;
    double fact2 = -(9.0 * invR3) / 6.0;
    double fact3 = (3.0 * invR3 * ez);
    double fact4 = (3.0 * invR3);

    M.xx += fact1 + fact2 * ex * ex + fact4 * ey * ey;
    M.xy += (fact2 - fact4) * ex * ey;
    M.xz += fact2 * ex * ez;
    M.yx += (fact2 - fact4) * ex * ey;
    M.yy += fact1 + fact2 * ey * ey + fact4 * ex * ex;
    M.yz += fact2 * ey * ez;
    M.zx += fact2 * ez * ex + fact3 * ex;
    M.zy += fact2 * ez * ey + fact3 * ey;
    M.zz += fact1 + fact2 * ez * ez + MUT(6, fact3 * ez, 0.0);
  }
  return M;
}

/* Unbounded RPY (mobility_numba.py:65-104). */
static inline mat3 uf_nowall_pair(double rx, double ry, double rz, int self)
{
  const double fourOverThree = 4.0 / 3.0;
  mat3 M;
  if (self) {
    M.xx = fourOverThree; M.xy = 0; M.xz = 0; M.yy = M.xx; M.yz = 0; M.zz = M.xx;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double invr = 1.0 / r;
    double invr2 = invr * invr;
    if (r > 2) {
      double c1 = 1.0 + 2.0 / (3.0 * r2);
      double c2 = (1.0 - 2.0 * invr2) * invr2;
      M.xx = (c1 + c2 * rx * rx) * invr;
      M.xy = (c2 * rx * ry) * invr;
      M.xz = (c2 * rx * rz) * invr;
      M.yy = (c1 + c2 * ry * ry) * invr;
      M.yz = (c2 * ry * rz) * invr;
      M.zz = (c1 + c2 * rz * rz) * invr;
    } else {
      double c1 = fourOverThree * (1.0 - 0.28125 * r);
      double c2 = fourOverThree * 0.09375 * invr;
      M.xx = c1 + c2 * rx * rx;
      M.xy = c2 * rx * ry;
      M.xz = c2 * rx * rz;
      M.yy = c1 + c2 * ry * ry;
      M.yz = c2 * ry * rz;
      M.zz = c1 + c2 * rz * rz;
    }
  }
  M.yx = M.xy; M.zx = M.xz; M.zy = M.yz;
  return M;
}

/* Unbounded rotational coupling, the same matrix for UT (mobility_numba.py
 * :500-535) and WF (:893-929): M_xy = rz c, M_xz = -ry c, M_yz = rx c,
 * antisymmetric, c = 1/r^3 (r >= 2) or (1 - 3r/8)/2; zero for the self pair. */
static inline mat3 rot_nowall_pair(double rx, double ry, double rz, int self)
{
  mat3 M = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (!self) {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double c = r >= 2 ? 1.0 / (r2 * r) : 0.5 * (1.0 - 0.375 * r);
    M.xy = rz * c;
    M.xz = -ry * c;
    M.yz = rx * c;
  }
  M.yx = -M.xy; M.zx = -M.xz; M.zy = -M.yz;
  return M;
}

/* Unbounded rotation from torque (mobility_numba.py:1141-1175). */
static inline mat3 wt_nowall_pair(double rx, double ry, double rz, int self)
{
  mat3 M;
  if (self) {
    M.xx = 1.0; M.xy = 0; M.xz = 0; M.yy = 1.0; M.yz = 0; M.zz = 1.0;
  } else {
    double r2 = rx * rx + ry * ry + rz * rz;
    double r = sqrt(r2);
    double r3 = r2 * r;
    double invr = 1.0 / r, invr2 = 1.0 / r2, invr3 = 1.0 / r3;
    if (r >= 2) {
      double c1 = -0.5, c2 = 1.5 * invr2;
      M.xx = (c1 + c2 * rx * rx) * invr3;
      M.xy = (c2 * rx * ry) * invr3;
      M.xz = (c2 * rx * rz) * invr3;
      M.yy = (c1 + c2 * ry * ry) * invr3;
      M.yz = (c2 * ry * rz) * invr3;
      M.zz = (c1 + c2 * rz * rz) * invr3;
    } else {
      double c1 = 1.0 - 0.84375 * r + 0.078125 * r3;
      double c2 = 0.28125 * invr - 0.046875 * r;
      M.xx = c1 + c2 * rx * rx;
      M.xy = c2 * rx * ry;
      M.xz = c2 * rx * rz;
      M.yy = c1 + c2 * ry * ry;
      M.yz = c2 * ry * rz;
      M.zz = c1 + c2 * rz * rz;
    }
  }
  M.yx = M.xy; M.zx = M.xz; M.zy = M.yz;
  return M;
}

/* In-plane blocks (mobility_numba.py:291-437, 690-830): the kernels differ
 * from the single-wall ones only by dropping every entry of the z row and
 * the z column (and the z components of the products), i.e. P M P. */
static inline mat3 in_plane(mat3 M)
{
  M.xz = 0; M.yz = 0; M.zx = 0; M.zy = 0; M.zz = 0;
  return M;
}

/* Free surface (mobility_numba.py:1840-1925): RPY of d (4/3 I for the self
 * pair) plus the RPY of the image vector R = (rx, ry, (z_i + z_j)/a), added
 * with the signs of M D, D = diag(1, 1, -1) (z column negated). */
static inline mat3 uf_free_surface_pair(double rx, double ry, double rz, double zi, double zj, int self)
{
  mat3 M = uf_nowall_pair(rx, ry, rz, self);
  if (self) { rx = 0; ry = 0; }
  mat3 I = uf_nowall_pair(rx, ry, zi + zj, 0);
  M.xx += I.xx; M.xy += I.xy; M.xz -= I.xz;
  M.yx += I.xy; M.yy += I.yy; M.yz -= I.yz;
  M.zx += I.xz; M.zy += I.yz; M.zz -= MUT(8, I.zz, -I.zz);
  return M;
}

/* Block ids used by the exported functions. */
enum {
  ORC_UF = 0, ORC_UT = 1, ORC_WF = 2, ORC_WT = 3, ORC_UF_NOWALL = 4,
  ORC_UT_NOWALL = 5, ORC_WF_NOWALL = 6, ORC_WT_NOWALL = 7,
  ORC_UF_INPLANE = 8, ORC_UT_INPLANE = 9, ORC_UF_FREE_SURFACE = 10,
  ORC_NBLOCKS = 11
};

/* Norm exponent p of 1/(8 pi eta a^p) per block (mobility_numba.py:133,559,949,1200). */
static int block_power(int block)
{
  switch (block) {
    case ORC_UT: case ORC_WF: case ORC_UT_NOWALL: case ORC_WF_NOWALL: case ORC_UT_INPLANE: return 2;
    case ORC_WT: case ORC_WT_NOWALL: return 3;
    default: return 1;
  }
}

/* Blocks whose wrapper shifts heights and damps (a wall bounds the fluid). */
static int block_has_wall(int block)
{
  return block <= ORC_WT || block == ORC_UF_INPLANE || block == ORC_UT_INPLANE;
}

static inline mat3 block_pair(int block, double rx, double ry, double rz, double zi, double zj, int self)
{
  switch (block) {
    case ORC_UF: return uf_pair(rx, ry, rz, zi, zj, self);
    case ORC_UT: return ut_pair(rx, ry, rz, zi, zj, self);
    case ORC_WF: return wf_pair(rx, ry, rz, zi, zj, self);
    case ORC_WT: return wt_pair(rx, ry, rz, zi, zj, self);
    case ORC_UT_NOWALL: case ORC_WF_NOWALL: return rot_nowall_pair(rx, ry, rz, self);
    case ORC_WT_NOWALL: return wt_nowall_pair(rx, ry, rz, self);
    case ORC_UF_INPLANE: return in_plane(uf_pair(rx, ry, rz, zi, zj, self));
    case ORC_UT_INPLANE: return in_plane(ut_pair(rx, ry, rz, zi, zj, self));
    case ORC_UF_FREE_SURFACE: return uf_free_surface_pair(rx, ry, rz, zi, zj, self);
    default: return uf_nowall_pair(rx, ry, rz, self);
  }
}

/*
 * Kernel-level product (the numba function itself): out = M(r) * vec for
 * targets t0 <= i < t1, r given with effective heights already applied.
 * r, vec: (n, 3) row-major; out: (n, 3) row-major, rows outside [t0,t1)
 * are not touched.
 */
int orc_block_kernel(int block, const double *r, const double *vec, double *out,
                     int64_t n, int64_t t0, int64_t t1, double eta, double a,
                     const double *L)
{
  if (block < 0 || block >= ORC_NBLOCKS || n < 0 || t0 < 0 || t1 > n || t0 > t1) return -1;
  const double inva = 1.0 / a;
  const int p = block_power(block);
  const double norm = 1.0 / (8.0 * M_PI * eta * pow(a, (double)p));
  const double Lx = L ? L[0] : 0.0, Ly = L ? L[1] : 0.0, Lz = L ? L[2] : 0.0;
  const int px = Lx > 0, py = Ly > 0, pz = Lz > 0;

#pragma omp parallel for schedule(dynamic, 16)
  for (int64_t i = t0; i < t1; ++i) {
    const double rxi = r[3 * i], ryi = r[3 * i + 1], rzi = r[3 * i + 2];
    double ux = 0.0, uy = 0.0, uz = 0.0;
    for (int bx = -px; bx <= px; ++bx)
      for (int by = -py; by <= py; ++by)
        for (int bz = -pz; bz <= pz; ++bz)
          for (int64_t j = 0; j < n; ++j) {
            double rx = project_image(rxi - r[3 * j], Lx, bx);
            double ry = project_image(ryi - r[3 * j + 1], Ly, by);
            double rz = project_image(rzi - r[3 * j + 2], Lz, bz);
            int central = (bx == 0 && by == 0 && bz == 0);
            int self = central && (MUT(7, i == j, i == j + 1));
            rx *= inva; ry *= inva; rz *= inva;
            double zi = rzi * inva, zj = r[3 * j + 2] * inva;
            mat3 M = block_pair(block, rx, ry, rz, zi, zj, self);
            const double fx = vec[3 * j], fy = vec[3 * j + 1], fz = vec[3 * j + 2];
            ux += (M.xx * fx + M.xy * fy + M.xz * fz) * norm;
            uy += (M.yx * fx + M.yy * fy + M.yz * fz) * norm;
            uz += (M.zx * fx + M.zy * fy + M.zz * fz) * norm;
          }
    out[3 * i] = ux;
    out[3 * i + 1] = uy;
    out[3 * i + 2] = uz;
  }
  return 0;
}

/*
 * shift_heights (mobility.py:52-64): z_eff = a where z <= a.
 * damping_matrix_B (mobility.py:67-84): B_i = z_i / a where z_i < a, else 1
 * (no clamp at z < 0).  Returns 1 when any blob overlaps the wall.
 */
int orc_shift_and_damping(const double *r, double *r_eff, double *B, int64_t n, double a)
{
  int overlap = 0;
  for (int64_t i = 0; i < n; ++i) {
    r_eff[3 * i] = r[3 * i];
    r_eff[3 * i + 1] = r[3 * i + 1];
    r_eff[3 * i + 2] = (r[3 * i + 2] <= a) ? a : r[3 * i + 2];
    if (r[3 * i + 2] < a) {
      B[i] = r[3 * i + 2] / a;
      overlap = 1;
    } else {
      B[i] = 1.0;
    }
  }
  return overlap;
}

/*
 * Wrapper-level product (mobility.py:1132-1163 pattern for every block):
 * out = B M(r_eff) B vec.  `work` must hold 4*n doubles (r_eff + B).
 * The no-wall block applies no shift or damping (mobility.py:1120-1130).
 */
int orc_block_apply(int block, const double *r, const double *vec, double *out,
                    int64_t n, int64_t t0, int64_t t1, double eta, double a,
                    const double *L, double *work, double *vec_work)
{
  if (!block_has_wall(block))
    return orc_block_kernel(block, r, vec, out, n, t0, t1, eta, a, L);
  double *r_eff = work;
  double *B = work + 3 * n;
  int overlap = orc_shift_and_damping(r, r_eff, B, n, a);
  const double *v = vec;
  if (overlap) {
    for (int64_t j = 0; j < n; ++j)
      for (int c = 0; c < 3; ++c) vec_work[3 * j + c] = B[j] * vec[3 * j + c];
    v = vec_work;
  }
  int rc = orc_block_kernel(block, r_eff, v, out, n, t0, t1, eta, a, L);
  if (rc) return rc;
  if (overlap)
    for (int64_t i = t0; i < t1; ++i)
      for (int c = 0; c < 3; ++c) out[3 * i + c] *= B[i];
  return 0;
}

/*
 * Dense 3x3 block M_ij (kernel level, effective heights, no periodicity),
 * already scaled by the block normalisation.  Used by the tests to build
 * small dense matrices for symmetry / SPD / transpose checks.
 */
int orc_block_pair_matrix(int block, const double *ri, const double *rj, int self,
                          double eta, double a, double *m9)
{
  const double inva = 1.0 / a;
  double rx = (ri[0] - rj[0]) * inva, ry = (ri[1] - rj[1]) * inva, rz = (ri[2] - rj[2]) * inva;
  double zi = ri[2] * inva, zj = rj[2] * inva;
  if (block < 0 || block >= ORC_NBLOCKS) return -1;
  mat3 M = block_pair(block, rx, ry, rz, zi, zj, self);
  const double norm = 1.0 / (8.0 * M_PI * eta * pow(a, (double)block_power(block)));
  double v[9] = {M.xx, M.xy, M.xz, M.yx, M.yy, M.yz, M.zx, M.zy, M.zz};
  for (int k = 0; k < 9; ++k) m9[k] = v[k] * norm;
  return 0;
}

#ifdef ORACLE_MUTANT
void orc_set_mutation(int m) { orc_mutation = m; }
#endif

int orc_num_threads(void)
{
#ifdef _OPENMP
  extern int omp_get_max_threads(void);
  return omp_get_max_threads();
#else
  return 1;
#endif
}

/* Threads of the following calls (timing of the CPU baseline). */
void orc_set_num_threads(int n)
{
#ifdef _OPENMP
  extern void omp_set_num_threads(int);
  omp_set_num_threads(n > 0 ? n : 1);
#else
  (void)n;
#endif
}

/* ---------------------------------------------------------------------- */
/* Blobs of different radii, source -> target:                             */
/* mobility_numba.py:1479-1658 (mobility_trans_times_force_source_target_  */
/* numba) in its expression order; lengths are NOT scaled by a radius, the */
/* normalisation is 1/(8 pi eta).  a = target radius, b = source radius;   */
/* x3, y3 = target / source heights.  wall = 0: the RPY part only;         */
/* wall = 1: plus the single-wall image; wall = 2: plus the free-surface   */
/* image of mobility_numba.py:1941-2093 (the RPY of R = (rx, ry, x3 + y3)  */
/* with its z column negated).                                             */
/* ---------------------------------------------------------------------- */
static inline mat3 radii_rpy(double rx, double ry, double rz, double a, double b)
{
  const double fourOverThree = 4.0 / 3.0;
  mat3 M;
  double r2 = rx * rx + ry * ry + rz * rz;
  double r = sqrt(r2);
  double C1, C2;
  if (r > (a + b)) {
    double a2 = a * a, b2 = b * b;
    C1 = (1 + (b2 + a2) / (3 * r2)) / r;
    C2 = ((1 - (b2 + a2) / r2) / r2) / r;
  } else if (r > fabs(b - a)) {
    double r3 = r2 * r;
    C1 = ((16 * (b + a) * r3 - pow(pow(b - a, 2) + 3 * r2, 2)) / (32 * r3)) * fourOverThree / (b * a);
    C2 = ((3 * pow(pow(b - a, 2) - r2, 2) / (32 * r3)) / r2) * fourOverThree / (b * a);
  } else {
    double largest_radius = a > b ? a : b;
    C1 = fourOverThree / largest_radius;
    C2 = 0;
  }
  M.xx = C1 + C2 * rx * rx;
  M.xy = C2 * rx * ry;
  M.xz = C2 * rx * rz;
  M.yy = C1 + C2 * ry * ry;
  M.yz = C2 * ry * rz;
  M.zz = C1 + C2 * rz * rz;
  M.yx = M.xy;
  M.zx = M.xz;
  M.zy = M.yz;
  return M;
}

static inline mat3 radii_pair(double rx, double ry, double rz, double x3, double y3, double a, double b, int wall)
{
  mat3 M = radii_rpy(rx, ry, rz, a, b);
  if (wall == 2) {
    mat3 I = radii_rpy(rx, ry, x3 + y3, a, b);
    M.xx = M.xx + I.xx;
    M.xy = M.xy + I.xy;
    M.xz = M.xz - I.xz;
    M.yx = M.yx + I.xy;
    M.yy = M.yy + I.yy;
    M.yz = M.yz - I.yz;
    M.zx = M.zx + I.xz;
    M.zy = M.zy + I.yz;
    M.zz = M.zz - MUT(9, I.zz, -I.zz);
  } else if (wall) {
    double rz = x3 + y3;
    double r2, r;
    r2 = rx * rx + ry * ry + rz * rz;
    r = sqrt(r2);
    double a2 = a * a, b2 = b * b;
    double r3 = r2 * r, r5 = r3 * r2, r7 = r5 * r2, r9 = r7 * r2;

    M.xx -= ((1 + (b2 + a2) / (3.0 * r2)) + (1 - (b2 + a2) / r2) * rx * rx / r2) / r;
    M.xy -= ((1 - (b2 + a2) / r2) * rx * ry / r2) / r;
    M.xz += ((1 - (b2 + a2) / r2) * rx * rz / r2) / r;
    M.yx -= ((1 - (b2 + a2) / r2) * ry * rx / r2) / r;
    M.yy -= ((1 + (b2 + a2) / (3.0 * r2)) + (1 - (b2 + a2) / r2) * ry * ry / r2) / r;
    M.yz += ((1 - (b2 + a2) / r2) * ry * rz / r2) / r;
    M.zx -= ((1 - (b2 + a2) / r2) * rz * rx / r2) / r;
    M.zy -= ((1 - (b2 + a2) / r2) * rz * ry / r2) / r;
    M.zz += ((1 + (b2 + a2) / (3.0 * r2)) + (1 - (b2 + a2) / r2) * rz * rz / r2) / r;

    M.xx -= 2 * (x3 * y3 * (1.0 / r3 - 3 * rx * rx / r5));
    M.xy -= 2 * (x3 * y3 * (-3 * rx * ry / r5));
    M.xz += 2 * (-rx * x3 / r3 + x3 * y3 * (-3 * rx * rz / r5));
    M.yx -= 2 * (x3 * y3 * (-3 * ry * rx / r5));
    M.yy -= 2 * (x3 * y3 * (1.0 / r3 - 3 * ry * ry / r5));
    M.yz += 2 * (-ry * x3 / r3 + x3 * y3 * (-3 * ry * rz / r5));
    M.zx -= 2 * (-y3 * rx / r3 + x3 * y3 * (-3 * rz * rx / r5));
    M.zy -= 2 * (-y3 * ry / r3 + x3 * y3 * (-3 * rz * ry / r5));
    M.zz += 2 * (-1.0 / r - rz * x3 / r3 - y3 * rz / r3 + x3 * y3 * (1.0 / r3 - 3 * rz * rz / r5));

    double ca = 2 * a2 / 3.0;
    M.xx -= ca * (-y3 * (3 * rz / r5 - 15 * rz * rx * rx / r7));
    M.xy -= ca * (-y3 * (-15 * rz * rx * ry / r7));
    M.xz += ca * (3 * rx * rz / r5 - y3 * (3 * rx / r5 - 15 * rz * rx * rz / r7));
    M.yx -= ca * (-y3 * (-15 * rz * ry * rx / r7));
    M.yy -= ca * (-y3 * (3 * rz / r5 - 15 * rz * ry * ry / r7));
    M.yz += ca * (3 * ry * rz / r5 - y3 * (3 * ry / r5 - 15 * rz * ry * rz / r7));
    M.zx -= ca * (-y3 * (3 * rx / r5 - 15 * rz * rz * rx / r7));
    M.zy -= ca * (-y3 * (3 * ry / r5 - 15 * rz * rz * ry / r7));
    M.zz += ca * (-1.0 / r3 + 3 * rz * rz / r5 - y3 * (3 * rz / r5 + 3 * rz / r5 + 3 * rz / r5 - 15 * rz * rz * rz / r7));

    double cb = 2 * b2 / 3.0;
    M.xx -= cb * (-x3 * (3 * rz / r5 - 15 * rz * rx * rx / r7));
    M.xy -= cb * (-x3 * (-15 * rz * rx * ry / r7));
    M.xz += cb * (-x3 * (3 * rx / r5 - 15 * rz * rx * rz / r7));
    M.yx -= cb * (-x3 * (-15 * rz * ry * rx / r7));
    M.yy -= cb * (-x3 * (3 * rz / r5 - 15 * rz * ry * ry / r7));
    M.yz += cb * (-x3 * (3 * ry / r5 - 15 * rz * ry * rz / r7));
    M.zx -= cb * (3 * rz * rx / r5 - x3 * (3 * rx / r5 - 15 * rz * rz * rx / r7));
    M.zy -= cb * (3 * rz * ry / r5 - x3 * (3 * ry / r5 - 15 * rz * rz * ry / r7));
    M.zz += cb * (-1.0 / r3 + 3 * rz * rz / r5 - x3 * (3 * rz / r5 + 3 * rz / r5 + 3 * rz / r5 - 15 * rz * rz * rz / r7));

    double cab = 2 * b2 * a2 / 3.0;
    M.xx -= cab * (-1.0 / r5 + 5 * rz * rz / r7 + 5 * rx * rx / r7 - 35 * rz * rz * rx * rx / r9);
    M.xy -= cab * (5 * rx * ry / r7 - 35 * rz * rz * rx * ry / r9);
    M.xz += cab * (5 * rx * rz / r7 + 5 * rx * rz / r7 + 5 * rx * rz / r7 - 35 * rz * rz * rx * rz / r9);
    M.yx -= cab * (5 * ry * rx / r7 - 35 * rz * rz * ry * rx / r9);
    M.yy -= cab * (-1.0 / r5 + 5 * rz * rz / r7 + 5 * ry * ry / r7 - 35 * rz * rz * ry * ry / r9);
    M.yz += cab * (5 * ry * rz / r7 + 5 * ry * rz / r7 + 5 * ry * rz / r7 - 35 * rz * rz * rz * ry / r9);
    M.zx -= cab * (5 * rz * rx / r7 + 5 * rz * rx / r7 + 5 * rz * rx / r7 - 35 * rz * rz * rz * rx / r9);
    M.zy -= cab * (5 * rz * ry / r7 + 5 * rz * ry / r7 + 5 * rz * ry / r7 - 35 * rz * rz * rz * ry / r9);
    M.zz += cab * (-1.0 / r5 + 5 * rz * rz / r7 - 1.0 / r5 + 5 * rz * rz / r7 - 1.0 / r5 + 5 * rz * rz / r7 +
                   5 * rz * rz / r7 + 5 * rz * rz / r7 + 5 * rz * rz / r7 - 35 * rz * rz * rz * rz / r9);
  }
  return M;
}

/* u = B_t M(x, y) B_s f for targets [t0, t1) (wall == 1: heights shifted to
 * max(z, radius) and damping B = z / radius below the radius, as the
 * wrapper mobility.py:551-591 with shift_heights_different_radius :87-99
 * and damping_matrix_B_different_radius :102-119; wall == 0: the bare
 * kernel as no_wall_mobility_trans_times_force_source_target_numba
 * :593-615; wall == 2: the free-surface kernel with the heights as given,
 * as the wrapper mobility.py:1409-1429).  out has nt*3 entries; rows
 * outside [t0, t1) are untouched. */
int orc_radii_apply(const double *src, const double *trg, const double *force, const double *rs,
                    const double *rt, int64_t ns, int64_t nt, int64_t t0, int64_t t1, double eta,
                    const double *L, int wall, double *out)
{
  if (t0 < 0 || t1 > nt || t0 > t1 || wall < 0 || wall > 2) return -1;
  const double norm_fact_f = 1.0 / (8.0 * M_PI * eta);
  const double Lx = L ? L[0] : 0.0, Ly = L ? L[1] : 0.0, Lz = L ? L[2] : 0.0;
  const int px = Lx > 0, py = Ly > 0, pz = Lz > 0;
#pragma omp parallel for schedule(static)
  for (int64_t i = t0; i < t1; ++i) {
    double a = rt[i];
    double xi = trg[3 * i], yi = trg[3 * i + 1], zi = trg[3 * i + 2];
    double Bt = 1.0;
    if (wall == 1) {
      if (!(zi > a)) { Bt = zi < a ? zi / a : 1.0; zi = a; }
    }
    double ux = 0, uy = 0, uz = 0;
    for (int bx = -px; bx <= px; ++bx)
      for (int by = -py; by <= py; ++by)
        for (int bz = -pz; bz <= pz; ++bz)
          for (int64_t j = 0; j < ns; ++j) {
            double b = rs[j];
            double zj = src[3 * j + 2], Bs = 1.0;
            if (wall == 1) {
              if (!(zj > b)) { Bs = zj < b ? zj / b : 1.0; zj = b; }
            }
            double rx = project_image(xi - src[3 * j], Lx, bx);
            double ry = project_image(yi - src[3 * j + 1], Ly, by);
            double rz = project_image(zi - zj, Lz, bz);
            mat3 M = radii_pair(rx, ry, rz, zi, zj, a, b, wall);
            double fx = Bs * force[3 * j], fy = Bs * force[3 * j + 1], fz = Bs * force[3 * j + 2];
            ux += (M.xx * fx + M.xy * fy + M.xz * fz) * norm_fact_f;
            uy += (M.yx * fx + M.yy * fy + M.yz * fz) * norm_fact_f;
            uz += (M.zx * fx + M.zy * fy + M.zz * fz) * norm_fact_f;
          }
    out[3 * i] = Bt * ux;
    out[3 * i + 1] = Bt * uy;
    out[3 * i + 2] = Bt * uz;
  }
  return 0;
}

/* The 3x3 block of one (target, source) pair, normalised, no wrapper. */
int orc_radii_pair_matrix(const double *ri, const double *rj, double a, double b, double eta, int wall, double *m9)
{
  mat3 M = radii_pair(ri[0] - rj[0], ri[1] - rj[1], ri[2] - rj[2], ri[2], rj[2], a, b, wall);
  const double n = 1.0 / (8.0 * M_PI * eta);
  m9[0] = M.xx * n; m9[1] = M.xy * n; m9[2] = M.xz * n;
  m9[3] = M.yx * n; m9[4] = M.yy * n; m9[5] = M.yz * n;
  m9[6] = M.zx * n; m9[7] = M.zy * n; m9[8] = M.zz * n;
  return 0;
}

/* ---------------------------------------------------------------------- */
/* Blob-blob soft repulsion (multi_bodies/forces_numba.py:12-56,           */
/* blob_blob_force_numba; radii variant :75-122 with a = (a_i + a_j)/2):   */
/* F_i = sum_{j != i} f0 dr, dr = r_j - r_i (minimum image for L_d > 0),   */
/* f0 = -(eps/b) exp(-(r - 2a)/b) / r for r > 2a, else -(eps/b)/max(r,     */
/* 1e-25).  radii may be NULL (one radius a).  cutoff > 0 keeps only pairs */
/* with r <= cutoff (the neighbour-list version :141-201 with cutoff       */
/* 2a + 30b).  Rows [t0, t1).                                              */
/* ---------------------------------------------------------------------- */
int orc_blob_forces(const double *r, const double *radii, int64_t n, int64_t t0, int64_t t1, double eps,
                    double b, double a, double cutoff, const double *L, double *out)
{
  if (t0 < 0 || t1 > n || t0 > t1) return -1;
#pragma omp parallel for schedule(static)
  for (int64_t i = t0; i < t1; ++i) {
    double fx = 0, fy = 0, fz = 0;
    for (int64_t j = 0; j < n; ++j) {
      if (i == j) continue;
      double dr[3];
      for (int k = 0; k < 3; ++k) {
        dr[k] = r[3 * j + k] - r[3 * i + k];
        double Lk = L ? L[k] : 0.0;
        if (Lk > 0) {
          double s = (double)((dr[k] > 0) - (dr[k] < 0));
          dr[k] -= (double)(long)(dr[k] / Lk + MUT(13, 0.5, 0.0) * s) * Lk;
        }
      }
      double aa = radii ? (radii[i] + radii[MUT(14, j, i)]) * 0.5 : a;
      double r_norm = sqrt(dr[0] * dr[0] + dr[1] * dr[1] + dr[2] * dr[2]);
      double f0;
      if (cutoff > 0 && r_norm > cutoff) continue;
      if (r_norm > 2 * aa)
        f0 = -((eps / b) * exp(-(r_norm - MUT(11, 2.0, 1.0) * aa) / b) / r_norm);
      else
        f0 = -((eps / b) / MUT(12, (r_norm > 1e-25 ? r_norm : 1e-25), 1.0));
      fx += f0 * dr[0];
      fy += f0 * dr[1];
      fz += f0 * dr[2];
    }
    out[3 * i] = fx;
    out[3 * i + 1] = fy;
    out[3 * i + 2] = fz;
  }
  return 0;
}
