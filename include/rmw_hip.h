/*
 * rmw_hip.h -- C ABI of the MI355X (gfx950) wall-corrected RPY blob mobility
 * library (librmw_hip.so).
 *
 * All array arguments are DEVICE pointers (any allocator; the Python layer
 * passes PyTorch-ROCm buffers), row-major, float64 unless stated.  Vectors of
 * blobs are (n, 3); k right-hand sides are stacked as (k, n, 3).  `stream` is
 * a hipStream_t passed as void* (NULL = the null stream).  Every call is
 * stream-ordered and asynchronous; nothing allocates device memory: callers
 * provide a workspace sized by the matching *_workspace_size query.
 *
 * Status: 0 on success, negative on error (RMW_E_*); rmw_last_error()
 * returns a thread-local message.  No C++ exception crosses this boundary.
 */
#ifndef RMW_HIP_H
#define RMW_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define RMW_OK 0
#define RMW_E_ARG (-1)       /* invalid argument */
#define RMW_E_HIP (-2)       /* HIP runtime error */
#define RMW_E_WORKSPACE (-3) /* workspace too small */
#define RMW_E_NOTFINITE (-4) /* non-finite value detected */
#define RMW_E_LIMIT (-5)     /* size above a compiled-in limit */

/* Blocks of the grand mobility [[UF, UT], [WF, WT]]. */
#define RMW_UF 1 /* velocity from force         (mobility_numba.py:123-287)  */
#define RMW_UT 2 /* velocity from torque        (mobility_numba.py:547-686)  */
#define RMW_WF 4 /* angular velocity from force (mobility_numba.py:937-1073) */
#define RMW_WT 8 /* angular velocity from torque(mobility_numba.py:1188-1328)*/

/* Arithmetic of the pair kernels. */
#define RMW_PREC_F32_F64ACC 0 /* FP32 pair terms, FP64 per-target sums     */
#define RMW_PREC_F32_KAHAN 1  /* FP32 pair terms, compensated FP32 sums     */
#define RMW_PREC_F64 2        /* FP64 throughout                           */
#define RMW_PREC_F32_BLOCKED 3 /* FP32 pair terms, FP32 sub-sums of <= 16
                                  sources folded into FP64 per-target sums */

/* Launch shape; pass NULL for the built-in heuristic (the Python layer
 * passes the autotuner's choice when its cache has one). */
typedef struct rmw_tuning {
  int64_t chunk_len;  /* sources per work item (rounded up to a multiple of 64; 0 = auto) */
  int32_t max_wgs;    /* cap on resident workgroups of persistent launches (0 = auto) */
  int32_t rhs_group;  /* right-hand sides per pair-kernel pass: 1, 2 or 4 (0 = auto) */
  int32_t symmetric;  /* each unordered pair evaluated once for both blobs
                         (UF_ji = UF_ij^T, UT_ji = WF_ij^T, WF_ji = UT_ij^T, WT_ji = WT_ij^T):
                         the UF block alone, all four blocks (wall, no wall), UF + UT
                         or one of UT, WF, WT alone (wall, one right-hand side per
                         pass); above the wall also with pseudo-periodic images, one
                         right-hand side per pass; the no-wall, in-plane and
                         free-surface domains without images.  0 = when the tile
                         pairs fill the device, 1 = required (error otherwise),
                         -1 = off */
  int32_t flags;      /* RMW_FLAG_* bits (0 = none) */
} rmw_tuning;

/* rmw_tuning.flags: after the product, scan the outputs and return
 * RMW_E_NOTFINITE if any entry is inf or NaN (e.g. two distinct blobs at
 * the same position, which the reference returns silently as NaN,
 * mobility/mobility_numba.py:88).  The call then waits for the stream to
 * finish; the workspace needs RMW_STATUS_BYTES more (the size queries
 * include them when they are given the same tuning). */
#define RMW_FLAG_CHECK_FINITE 1
#define RMW_STATUS_BYTES 256

const char* rmw_last_error(void);
const char* rmw_version(void);

/*
 * Workspace bytes needed by rmw_wall_rpy_apply for these sizes.
 */
int rmw_wall_rpy_workspace_size(int mask, int prec, int64_t n, int k,
                                const rmw_tuning* tuning, size_t* bytes);

/*
 * Grand wall-RPY product on one suspension of n blobs:
 *     velocity = UF * force + UT * torque     (blocks present in mask)
 *     omega    = WF * force + WT * torque
 * with the wall-overlap regularisation B M(z_eff) B of the reference
 * wrappers (z_eff = max(z, a), B_ii = z_i/a below z = a).  r_vectors (n,3);
 * force/torque: k right-hand sides of (n,3), `ld` doubles apart (0 = 3n),
 * or NULL when no block of the mask reads them; velocity/omega likewise
 * (same ld), or NULL when no block writes them.  L[3] are the
 * pseudo-periodic lengths (host array; NULL or entries <= 0 = unbounded).
 *
 * With tuning->flags & RMW_FLAG_CHECK_FINITE the outputs are scanned and a
 * non-finite entry returns RMW_E_NOTFINITE (the outputs are still written).
 *
 * Replaces, for mask = RMW_UF / RMW_UT / RMW_WF / RMW_WT:
 *   single_wall_mobility_trans_times_force_numba   mobility/mobility.py:1132-1163
 *   single_wall_mobility_trans_times_torque_numba  mobility/mobility.py:1213-1232
 *   single_wall_mobility_rot_times_force_numba     mobility/mobility.py:1273-1298
 *   single_wall_mobility_rot_times_torque_numba    mobility/mobility.py:1316-1342
 * for RMW_UF|RMW_UT: single_wall_mobility_trans_times_force_torque_pycuda
 *   mobility/mobility.py:384-410, and for all four the rollers' grand
 *   mobility closure quaternion_integrator/quaternion_integrator_rollers.py:1114-1122.
 */
int rmw_wall_rpy_apply(int mask, int prec, const double* r_vectors, const double* force,
                       const double* torque, double* velocity, double* omega, int64_t n, int k,
                       int64_t ld, double eta, double a, const double* L, const rmw_tuning* tuning,
                       void* workspace, size_t workspace_bytes, void* stream);

/*
 * rmw_wall_rpy_apply without a workspace argument and without any device,
 * pinned or event allocation: the partial sums go to a fixed array of
 * RMW_NOALLOC_SCRATCH_BYTES in device memory that the library holds on
 * every device it runs on until the process ends.  No counterpart in the
 * reference: this entry point is for callers that cannot allocate (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * Same arguments and results as rmw_wall_rpy_apply; when the launch shape
 * rmw_wall_rpy_apply would choose needs more partial-sum space than the
 * array holds, the ordered kernel runs with longer source chunks (and
 * fewer right-hand sides per pass) so that it fits; sizes for which even
 * that does not fit (about 2.7 million blobs per right-hand side and
 * output block) return RMW_E_LIMIT before any launch.  Every call on one
 * device shares that device's array: calls that overlap in time on one
 * device (two streams, two threads) race and must be serialised by the
 * caller.  Pair-kernel timing records (rmw_profile_enable) are not taken.
 */
#define RMW_NOALLOC_SCRATCH_BYTES (64u << 20)
int rmw_wall_rpy_apply_noalloc(int mask, int prec, const double* r_vectors, const double* force,
                               const double* torque, double* velocity, double* omega, int64_t n, int k,
                               int64_t ld, double eta, double a, const double* L, const rmw_tuning* tuning,
                               void* stream);

/*
 * The same products for the other boundaries of the fluid.  domain:
 *   RMW_DOMAIN_WALL          = rmw_wall_rpy_apply
 *   RMW_DOMAIN_NO_WALL       unbounded RPY, no height shift or damping, any
 *                            mask: no_wall_mobility_{trans,rot}_times_
 *                            {force,torque}_numba, mobility/mobility_numba.py
 *                            :12-120, 440-545, 832-936, 1077-1186 (wrappers
 *                            mobility.py:1119-1129, 1200-1210, 1258-1270,
 *                            1301-1313)
 *   RMW_DOMAIN_IN_PLANE      the wall blocks with only the x-y components of
 *                            the sources and outputs (P M P, P = diag(1,1,0)),
 *                            masks UF, UT, UF|UT: in_plane_mobility_trans_
 *                            times_{force,torque}_numba, :291-437, 690-830
 *                            (wrappers mobility.py:1166-1198, 1235-1256)
 *   RMW_DOMAIN_FREE_SURFACE  RPY(r_i - r_j) + RPY(r_i - r_j*) D with the image
 *                            r_j* = (x_j, y_j, -z_j), D = diag(1,1,-1), UF
 *                            only: free_surface_mobility_trans_times_force_
 *                            numba, :1770-1938 (wrapper mobility.py:1390-1406)
 * One right-hand side per pass (k > 1 runs k passes).  The symmetric
 * kernels serve the UF block of every domain and the grand mobility of the
 * unbounded one (without periodic images); the other masks and periodic
 * calls outside the wall domain run the ordered kernels.
 */
#define RMW_DOMAIN_WALL 0
#define RMW_DOMAIN_NO_WALL 1
#define RMW_DOMAIN_IN_PLANE 2
#define RMW_DOMAIN_FREE_SURFACE 3
int rmw_mobility_workspace_size(int domain, int mask, int prec, int64_t n, int k, const rmw_tuning* tuning,
                                size_t* bytes);
int rmw_mobility_apply(int domain, int mask, int prec, const double* r_vectors, const double* force,
                       const double* torque, double* velocity, double* omega, int64_t n, int k, int64_t ld,
                       double eta, double a, const double* L, const rmw_tuning* tuning, void* workspace,
                       size_t workspace_bytes, void* stream);

/*
 * The launch plan rmw_wall_rpy_apply would use for these sizes (periodic:
 * whether L has a positive entry): which kernel, the right-hand sides per
 * pass, the source chunk and the number of work items.  No reference
 * counterpart; for benchmarks and tuning reports.
 */
typedef struct rmw_plan_info {
  int32_t symmetric;       /* 1: each unordered pair once for both blobs */
  int32_t rhs_group;       /* right-hand sides per pair-kernel pass */
  int64_t chunk_len;       /* sources per work item */
  int64_t work_items;      /* workgroup work items of one pass */
  int64_t workspace_bytes;
} rmw_plan_info;
int rmw_wall_rpy_describe(int mask, int prec, int64_t n, int k, const rmw_tuning* tuning, int periodic,
                          rmw_plan_info* info);

/*
 * Ragged batch: nsys independent suspensions stored back to back, system s
 * owning blobs [offsets[s], offsets[s+1]) (offsets is a HOST array of nsys+1
 * entries, offsets[0] = 0).  Same arrays and semantics as rmw_wall_rpy_apply
 * with n = offsets[nsys]; no interaction between systems; one persistent
 * launch serves all systems and all k right-hand sides.
 * No single reference counterpart: the reference evaluates one suspension
 * per call (mobility/mobility.py:1132-1342).
 */
int rmw_wall_rpy_ragged_workspace_size(int mask, int prec, const int64_t* offsets, int nsys,
                                       int k, const rmw_tuning* tuning, size_t* bytes);
int rmw_wall_rpy_ragged_apply(int mask, int prec, const double* r_vectors, const double* force,
                              const double* torque, double* velocity, double* omega,
                              const int64_t* offsets, int nsys, int k, int64_t ld, double eta,
                              double a, const double* L, const rmw_tuning* tuning, void* workspace,
                              size_t workspace_bytes, void* stream);

/* boundary argument of the different-radii products */
#define RMW_RADII_UNBOUNDED 0
#define RMW_RADII_WALL 1
#define RMW_RADII_FREE_SURFACE 2

/*
 * Blobs of different radii, source -> target: velocity (nt, 3) =
 * B_t M(x, y) B_s force, force (ns, 3), radii per blob, lengths in absolute
 * units (not scaled by a radius), normalisation 1/(8 pi eta).  boundary
 * RMW_RADII_WALL: single-wall Rotne-Prager-Blake with heights shifted to
 * max(z, radius) and damping B = z/radius below the radius;
 * RMW_RADII_UNBOUNDED: unbounded RPY of different radii;
 * RMW_RADII_FREE_SURFACE: RPY of d plus the RPY of the image vector
 * (d_x, d_y, z_i + z_j) times diag(1, 1, -1), heights as given; other
 * values return RMW_E_ARG.  L: pseudo-periodic lengths as
 * rmw_wall_rpy_apply.  Sources and targets may be the same arrays (then
 * the self pair is the RPY self mobility 1/(6 pi eta a) plus its image).
 * Replaces
 *   mobility_trans_times_force_source_target_numba  mobility/mobility_numba.py:1479-1658
 *   single_wall_mobility_trans_times_force_source_target_numba  mobility/mobility.py:551-591
 *   no_wall_mobility_trans_times_force_source_target_numba      mobility/mobility.py:593-615
 *   free_surface_mobility_trans_times_force_source_target_numba mobility/mobility.py:1409-1429
 *     (kernel mobility/mobility_numba.py:1941-2093)
 * as selected by the 'radii*' options of multi_bodies/multi_bodies.py:266-287.
 */
int rmw_radii_workspace_size(int prec, int64_t ns, int64_t nt, const rmw_tuning* tuning, size_t* bytes);
int rmw_radii_apply(int prec, int boundary, const double* source, const double* target, const double* force,
                    const double* radius_source, const double* radius_target, double* velocity, int64_t ns,
                    int64_t nt, double eta, const double* L, const rmw_tuning* tuning, void* workspace,
                    size_t workspace_bytes, void* stream);

/*
 * Blobs of different radii, sources = targets: velocity (n, 3) =
 * B M(r, r) B force with the conventions of rmw_radii_apply.  The pair
 * tensor is its own transpose under exchange of the two blobs, so when the
 * unordered tile pairs fill the device (tuning->symmetric > 0 forces it,
 * < 0 forbids it) and no L[d] > 0, each unordered pair is evaluated once
 * and applied to both blobs; otherwise this is rmw_radii_apply with
 * source = target.  describe reports the choice (symmetric 1/0), the work
 * items and the workspace bytes.  Replaces
 *   mobility_radii_trans_times_force  mobility/mobility.py:1369-1374
 * (sources = targets of the source-target kernels above).
 */
int rmw_radii_self_workspace_size(int prec, int64_t n, const double* L, const rmw_tuning* tuning,
                                  size_t* bytes);
int rmw_radii_self_describe(int prec, int64_t n, const double* L, const rmw_tuning* tuning, int* symmetric,
                            int64_t* work_items, size_t* bytes);
int rmw_radii_self_apply(int prec, int boundary, const double* r_vectors, const double* force,
                         const double* radius, double* velocity, int64_t n, double eta, const double* L,
                         const rmw_tuning* tuning, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Blob-blob soft repulsion (FP64 positions, radii and forces, n x 3 row
 * major).  force_i = sum_{j != i} f0(r) (r_j - r_i), dr taken to its
 * minimum image where L[d] > 0 (L may be null), with
 *   f0 = -(eps/b) exp(-(r - 2a)/b) / r  for r > 2a,  -(eps/b) / r  otherwise,
 * a = blob_radius, or (a_i + a_j)/2 when radius_blobs is given.  cutoff > 0
 * drops the pairs with r > cutoff (the neighbour-list form uses
 * 2a + 30b); 0 keeps every pair.  prec RMW_PREC_F64: FP64 1/r (hardware
 * estimate and one third-order Newton step, within 1.4e-16) and FP64 exp;
 * any other mode: FP64 geometry and sums with a hardware reciprocal square
 * root (one Newton step) and an FP32 exponential.
 * Replaces
 *   blob_blob_force_numba / calc_blob_blob_forces_numba
 *                                  multi_bodies/forces_numba.py:12-72
 *   blob_blob_force_radii_numba    multi_bodies/forces_numba.py:75-138
 *   blob_blob_force_tree_numba     multi_bodies/forces_numba.py:141-241
 *   blob_blob_force (C++)          multi_bodies/forces.cpp:11-42
 *   calc_blob_blob_forces_pycuda   multi_bodies/forces_pycuda.py:148-180
 * as selected by set_blob_blob_forces, multi_bodies_functions.py:249-279.
 */
int rmw_blob_forces_workspace_size(int prec, int64_t n, const rmw_tuning* tuning, size_t* bytes);
/* Launch plan of rmw_blob_forces: symmetric (each unordered pair once, the
 * default when there are at least as many tile pairs as CUs and the
 * workspace stays within 2 GiB; tuning->symmetric 1 requires it, -1 turns
 * it off) or ordered, chunk length, work items, workspace bytes. */
int rmw_blob_forces_describe(int prec, int64_t n, const rmw_tuning* tuning, rmw_plan_info* info);
int rmw_blob_forces(int prec, const double* r_vectors, const double* radius_blobs, double* force, int64_t n,
                    double repulsion_strength, double debye_length, double blob_radius, double cutoff,
                    const double* L, const rmw_tuning* tuning, void* workspace, size_t workspace_bytes,
                    void* stream);

/* ------------------------------------------------------------------------
 * Per-roller parts of the rollers time step (FP64, n x 3 row-major device
 * arrays; omega is a HOST array of 3 values).
 * ---------------------------------------------------------------------- */
/* force_z = -g m + (eps_w/b_w) exp(-(z - a)/b_w) (z > a) or + eps_w/b_w,
 * force_x = force_y = 0; accumulate != 0 adds to force instead.  Replaces
 * blob_external_forces, multi_bodies/multi_bodies_functions.py:151-196. */
int rmw_rollers_one_blob_forces(const double* r, int64_t n, double blob_mass, double g, double blob_radius,
                                double repulsion_strength_wall, double debye_length_wall, int accumulate,
                                double* force, void* stream);
/* Velocity (and torque, if torque != NULL) without hydrodynamic
 * interactions: compute_deterministic_velocity_and_torque_uncorrelated,
 * quaternion_integrator_rollers.py:985-1080 (free_kinematics != 0: the
 * torque of get_torque, :1586-1597). */
int rmw_rollers_uncorrelated_velocity(const double* r, const double* force, int64_t n, const double* omega,
                                      int free_kinematics, double eta, double a, double* velocity, double* torque,
                                      void* stream);
/* sqrt(2 kT/dt) sqrt(mu_self) z + kT d(mu_perp)/dh e_z:
 * compute_stochastic_linear_velocity_uncorrelated, :1263-1313. */
int rmw_rollers_uncorrelated_noise(const double* r, const double* z, int64_t n, double eta, double a, double kT,
                                   double dt, double* velocity, void* stream);
/* r_new = r_old + dt (c_det det - c_prev prev + stoch) (prev, stoch may be
 * NULL) and counts[0..1] (device int64) = number of new heights below 0
 * and below a: the position update of the rollers schemes,
 * quaternion_integrator_rollers.py:119-302 (Adams-Bashforth c_det = 1.5,
 * c_prev = 0.5; forward Euler c_det = 1). */
int rmw_rollers_advance_workspace_size(int64_t n, size_t* bytes);
int rmw_rollers_advance(const double* r_old, const double* det_velocity, const double* prev_velocity,
                        const double* stoch_velocity, int64_t n, double c_det, double c_prev, double dt, double a,
                        double* r_new, int64_t* counts, void* workspace, size_t workspace_bytes, void* stream);

/*
 * Optional timing of the pair kernels (wall RPY, different radii and
 * blob-blob forces) with HIP events recorded on the launch stream.  rmw_profile_collect synchronises on the recorded events,
 * returns the number of pair-kernel launches since the last collect and
 * their total and mean duration in ms, and clears the record.
 */
int rmw_profile_enable(int on);
int rmw_profile_collect(int64_t* launches, double* total_ms, double* mean_ms);
/* The same record split by kernel family (arrays of 3: wall RPY, different
 * radii, blob-blob forces); also clears the record. */
int rmw_profile_collect_kinds(int64_t* launches, double* total_ms);
/* The record launch by launch, in launch order: the number of launches
 * recorded so far (rmw_profile_pending), then each one's kernel family
 * (0 wall RPY, 1 different radii, 2 blob-blob forces) and duration in ms,
 * the first cap of them written; *launches = all of them; clears the
 * record.  Medians and spreads of the launch times come from this (no
 * counterpart in the reference: timing of this library's own launches). */
int rmw_profile_pending(int64_t* launches);
int rmw_profile_collect_launches(int32_t* kinds, double* ms, int64_t cap, int64_t* launches);

/* ------------------------------------------------------------------------
 * Rigid multiblob bodies (FP64).  body_offsets is a HOST array of nbodies+1
 * blob offsets (body b owns blobs [off[b], off[b+1])).  rmw_rigid_setup
 * uploads the offsets into the workspace; every other rigid call reads them
 * from that workspace (same offsets, stream-ordered after the setup).
 * Vectors of the rigid system are [lambda (3N); U (6B)] with U_b = (v, w).
 * ---------------------------------------------------------------------- */
int rmw_rigid_workspace_size(const int64_t* body_offsets, int nbodies, size_t* bytes,
                             size_t* minv_doubles, size_t* pc_scratch_doubles);
int rmw_rigid_setup(const int64_t* body_offsets, int nbodies, void* workspace,
                    size_t workspace_bytes, void* stream);

/* Blob positions r = R(q_b) ref + c_b and r_rel = r - c_b (quaternions (s,p),
 * rotation of quaternion_integrator/quaternion.py:42-52).  Replaces
 * Body.get_r_vectors, body/body.py:64-77, for all bodies at once. */
int rmw_body_blobs(const double* centres, const double* quats, const double* ref,
                   const int64_t* body_offsets, int nbodies, double* r_blobs, double* r_rel,
                   void* workspace, size_t workspace_bytes, void* stream);

/* Second half of the rigid operator: given m_lambda = M lambda,
 *   y_lambda = m_lambda - K U,   y_U = -K^T lambda
 * (K_b = [J, rot], body/body.py:81-113).  With rmw_wall_rpy_apply for
 * M lambda this replaces linear_operator_rigid, multi_bodies/multi_bodies.py:
 * 424-477 (K U and K^T lambda of :328-370); constraints and prescribed
 * kinematics are not supported. */
int rmw_rigid_operator_post(const double* r_rel, const int64_t* body_offsets, int nbodies,
                            const double* x, const double* m_lambda, double* y,
                            void* workspace, size_t workspace_bytes, void* stream);

/* Dense per-body wall mobility B M(z_eff) B, row-major 3n x 3n blocks packed
 * back to back (block b at offset sum_{b'<b} (3 n_b')^2).  Replaces
 * single_wall_fluid_mobility, mobility/mobility.py:1018-1116, as selected by
 * mobility_blobs_implementation (multi_bodies/multi_bodies.py:207-230). */
/* force_torque (B, 6) = [sum_i f_i, sum_i r_rel_i x f_i] per body from the
 * blob forces (N, 3): force_torque_calculator_sort_by_bodies,
 * multi_bodies/multi_bodies_functions.py:411-445. */
int rmw_rigid_force_torque(const double* r_rel, const int64_t* body_offsets, int nbodies, const double* blob_force,
                           double* force_torque, void* workspace, size_t workspace_bytes, void* stream);
/* Body update of the deterministic rigid schemes: V = c_det U - c_prev U_prev
 * (U_prev may be NULL), centre += V_v dt, quaternion = q(V_w dt) * quaternion
 * (quaternion_integrator_multi_bodies.py:75-186, quaternion.py:15-49). */
int rmw_bodies_advance(const double* centres, const double* quats, const double* velocities,
                       const double* prev_velocities, int nbodies, double c_det, double c_prev, double dt,
                       double* centres_new, double* quats_new, void* stream);
/* count (device int64) = number of rows of r (n, 3) with r_z < level
 * (body.check_function, body/body.py:118-140). */
int rmw_count_below_workspace_size(int64_t n, size_t* bytes);
int rmw_count_below(const double* r, int64_t n, double level, int64_t* count, void* workspace,
                    size_t workspace_bytes, void* stream);
int rmw_dense_wall_mobility(const double* r_blobs, const int64_t* body_offsets, int nbodies,
                            double eta, double a, double* out, void* workspace,
                            size_t workspace_bytes, void* stream);
/* The same for domain RMW_DOMAIN_WALL (= rmw_dense_wall_mobility) or
 * RMW_DOMAIN_NO_WALL: the unbounded RPY tensor of each body, no height
 * shift or damping, replacing rotne_prager_tensor, mobility/mobility.py
 * :967-1012 ('python_no_wall' / 'C++_no_wall' of set_mobility_blobs,
 * multi_bodies/multi_bodies.py:218-221). */
int rmw_dense_mobility(int domain, const double* r_blobs, const int64_t* body_offsets, int nbodies, double eta,
                       double a, double* out, void* workspace, size_t workspace_bytes, void* stream);

/* Block-diagonal body preconditioner build: per body M_b^-1 (via Cholesky,
 * minv: packed blocks as above, minv_doubles long), M_b^-1 K (minvk:
 * (3N, 6)) and N_b = pinv(K^T M_b^-1 K) (nbody: (B, 6, 6)); status[b] = 1
 * when M_b is not positive definite (may be NULL).  Bodies of up to 45 blobs
 * factorise in LDS; larger ones need `scratch` of pc_scratch_doubles
 * (else NULL, 0).  Replaces the factorisation loop of
 * build_block_diagonal_preconditioner, multi_bodies/multi_bodies.py:752-822. */
int rmw_rigid_pc_build(const double* r_blobs, const double* r_rel, const int64_t* body_offsets,
                       int nbodies, double eta, double a, double* minv, double* minvk,
                       double* nbody, int* status, double* scratch, size_t scratch_doubles,
                       void* workspace, size_t workspace_bytes, void* stream);

/* The general build: rmw_rigid_pc_build for domain RMW_DOMAIN_WALL or
 * RMW_DOMAIN_NO_WALL (M_b the unbounded RPY tensor: rotne_prager_tensor,
 * mobility/mobility.py:967-1012, as mobility_blobs_implementation
 * 'python_no_wall' / 'C++_no_wall' selects it, multi_bodies.py:218-221),
 * and with chol, chol_inv non-NULL the stochastic variant: besides minv,
 * minvk and nbody, per body the lower Cholesky factor L_b of M_b
 * (M_b = L_b L_b^T; chol) and L_b^-1 (chol_inv), row-major blocks packed as
 * minv with zero upper triangles.  That variant replaces the factorisation
 * loop of build_block_diagonal_preconditioners_det_stoch, multi_bodies/
 * multi_bodies.py:480-560 (its M_factorization_blobs = L_b; its
 * M_factorization_blobs_inv = U_b^-1 = L_b^-T for scipy's M_b = U_b^T U_b). */
int rmw_rigid_pc_build_ex(int domain, const double* r_blobs, const double* r_rel, const int64_t* body_offsets,
                          int nbodies, double eta, double a, double* minv, double* minvk, double* nbody, double* chol,
                          double* chol_inv, int* status, double* scratch, size_t scratch_doubles, void* workspace,
                          size_t workspace_bytes, void* stream);
/* y_b = B_b x_b (trans = 0) or B_b^T x_b (trans = 1) for every body b and
 * each of k vectors ld doubles apart (0 = 3N), B packed as minv; lower != 0
 * reads only the lower triangle of each block.  With chol / chol_inv these
 * are the per-body products of the mobility_pc and P_inv_mult closures,
 * multi_bodies/multi_bodies.py:591-615 (x != y). */
int rmw_rigid_block_apply(const int64_t* body_offsets, int nbodies, const double* blocks, int trans, int lower,
                          const double* x, double* y, int k, int64_t ld, void* workspace, size_t workspace_bytes,
                          void* stream);

/* y = P^-1 x, the body preconditioner apply of multi_bodies/multi_bodies.py:
 * 825-849 (x != y). */
/* Velocity (nblobs, 3) at every blob induced by the net forces (first three
 * entries of each force_stride-long row of `force`) of all other bodies,
 * each body a sphere of radius `radius` at its centre, through the
 * different-radii wall RPY of rmw_radii_apply (wall != 0) or the unbounded
 * one: u_i = B_i sum_{c != body(i)} M(r_i, centre_c; a, radius) B_c F_c.
 * The far field of the off-diagonal blocks of the rigid operator
 * (multi_bodies.py:424-477) on any iterate whose body force sums are F:
 * P^-1 (b - [u; 0]) is a block-Jacobi step from P^-1 b without the blob
 * product.  No reference counterpart (an initial-guess helper). */
int rmw_rigid_point_field(const double* r_blobs, const double* centres, const double* force,
                          int64_t force_stride, const int64_t* body_offsets, int nbodies, double radius,
                          double a, double eta, int wall, double* velocity, void* workspace,
                          size_t workspace_bytes, void* stream);
int rmw_rigid_pc_apply(const int64_t* body_offsets, int nbodies, const double* minv,
                       const double* minvk, const double* nbody, const double* r_rel,
                       const double* x, double* y, void* workspace, size_t workspace_bytes,
                       void* stream);

/* ------------------------------------------------------------------------
 * FP64 Krylov vector kernels (GMRES Arnoldi and Lanczos).  Bases are nvec
 * rows of stride ldv; coefficient vectors live on the device.  The scipy
 * gmres / numpy vector algebra they replace is called from
 * general_application_utils.py:514-635 and
 * stochastic_forcing/stochastic_forcing.py:112-262.
 * ---------------------------------------------------------------------- */
int rmw_krylov_workspace_size(int64_t n, int max_vectors, size_t* bytes);
int rmw_dot_many(const double* V, int64_t ldv, int nvec, const double* w, int64_t n, double* out,
                 int sqrt_out, void* workspace, size_t workspace_bytes, void* stream);
int rmw_sub_combination(const double* V, int64_t ldv, int nvec, const double* coef, double* w,
                        int64_t n, void* stream);
int rmw_combination(const double* V, int64_t ldv, int nvec, const double* coef, double* y,
                    int64_t n, int accumulate, void* stream);
int rmw_axpby(double alpha, const double* x, double beta, double* y, int64_t n, const double* div,
              void* stream);
/* Arnoldi step, classical Gram-Schmidt with one re-orthogonalisation:
 * h[0..j] = coefficients, h[j+1] = ||w||, V[j+1] = w / ||w||. */
int rmw_arnoldi_cgs2(double* V, int64_t ldv, int j, double* w, int64_t n, double* h,
                     double* hscratch, void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------------------------------
 * Segmented (ragged-batch) FP64 Krylov kernels: the vector algebra of many
 * independent suspensions stacked in one vector, one launch sequence per
 * step for the whole batch.  System s owns blobs [offsets[s], offsets[s+1])
 * (offsets: DEVICE int64 array of nsys + 1 entries, nblobs = offsets[nsys]);
 * a vector holds nparts parts of stride part_stride with comp values per
 * blob.  The reference runs one suspension per process, so these replace
 * calling the per-system algebra of stochastic_forcing_lanczos
 * (stochastic_forcing/stochastic_forcing.py:112-262) once per replica.
 * Results are bit-reproducible (fixed-order two-pass reductions).
 * ---------------------------------------------------------------------- */
typedef struct rmw_segments {
  const int64_t* offsets; /* device, nsys + 1 */
  int32_t nsys;
  int32_t comp;           /* values per blob in each part (3) */
  int32_t nparts;         /* 2 for [force; torque] grand vectors */
  int32_t reserved;
  int64_t nblobs;         /* offsets[nsys], known to the host */
  int64_t part_stride;    /* element offset between parts (>= comp * nblobs) */
} rmw_segments;

int rmw_seg_workspace_size(const rmw_segments* seg, int max_vectors, size_t* bytes);
/* out[s*ldo + j] = (V_j . w) restricted to system s (sqrt when sqrt_out) */
int rmw_seg_dot_many(const double* V, int64_t ldv, int nvec, const double* w, const rmw_segments* seg,
                     double* out, int64_t ldo, int sqrt_out, void* workspace, size_t workspace_bytes,
                     void* stream);
/* y|s = sum_j coef[s*ldc + j] V_j|s (+ y|s when accumulate) */
int rmw_seg_combination(const double* V, int64_t ldv, int nvec, const double* coef, int64_t ldc, double* y,
                        const rmw_segments* seg, int accumulate, void* stream);
/* y|s = x|s / div[s*ld_div] (a non-positive divisor writes e_0 of system s) */
int rmw_seg_scale(const double* x, double* y, const double* div, int64_t ld_div, const rmw_segments* seg,
                  void* stream);
/* Arnoldi/Lanczos step with CGS2 for every system at once: with
 * h_s = H + s*h_sys_stride, h_s[0..j] = coefficients, h_s[j+1] = ||w|s||,
 * V[j+1]|s = w|s / ||w|s|| (e_0 of the system when zero); hscratch holds
 * nsys*(j+1) doubles. */
int rmw_seg_arnoldi_cgs2(double* V, int64_t ldv, int j, double* w, const rmw_segments* seg, double* H,
                         int64_t h_sys_stride, double* hscratch, void* workspace, size_t workspace_bytes,
                         void* stream);

/* ------------------------------------------------------------------------
 * Lanczos square root (FP64).  For k tridiagonals T_r of order m whose
 * entries the Lanczos steps left in H (alpha_i = H_r[i*ldh + i], beta_i =
 * H_r[i*ldh + i + 1], H_r = H + r*h_rhs_stride), writes
 *   c_out[r*c_stride + j] = scale[r] * (sqrt+(T_r) e_1)_j,   j < m,
 * sqrt+ taking the square root of the non-negative eigenvalues (negative
 * ones -> 0) of an implicit-QL eigen-decomposition; status[r] = 1 when QL
 * did not converge.  This is the eigh step of stochastic_forcing_lanczos,
 * stochastic_forcing/stochastic_forcing.py:208-227 (noise = V c).
 * ---------------------------------------------------------------------- */
int rmw_lanczos_scratch_doubles(int m, int k, size_t* doubles);
int rmw_lanczos_sqrt_coeffs(const double* H, int64_t h_rhs_stride, int64_t ldh, int m, int k,
                            const double* scale, double* c_out, int64_t c_stride, double* scratch,
                            size_t scratch_doubles, int* status, void* stream);

/* ------------------------------------------------------------------------
 * Learned GMRES warm start (no reference counterpart: the network only
 * supplies the initial guess x0 of the rigid-body solve, GMRES to the same
 * tolerance stays the judge).  BF16 operands are uint16 arrays.
 * ---------------------------------------------------------------------- */
/* C = op(A) op(B) + bias (+ beta C), H = act(op(A) op(B) + bias) in BF16;
 * op(A) is M x K (A stored K x M when transA), op(B) is K x N (B stored N x K
 * when transB); FP32 accumulation on MFMA (v_mfma_f32_16x16x16bf16_1k).
 * C or H may be NULL; act 0 = identity, 1 = SiLU. */
int rmw_gemm_bf16(int transA, int transB, int M, int N, int K, const uint16_t* A, int64_t lda,
                  const uint16_t* B, int64_t ldb, float* C, int64_t ldc, float beta, uint16_t* H,
                  int64_t ldh, const float* bias, int act, void* stream);
/* C = op(A) op(B) (+ beta C) with K split into `splits` slices of whole
 * 32-wide steps, each slice's FP32 sums written to `partial` ([slice][M][N],
 * rmw_gemm_bf16_splitk_workspace bytes) and added in slice order: the
 * weight gradients, whose K is the batch (M, N <= 128, K up to 16384). */
int rmw_gemm_bf16_splitk_workspace(int M, int N, int K, int splits, size_t* bytes);
int rmw_gemm_bf16_splitk(int transA, int transB, int M, int N, int K, const uint16_t* A, int64_t lda,
                         const uint16_t* B, int64_t ldb, float* C, int64_t ldc, float beta, int splits,
                         float* partial, size_t partial_bytes, void* stream);
int rmw_f32_to_bf16(const float* x, uint16_t* y, int64_t n, void* stream);
/* dZ = dH * SiLU'(Z), BF16 out */
int rmw_silu_backward(const float* dH, const float* Z, uint16_t* dZ, int64_t n, void* stream);
/* out[j] = sum_i X[i, j] (+ beta out[j]) */
int rmw_colsum_bf16(const uint16_t* X, int64_t ldx, int rows, int cols, float* out, float beta,
                    void* stream);
/* node input rows [Xn | sum of the deg edge messages | 0] and its adjoint */
int rmw_gnn_gather(const float* Xn, int nin, const float* Msg, int nmsg, int deg, int nb, uint16_t* Xc,
                   int ldc, void* stream);
int rmw_gnn_scatter(const float* dXc, int ldc, int nin, int nmsg, int deg, int nb, uint16_t* dE,
                    void* stream);
/* weighted mean-square error: per-block partial losses and dY = 2 w (Y - T) / rows */
int rmw_mse(const float* Y, const float* T, const float* w, int rows, int cols, uint16_t* dY,
            float* loss_part, int nparts, void* stream);
/* Adam step (bias-corrected, step >= 1) on FP32 weights; wb (BF16 copy) may be NULL */
int rmw_adam(float* w, const float* g, float* m, float* v, uint16_t* wb, int64_t n, float lr, float b1,
             float b2, float eps, int step, float weight_decay, void* stream);
/* deg nearest neighbours of every body centre (brute force) */
int rmw_gnn_knn(const double* centres, int nb, int deg, int* nbr, void* stream);
/* node (10) and edge (16 per neighbour) features in each body's frame */
int rmw_gnn_features(const double* centres, const double* quats, const double* force_torque,
                     const int* nbr, int nb, int deg, double fscale, float* Xn, float* Xe, void* stream);
/* node (22) and edge (16) features of the network that corrects the previous
 * step's solution: those of rmw_gnn_features plus the body-frame motion
 * since (centres0, quats0, force_torque0) -- displacement / dscale, rotation
 * vector / rscale, force and torque changes / fscale -- with the edge's
 * relative orientation replaced by the change of the neighbour offset */
int rmw_gnn_motion_features(const double* centres, const double* quats, const double* force_torque,
                            const double* centres0, const double* quats0, const double* force_torque0, const int* nbr,
                            int nb, int deg, double fscale, double dscale, double rscale, float* Xn, float* Xe,
                            void* stream);
/* mode 0: T = standardised body-frame (sol - pc) / fscale; mode 1: sol = pc + R (T sd + mean) fscale */
int rmw_gnn_body_frame(const double* quats, const int64_t* body_offsets_dev, int nb, int64_t nblobs,
                       int blobs_per_body, double* sol, const double* pc, float* T, const float* mean,
                       const float* sd, double fscale, int mode, void* stream);

/* Device properties used for launch sizing and roofline reporting. */
int rmw_device_info(int* cu_count, int* clock_khz, int* wavefront, char* arch_name, int arch_len);

#ifdef __cplusplus
}
#endif

#endif /* RMW_HIP_H */
