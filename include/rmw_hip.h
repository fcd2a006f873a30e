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

/* Launch shape; pass NULL for the built-in heuristic. */
typedef struct rmw_tuning {
  int64_t chunk_len;  /* sources per work item (multiple of 256; 0 = auto) */
  int32_t max_wgs;    /* cap on resident workgroups of persistent launches (0 = auto) */
  int32_t reserved;
} rmw_tuning;

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
 * force/torque (k,n,3) or NULL when no block of the mask reads them;
 * velocity/omega (k,n,3) or NULL when no block writes them.  L[3] are the
 * pseudo-periodic lengths (host array; NULL or entries <= 0 = unbounded).
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
                       double eta, double a, const double* L, const rmw_tuning* tuning,
                       void* workspace, size_t workspace_bytes, void* stream);

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
                              const int64_t* offsets, int nsys, int k, double eta, double a,
                              const double* L, const rmw_tuning* tuning, void* workspace,
                              size_t workspace_bytes, void* stream);

/*
 * Optional timing of the pair kernels with HIP events recorded on the
 * launch stream.  rmw_profile_collect synchronises on the recorded events,
 * returns the number of pair-kernel launches since the last collect and
 * their total and mean duration in ms, and clears the record.
 */
int rmw_profile_enable(int on);
int rmw_profile_collect(int64_t* launches, double* total_ms, double* mean_ms);

/* Device properties used for launch sizing and roofline reporting. */
int rmw_device_info(int* cu_count, int* clock_khz, int* wavefront, char* arch_name, int arch_len);

#ifdef __cplusplus
}
#endif

#endif /* RMW_HIP_H */
