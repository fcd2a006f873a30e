"""Implementation selection with the reference's option names.

  set_mobility_vector_prod(implementation)   multi_bodies/multi_bodies.py:233-287
  set_mobility_blobs(implementation)         multi_bodies/multi_bodies.py:207-230
  set_rollers_blocks(implementation)         the four block functions the rollers
                                             integrator binds by substring
                                             (quaternion_integrator_rollers.py:59-92)

New implementation values served by librmw_hip: 'hip' (FP32 pair terms, FP64
per-target sums), 'hip_fp32_blocked', 'hip_fp32_kahan' and 'hip_fp64'.  None
contains 'numba' or 'pycuda', so the reference's substring dispatch never
mistakes them for its own backends.
"""
from __future__ import annotations

from functools import partial

from . import mobility as mb

IMPLEMENTATIONS = {
    "hip": "fp32",
    "hip_fp32": "fp32",
    "hip_fp32_blocked": "fp32_blocked",
    "hip_fp32_kahan": "fp32_kahan",
    "hip_fp64": "fp64",
}


def _precision(implementation: str) -> str:
    try:
        return IMPLEMENTATIONS[implementation]
    except KeyError:
        raise ValueError(f"unknown implementation {implementation!r}; this package provides "
                         f"{sorted(IMPLEMENTATIONS)}") from None


def set_mobility_vector_prod(implementation, *args, **kwargs):
    """M*F at the blob level, signature f(r_vectors, force, eta, a, *args, **kwargs)."""
    return partial(mb.single_wall_mobility_trans_times_force_hip, precision=_precision(implementation))


def set_mobility_blobs(implementation):
    """Dense (3N, 3N) blob mobility, signature g(r_vectors, eta, a).  The dense
    builder is always FP64."""
    _precision(implementation)
    return mb.single_wall_fluid_mobility_hip


def set_rollers_blocks(implementation):
    """(UF, UT, WF, WT, UF+UT) block functions for a roller integrator."""
    p = _precision(implementation)
    return (partial(mb.single_wall_mobility_trans_times_force_hip, precision=p),
            partial(mb.single_wall_mobility_trans_times_torque_hip, precision=p),
            partial(mb.single_wall_mobility_rot_times_force_hip, precision=p),
            partial(mb.single_wall_mobility_rot_times_torque_hip, precision=p),
            partial(mb.single_wall_mobility_trans_times_force_torque_hip, precision=p))
