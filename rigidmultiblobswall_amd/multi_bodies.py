"""Implementation selection with the reference's option names.

  set_mobility_vector_prod(implementation)   multi_bodies/multi_bodies.py:233-287
  set_mobility_blobs(implementation)         multi_bodies/multi_bodies.py:207-230
  set_rollers_blocks(implementation)         the four block functions the rollers
                                             integrator binds by substring
                                             (quaternion_integrator_rollers.py:59-92)

New implementation values served by librmw_hip: 'hip' (FP32 pair terms, FP64
per-target sums), 'hip_fp32_blocked', 'hip_fp32_kahan' and 'hip_fp64', the
other boundaries 'hip_no_wall[_<precision>]' and
'hip_free_surface[_<precision>]', and for blobs of different radii
'radii_hip[_no_wall|_free_surface][_<precision>]'.  None
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
    """M*F at the blob level, signature f(r_vectors, force, eta, a, *args, **kwargs).

    'radii_hip', 'radii_hip_no_wall' and 'radii_hip_free_surface'
    (optionally with a precision suffix, e.g. 'radii_hip_fp64') select the
    blobs-of-different-radii product as the reference's 'radii_numba',
    'radii_numba_no_wall' and 'radii_numba_free_surface' do
    (multi_bodies.py:266-287); the radii come from kwargs['radius_blobs']
    (the reference reads them from the bodies)."""
    if implementation.startswith("radii_hip"):
        rest = implementation[len("radii_hip"):]
        fn = mb.single_wall_mobility_trans_times_force_source_target_hip
        for dom, dom_fn in (("_no_wall", mb.no_wall_mobility_trans_times_force_source_target_hip),
                            ("_free_surface", mb.free_surface_mobility_trans_times_force_source_target_hip)):
            if rest.startswith(dom):
                rest, fn = rest[len(dom):], dom_fn
        prec = _precision("hip" + rest)
        radius_blobs = kwargs.get("radius_blobs")
        if radius_blobs is None:
            raise ValueError("radii implementations need radius_blobs=")
        return partial(mb.mobility_radii_trans_times_force_hip, radius_blobs=radius_blobs,
                       function=partial(fn, precision=prec))
    for dom, fn in (("_no_wall", mb.no_wall_mobility_trans_times_force_hip),
                    ("_free_surface", mb.free_surface_mobility_trans_times_force_hip)):
        # 'hip_no_wall[_<precision>]', 'hip_free_surface[_<precision>]'
        # (the reference's 'numba_no_wall' / 'numba_free_surface', :250-265)
        if implementation.startswith("hip" + dom):
            return partial(fn, precision=_precision("hip" + implementation[len("hip" + dom):]))
    return partial(mb.single_wall_mobility_trans_times_force_hip, precision=_precision(implementation))


def set_mobility_blobs(implementation):
    """Dense (3N, 3N) blob mobility, signature g(r_vectors, eta, a), as
    multi_bodies.py:207-230 selects it: 'hip[_<precision>]' above the wall
    (the reference's 'python' / 'C++'), 'hip_no_wall[_<precision>]' the
    unbounded RPY tensor ('python_no_wall' / 'C++_no_wall').  The dense
    builder is always FP64."""
    if implementation.startswith("hip_no_wall"):
        _precision("hip" + implementation[len("hip_no_wall"):])
        return mb.rotne_prager_tensor_hip
    _precision(implementation)
    return mb.single_wall_fluid_mobility_hip


def set_rollers_blocks(implementation, domain: str = "single_wall"):
    """(UF, UT, WF, WT, UF+UT) block functions for a roller integrator, per
    domain as quaternion_integrator_rollers.py:59-92 binds them ('in_plane'
    has no rotational blocks: None there)."""
    p = _precision(implementation)
    if domain == "single_wall":
        fns = (mb.single_wall_mobility_trans_times_force_hip, mb.single_wall_mobility_trans_times_torque_hip,
               mb.single_wall_mobility_rot_times_force_hip, mb.single_wall_mobility_rot_times_torque_hip,
               mb.single_wall_mobility_trans_times_force_torque_hip)
    elif domain == "no_wall":
        fns = (mb.no_wall_mobility_trans_times_force_hip, mb.no_wall_mobility_trans_times_torque_hip,
               mb.no_wall_mobility_rot_times_force_hip, mb.no_wall_mobility_rot_times_torque_hip,
               mb.no_wall_mobility_trans_times_force_torque_hip)
    elif domain == "in_plane":
        fns = (mb.in_plane_mobility_trans_times_force_hip, mb.in_plane_mobility_trans_times_torque_hip,
               None, None, None)
    else:
        raise ValueError(f"unknown domain {domain!r}")
    return tuple(None if f is None else partial(f, precision=p) for f in fns)
