"""Drop-in wall-RPY mobility products with the reference signatures.

Every function has the calling convention of its reference counterpart in
mobility/mobility.py -- ``f(r_vectors, vector, eta, a, *args, **kwargs)``
with ``periodic_length=np.array([Lx, Ly, Lz])`` (zeros = unbounded) -- and
returns a new (3N,) float64 numpy array, so it can be bound wherever the
reference binds ``single_wall_mobility_*_numba`` / ``*_pycuda``.

Extra keyword arguments understood here:
  precision: 'fp32' (FP32 pair terms, FP64 per-target sums; default),
             'fp32_kahan' (compensated FP32 sums) or 'fp64'.
  tuning:    {'chunk_len': int, 'max_wgs': int} launch shape override.
  check_finite: raise RmwError (code RMW_E_NOTFINITE, -4) when an output
             entry is inf or NaN -- e.g. two distinct blobs at one position,
             which the reference returns silently as NaN
             (mobility/mobility_numba.py:88).  Off by default, as there.

The computation always runs on the GPU through librmw_hip; inputs given as
numpy arrays are copied to the device and results copied back.  Inputs
given as torch device tensors stay on the device (the copy back to the
host is the only transfer).
"""
from __future__ import annotations

import functools

import numpy as np
import torch

from . import wall_rpy
from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

DEFAULT_PRECISION = "fp32"


def shift_heights(r_vectors, blob_radius, *args, **kwargs):
    """z_eff = max(z, a); mobility/mobility.py:52-64 (host helper)."""
    r_effective = np.copy(r_vectors)
    r_effective[r_vectors[:, 2] <= blob_radius, 2] = blob_radius
    return r_effective


def damping_matrix_B(r_vectors, blob_radius, *args, **kwargs):
    """Diagonal B_ii = z_i/a below the wall-overlap height, else 1, and the
    overlap flag; mobility/mobility.py:67-84 (host helper, returned as the
    diagonal vector of length 3N rather than a scipy dia_matrix)."""
    r = np.asarray(r_vectors).reshape(-1, 3)
    B = np.ones(r.size)
    low = r[:, 2] < blob_radius
    B.reshape(-1, 3)[low] = (r[low, 2] / blob_radius)[:, None]
    return B, bool(low.any())


def _opts(kwargs):
    return (kwargs.get("periodic_length", np.array([0.0, 0.0, 0.0])),
            kwargs.get("precision", DEFAULT_PRECISION), kwargs.get("tuning"))


def _apply_opts(kwargs):
    """Keyword arguments of wall_rpy.apply from a drop-in's **kwargs."""
    L, prec, tun = _opts(kwargs)
    return dict(periodic_length=L, precision=prec, tuning=tun, check_finite=bool(kwargs.get("check_finite", False)))


def _host(t):
    return t.reshape(-1).cpu().numpy()


def single_wall_mobility_trans_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """u = B M_UF(z_eff) B f; drop-in for mobility.py:1132-1163."""
    u, _ = wall_rpy.apply(RMW_UF, r_vectors, force=force, eta=eta, a=a, **_apply_opts(kwargs))
    return _host(u)


def single_wall_mobility_trans_times_torque_hip(r_vectors, torque, eta, a, *args, **kwargs):
    """u = B M_UT(z_eff) B t; drop-in for mobility.py:1213-1232."""
    u, _ = wall_rpy.apply(RMW_UT, r_vectors, torque=torque, eta=eta, a=a, **_apply_opts(kwargs))
    return _host(u)


def single_wall_mobility_rot_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """w = B M_WF(z_eff) B f; drop-in for mobility.py:1273-1298."""
    _, w = wall_rpy.apply(RMW_WF, r_vectors, force=force, eta=eta, a=a, **_apply_opts(kwargs))
    return _host(w)


def single_wall_mobility_rot_times_torque_hip(r_vectors, torque, eta, a, *args, **kwargs):
    """w = B M_WT(z_eff) B t; drop-in for mobility.py:1316-1342."""
    _, w = wall_rpy.apply(RMW_WT, r_vectors, torque=torque, eta=eta, a=a, **_apply_opts(kwargs))
    return _host(w)


def single_wall_mobility_trans_times_force_torque_hip(r_vectors, force, torque, eta, a, *args, **kwargs):
    """u = M_UF f + M_UT t in one pass; drop-in for mobility.py:384-410."""
    u, _ = wall_rpy.apply(RMW_UF | RMW_UT, r_vectors, force=force, torque=torque, eta=eta, a=a, **_apply_opts(kwargs))
    return _host(u)


def single_wall_grand_mobility_hip(r_vectors, force_torque, eta, a, *args, **kwargs):
    """[UF UT; WF WT] [f; t] -> [u; w] in one fused pass; drop-in for the
    rollers' grand_mobility_matrix closure
    (quaternion_integrator/quaternion_integrator_rollers.py:1114-1122)."""
    ft = np.asarray(force_torque, dtype=np.float64).reshape(-1) if not torch.is_tensor(force_torque) \
        else force_torque.reshape(-1)
    half = ft.shape[0] // 2
    u, w = wall_rpy.apply(RMW_UF | RMW_UT | RMW_WF | RMW_WT, r_vectors, force=ft[:half], torque=ft[half:],
                          eta=eta, a=a, **_apply_opts(kwargs))
    return np.concatenate([_host(u), _host(w)])


def _product(domain, mask, r_vectors, force, torque, eta, a, kwargs, out="u"):
    u, w = wall_rpy.apply(mask, r_vectors, force=force, torque=torque, eta=eta, a=a, domain=domain,
                          **_apply_opts(kwargs))
    return _host(u if out == "u" else w)


def no_wall_mobility_trans_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """Unbounded RPY u = M_UF f; drop-in for mobility.py:1119-1129."""
    return _product("no_wall", RMW_UF, r_vectors, force, None, eta, a, kwargs)


def no_wall_mobility_trans_times_torque_hip(r_vectors, torque, eta, a, *args, **kwargs):
    """Unbounded u = M_UT t; drop-in for mobility.py:1200-1210."""
    return _product("no_wall", RMW_UT, r_vectors, None, torque, eta, a, kwargs)


def no_wall_mobility_rot_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """Unbounded w = M_WF f; drop-in for mobility.py:1258-1270."""
    return _product("no_wall", RMW_WF, r_vectors, force, None, eta, a, kwargs, out="w")


def no_wall_mobility_rot_times_torque_hip(r_vectors, torque, eta, a, *args, **kwargs):
    """Unbounded w = M_WT t; drop-in for mobility.py:1301-1313."""
    return _product("no_wall", RMW_WT, r_vectors, None, torque, eta, a, kwargs, out="w")


def no_wall_mobility_trans_times_force_torque_hip(r_vectors, force, torque, eta, a, *args, **kwargs):
    """Unbounded u = M_UF f + M_UT t in one pass; drop-in for mobility.py:413-425."""
    return _product("no_wall", RMW_UF | RMW_UT, r_vectors, force, torque, eta, a, kwargs)


def in_plane_mobility_trans_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """u = P B M_UF(z_eff) B P f (x-y components only, u_z = 0); drop-in for
    mobility.py:1166-1198."""
    return _product("in_plane", RMW_UF, r_vectors, force, None, eta, a, kwargs)


def in_plane_mobility_trans_times_torque_hip(r_vectors, torque, eta, a, *args, **kwargs):
    """u = P B M_UT(z_eff) B P t; drop-in for mobility.py:1235-1256."""
    return _product("in_plane", RMW_UT, r_vectors, None, torque, eta, a, kwargs)


def free_surface_mobility_trans_times_force_hip(r_vectors, force, eta, a, *args, **kwargs):
    """u = [RPY(r_i - r_j) + RPY(r_i - r_j*) D] f below a free surface at
    z = 0 (image r_j* = (x_j, y_j, -z_j)); drop-in for mobility.py:1390-1406."""
    return _product("free_surface", RMW_UF, r_vectors, force, None, eta, a, kwargs)


__all__ = [
    "shift_heights", "damping_matrix_B",
    "single_wall_mobility_trans_times_force_hip", "single_wall_mobility_trans_times_torque_hip",
    "single_wall_mobility_rot_times_force_hip", "single_wall_mobility_rot_times_torque_hip",
    "single_wall_mobility_trans_times_force_torque_hip", "single_wall_grand_mobility_hip",
    "no_wall_mobility_trans_times_force_hip", "no_wall_mobility_trans_times_torque_hip",
    "no_wall_mobility_rot_times_force_hip", "no_wall_mobility_rot_times_torque_hip",
    "no_wall_mobility_trans_times_force_torque_hip", "in_plane_mobility_trans_times_force_hip",
    "in_plane_mobility_trans_times_torque_hip", "free_surface_mobility_trans_times_force_hip",
]


def _dense_mobility(domain: int, r_vectors, eta, a):
    import ctypes
    from . import _lib
    lib = _lib.load()
    r = torch.as_tensor(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3) if not torch.is_tensor(r_vectors)
                        else r_vectors.reshape(-1, 3)).to(dtype=torch.float64)
    device = torch.device("cuda", torch.cuda.current_device())
    r = r.to(device).contiguous()
    n = r.shape[0]
    offs = np.array([0, n], dtype=np.int64)
    offp = offs.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_rigid_workspace_size(offp, 1, ctypes.byref(nbytes), None, None))
    ws = wall_rpy.workspace(nbytes.value, device, slot="dense")
    out = torch.empty((3 * n, 3 * n), dtype=torch.float64, device=device)
    stream = torch.cuda.current_stream(device).cuda_stream
    _lib.check(lib.rmw_rigid_setup(offp, 1, ws.data_ptr(), ws.numel(), stream))
    _lib.check(lib.rmw_dense_mobility(domain, r.data_ptr(), offp, 1, float(eta), float(a), out.data_ptr(),
                                      ws.data_ptr(), ws.numel(), stream))
    return out.cpu().numpy()


def single_wall_fluid_mobility_hip(r_vectors, eta, a, *args, **kwargs):
    """Dense 3N x 3N wall mobility B M(z_eff) B (FP64, built on the GPU);
    drop-in for single_wall_fluid_mobility, mobility/mobility.py:1018-1116,
    as selected by mobility_blobs_implementation."""
    return _dense_mobility(wall_rpy.DOMAINS["single_wall"], r_vectors, eta, a)


def rotne_prager_tensor_hip(r_vectors, eta, a, *args, **kwargs):
    """Dense 3N x 3N unbounded RPY mobility (FP64, built on the GPU); drop-in
    for rotne_prager_tensor, mobility/mobility.py:967-1012, as selected by
    mobility_blobs_implementation 'python_no_wall' / 'C++_no_wall'."""
    return _dense_mobility(wall_rpy.DOMAINS["no_wall"], r_vectors, eta, a)


__all__ += ["single_wall_fluid_mobility_hip", "rotne_prager_tensor_hip"]


# ---------------------------------------------------------------------------
# Blobs of different radii (source -> target), the 'radii*' implementations
# ---------------------------------------------------------------------------
def single_wall_mobility_trans_times_force_source_target_hip(source, target, force, radius_source, radius_target,
                                                             eta, *args, **kwargs):
    """u = B_t M(z_eff) B_s f for blobs of different radii above the wall;
    drop-in for mobility.py:551-591 (kernel mobility_numba.py:1479-1658)."""
    L, prec, tun = _opts(kwargs)
    u = wall_rpy.apply_radii(source, target, force, radius_source, radius_target, eta, periodic_length=L, wall=True,
                             precision=prec, tuning=tun)
    return _host(u)


def no_wall_mobility_trans_times_force_source_target_hip(source, target, force, radius_source, radius_target,
                                                         eta, *args, **kwargs):
    """Unbounded RPY of different radii; drop-in for mobility.py:593-615."""
    L, prec, tun = _opts(kwargs)
    u = wall_rpy.apply_radii(source, target, force, radius_source, radius_target, eta, periodic_length=L,
                             wall=False, precision=prec, tuning=tun)
    return _host(u)


def free_surface_mobility_trans_times_force_source_target_hip(source, target, force, radius_source, radius_target,
                                                              eta, *args, **kwargs):
    """Blobs of different radii below a free surface: RPY of d plus the RPY
    of the image vector (d_x, d_y, z_i + z_j) times diag(1, 1, -1), heights
    as given; drop-in for mobility.py:1409-1429 (kernel
    mobility_numba.py:1941-2093)."""
    L, prec, tun = _opts(kwargs)
    u = wall_rpy.apply_radii(source, target, force, radius_source, radius_target, eta, periodic_length=L,
                             boundary="free_surface", precision=prec, tuning=tun)
    return _host(u)


def mobility_radii_trans_times_force_hip(r_vectors, force, eta, a, radius_blobs, function=None, *args, **kwargs):
    """M f with per-blob radii, sources = targets; drop-in for
    mobility.py:1369-1374 (``function`` defaults to the wall kernel).  With
    the wall, unbounded or free-surface kernel of this module the product
    runs as one sources-equal-targets call (each unordered pair evaluated
    once)."""
    fn = function or single_wall_mobility_trans_times_force_source_target_hip
    own = {single_wall_mobility_trans_times_force_source_target_hip: "wall",
           no_wall_mobility_trans_times_force_source_target_hip: "unbounded",
           free_surface_mobility_trans_times_force_source_target_hip: "free_surface"}
    base, opts = fn, dict(kwargs)
    if isinstance(fn, functools.partial) and not fn.args:
        base, opts = fn.func, {**fn.keywords, **kwargs}
    if base in own and not args:
        L, prec, tun = _opts(opts)
        u = wall_rpy.apply_radii_self(r_vectors, force, radius_blobs, eta, periodic_length=L, boundary=own[base],
                                      precision=prec, tuning=tun)
        return _host(u)
    return fn(r_vectors, r_vectors, force, radius_blobs, radius_blobs, eta, *args, **kwargs)
