"""Blob-blob soft repulsion on the GPU (``rmw_blob_forces``) and drop-ins
for the reference's force functions.

  blob_forces(...)                       device tensors in and out
  calc_blob_blob_forces_hip              forces_numba.py:59-72 (calc_blob_blob_forces_numba)
  calc_blob_blob_forces_radii_hip        forces_numba.py:125-138 (different radii)
  calc_blob_blob_forces_tree_hip         forces_numba.py:200-241 (cutoff 2a + 30b)
  set_blob_blob_forces(implementation)   multi_bodies_functions.py:249-279

The force on blob i is sum_{j != i} f0(r) (r_j - r_i) with
f0 = -(eps/b) exp(-(r - 2a)/b) / r beyond contact (r > 2a) and -(eps/b)/r
inside it; the drop-ins read periodic_length, repulsion_strength,
debye_length and blob_radius from their keyword arguments exactly as the
reference's functions do and return an (N, 3) float64 numpy array.
"""
from __future__ import annotations

import ctypes
from functools import partial
from typing import Optional

import numpy as np
import torch

from . import _lib
from .wall_rpy import _L_array, _as_device_f64, _device, _stream_ptr, workspace

# implementation suffix -> precision mode of rmw_blob_forces
_PRECISIONS = {"": "fp32", "_fp32": "fp32", "_fp64": "fp64"}


def blob_forces(r_vectors, repulsion_strength: float, debye_length: float, blob_radius: float = 0.0,
                periodic_length=None, radius_blobs=None, cutoff: float = 0.0, precision: str = "fp32",
                tuning=None, out: Optional[torch.Tensor] = None) -> torch.Tensor:
    """Soft-repulsion forces (N, 3) float64 on the current GPU.

    radius_blobs (N,) switches to the different-radii form (a = (a_i+a_j)/2);
    cutoff > 0 keeps only pairs with r <= cutoff.  precision "fp64": FP64
    1/r (to 1.4e-16) and exp; "fp32" (any FP32 mode): FP64 geometry and
    sums, FP32 exponential."""
    lib = _lib.load()
    device = _device()
    r = torch.as_tensor(r_vectors)
    n = r.numel() // 3
    r = _as_device_f64(r, n, None, device, "r_vectors")
    rad = None
    if radius_blobs is not None:
        rad = torch.as_tensor(radius_blobs).to(device=device, dtype=torch.float64).contiguous().reshape(-1)
        if rad.numel() != n:
            raise ValueError(f"radius_blobs must have N = {n} entries, got {rad.numel()}")
    F = out if out is not None else torch.empty((n, 3), dtype=torch.float64, device=device)
    if F.numel() != 3 * n or F.dtype != torch.float64 or not F.is_contiguous() or F.device != device:
        raise ValueError("out must be a contiguous float64 (N, 3) tensor on the current GPU")
    prec = _lib.precision_code(precision)
    tp = _lib.tuning_ptr(tuning)
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_blob_forces_workspace_size(prec, n, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="forces")
    _lib.check(lib.rmw_blob_forces(prec, r.data_ptr(), rad.data_ptr() if rad is not None else None, F.data_ptr(),
                                   n, float(repulsion_strength), float(debye_length), float(blob_radius),
                                   float(cutoff), _L_array(periodic_length), tp, ws.data_ptr(), ws.numel(),
                                   _stream_ptr(device)))
    return F


def describe(n: int, precision: str = "fp32", tuning=None) -> dict:
    """The launch plan ``blob_forces`` would use for n blobs."""
    lib = _lib.load()
    _device()
    info = _lib.PlanInfo()
    _lib.check(lib.rmw_blob_forces_describe(_lib.precision_code(precision), int(n), _lib.tuning_ptr(tuning),
                                            ctypes.byref(info)))
    return {"kernel": "symmetric" if info.symmetric else "ordered", "chunk_len": info.chunk_len,
            "work_items": info.work_items, "workspace_bytes": info.workspace_bytes}


def _kw(kwargs):
    L = kwargs.get("periodic_length")
    return dict(periodic_length=L, repulsion_strength=kwargs.get("repulsion_strength"),
                debye_length=kwargs.get("debye_length"), blob_radius=kwargs.get("blob_radius"))


def calc_blob_blob_forces_hip(r_vectors, *args, precision: str = "fp32", **kwargs) -> np.ndarray:
    """Drop-in for calc_blob_blob_forces_numba (forces_numba.py:59-72)."""
    k = _kw(kwargs)
    return blob_forces(np.asarray(r_vectors, dtype=np.float64), k["repulsion_strength"], k["debye_length"],
                       k["blob_radius"], periodic_length=k["periodic_length"],
                       precision=precision).cpu().numpy()


def calc_blob_blob_forces_radii_hip(r_vectors, radius_blobs, *args, precision: str = "fp32", **kwargs) -> np.ndarray:
    """Drop-in for calc_blob_blob_forces_radii_numba (forces_numba.py:125-138):
    a = (a_i + a_j)/2 per pair."""
    k = _kw(kwargs)
    return blob_forces(np.asarray(r_vectors, dtype=np.float64), k["repulsion_strength"], k["debye_length"],
                       0.0, periodic_length=k["periodic_length"], radius_blobs=radius_blobs,
                       precision=precision).cpu().numpy()


def calc_blob_blob_forces_tree_hip(r_vectors, *args, precision: str = "fp32", **kwargs) -> np.ndarray:
    """Drop-in for calc_blob_blob_forces_tree_numba (forces_numba.py:200-241):
    only pairs closer than d_max = 2a + 30b interact.  Every pair is
    visited on the GPU, so no neighbour list is built or cached."""
    k = _kw(kwargs)
    d_max = 2 * k["blob_radius"] + 30 * k["debye_length"]
    return blob_forces(np.asarray(r_vectors, dtype=np.float64), k["repulsion_strength"], k["debye_length"],
                       k["blob_radius"], periodic_length=k["periodic_length"], cutoff=d_max,
                       precision=precision).cpu().numpy()


def default_zero_r_vectors(r_vectors, *args, **kwargs) -> np.ndarray:
    """No blob-blob interaction (multi_bodies_functions.py:84-90)."""
    return np.zeros((np.asarray(r_vectors).size // 3, 3))


def set_blob_blob_forces(implementation, *args, **kwargs):
    """Select the blob-blob force function by name, as
    multi_bodies_functions.py:249-279 does.  Served here: 'None', 'hip',
    'tree_hip' and 'radii_hip', each optionally suffixed '_fp32' or
    '_fp64'.  'radii_hip' takes the radii from kwargs['bodies'] (their
    blobs_radius, concatenated) or kwargs['radius_blobs']."""
    if implementation == "None":
        return default_zero_r_vectors
    for base, fn in (("tree_hip", calc_blob_blob_forces_tree_hip), ("radii_hip", calc_blob_blob_forces_radii_hip),
                     ("hip", calc_blob_blob_forces_hip)):
        if implementation.startswith(base) and implementation[len(base):] in _PRECISIONS:
            prec = _PRECISIONS[implementation[len(base):]]
            if base != "radii_hip":
                return partial(fn, precision=prec)
            radius_blobs = kwargs.get("radius_blobs")
            if radius_blobs is None:
                bodies = kwargs.get("bodies")
                if bodies is None:
                    raise ValueError("radii_hip needs bodies= or radius_blobs=")
                radius_blobs = np.concatenate([np.asarray(b.blobs_radius, dtype=np.float64).reshape(-1)
                                               for b in bodies])
            return partial(fn, radius_blobs=np.asarray(radius_blobs, dtype=np.float64), precision=prec)
    raise ValueError(f"unknown blob-blob force implementation {implementation!r}; this package provides "
                     "'None', 'hip', 'tree_hip', 'radii_hip' (suffix '_fp32' or '_fp64')")
