"""Device-level wall-RPY products (torch ROCm tensors in, torch tensors out).

Thin host layer over ``rmw_wall_rpy_apply`` / ``rmw_wall_rpy_ragged_apply``:
it validates shapes, owns a per-device workspace that only grows, and
passes raw device pointers and the current HIP stream.  PyTorch is used
only to own device memory and streams.
"""
from __future__ import annotations

import ctypes
import threading
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib
from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

_ws_lock = threading.Lock()
_workspaces: dict = {}


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("librmw_hip needs a ROCm GPU; none is visible (there is no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def workspace(nbytes: int, device: Optional[torch.device] = None, slot: str = "default") -> torch.Tensor:
    """A persistent uint8 device buffer of at least nbytes (grown, never shrunk)."""
    device = device or _device()
    key = (device.index, slot)
    with _ws_lock:
        buf = _workspaces.get(key)
        if buf is None or buf.numel() < nbytes:
            buf = torch.empty(max(int(nbytes), 256), dtype=torch.uint8, device=device)
            _workspaces[key] = buf
        return buf


def _as_device_f64(x, n: int, k: Optional[int], device, name: str) -> torch.Tensor:
    t = torch.as_tensor(x)
    if t.device != device or t.dtype != torch.float64:
        t = t.to(device=device, dtype=torch.float64)
    t = t.contiguous()
    if k is None:
        if t.numel() != 3 * n:
            raise ValueError(f"{name} must have 3*N = {3 * n} entries, got {t.numel()}")
    else:
        if t.numel() != 3 * n * k:
            raise ValueError(f"{name} must have k*3*N = {3 * n * k} entries, got {t.numel()}")
    return t


def _L_array(periodic_length) -> ctypes.Array:
    L = np.zeros(3) if periodic_length is None else np.asarray(periodic_length, dtype=np.float64).reshape(-1)
    if L.size != 3:
        raise ValueError("periodic_length must have 3 entries")
    return (ctypes.c_double * 3)(*[float(v) for v in L])


def _stream_ptr(device) -> int:
    return torch.cuda.current_stream(device).cuda_stream


def _tuning_for(tuning, mask: int, precision: str, k: int, pairs: float, ragged: bool):
    """Launch shape for a call: the caller's, the autotuner's cached one
    (tuning=None), or the library's built-in heuristic (tuning="heuristic")."""
    if isinstance(tuning, str):
        if tuning != "heuristic":
            raise ValueError("tuning must be None, 'heuristic', a dict or an rmw_tuning")
        return None
    if tuning is not None:
        return tuning
    from . import autotune
    return autotune.lookup(mask, precision if isinstance(precision, str) else str(precision), k, pairs, ragged)


# Boundaries of the fluid (the reference's `domain` values, plus the free
# surface of its 'numba_free_surface' implementations).
DOMAINS = {"single_wall": 0, "no_wall": 1, "in_plane": 2, "free_surface": 3}


def apply(mask: int, r_vectors, force=None, torque=None, eta: float = 1.0, a: float = 1.0,
          periodic_length=None, precision="fp32", k: int = 1, tuning=None,
          out_velocity: Optional[torch.Tensor] = None, out_omega: Optional[torch.Tensor] = None,
          domain: str = "single_wall"):
    """Evaluate the blocks of ``mask`` on one suspension.

    r_vectors: (N, 3); force / torque: (k, N, 3) (or (N, 3) for k = 1).
    Returns (velocity, omega) as float64 device tensors of shape (k, N, 3)
    (squeezed to (N, 3) when k == 1), None for outputs the mask does not write.
    domain: 'single_wall' (default), 'no_wall' (all blocks), 'in_plane'
    (UF, UT) or 'free_surface' (UF).
    """
    if domain not in DOMAINS:
        raise ValueError(f"unknown domain {domain!r}; use one of {sorted(DOMAINS)}")
    lib = _lib.load()
    device = _device()
    r = torch.as_tensor(r_vectors)
    n = r.numel() // 3
    r = _as_device_f64(r, n, None, device, "r_vectors")
    has_u = bool(mask & (RMW_UF | RMW_UT))
    has_w = bool(mask & (RMW_WF | RMW_WT))
    f = _as_device_f64(force, n, k, device, "force") if mask & (RMW_UF | RMW_WF) else None
    t = _as_device_f64(torque, n, k, device, "torque") if mask & (RMW_UT | RMW_WT) else None
    u = out_velocity if out_velocity is not None else (
        torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_u else None)
    w = out_omega if out_omega is not None else (
        torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_w else None)
    prec = _lib.precision_code(precision)
    dom = DOMAINS[domain]
    if dom == 0:
        tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, float(n) * n, False))
    else:
        # the other domains run the ordered kernel; the cache holds wall shapes
        tp = _lib.tuning_ptr(None if tuning is None or isinstance(tuning, str) else tuning)
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_mobility_workspace_size(dom, mask, prec, n, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device)
    _lib.check(lib.rmw_mobility_apply(
        dom, mask, prec, r.data_ptr(), f.data_ptr() if f is not None else None,
        t.data_ptr() if t is not None else None, u.data_ptr() if u is not None else None,
        w.data_ptr() if w is not None else None, n, k, 0, float(eta), float(a), _L_array(periodic_length),
        tp, ws.data_ptr(), ws.numel(), _stream_ptr(device)))
    if k == 1:
        u = u.view(n, 3) if u is not None else None
        w = w.view(n, 3) if w is not None else None
    return u, w


def apply_strided(mask: int, r: torch.Tensor, n: int, k: int, ld: int, force_ptr, torque_ptr, vel_ptr, omg_ptr,
                  eta: float, a: float, periodic_length=None, precision="fp32", tuning=None):
    """Low-level product on raw device pointers with k right-hand sides `ld`
    doubles apart (e.g. force and torque halves of stacked 6N vectors)."""
    lib = _lib.load()
    device = _device()
    prec = _lib.precision_code(precision)
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, float(n) * n, False))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_wall_rpy_workspace_size(mask, prec, n, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device)
    _lib.check(lib.rmw_wall_rpy_apply(mask, prec, r.data_ptr(), force_ptr, torque_ptr, vel_ptr, omg_ptr, n, k, ld,
                                      float(eta), float(a), _L_array(periodic_length), tp, ws.data_ptr(),
                                      ws.numel(), _stream_ptr(device)))


def apply_ragged(mask: int, r_vectors, offsets: Sequence[int], force=None, torque=None, eta: float = 1.0,
                 a: float = 1.0, periodic_length=None, precision="fp32", k: int = 1, tuning=None):
    """Independent suspensions stored back to back; system s owns blobs
    [offsets[s], offsets[s+1]).  One persistent launch for all systems."""
    lib = _lib.load()
    device = _device()
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    nsys = off.size - 1
    n = int(off[-1])
    r = _as_device_f64(r_vectors, n, None, device, "r_vectors")
    has_u = bool(mask & (RMW_UF | RMW_UT))
    has_w = bool(mask & (RMW_WF | RMW_WT))
    f = _as_device_f64(force, n, k, device, "force") if mask & (RMW_UF | RMW_WF) else None
    t = _as_device_f64(torque, n, k, device, "torque") if mask & (RMW_UT | RMW_WT) else None
    u = torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_u else None
    w = torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_w else None
    prec = _lib.precision_code(precision)
    pairs = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, pairs, True))
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_wall_rpy_ragged_workspace_size(mask, prec, offp, nsys, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="ragged")
    _lib.check(lib.rmw_wall_rpy_ragged_apply(
        mask, prec, r.data_ptr(), f.data_ptr() if f is not None else None,
        t.data_ptr() if t is not None else None, u.data_ptr() if u is not None else None,
        w.data_ptr() if w is not None else None, offp, nsys, k, 0, float(eta), float(a),
        _L_array(periodic_length), tp, ws.data_ptr(), ws.numel(), _stream_ptr(device)))
    if k == 1:
        u = u.view(n, 3) if u is not None else None
        w = w.view(n, 3) if w is not None else None
    return u, w


def apply_ragged_strided(mask: int, r: torch.Tensor, offsets: np.ndarray, k: int, ld: int, force_ptr, torque_ptr,
                         vel_ptr, omg_ptr, eta: float, a: float, periodic_length=None, precision="fp32_blocked",
                         tuning=None):
    """Ragged product on raw device pointers, k right-hand sides `ld` doubles
    apart (force and torque parts of stacked grand vectors of a batch)."""
    lib = _lib.load()
    device = _device()
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    prec = _lib.precision_code(precision)
    pairs = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, pairs, True))
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_wall_rpy_ragged_workspace_size(mask, prec, offp, off.size - 1, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="ragged")
    _lib.check(lib.rmw_wall_rpy_ragged_apply(mask, prec, r.data_ptr(), force_ptr, torque_ptr, vel_ptr, omg_ptr, offp,
                                             off.size - 1, k, ld, float(eta), float(a), _L_array(periodic_length),
                                             tp, ws.data_ptr(), ws.numel(), _stream_ptr(device)))


# boundary argument of the different-radii products (rmw_radii_apply)
RADII_BOUNDARIES = {"unbounded": 0, "wall": 1, "free_surface": 2}


def _radii_boundary(wall, boundary) -> int:
    if boundary is None:
        return RADII_BOUNDARIES["wall" if wall else "unbounded"]
    if boundary not in RADII_BOUNDARIES:
        raise ValueError(f"unknown boundary {boundary!r}; use one of {sorted(RADII_BOUNDARIES)}")
    return RADII_BOUNDARIES[boundary]


def apply_radii(source, target, force, radius_source, radius_target, eta: float, periodic_length=None,
                wall: bool = True, precision="fp32_blocked", tuning=None, out: Optional[torch.Tensor] = None,
                boundary: Optional[str] = None):
    """Different radii, source -> target: u (N_t, 3) = B_t M B_s f (wall) or
    the unbounded RPY of different radii (wall=False).  boundary
    ('wall', 'unbounded' or 'free_surface') overrides wall; 'free_surface'
    is the RPY of d plus the RPY of the image vector (d_x, d_y, z_i + z_j)
    times diag(1, 1, -1), heights used as given."""
    bnd = _radii_boundary(wall, boundary)
    lib = _lib.load()
    device = _device()
    src = torch.as_tensor(source)
    ns = src.numel() // 3
    trg = torch.as_tensor(target)
    nt = trg.numel() // 3
    src = _as_device_f64(src, ns, None, device, "source")
    trg = _as_device_f64(trg, nt, None, device, "target")
    f = _as_device_f64(force, ns, None, device, "force")
    rs = torch.as_tensor(radius_source).to(device=device, dtype=torch.float64).contiguous().reshape(-1)
    rt = torch.as_tensor(radius_target).to(device=device, dtype=torch.float64).contiguous().reshape(-1)
    if rs.numel() != ns or rt.numel() != nt:
        raise ValueError("one radius per source and per target blob is required")
    u = out if out is not None else torch.empty((nt, 3), dtype=torch.float64, device=device)
    prec = _lib.precision_code(precision)
    tp = _lib.tuning_ptr(None if isinstance(tuning, str) else tuning)
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_radii_workspace_size(prec, ns, nt, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="radii")
    _lib.check(lib.rmw_radii_apply(prec, bnd, src.data_ptr(), trg.data_ptr(), f.data_ptr(),
                                   rs.data_ptr(), rt.data_ptr(), u.data_ptr(), ns, nt, float(eta),
                                   _L_array(periodic_length), tp, ws.data_ptr(), ws.numel(), _stream_ptr(device)))
    return u


def apply_radii_self(r_vectors, force, radius, eta: float, periodic_length=None, wall: bool = True,
                     precision="fp32_blocked", tuning=None, out: Optional[torch.Tensor] = None,
                     boundary: Optional[str] = None):
    """Different radii, sources = targets: u (N, 3) = B M B f (wall) or the
    unbounded RPY of different radii (wall=False); boundary as apply_radii.
    Each unordered pair is evaluated once when the symmetric kernel is
    chosen (no periodic box)."""
    bnd = _radii_boundary(wall, boundary)
    lib = _lib.load()
    device = _device()
    r = torch.as_tensor(r_vectors)
    n = r.numel() // 3
    r = _as_device_f64(r, n, None, device, "r_vectors")
    f = _as_device_f64(force, n, None, device, "force")
    rad = torch.as_tensor(radius).to(device=device, dtype=torch.float64).contiguous().reshape(-1)
    if rad.numel() != n:
        raise ValueError("one radius per blob is required")
    u = out if out is not None else torch.empty((n, 3), dtype=torch.float64, device=device)
    prec = _lib.precision_code(precision)
    tp = _lib.tuning_ptr(None if isinstance(tuning, str) else tuning)
    L = _L_array(periodic_length)
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_radii_self_workspace_size(prec, n, L, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="radii")
    _lib.check(lib.rmw_radii_self_apply(prec, bnd, r.data_ptr(), f.data_ptr(), rad.data_ptr(),
                                        u.data_ptr(), n, float(eta), L, tp, ws.data_ptr(), ws.numel(),
                                        _stream_ptr(device)))
    return u


def describe_radii_self(n: int, precision="fp32_blocked", periodic_length=None, tuning=None) -> dict:
    """The launch apply_radii_self would use."""
    lib = _lib.load()
    _device()
    tp = _lib.tuning_ptr(None if isinstance(tuning, str) else tuning)
    sym, items, nbytes = ctypes.c_int(), ctypes.c_int64(), ctypes.c_size_t()
    _lib.check(lib.rmw_radii_self_describe(_lib.precision_code(precision), n, _L_array(periodic_length), tp,
                                           ctypes.byref(sym), ctypes.byref(items), ctypes.byref(nbytes)))
    return {"kernel": "symmetric" if sym.value else "ordered", "work_items": items.value,
            "workspace_bytes": nbytes.value}


def describe(mask: int, n: int, k: int = 1, precision="fp32", periodic_length=None, tuning=None) -> dict:
    """The launch plan ``apply`` would use (kernel, RHS per pass, chunk,
    work items), with the same tuning resolution as ``apply``."""
    lib = _lib.load()
    _device()
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, float(n) * n, False))
    L = np.zeros(3) if periodic_length is None else np.asarray(periodic_length, dtype=np.float64)
    info = _lib.PlanInfo()
    _lib.check(lib.rmw_wall_rpy_describe(mask, _lib.precision_code(precision), n, k, tp, int(bool(np.any(L > 0))),
                                         ctypes.byref(info)))
    return {"kernel": "symmetric" if info.symmetric else "ordered", "rhs_group": info.rhs_group,
            "chunk_len": info.chunk_len, "work_items": info.work_items, "workspace_bytes": info.workspace_bytes}


_info_cache: dict = {}


def device_info() -> dict:
    dev = _device()
    if dev.index in _info_cache:
        return dict(_info_cache[dev.index])
    lib = _lib.load()
    cu, clk, wf = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
    name = ctypes.create_string_buffer(64)
    _lib.check(lib.rmw_device_info(ctypes.byref(cu), ctypes.byref(clk), ctypes.byref(wf), name, 64))
    info = {"cu_count": cu.value, "clock_khz": clk.value, "wavefront": wf.value, "arch": name.value.decode()}
    _info_cache[dev.index] = info
    return dict(info)


PROFILE_KINDS = ("wall_rpy", "radii", "forces")


class PairKernelTimer:
    """Context manager timing every pair-kernel launch with HIP events.

    After the block: launches / total_ms / mean_ms over all pair kernels and
    ``by_kind[name] = (launches, total_ms, mean_ms)`` per kernel family
    (wall RPY, different radii, blob-blob forces)."""

    def __enter__(self):
        lib = _lib.load()
        _lib.check(lib.rmw_profile_collect(None, None, None))
        _lib.check(lib.rmw_profile_enable(1))
        return self

    def __exit__(self, *exc):
        lib = _lib.load()
        _lib.check(lib.rmw_profile_enable(0))
        n = (ctypes.c_int64 * 3)()
        tot = (ctypes.c_double * 3)()
        _lib.check(lib.rmw_profile_collect_kinds(n, tot))
        self.by_kind = {k: (n[i], tot[i], tot[i] / n[i] if n[i] else 0.0) for i, k in enumerate(PROFILE_KINDS)}
        self.launches = sum(n)
        self.total_ms = sum(tot)
        self.mean_ms = self.total_ms / self.launches if self.launches else 0.0
        return False
