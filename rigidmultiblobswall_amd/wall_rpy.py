"""Device-level wall-RPY products (torch ROCm tensors in, torch tensors out).

Thin host layer over ``rmw_wall_rpy_apply`` / ``rmw_wall_rpy_ragged_apply``:
it validates shapes, owns a workspace per (device, stream) that only grows,
and passes raw device pointers and the current HIP stream.  PyTorch is used
only to own device memory and streams.
"""
from __future__ import annotations

import ctypes
import threading
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib, _stream
from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

_ws_lock = threading.Lock()
_workspaces: dict = {}


def _device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("librmw_hip needs a ROCm GPU; none is visible (there is no CPU fallback)")
    return torch.device("cuda", torch.cuda.current_device())


def workspace(nbytes: int, device: Optional[torch.device] = None, slot: str = "default") -> torch.Tensor:
    """A persistent uint8 device buffer of at least nbytes (grown, never
    shrunk), one per (device, current stream, slot): calls issued on
    different streams never share scratch, so they may overlap.  A buffer is
    allocated on, used on and (when grown) released on its own stream."""
    device = device or _device()
    return _workspace_of((device.index, _stream.current(device), slot), nbytes, device)


def _workspace_of(key, nbytes: int, device) -> torch.Tensor:
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


def _check_out(t: torch.Tensor, count: int, device, name: str) -> torch.Tensor:
    """A caller-supplied output buffer must be a contiguous float64 tensor on
    the launch device with exactly `count` entries: the kernels write
    through its raw pointer."""
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{name} must be a torch tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64, got {t.dtype}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, the product runs on {device}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if t.numel() != count:
        raise ValueError(f"{name} must have {count} entries, got {t.numel()}")
    return t


def _L_array(periodic_length) -> ctypes.Array:
    L = np.zeros(3) if periodic_length is None else np.asarray(periodic_length, dtype=np.float64).reshape(-1)
    if L.size != 3:
        raise ValueError("periodic_length must have 3 entries")
    return (ctypes.c_double * 3)(*[float(v) for v in L])


def _stream_ptr(device) -> int:
    return _stream.current(device)


def _periodic(periodic_length) -> bool:
    return periodic_length is not None and bool(np.any(np.asarray(periodic_length, dtype=np.float64) > 0))


def _tuning_for(tuning, mask: int, precision: str, k: int, pairs: float, ragged: bool, periodic: bool = False):
    """Launch shape for a call: the caller's, the autotuner's cached one
    (tuning=None), or the library's built-in heuristic (tuning="heuristic").
    The cache holds shapes timed without periodic images (each pair then
    costs 9-27 image evaluations), so calls with images use the built-in
    heuristic unless the caller passes a shape."""
    if isinstance(tuning, str):
        if tuning != "heuristic":
            raise ValueError("tuning must be None, 'heuristic', a dict or an rmw_tuning")
        return None
    if tuning is not None:
        return tuning
    if periodic:
        return None
    from . import autotune
    return autotune.lookup(mask, precision if isinstance(precision, str) else str(precision), k, pairs, ragged)


# Boundaries of the fluid (the reference's `domain` values, plus the free
# surface of its 'numba_free_surface' implementations).
DOMAINS = {"single_wall": 0, "no_wall": 1, "in_plane": 2, "free_surface": 3}


def apply(mask: int, r_vectors, force=None, torque=None, eta: float = 1.0, a: float = 1.0,
          periodic_length=None, precision="fp32", k: int = 1, tuning=None,
          out_velocity: Optional[torch.Tensor] = None, out_omega: Optional[torch.Tensor] = None,
          domain: str = "single_wall", check_finite: bool = False):
    """Evaluate the blocks of ``mask`` on one suspension.

    r_vectors: (N, 3); force / torque: (k, N, 3) (or (N, 3) for k = 1).
    Returns (velocity, omega) as float64 device tensors of shape (k, N, 3)
    (squeezed to (N, 3) when k == 1), None for outputs the mask does not write.
    domain: 'single_wall' (default), 'no_wall' (all blocks), 'in_plane'
    (UF, UT) or 'free_surface' (UF).
    check_finite: scan the outputs after the product (the call then waits for
    the stream) and raise RmwError with code RMW_E_NOTFINITE if an entry is
    inf or NaN.
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
    if out_velocity is not None and has_u:
        u = _check_out(out_velocity, 3 * n * k, device, "out_velocity")
    else:
        u = torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_u else None
    if out_omega is not None and has_w:
        w = _check_out(out_omega, 3 * n * k, device, "out_omega")
    else:
        w = torch.empty((k, n, 3), dtype=torch.float64, device=device) if has_w else None
    prec = _lib.precision_code(precision)
    dom = DOMAINS[domain]
    if dom == 0:
        tun = _tuning_for(tuning, mask, precision, k, float(n) * n, False, _periodic(periodic_length))
    else:
        # the cache holds wall shapes
        tun = None if tuning is None or isinstance(tuning, str) else tuning
    if check_finite:
        tun = _lib.with_flags(tun, _lib.RMW_FLAG_CHECK_FINITE)
    tp = _lib.tuning_ptr(tun)
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


class Product:
    """A prepared product of the blocks of ``mask`` on the positions held by
    ``r`` (a contiguous float64 (N, 3) device tensor, read at every call, so
    it may be rewritten in place between calls): the shape checks, the launch
    shape lookup and the workspace size are settled once, and each call
    passes the right-hand side and output pointers and the current stream.
    Products of a Krylov operator run one after another from Python; this
    keeps the host time between them short.  Same arguments and results as
    apply() with k = 1 and caller-supplied outputs."""

    def __init__(self, mask: int, r: torch.Tensor, eta: float = 1.0, a: float = 1.0, periodic_length=None,
                 precision="fp32", tuning=None, domain: str = "single_wall"):
        if domain not in DOMAINS:
            raise ValueError(f"unknown domain {domain!r}; use one of {sorted(DOMAINS)}")
        self.lib = _lib.load()
        self.device = _device()
        if not isinstance(r, torch.Tensor) or r.dtype != torch.float64 or r.device != self.device or \
                not r.is_contiguous() or r.numel() % 3:
            raise ValueError(f"r must be a contiguous float64 (N, 3) tensor on {self.device}")
        self.r = r
        self.n = n = r.numel() // 3
        self.mask = int(mask)
        self.has_u = bool(mask & (RMW_UF | RMW_UT))
        self.has_w = bool(mask & (RMW_WF | RMW_WT))
        self.dom = DOMAINS[domain]
        self.prec = _lib.precision_code(precision)
        if self.dom == 0:
            self.tuning = _tuning_for(tuning, mask, precision, 1, float(n) * n, False, _periodic(periodic_length))
        else:
            self.tuning = None if tuning is None or isinstance(tuning, str) else tuning
        if isinstance(self.tuning, dict):
            self.tuning = _lib.with_flags(self.tuning, 0)
        self.tp = _lib.tuning_ptr(self.tuning)
        self.L = _L_array(periodic_length)
        self.eta, self.a = float(eta), float(a)
        nbytes = ctypes.c_size_t()
        _lib.check(self.lib.rmw_mobility_workspace_size(self.dom, self.mask, self.prec, n, 1, self.tp,
                                                        ctypes.byref(nbytes)))
        self.nbytes = nbytes.value

    def _ptr(self, t, name: str, needed: bool):
        if not needed:
            return None
        if t is None:
            raise ValueError(f"{name} is required by this block mask")
        return _check_out(t, 3 * self.n, self.device, name).data_ptr()

    def __call__(self, force=None, torque=None, out_velocity=None, out_omega=None) -> None:
        m = self.mask
        fp = self._ptr(force, "force", bool(m & (RMW_UF | RMW_WF)))
        tp_ = self._ptr(torque, "torque", bool(m & (RMW_UT | RMW_WT)))
        up = self._ptr(out_velocity, "out_velocity", self.has_u)
        wp = self._ptr(out_omega, "out_omega", self.has_w)
        st = _stream.current(self.device)
        ws = _workspace_of((self.device.index, st, "default"), self.nbytes, self.device)
        _lib.check(self.lib.rmw_mobility_apply(self.dom, m, self.prec, self.r.data_ptr(), fp, tp_, up, wp, self.n, 1,
                                               0, self.eta, self.a, self.L, self.tp, ws.data_ptr(), ws.numel(), st))


class StridedProduct:
    """A prepared apply_strided (offsets None: one suspension) or
    apply_ragged_strided (offsets: the batch's system boundaries) on the
    positions held by ``r`` (a contiguous float64 (N, 3) device tensor, read
    at every call): the launch shape and workspace size are settled once per
    number of rows k.  Same arguments and results as those functions; the
    Lanczos loops call it once per step, between read-backs of their
    coefficients, where host time leaves the GPU idle."""

    def __init__(self, mask: int, r: torch.Tensor, eta: float, a: float, periodic_length=None,
                 precision="fp32_blocked", tuning=None, offsets=None):
        self.lib = _lib.load()
        self.device = _device()
        if not isinstance(r, torch.Tensor) or r.dtype != torch.float64 or r.device != self.device or \
                not r.is_contiguous() or r.numel() % 3:
            raise ValueError(f"r must be a contiguous float64 (N, 3) tensor on {self.device}")
        self.r = r
        self.n = r.numel() // 3
        self.mask = int(mask)
        self.precision = precision if isinstance(precision, str) else str(precision)
        self.prec = _lib.precision_code(precision)
        self.tuning = tuning
        self.periodic = _periodic(periodic_length)
        self.L = _L_array(periodic_length)
        self.eta, self.a = float(eta), float(a)
        if offsets is None:
            self.off = None
            self.pairs = float(self.n) * self.n
        else:
            self.off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
            if self.off.size < 2 or int(self.off[-1]) != self.n:
                raise ValueError(f"offsets must end at N = {self.n}")
            self.offp = self.off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
            self.pairs = float(np.sum(np.diff(self.off).astype(np.float64) ** 2))
        self._shapes: dict = {}

    def _shape(self, k: int):
        sh = self._shapes.get(k)
        if sh is None:
            tun = _tuning_for(self.tuning, self.mask, self.precision, k, self.pairs, self.off is not None,
                              self.periodic)
            if isinstance(tun, dict):
                tun = _lib.with_flags(tun, 0)
            tp = _lib.tuning_ptr(tun)
            nbytes = ctypes.c_size_t()
            if self.off is None:
                _lib.check(self.lib.rmw_wall_rpy_workspace_size(self.mask, self.prec, self.n, k, tp,
                                                                ctypes.byref(nbytes)))
            else:
                _lib.check(self.lib.rmw_wall_rpy_ragged_workspace_size(self.mask, self.prec, self.offp,
                                                                       self.off.size - 1, k, tp,
                                                                       ctypes.byref(nbytes)))
            sh = (tun, tp, nbytes.value)
            self._shapes[k] = sh
        return sh

    def __call__(self, X: torch.Tensor, Y: torch.Tensor) -> None:
        k, ld, fp, tp_, vp, wp = _grand_rows(X, Y, self.n, self.device)
        _, tp, nbytes = self._shape(k)
        st = _stream.current(self.device)
        if self.off is None:
            ws = _workspace_of((self.device.index, st, "default"), nbytes, self.device)
            _lib.check(self.lib.rmw_wall_rpy_apply(self.mask, self.prec, self.r.data_ptr(), fp, tp_, vp, wp, self.n,
                                                   k, ld, self.eta, self.a, self.L, tp, ws.data_ptr(), ws.numel(),
                                                   st))
        else:
            ws = _workspace_of((self.device.index, st, "ragged"), nbytes, self.device)
            _lib.check(self.lib.rmw_wall_rpy_ragged_apply(self.mask, self.prec, self.r.data_ptr(), fp, tp_, vp, wp,
                                                          self.offp, self.off.size - 1, k, ld, self.eta, self.a,
                                                          self.L, tp, ws.data_ptr(), ws.numel(), st))


def _grand_rows(X: torch.Tensor, Y: torch.Tensor, n: int, device):
    """Pointers of the [force; torque] halves of the k rows of stacked grand
    vectors X, Y (k, >= 6n), after checking that every row the kernels read
    or write lies inside the tensors."""
    for name, t in (("X", X), ("Y", Y)):
        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or t.device != device:
            raise ValueError(f"{name} must be a float64 tensor on {device}")
        if t.dim() != 2 or t.stride(1) != 1:
            raise ValueError(f"{name} must be (k, ld) with unit column stride")
    if X.shape[0] != Y.shape[0] or X.stride(0) != Y.stride(0):
        raise ValueError("X and Y must have the same rows and row stride")
    k, ld = X.shape[0], X.stride(0)
    if min(X.shape[1], Y.shape[1]) < 6 * n or (k > 1 and ld < 6 * n):
        raise ValueError(f"rows of X and Y must hold 6N = {6 * n} values")
    return k, ld, X.data_ptr(), X.data_ptr() + 24 * n, Y.data_ptr(), Y.data_ptr() + 24 * n


def apply_strided(mask: int, r: torch.Tensor, n: int, X: torch.Tensor, Y: torch.Tensor, eta: float, a: float,
                  periodic_length=None, precision="fp32", tuning=None):
    """Y[r] = [[UF, UT], [WF, WT]] X[r] (blocks of ``mask``) for the k rows of
    stacked grand vectors X, Y (k, 6N): forces then torques in each row, rows
    X.stride(0) doubles apart.  One launch per pass serves every row."""
    lib = _lib.load()
    device = _device()
    if not isinstance(r, torch.Tensor) or r.dtype != torch.float64 or r.device != device or \
            not r.is_contiguous() or r.numel() != 3 * n:
        raise ValueError(f"r must be a contiguous float64 (N, 3) tensor on {device} with N = {n}")
    k, ld, fp, tp_, vp, wp = _grand_rows(X, Y, n, device)
    prec = _lib.precision_code(precision)
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, float(n) * n, False, _periodic(periodic_length)))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_wall_rpy_workspace_size(mask, prec, n, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device)
    _lib.check(lib.rmw_wall_rpy_apply(mask, prec, r.data_ptr(), fp, tp_, vp, wp, n, k, ld,
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
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, pairs, True, _periodic(periodic_length)))
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


def apply_ragged_strided(mask: int, r: torch.Tensor, offsets: np.ndarray, X: torch.Tensor, Y: torch.Tensor,
                         eta: float, a: float, periodic_length=None, precision="fp32_blocked", tuning=None):
    """Ragged product on stacked grand vectors of a batch: X, Y (k, 6N) hold
    the forces of all blobs, then their torques; system s owns blobs
    [offsets[s], offsets[s+1])."""
    lib = _lib.load()
    device = _device()
    off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
    n = int(off[-1])
    if not isinstance(r, torch.Tensor) or r.dtype != torch.float64 or r.device != device or \
            not r.is_contiguous() or r.numel() != 3 * n:
        raise ValueError(f"r must be a contiguous float64 (N, 3) tensor on {device} with N = offsets[-1] = {n}")
    k, ld, fp, tp_, vp, wp = _grand_rows(X, Y, n, device)
    prec = _lib.precision_code(precision)
    pairs = float(np.sum(np.diff(off).astype(np.float64) ** 2))
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, pairs, True, _periodic(periodic_length)))
    offp = off.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
    nbytes = ctypes.c_size_t()
    _lib.check(lib.rmw_wall_rpy_ragged_workspace_size(mask, prec, offp, off.size - 1, k, tp, ctypes.byref(nbytes)))
    ws = workspace(nbytes.value, device, slot="ragged")
    _lib.check(lib.rmw_wall_rpy_ragged_apply(mask, prec, r.data_ptr(), fp, tp_, vp, wp, offp,
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
    u = _check_out(out, 3 * nt, device, "out") if out is not None else \
        torch.empty((nt, 3), dtype=torch.float64, device=device)
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
    u = _check_out(out, 3 * n, device, "out") if out is not None else \
        torch.empty((n, 3), dtype=torch.float64, device=device)
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
    tp = _lib.tuning_ptr(_tuning_for(tuning, mask, precision, k, float(n) * n, False, _periodic(periodic_length)))
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

    After the block: launches / total_ms / mean_ms / median_ms over all pair
    kernels, ``by_kind[name] = (launches, total_ms, mean_ms)`` per kernel
    family (wall RPY, different radii, blob-blob forces) and
    ``durations[name]``, each launch's time in ms in launch order."""

    def __enter__(self):
        lib = _lib.load()
        _lib.check(lib.rmw_profile_collect(None, None, None))
        _lib.check(lib.rmw_profile_enable(1))
        return self

    def __exit__(self, *exc):
        lib = _lib.load()
        _lib.check(lib.rmw_profile_enable(0))
        pending = ctypes.c_int64()
        _lib.check(lib.rmw_profile_pending(ctypes.byref(pending)))
        cap = pending.value
        kinds = (ctypes.c_int32 * max(cap, 1))()
        ms = (ctypes.c_double * max(cap, 1))()
        got = ctypes.c_int64()
        _lib.check(lib.rmw_profile_collect_launches(kinds, ms, cap, ctypes.byref(got)))
        if got.value != cap:  # recording is off, so nothing can be added in between
            raise RuntimeError(f"pair-kernel record changed while collected ({cap} -> {got.value})")
        self.durations = {k: [ms[j] for j in range(cap) if kinds[j] == i] for i, k in enumerate(PROFILE_KINDS)}
        self.by_kind = {k: (len(d), sum(d), sum(d) / len(d) if d else 0.0) for k, d in self.durations.items()}
        every = [ms[j] for j in range(cap)]
        self.launches = cap
        self.total_ms = sum(every)
        self.mean_ms = self.total_ms / cap if cap else 0.0
        self.median_ms = float(np.median(every)) if cap else 0.0
        return False
