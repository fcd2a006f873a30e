"""Brownian noise factor * M^1/2 z by block Lanczos on the GPU (FP64).

Reference: stochastic_forcing_lanczos, stochastic_forcing/stochastic_forcing.py:
112-262 (Ando, Chow, Saad & Skolnick, J. Chem. Phys. 137, 064106 (2012)).

k independent right-hand sides z_r advance together so that every Lanczos
step needs ONE k-right-hand-side matvec (for the rollers' grand mobility
that is one fused UF+UT+WF+WT launch).  Per right-hand side and step:

  w = M v_i;  full re-orthogonalisation of w against v_0..v_i (classical
  Gram-Schmidt twice, FP64 kernels) gives alpha_i = h_i and beta_i = ||w||,
  v_{i+1} = w / beta_i (e_0 if beta_i = 0, as the reference);
  c_i = |z| factor sqrt+(T_i) e_1 from the GPU tridiagonal QL kernel.

The reference forms noise_i = V_i c_i every step and stops when
||noise_i - noise_{i-1}|| / max(||noise_{i-1}||, eps) < tol.  With an
orthonormal basis ||V_i c_i - V_{i-1} c_{i-1}|| = ||c_i - [c_{i-1}; 0]||, so
the same test is made on the coefficient vectors and the noise V c is formed
once, at convergence.  All right-hand sides continue until every one has
converged (extra steps only refine the converged ones).
"""
from __future__ import annotations

import ctypes
import warnings
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib, krylov
from . import _stream as _stream_mod


def _stream():
    return _stream_mod.current()


class LanczosNotConverged(RuntimeWarning):
    """The noise difference between the last two Lanczos steps stayed above
    the tolerance for max_iter steps.  The reference returns that last,
    unconverged noise without a status (stochastic_forcing.py:257-262);
    here it is returned too, with this warning, or raised as an error when
    the caller passes strict=True."""


def _unconverged(what: str, strict: bool):
    msg = f"Lanczos did not reach the tolerance: {what}"
    if strict:
        raise RuntimeError(msg)
    warnings.warn(msg, LanczosNotConverged, stacklevel=3)


def lanczos_block(matvec: Callable[[torch.Tensor, torch.Tensor], None], Z: torch.Tensor, factor: float = 1.0,
                  tolerance: float = 1e-6, max_iter: int = 1000,
                  L_mult: Optional[Callable[[torch.Tensor], torch.Tensor]] = None, strict: bool = False):
    """noise (k, dim) = factor * M^1/2 Z for Z (k, dim) on the device.

    matvec(X, Y) must write Y[r] = M X[r] for the k rows (X, Y (k, dim)
    contiguous float64).  Returns (noise, iterations) with iterations as the
    reference counts them (the step index at which the test passed, or
    max_iter when it did not).  A row that never passes the test gets its
    last noise and a LanczosNotConverged warning (strict=True: an error).
    L_mult(x) maps a (dim,) device vector to L x (noise = L (P M P^T)^1/2 z
    for a preconditioned operator, stochastic_forcing.py:253-256)."""
    lib = _lib.load()
    if Z.dim() == 1:
        Z = Z.reshape(1, -1)
    k, dim = Z.shape
    dev = Z.device
    if factor == 0.0:
        return torch.zeros((k, dim), dtype=torch.float64, device=dev), 0
    mmax = max_iter + 2
    if mmax > 1024:
        raise ValueError("max_iter above 1022 is not supported")
    ops = krylov.KrylovOps(dim, mmax + 1, dev)
    # iteration-major basis and the tridiagonals, grown geometrically (the
    # solves of a well-conditioned operator stop after a few tens of steps)
    mcap = min(mmax, 32)
    V = torch.empty((mcap, k, dim), dtype=torch.float64, device=dev)
    W = torch.empty((k, dim), dtype=torch.float64, device=dev)
    H = torch.zeros((k, mcap, mcap + 1), dtype=torch.float64, device=dev)
    hs = torch.empty(mmax + 1, dtype=torch.float64, device=dev)
    C = torch.empty((k, mmax + 1), dtype=torch.float64, device=dev)   # last column: QL status
    scale = torch.empty(k, dtype=torch.float64, device=dev)
    zsrc = Z.to(torch.float64).contiguous()
    znorm = np.zeros(k)
    for r in range(k):
        znorm[r] = ops.norm(zsrc[r])
        if znorm[r] == 0.0:
            raise ValueError("a zero right-hand side has no Lanczos basis")
        ops.axpby(1.0 / znorm[r], zsrc[r], 0.0, V[0, r])
    scale.copy_(torch.from_numpy(znorm * factor))
    ldv = k * dim
    status = torch.zeros(k, dtype=torch.int32, device=dev)
    nscr = ctypes.c_size_t()
    # scratch of the tridiagonal solve (k m^2 doubles), grown geometrically,
    # and a pinned host buffer for the coefficients the stop test reads
    zscr = None
    c_host = torch.empty((k, mmax + 1), dtype=torch.float64, pin_memory=True)
    c_prev = [None] * k
    converged = np.zeros(k, dtype=bool)
    it_conv = np.full(k, max_iter)
    i_last = 0
    for i in range(max_iter + 1):
        if i + 2 > mcap:
            m2 = min(mmax, 2 * mcap)
            V2 = torch.empty((m2, k, dim), dtype=torch.float64, device=dev)
            V2[:mcap] = V
            H2 = torch.zeros((k, m2, m2 + 1), dtype=torch.float64, device=dev)
            H2[:, :mcap, :mcap + 1] = H
            V, H, mcap = V2, H2, m2
        matvec(V[i], W)
        for r in range(k):
            _lib.check(lib.rmw_arnoldi_cgs2(V[0, r].data_ptr(), ldv, i, W[r].data_ptr(), dim, H[r, i].data_ptr(),
                                            hs.data_ptr(), ops.ws.data_ptr(), ops.ws.numel(), _stream()))
        m = i + 1
        _lib.check(lib.rmw_lanczos_scratch_doubles(m, k, ctypes.byref(nscr)))
        if zscr is None or zscr.numel() < nscr.value:
            _lib.check(lib.rmw_lanczos_scratch_doubles(min(mmax, max(2 * m, 32)), k, ctypes.byref(nscr)))
            zscr = torch.empty(nscr.value, dtype=torch.float64, device=dev)
        _lib.check(lib.rmw_lanczos_sqrt_coeffs(H.data_ptr(), H.stride(0), H.stride(1), m, k, scale.data_ptr(),
                                               C.data_ptr(), C.stride(0), zscr.data_ptr(), zscr.numel(),
                                               status.data_ptr(), _stream()))
        # one read-back per step: the coefficients and the QL status together
        C[:, mmax] = status.to(torch.float64)
        c_host[:, :m].copy_(C[:, :m], non_blocking=True)
        c_host[:, mmax].copy_(C[:, mmax], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        if float(c_host[:, mmax].sum()) != 0.0:
            raise RuntimeError("tridiagonal eigen-solve did not converge")
        c_now = c_host[:, :m].numpy().copy()
        i_last = i
        if i > 0:
            for r in range(k):
                if converged[r]:
                    continue
                prev = np.concatenate([c_prev[r], [0.0]])
                diff = np.linalg.norm(c_now[r] - prev)
                if diff / max(np.linalg.norm(c_prev[r]), np.finfo(float).eps) < tolerance:
                    converged[r] = True
                    it_conv[r] = i
        c_prev = [c_now[r].copy() for r in range(k)]
        if converged.all():
            break
    m = i_last + 1
    noise = torch.empty((k, dim), dtype=torch.float64, device=dev)
    coef = torch.empty(m, dtype=torch.float64, device=dev)
    for r in range(k):
        coef.copy_(C[r, :m])
        ops.combination(V[:, r], m, coef, noise[r], accumulate=False)
    if not converged.all():
        _unconverged(f"{int((~converged).sum())} of {k} right-hand sides after {max_iter} steps", strict)
    if L_mult is not None:
        noise = torch.stack([L_mult(noise[r]) for r in range(k)])
    return noise, int(it_conv.max())


class Segments:
    """A ragged batch of independent suspensions for the segmented kernels:
    system s owns blobs [offsets[s], offsets[s+1]); a batch vector holds
    `nparts` parts of comp * nblobs values (for the grand mobility the
    forces of all blobs, then their torques)."""

    def __init__(self, offsets, comp: int = 3, nparts: int = 2, device=None):
        off = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        if off.ndim != 1 or off.size < 2 or off[0] != 0 or np.any(np.diff(off) < 0):
            raise ValueError("offsets must start at 0 and be non-decreasing")
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.offsets = off
        self.nsys = off.size - 1
        self.nblobs = int(off[-1])
        self.comp, self.nparts = comp, nparts
        self.part_stride = comp * self.nblobs
        self.dim = nparts * self.part_stride
        self.off_dev = torch.from_numpy(off).to(self.device)
        self.c = _lib.Segments(self.off_dev.data_ptr(), self.nsys, comp, nparts, 0, self.nblobs, self.part_stride)

    @property
    def ref(self):
        return ctypes.byref(self.c)

    def workspace(self, max_vectors: int) -> torch.Tensor:
        nb = ctypes.c_size_t()
        _lib.check(_lib.load().rmw_seg_workspace_size(self.ref, max_vectors, ctypes.byref(nb)))
        return torch.empty(max(nb.value, 8), dtype=torch.uint8, device=self.device)

    def dot(self, x: torch.Tensor, y: torch.Tensor, sqrt: bool = False) -> torch.Tensor:
        """Per-system x|s . y|s (nsys,) on the device."""
        ws = self.workspace(1)
        out = torch.empty(self.nsys, dtype=torch.float64, device=self.device)
        _lib.check(_lib.load().rmw_seg_dot_many(x.data_ptr(), 0, 1, y.data_ptr(), self.ref, out.data_ptr(), 1,
                                                int(sqrt), ws.data_ptr(), ws.numel(), _stream()))
        return out

    def element_system(self) -> np.ndarray:
        """System index of every element of a batch vector (host, for checks)."""
        sys_of_blob = np.repeat(np.arange(self.nsys), np.diff(self.offsets))
        return np.tile(np.repeat(sys_of_blob, self.comp), self.nparts)


def lanczos_ragged(matvec: Callable[[torch.Tensor, torch.Tensor], None], Z: torch.Tensor, seg: Segments,
                   factor: float = 1.0, tolerance: float = 1e-6, max_iter: int = 1000, strict: bool = False):
    """factor * M_s^1/2 z_s for every system s of a ragged batch at once.

    Z (k, seg.dim): k batch vectors; matvec(X, Y) writes Y[r] = M X[r] for
    all systems (one ragged launch).  Every (row, system) pair is its own
    Lanczos recurrence with its own convergence test, as the reference runs
    one suspension at a time (stochastic_forcing.py:112-262); the Gram-
    Schmidt, norms and scalings of all of them are the segmented kernels,
    one launch sequence per step for the whole batch.  A system's
    coefficients are frozen at the step its test passes, so its noise is
    the one a separate run would return.  Returns (noise (k, dim),
    iterations (k, nsys))."""
    lib = _lib.load()
    if Z.dim() == 1:
        Z = Z.reshape(1, -1)
    k, dim = Z.shape
    if dim != seg.dim:
        raise ValueError(f"Z has {dim} columns, the batch needs {seg.dim}")
    dev = Z.device
    nsys = seg.nsys
    R = k * nsys
    sizes = np.diff(seg.offsets)
    if factor == 0.0:
        return torch.zeros((k, dim), dtype=torch.float64, device=dev), np.zeros((k, nsys), dtype=int)
    mmax = max_iter + 2
    if mmax > 1024:
        raise ValueError("max_iter above 1022 is not supported")
    mcap = min(mmax, 32)
    V = torch.empty((mcap, k, dim), dtype=torch.float64, device=dev)
    H = torch.zeros((k, nsys, mcap, mcap + 1), dtype=torch.float64, device=dev)
    W = torch.empty((k, dim), dtype=torch.float64, device=dev)
    ws = seg.workspace(mmax)
    hs = torch.empty(nsys * (mmax + 1), dtype=torch.float64, device=dev)
    zsrc = Z.to(torch.float64).contiguous()
    znorm = torch.empty((k, nsys), dtype=torch.float64, device=dev)
    for r in range(k):
        _lib.check(lib.rmw_seg_dot_many(zsrc[r].data_ptr(), 0, 1, zsrc[r].data_ptr(), seg.ref, znorm[r].data_ptr(), 1,
                                        1, ws.data_ptr(), ws.numel(), _stream()))
        _lib.check(lib.rmw_seg_scale(zsrc[r].data_ptr(), V[0, r].data_ptr(), znorm[r].data_ptr(), 1, seg.ref,
                                     _stream()))
    zn = znorm.cpu().numpy()
    live = np.broadcast_to(sizes > 0, (k, nsys))
    if np.any((zn == 0.0) & live):
        raise ValueError("a zero right-hand side has no Lanczos basis")
    scale = (znorm * factor).reshape(R).contiguous()
    status = torch.zeros(R, dtype=torch.int32, device=dev)
    nscr = ctypes.c_size_t()
    converged = ~live.copy()
    it_conv = np.where(live, max_iter, 0)
    c_final = np.zeros((k, nsys, mmax))
    m_used = np.ones((k, nsys), dtype=int)
    c_prev = None
    i_last = 0
    zscr = None
    C = torch.empty((R, mmax + 1), dtype=torch.float64, device=dev)       # last column: QL status
    c_host = torch.empty((R, mmax + 1), dtype=torch.float64, pin_memory=True)
    for i in range(max_iter + 1):
        if i + 2 > mcap:                      # grow the basis and H geometrically
            m2 = min(mmax, 2 * mcap)
            V2 = torch.empty((m2, k, dim), dtype=torch.float64, device=dev)
            V2[:mcap] = V
            H2 = torch.zeros((k, nsys, m2, m2 + 1), dtype=torch.float64, device=dev)
            H2[:, :, :mcap, :mcap + 1] = H
            V, H, mcap = V2, H2, m2
        matvec(V[i], W)
        for r in range(k):
            _lib.check(lib.rmw_seg_arnoldi_cgs2(V[0, r].data_ptr(), k * dim, i, W[r].data_ptr(), seg.ref,
                                                H[r, 0, i].data_ptr(), H.stride(1), hs.data_ptr(), ws.data_ptr(),
                                                ws.numel(), _stream()))
        m = i + 1
        _lib.check(lib.rmw_lanczos_scratch_doubles(m, R, ctypes.byref(nscr)))
        if zscr is None or zscr.numel() < nscr.value:
            _lib.check(lib.rmw_lanczos_scratch_doubles(min(mmax, max(2 * m, 32)), R, ctypes.byref(nscr)))
            zscr = torch.empty(nscr.value, dtype=torch.float64, device=dev)
        _lib.check(lib.rmw_lanczos_sqrt_coeffs(H.data_ptr(), H.stride(1), H.stride(2), m, R, scale.data_ptr(),
                                               C.data_ptr(), C.stride(0), zscr.data_ptr(), zscr.numel(),
                                               status.data_ptr(), _stream()))
        # one read-back per step: the coefficients and the QL status together
        C[:, mmax] = status.to(torch.float64)
        c_host[:, :m].copy_(C[:, :m], non_blocking=True)
        c_host[:, mmax].copy_(C[:, mmax], non_blocking=True)
        torch.cuda.current_stream().synchronize()
        if float(c_host[:, mmax].sum()) != 0.0:
            raise RuntimeError("tridiagonal eigen-solve did not converge")
        c_now = c_host[:, :m].numpy().reshape(k, nsys, m).copy()
        i_last = i
        if i > 0:
            prev = np.concatenate([c_prev, np.zeros((k, nsys, 1))], axis=2)
            diff = np.linalg.norm(c_now - prev, axis=2)
            den = np.maximum(np.linalg.norm(c_prev, axis=2), np.finfo(float).eps)
            newly = (~converged) & (diff / den < tolerance)
            for r, s in zip(*np.nonzero(newly)):
                c_final[r, s, :m] = c_now[r, s]
                m_used[r, s] = m
                it_conv[r, s] = i
            converged |= newly
        c_prev = c_now
        if converged.all():
            break
    # systems that never passed keep their last coefficients (as the reference)
    if not converged.all():
        _unconverged(f"{int((~converged).sum())} (right-hand side, system) pairs after {max_iter} steps", strict)
    for r, s in zip(*np.nonzero(~converged)):
        c_final[r, s, :i_last + 1] = c_prev[r, s]
        m_used[r, s] = i_last + 1
    mfin = int(m_used.max())
    coef = torch.from_numpy(np.ascontiguousarray(c_final[:, :, :mfin])).to(dev)
    noise = torch.empty((k, dim), dtype=torch.float64, device=dev)
    for r in range(k):
        _lib.check(lib.rmw_seg_combination(V[0, r].data_ptr(), k * dim, mfin, coef[r].data_ptr(), mfin,
                                           noise[r].data_ptr(), seg.ref, 0, _stream()))
    return noise, it_conv


def stochastic_forcing_lanczos(factor=1.0, tolerance=1e-06, max_iter=1000, dim=None, mobility=None,
                               mobility_mult=None, L_mult=None, z=None, print_residual=False):
    """Drop-in for stochastic_forcing.stochastic_forcing_lanczos
    (stochastic_forcing/stochastic_forcing.py:112-262): returns
    (noise, iterations) as numpy.  `mobility_mult` takes and returns numpy
    vectors like the reference's closures; the Lanczos algebra runs on the
    GPU.  For device-resident operators use lanczos_block."""
    if z is None:
        if dim is None:
            raise ValueError("dim or z is required")
        z = np.random.randn(dim)
    z = np.asarray(z, dtype=np.float64).reshape(-1) if not torch.is_tensor(z) else z.reshape(-1)
    dim = z.shape[0] if dim is None else dim
    if factor == 0.0:
        return np.zeros(dim), 0
    dev = torch.device("cuda", torch.cuda.current_device())
    Zd = torch.as_tensor(z, dtype=torch.float64).to(dev).reshape(1, dim)
    if mobility is not None:
        # a dense matrix given by the caller is applied on the host, as the
        # reference does with np.dot (stochastic_forcing.py:199-202)
        Mh = np.asarray(mobility, dtype=np.float64)
        mobility_mult = lambda v: Mh @ v  # noqa: E731

    def mv(X, Y):
        # reference convention: numpy vector in, numpy vector out
        out = np.asarray(mobility_mult(X[0].cpu().numpy()), dtype=np.float64).reshape(-1)
        Y[0].copy_(torch.from_numpy(out).to(dev))
    noise, it = lanczos_block(mv, Zd, factor=factor, tolerance=tolerance, max_iter=max_iter)
    out = noise[0].cpu().numpy()
    if L_mult is not None:
        out = np.reshape(L_mult(out), dim)
    return out, it
