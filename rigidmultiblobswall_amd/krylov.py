"""FP64 Krylov solvers on the GPU: right-preconditioned restarted GMRES.

All vector algebra runs in librmw_hip's FP64 kernels (Arnoldi with
classical Gram-Schmidt and one re-orthogonalisation); only the (m+1) x m
Hessenberg least-squares problem -- a few hundred scalars -- is solved on
the host, in FP64, with Givens rotations.

Semantics follow general_application_utils.gmres (general_application_utils.
py:514-635) with PC_side='right' as used by solve_mobility_problem
(quaternion_integrator/quaternion_integrator_multi_bodies.py:1441-1536): solve
(A P^-1) y = b and return x = P^-1 y, stopping when ||b - A x|| <= tol ||b||
(scipy's rtol=tol, atol=0).  Unlike the reference wrapper (which passes
x0=None to scipy at :627), a warm start x0 is honoured: the solver iterates
on the residual r0 = b - A x0 and returns x0 + P^-1 y.
"""
from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field
from typing import Callable, List, Optional

import numpy as np
import torch

from . import _lib
from . import _stream as _stream_mod


@dataclass
class GmresResult:
    x: torch.Tensor
    info: int                    # 0 converged, >0 iterations done without convergence
    iterations: int              # Arnoldi steps (matvecs inside cycles)
    matvecs: int                 # all operator applications (incl. residual checks)
    residuals: List[float] = field(default_factory=list)  # relative residual per step
    final_residual: float = float("nan")                  # true ||b - A x|| / ||b||


def _stream():
    return _stream_mod.current()


# pinned host buffers for the Arnoldi column, one per (length, stream): a
# solve reads through it only between its own stream synchronisations
_pinned: dict = {}


def _pinned_column(length: int, stream_handle: int) -> torch.Tensor:
    key = (length, stream_handle)
    buf = _pinned.get(key)
    if buf is None:
        buf = torch.empty(length, dtype=torch.float64, pin_memory=True)
        _pinned[key] = buf
    return buf


class KrylovOps:
    """Thin wrappers of the FP64 vector kernels on torch device tensors."""

    def __init__(self, n: int, max_vectors: int, device):
        self.lib = _lib.load()
        self.n = n
        nb = ctypes.c_size_t()
        _lib.check(self.lib.rmw_krylov_workspace_size(n, max_vectors, ctypes.byref(nb)))
        self.ws = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=device)
        self.scal = torch.empty(2, dtype=torch.float64, device=device)

    def norm(self, v: torch.Tensor) -> float:
        _lib.check(self.lib.rmw_dot_many(v.data_ptr(), 0, 1, v.data_ptr(), self.n, self.scal.data_ptr(), 1,
                                         self.ws.data_ptr(), self.ws.numel(), _stream()))
        return float(self.scal[0].item())

    def dot_many(self, V: torch.Tensor, nvec: int, w: torch.Tensor, out: torch.Tensor):
        _lib.check(self.lib.rmw_dot_many(V.data_ptr(), V.stride(0), nvec, w.data_ptr(), self.n, out.data_ptr(), 0,
                                         self.ws.data_ptr(), self.ws.numel(), _stream()))

    def axpby(self, alpha: float, x: torch.Tensor, beta: float, y: torch.Tensor, div: Optional[torch.Tensor] = None):
        _lib.check(self.lib.rmw_axpby(float(alpha), x.data_ptr(), float(beta), y.data_ptr(), self.n,
                                      div.data_ptr() if div is not None else None, _stream()))

    def combination(self, V: torch.Tensor, nvec: int, coef: torch.Tensor, y: torch.Tensor, accumulate: bool):
        _lib.check(self.lib.rmw_combination(V.data_ptr(), V.stride(0), nvec, coef.data_ptr(), y.data_ptr(), self.n,
                                            int(accumulate), _stream()))

    def sub_combination(self, V: torch.Tensor, nvec: int, coef: torch.Tensor, w: torch.Tensor):
        _lib.check(self.lib.rmw_sub_combination(V.data_ptr(), V.stride(0), nvec, coef.data_ptr(), w.data_ptr(),
                                                self.n, _stream()))

    def arnoldi(self, V: torch.Tensor, j: int, w: torch.Tensor, h: torch.Tensor, hs: torch.Tensor):
        _lib.check(self.lib.rmw_arnoldi_cgs2(V.data_ptr(), V.stride(0), j, w.data_ptr(), self.n, h.data_ptr(),
                                             hs.data_ptr(), self.ws.data_ptr(), self.ws.numel(), _stream()))


def gmres(A: Callable[[torch.Tensor, torch.Tensor], None], b: torch.Tensor,
          Pinv: Optional[Callable[[torch.Tensor, torch.Tensor], None]] = None,
          x0: Optional[torch.Tensor] = None, tol: float = 1e-5, restart: int = 20, maxiter: int = 1000,
          callback: Optional[Callable[[float], None]] = None, verify: bool = True) -> GmresResult:
    """Solve A x = b with right preconditioning.

    A(x, out) and Pinv(x, out) write their product into `out` (device FP64
    vectors of length n).  Returns a GmresResult; info = 0 on convergence.
    verify=False returns as soon as the Arnoldi residual estimate meets the
    tolerance, without the product that recomputes b - A x (for inner solves
    whose caller forms the true residual itself); final_residual is then
    that estimate.
    """
    n = b.numel()
    device = b.device
    m = int(restart)
    ops = KrylovOps(n, m + 2, device)
    V = torch.empty((m + 1, n), dtype=torch.float64, device=device)
    z = torch.empty(n, dtype=torch.float64, device=device)
    w = torch.empty(n, dtype=torch.float64, device=device)
    r = torch.empty(n, dtype=torch.float64, device=device)
    h = torch.empty(m + 2, dtype=torch.float64, device=device)
    hs = torch.empty(m + 2, dtype=torch.float64, device=device)
    ycoef = torch.empty(m + 1, dtype=torch.float64, device=device)
    # pinned host copy of the Arnoldi column: one short copy and a stream
    # synchronisation per step
    stream = torch.cuda.current_stream(device)
    h_host = _pinned_column(m + 2, stream.cuda_stream)
    x = torch.zeros(n, dtype=torch.float64, device=device) if x0 is None else x0.clone()
    ident = Pinv is None
    bnorm = ops.norm(b)
    res = GmresResult(x=x, info=0, iterations=0, matvecs=0)
    if bnorm == 0.0:
        x.zero_()
        res.final_residual = 0.0
        return res
    target = tol * bnorm
    while True:
        # true residual r = b - A x
        if x0 is None and res.matvecs == 0:
            ops.axpby(1.0, b, 0.0, r)
            beta = bnorm
        else:
            A(x, r)
            res.matvecs += 1
            ops.axpby(1.0, b, -1.0, r)
            beta = ops.norm(r)
        res.final_residual = beta / bnorm
        if beta <= target:
            res.info = 0
            break
        if res.iterations >= maxiter:
            res.info = res.iterations
            break
        ops.axpby(1.0 / beta, r, 0.0, V[0])
        H = np.zeros((m + 1, m))
        g = [0.0] * (m + 1)        # plain floats: the Givens sweep runs between GPU steps
        g[0] = beta
        cs = [0.0] * m
        sn = [0.0] * m
        jj = 0
        for j in range(m):
            if ident:
                A(V[j], w)
            else:
                Pinv(V[j], z)
                A(z, w)
            res.matvecs += 1
            ops.arnoldi(V, j, w, h, hs)
            h_host[: j + 2].copy_(h[: j + 2], non_blocking=True)
            stream.synchronize()
            col = h_host[: j + 2].tolist()
            hnext = col[j + 1]
            for i in range(j):
                c, s_ = cs[i], sn[i]
                t = c * col[i] + s_ * col[i + 1]
                col[i + 1] = -s_ * col[i] + c * col[i + 1]
                col[i] = t
            den = math.hypot(col[j], col[j + 1])
            if den == 0.0:
                cs[j], sn[j] = 1.0, 0.0
            else:
                cs[j], sn[j] = col[j] / den, col[j + 1] / den
            col[j] = den
            col[j + 1] = 0.0
            H[: j + 2, j] = col
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            res.iterations += 1
            jj = j + 1
            rel = abs(g[j + 1]) / bnorm
            res.residuals.append(rel)
            if callback is not None:
                callback(rel)
            if abs(g[j + 1]) <= target or res.iterations >= maxiter or hnext == 0.0:
                break
        # x += P^-1 V y
        R = np.triu(H[:jj, :jj])
        y = np.linalg.solve(R, np.asarray(g[:jj])) if jj > 0 else np.zeros(0)
        ycoef[:jj].copy_(torch.from_numpy(y))
        if ident:
            ops.combination(V, jj, ycoef, x, accumulate=True)
        else:
            ops.combination(V, jj, ycoef, w, accumulate=False)
            Pinv(w, z)
            ops.axpby(1.0, z, 1.0, x)
        if not verify and jj > 0 and abs(g[jj]) <= target:
            res.final_residual = abs(g[jj]) / bnorm
            break
    res.x = x
    return res
