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
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib, krylov


def _stream():
    return torch.cuda.current_stream().cuda_stream


def lanczos_block(matvec: Callable[[torch.Tensor, torch.Tensor], None], Z: torch.Tensor, factor: float = 1.0,
                  tolerance: float = 1e-6, max_iter: int = 1000,
                  L_mult: Optional[Callable[[torch.Tensor], torch.Tensor]] = None):
    """noise (k, dim) = factor * M^1/2 Z for Z (k, dim) on the device.

    matvec(X, Y) must write Y[r] = M X[r] for the k rows (X, Y (k, dim)
    contiguous float64).  Returns (noise, iterations) with iterations as the
    reference counts them (the step index at which the test passed, or
    max_iter when it did not)."""
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
    V = torch.empty((mmax, k, dim), dtype=torch.float64, device=dev)   # iteration-major basis
    W = torch.empty((k, dim), dtype=torch.float64, device=dev)
    H = torch.zeros((k, mmax, mmax + 1), dtype=torch.float64, device=dev)
    hs = torch.empty(mmax + 1, dtype=torch.float64, device=dev)
    C = torch.empty((k, mmax), dtype=torch.float64, device=dev)
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
    c_prev = [None] * k
    converged = np.zeros(k, dtype=bool)
    it_conv = np.full(k, max_iter)
    i_last = 0
    for i in range(max_iter + 1):
        matvec(V[i], W)
        for r in range(k):
            _lib.check(lib.rmw_arnoldi_cgs2(V[0, r].data_ptr(), ldv, i, W[r].data_ptr(), dim, H[r, i].data_ptr(),
                                            hs.data_ptr(), ops.ws.data_ptr(), ops.ws.numel(), _stream()))
        m = i + 1
        _lib.check(lib.rmw_lanczos_scratch_doubles(m, k, ctypes.byref(nscr)))
        zscr = torch.empty(max(nscr.value, 1), dtype=torch.float64, device=dev)
        _lib.check(lib.rmw_lanczos_sqrt_coeffs(H.data_ptr(), H.stride(0), H.stride(1), m, k, scale.data_ptr(),
                                               C.data_ptr(), C.stride(0), zscr.data_ptr(), nscr.value,
                                               status.data_ptr(), _stream()))
        c_now = C[:, :m].cpu().numpy()
        if int(status.sum().item()):
            raise RuntimeError("tridiagonal eigen-solve did not converge")
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
    if L_mult is not None:
        noise = torch.stack([L_mult(noise[r]) for r in range(k)])
    return noise, int(it_conv.max())


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
