"""Single-blob rollers: the grand (translation + rotation) mobility of the
rollers integrator and its Brownian slip by block Lanczos.

Reference: quaternion_integrator/quaternion_integrator_rollers.py, the
grand_mobility_matrix closure (:1114-1122) -- four separate block products
UF, UT, WF, WT per application -- and compute_stochastic_velocity (:1082-
1201), which draws factor * N^1/2 z by stochastic_forcing_lanczos.  Here one
fused launch evaluates all four blocks for k right-hand sides at once.
"""
from __future__ import annotations

import numpy as np
import torch

from . import lanczos, wall_rpy
from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

GRAND = RMW_UF | RMW_UT | RMW_WF | RMW_WT


class Rollers:
    def __init__(self, r_vectors, eta: float, a: float, periodic_length=None, precision: str = "fp32_blocked",
                 device=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.r = torch.as_tensor(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
                                 if not torch.is_tensor(r_vectors) else r_vectors).to(self.device,
                                                                                      torch.float64).contiguous()
        self.n = self.r.shape[0]
        self.eta, self.a = float(eta), float(a)
        self.L = periodic_length
        self.precision = precision
        self._products = {}      # precision -> wall_rpy.StridedProduct on self.r

    def grand_matvec(self, X: torch.Tensor, Y: torch.Tensor, precision=None):
        """Y[r] = [UF UT; WF WT] X[r] for X, Y (k, 6N) with X[r] = [f; t]."""
        prec = precision or self.precision
        prod = self._products.get(prec)
        if prod is None:
            prod = wall_rpy.StridedProduct(GRAND, self.r, self.eta, self.a, self.L, prec)
            self._products[prec] = prod
        prod(X, Y)

    def brownian_noise(self, Z: torch.Tensor, factor: float = 1.0, tolerance: float = 1e-3, max_iter: int = 1000,
                       precision=None):
        """factor * N^1/2 Z[r] for the k rows of Z (k, 6N), one fused grand
        matvec per Lanczos step for all rows."""
        return lanczos.lanczos_block(lambda X, Y: self.grand_matvec(X, Y, precision), Z, factor=factor,
                                     tolerance=tolerance, max_iter=max_iter)


class RaggedRollers:
    """A batch of independent roller suspensions stored back to back (system
    s owns rollers [offsets[s], offsets[s+1])).  The grand product of every
    system is one persistent ragged launch, and the Brownian slip of all
    systems advances together through the segmented Lanczos kernels.

    Batch vectors are (k, 6 N_total): the forces of all rollers, then their
    torques (system s's part of each is contiguous)."""

    def __init__(self, r_vectors, offsets, eta: float, a: float, periodic_length=None,
                 precision: str = "fp32_blocked", device=None, tuning=None):
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        self.offsets = np.ascontiguousarray(np.asarray(offsets, dtype=np.int64))
        self.r = torch.as_tensor(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
                                 if not torch.is_tensor(r_vectors) else r_vectors).to(self.device,
                                                                                      torch.float64).contiguous()
        self.n = self.r.shape[0]
        if self.n != int(self.offsets[-1]):
            raise ValueError("offsets[-1] must equal the number of rollers")
        self.eta, self.a = float(eta), float(a)
        self.L = periodic_length
        self.precision = precision
        self.tuning = tuning
        self.seg = lanczos.Segments(self.offsets, comp=3, nparts=2, device=self.device)
        self._products = {}      # precision -> wall_rpy.StridedProduct on self.r

    def grand_matvec(self, X: torch.Tensor, Y: torch.Tensor, precision=None):
        """Y[r] = blockdiag_s([UF UT; WF WT]_s) X[r] for X, Y (k, 6N)."""
        prec = precision or self.precision
        prod = self._products.get(prec)
        if prod is None:
            prod = wall_rpy.StridedProduct(GRAND, self.r, self.eta, self.a, self.L, prec, tuning=self.tuning,
                                           offsets=self.offsets)
            self._products[prec] = prod
        prod(X, Y)

    def brownian_noise(self, Z: torch.Tensor, factor: float = 1.0, tolerance: float = 1e-3, max_iter: int = 1000,
                       precision=None):
        """factor * N_s^1/2 z_s for every system; returns (noise (k, 6N),
        iterations (k, nsys))."""
        return lanczos.lanczos_ragged(lambda X, Y: self.grand_matvec(X, Y, precision), Z, self.seg, factor=factor,
                                      tolerance=tolerance, max_iter=max_iter)

    def system_slices(self, s: int):
        """Index arrays of system s's forces and torques in a batch vector."""
        b0, b1 = int(self.offsets[s]), int(self.offsets[s + 1])
        return np.r_[3 * b0:3 * b1, 3 * self.n + 3 * b0:3 * self.n + 3 * b1]
