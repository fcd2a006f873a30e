"""Time stepping of single-blob rollers above a wall on the GPU.

Mirror of QuaternionIntegratorRollers (quaternion_integrator/
quaternion_integrator_rollers.py:23-1600) for its first-order schemes:

  deterministic_forward_euler     :119-151
  stochastic_first_order          :154-197
  deterministic_adams_bashforth   :199-248
  stochastic_adams_bashforth      :251-302

with the velocities of compute_deterministic_velocity_and_torque (:905-983),
compute_stochastic_linear_velocity (:1203-1261) and their uncorrelated forms
(:985-1080, :1263-1313).  The whole step stays on the device: positions,
forces (one-blob kernel + blob-blob pair kernel), the fused UF+UT product,
the Lanczos noise M_tt^1/2 z (FP64 algebra, pair products in the chosen
precision), the random finite difference drift (two UF products at
displaced positions) and the position update with the wall test.  The host
reads back two integers per attempt (the wall test), as the reference's
loop over bodies does.

Random numbers come from a torch generator on the device (the reference
draws from numpy's global generator; the streams differ).  A caller can
supply its own draws through ``noise_source``: a callable
``(kind, shape) -> tensor`` with kind "z" (Lanczos right-hand side) or
"dx" (finite-difference direction), called in the reference's order.
"""
from __future__ import annotations

import ctypes
import math
from typing import Callable, Optional

import numpy as np
import torch

from . import _lib, forces, krylov, lanczos, wall_rpy
from . import _stream as _stream_mod
from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

SCHEMES = ("deterministic_forward_euler", "stochastic_first_order", "deterministic_adams_bashforth",
           "stochastic_adams_bashforth")


class RollersIntegrator:
    """State and schemes of the rollers integrator (one blob per body)."""

    def __init__(self, r_vectors, scheme: str = "stochastic_adams_bashforth_rollers", eta: float = 1.0,
                 a: float = 1.0, kT: float = 0.0, g: float = 1.0, blob_mass: float = 1.0,
                 repulsion_strength_wall: float = 1.0, debye_length_wall: float = 1.0,
                 blob_blob_force_implementation: str = "None", repulsion_strength: float = 1.0,
                 debye_length: float = 1.0, periodic_length=None, omega_one_roller=(0.0, 0.0, 0.0),
                 free_kinematics: str = "True", hydro_interactions: int = 1, tolerance: float = 1e-8,
                 rf_delta: float = 1e-3, domain: str = "single_wall", precision: str = "fp32_blocked",
                 seed: Optional[int] = None, noise_source: Optional[Callable] = None, max_invalid: int = 1000,
                 gmres_restart: int = 20):
        self.scheme = scheme.replace("_rollers", "")
        if self.scheme not in SCHEMES:
            raise ValueError(f"scheme {scheme!r} is not provided; use one of {[s + '_rollers' for s in SCHEMES]}")
        if domain not in ("single_wall", "no_wall", "in_plane"):
            raise ValueError(f"unknown domain {domain!r}; use single_wall, no_wall or in_plane")
        if domain == "in_plane" and str(free_kinematics) == "False" and int(hydro_interactions) == 1:
            raise ValueError("the in-plane mobility has no rotational blocks: constrained kinematics needs "
                             "single_wall or no_wall")
        self.device = torch.device("cuda", torch.cuda.current_device())
        self.lib = _lib.load()
        r = torch.as_tensor(r_vectors) if torch.is_tensor(r_vectors) else torch.from_numpy(
            np.ascontiguousarray(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)))
        self.r = r.to(self.device, torch.float64).reshape(-1, 3).contiguous().clone()
        self.n = self.r.shape[0]
        self.eta, self.a, self.kT, self.g = float(eta), float(a), float(kT), float(g)
        self.blob_mass = float(blob_mass)
        self.repulsion_strength_wall, self.debye_length_wall = float(repulsion_strength_wall), float(debye_length_wall)
        self.repulsion_strength, self.debye_length = float(repulsion_strength), float(debye_length)
        self.periodic_length = None if periodic_length is None else np.asarray(periodic_length, dtype=np.float64)
        self.omega_one_roller = np.asarray(omega_one_roller, dtype=np.float64).reshape(3)
        self.free_kinematics = str(free_kinematics)
        self.hydro_interactions = int(hydro_interactions)
        self.tolerance, self.rf_delta = float(tolerance), float(rf_delta)
        self.domain = domain
        self.precision = precision
        self.max_invalid = int(max_invalid)
        self.gmres_restart = int(gmres_restart)
        self.blob_blob = None
        if blob_blob_force_implementation != "None":
            self.blob_blob = forces.set_blob_blob_forces(blob_blob_force_implementation).keywords
            self._bb_precision = self.blob_blob["precision"]
            self._bb_cutoff = (2 * self.a + 30 * self.debye_length
                               if blob_blob_force_implementation.startswith("tree_hip") else 0.0)
            if blob_blob_force_implementation.startswith("radii_hip"):
                raise ValueError("rollers have one blob radius; use 'hip' or 'tree_hip'")
        self.gen = torch.Generator(device=self.device)
        if seed is not None:
            self.gen.manual_seed(int(seed))
        else:
            self.gen.seed()
        self.noise_source = noise_source
        # integrator state (quaternion_integrator_rollers.py:34-57)
        self.first_step = True
        self.velocities_previous_step: Optional[torch.Tensor] = None
        self.deterministic_torque_previous_step: Optional[torch.Tensor] = None
        self.invalid_configuration_count = 0
        self.wall_overlaps = 0
        self.det_iterations_count = 0
        self.stoch_iterations_count = 0
        # mobility products issued, by block mask ("UF", "UF+UT", "WF", "WT")
        self.products = {"UF": 0, "UF+UT": 0, "WF": 0, "WT": 0}
        # device buffers
        dev, n = self.device, self.n
        self._F = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self._T = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self._r_new = torch.empty((n, 3), dtype=torch.float64, device=dev)
        self._counts = torch.zeros(2, dtype=torch.int64, device=dev)
        nb = ctypes.c_size_t()
        _lib.check(self.lib.rmw_rollers_advance_workspace_size(n, ctypes.byref(nb)))
        self._adv_ws = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=dev)
        self._omega_c = (ctypes.c_double * 3)(*[float(v) for v in self.omega_one_roller])

    # ------------------------------------------------------------------
    def _stream(self):
        return _stream_mod.current(self.device)

    def _draw(self, kind: str, shape) -> torch.Tensor:
        if self.noise_source is not None:
            t = torch.as_tensor(self.noise_source(kind, shape))
            return t.to(self.device, torch.float64).reshape(shape).contiguous()
        return torch.randn(shape, dtype=torch.float64, device=self.device, generator=self.gen)

    def _uf(self, r: torch.Tensor, f: torch.Tensor, out: torch.Tensor) -> None:
        self.products["UF"] += 1
        wall_rpy.apply(RMW_UF, r, force=f, eta=self.eta, a=self.a, periodic_length=self.periodic_length,
                       precision=self.precision, out_velocity=out.reshape(1, self.n, 3), domain=self.domain)

    # ------------------------------------------------------------------
    def blob_forces(self, r: torch.Tensor) -> torch.Tensor:
        """One-blob forces (gravity + wall) plus blob-blob forces (N, 3):
        calc_one_blob_forces + calc_blob_blob_forces as bound by
        multi_bodies.py:1325-1336."""
        F = self._F
        accumulate = 0
        if self.blob_blob is not None:
            forces.blob_forces(r, self.repulsion_strength, self.debye_length, self.a,
                               periodic_length=self.periodic_length, cutoff=self._bb_cutoff,
                               precision=self._bb_precision, out=F)
            accumulate = 1
        _lib.check(self.lib.rmw_rollers_one_blob_forces(r.data_ptr(), self.n, self.blob_mass, self.g, self.a,
                                                        self.repulsion_strength_wall, self.debye_length_wall,
                                                        accumulate, F.data_ptr(), self._stream()))
        return F

    def get_torque(self) -> torch.Tensor:
        """Free kinematics: T = omega 8 pi eta a^3 on every roller (:1586-1597)."""
        t = torch.from_numpy(self.omega_one_roller * 8.0 * math.pi * self.eta * self.a ** 3).to(self.device)
        self._T.copy_(t.expand(self.n, 3))
        return self._T

    def compute_deterministic_velocity_and_torque(self, r: Optional[torch.Tensor] = None):
        """(v, T) (:905-983): T from free kinematics or from M_rr T =
        omega - M_rt F by GMRES (RHS scaled to norm 1, warm start from the
        previous scaled solution); v = M_tt F + M_tr T in one launch."""
        r = self.r if r is None else r
        F = self.blob_forces(r)
        if self.free_kinematics == "False":
            rhs = torch.from_numpy(np.tile(self.omega_one_roller, self.n)).to(self.device)
            self.products["WF"] += 1
            _, wf = wall_rpy.apply(RMW_WF, r, force=F, eta=self.eta, a=self.a, periodic_length=self.periodic_length,
                                   precision=self.precision, domain=self.domain)
            rhs = rhs - wf.reshape(-1)
            rhs_norm = float(torch.linalg.norm(rhs))
            if rhs_norm > 0:
                rhs = rhs / rhs_norm

            def A(x, out):
                self.products["WT"] += 1
                wall_rpy.apply(RMW_WT, r, torque=x.reshape(self.n, 3), eta=self.eta, a=self.a,
                               periodic_length=self.periodic_length, precision=self.precision,
                               out_omega=out.reshape(1, self.n, 3), domain=self.domain)
            res = krylov.gmres(A, rhs.contiguous(), x0=self.deterministic_torque_previous_step, tol=self.tolerance,
                               restart=self.gmres_restart, maxiter=1000)
            self.det_iterations_count += res.iterations
            self.deterministic_torque_previous_step = res.x.clone()
            T = res.x * rhs_norm if rhs_norm > 0 else res.x
            T = T.reshape(self.n, 3)
        else:
            T = self.get_torque()
        v = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        if bool(torch.any(T != 0)):
            self.products["UF+UT"] += 1
            wall_rpy.apply(RMW_UF | RMW_UT, r, force=F, torque=T, eta=self.eta, a=self.a,
                           periodic_length=self.periodic_length, precision=self.precision,
                           out_velocity=v.reshape(1, self.n, 3), domain=self.domain)
        else:
            self._uf(r, F, v)
        return v, T

    def compute_deterministic_velocity_and_torque_uncorrelated(self, r: Optional[torch.Tensor] = None):
        """(v, T) from the self mobility only (:985-1080)."""
        r = self.r if r is None else r
        F = self.blob_forces(r)
        v = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        T = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.rmw_rollers_uncorrelated_velocity(r.data_ptr(), F.data_ptr(), self.n, self._omega_c,
                                                              int(self.free_kinematics != "False"), self.eta,
                                                              self.a, v.data_ptr(), T.data_ptr(), self._stream()))
        return v, T

    def compute_stochastic_linear_velocity(self, dt: float, r: Optional[torch.Tensor] = None) -> torch.Tensor:
        """sqrt(2 kT/dt) M_tt^1/2 z + kT/(delta a) [M_tt(r + delta a dx/2)
        - M_tt(r - delta a dx/2)] dx (:1203-1261)."""
        r = self.r if r is None else r
        n = self.n
        z = self._draw("z", (3 * n,))
        # one prepared product serves every Lanczos step on these positions
        prod = wall_rpy.Product(RMW_UF, r.contiguous(), eta=self.eta, a=self.a, periodic_length=self.periodic_length,
                                precision=self.precision, domain=self.domain)

        def mv(X, Y):
            self.products["UF"] += 1
            prod(force=X[0], out_velocity=Y[0])

        noise, it = lanczos.lanczos_block(mv, z.reshape(1, -1), factor=math.sqrt(2 * self.kT / dt),
                                          tolerance=self.tolerance, max_iter=1000)
        self.stoch_iterations_count += it
        v = noise.reshape(n, 3)
        if self.kT > 0.0 and self.domain != "no_wall":
            dx = self._draw("dx", (n, 3))
            r1 = r + dx * (self.rf_delta * self.a * 0.5)
            div = torch.empty((n, 3), dtype=torch.float64, device=self.device)
            tmp = torch.empty((n, 3), dtype=torch.float64, device=self.device)
            self._uf(r1, dx, div)
            r2 = r1 - dx * (self.rf_delta * self.a)
            self._uf(r2, dx, tmp)
            div -= tmp
            v = v + (self.kT / (self.rf_delta * self.a)) * div
        return v

    def compute_stochastic_linear_velocity_uncorrelated(self, dt: float, r: Optional[torch.Tensor] = None):
        """Self-mobility noise and drift (:1263-1313)."""
        r = self.r if r is None else r
        z = self._draw("z", (3 * self.n,))
        v = torch.empty((self.n, 3), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.rmw_rollers_uncorrelated_noise(r.data_ptr(), z.data_ptr(), self.n, self.eta, self.a,
                                                           self.kT, float(dt), v.data_ptr(), self._stream()))
        return v

    # ------------------------------------------------------------------
    def _try_update(self, dt: float, det, prev, stoch, c_det: float, c_prev: float):
        """r_new = r + dt v; returns (valid, heights below a)."""
        _lib.check(self.lib.rmw_rollers_advance(self.r.data_ptr(), det.data_ptr(),
                                                prev.data_ptr() if prev is not None else None,
                                                stoch.data_ptr() if stoch is not None else None, self.n,
                                                c_det, c_prev, float(dt), self.a, self._r_new.data_ptr(),
                                                self._counts.data_ptr(), self._adv_ws.data_ptr(),
                                                self._adv_ws.numel(), self._stream()))
        below0, below_a = (int(v) for v in self._counts.cpu())
        valid = not (self.domain == "single_wall" and below0 > 0)
        return valid, below_a

    def _accept(self, below_a: int, count_overlaps: bool) -> None:
        self.r, self._r_new = self._r_new, self.r
        if count_overlaps:
            self.wall_overlaps += below_a

    def _velocities(self, dt: float, stochastic: bool):
        if self.hydro_interactions == 1:
            det, _ = self.compute_deterministic_velocity_and_torque()
            stoch = self.compute_stochastic_linear_velocity(dt) if stochastic else None
        else:
            det, _ = self.compute_deterministic_velocity_and_torque_uncorrelated()
            stoch = self.compute_stochastic_linear_velocity_uncorrelated(dt) if stochastic else None
        return det, stoch

    def _loop(self, dt: float, stochastic: bool, adams_bashforth: bool) -> None:
        """The retry loop shared by the four schemes.  Two departures from
        the reference, both on paths where it cannot proceed: a
        deterministic step that leaves a roller below the wall raises (the
        reference recomputes the same step forever), and a retried first
        Adams-Bashforth step stays forward Euler (the reference has no
        previous velocity there)."""
        while True:
            det, stoch = self._velocities(dt, stochastic)
            if adams_bashforth and not self.first_step:
                valid, below_a = self._try_update(dt, det, self.velocities_previous_step, stoch, 1.5, 0.5)
            else:
                valid, below_a = self._try_update(dt, det, None, stoch, 1.0, 0.0)
            if valid:
                if adams_bashforth:
                    self.first_step = False
                    self.velocities_previous_step = det
                # stochastic_first_order counts overlaps whatever the domain
                self._accept(below_a, self.domain == "single_wall" or self.scheme == "stochastic_first_order")
                return
            self.invalid_configuration_count += 1
            if not stochastic:
                raise RuntimeError("the deterministic step leaves a roller below the wall")
            if self.invalid_configuration_count >= self.max_invalid:
                raise RuntimeError(f"{self.invalid_configuration_count} invalid configurations (a roller below the "
                                   "wall after every attempt)")

    def deterministic_forward_euler(self, dt: float, *args, **kwargs):
        self._loop(dt, stochastic=False, adams_bashforth=False)

    def stochastic_first_order(self, dt: float, *args, **kwargs):
        self._loop(dt, stochastic=True, adams_bashforth=False)

    def deterministic_adams_bashforth(self, dt: float, *args, **kwargs):
        self._loop(dt, stochastic=False, adams_bashforth=True)

    def stochastic_adams_bashforth(self, dt: float, *args, **kwargs):
        self._loop(dt, stochastic=True, adams_bashforth=True)

    def advance_time_step(self, dt: float, *args, **kwargs):
        """One step of the selected scheme (:112-117)."""
        return getattr(self, self.scheme)(dt, *args, **kwargs)

    def locations(self) -> np.ndarray:
        return self.r.cpu().numpy()
