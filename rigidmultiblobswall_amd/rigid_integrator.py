"""Time stepping of rigid multiblob bodies above a wall.

Mirror of QuaternionIntegrator (quaternion_integrator/
quaternion_integrator_multi_bodies.py) for its GMRES schemes

  deterministic_forward_euler     :75-107
  deterministic_adams_bashforth   :142-186
  stochastic_EM                   :262-300  (Euler-Maruyama: blob noise
      factor L (P^T M P)^1/2 z with factor sqrt(2 kT / dt) from Lanczos on
      the stochastic block-diagonal preconditioner, multi_bodies.py:480-623,
      subtracted from the slip of the GMRES right-hand side; a step that
      moves a blob below the wall is redrawn)

with solve_mobility_problem (:1441-1548), the body forces of
force_torque_calculator_sort_by_bodies (multi_bodies/multi_bodies_functions.py
:411-445: one-blob gravity and wall repulsion, blob-blob soft repulsion,
torque sum r x f) and check_positions (:1779-1829: every blob of every body
above the wall).  Everything runs on the device: blob positions, forces,
the body force-torque sums, the preconditioned FP64 GMRES solve and the
position / quaternion update.  The previous step's scaled solution is the
GMRES initial guess (the reference passes it as x0, which its
right-preconditioned wrapper then ignores, general_application_utils.py
:627; honouring it changes only the iteration count).

Active slip: ``slip_reference`` gives each body's blob slip in its
reference frame (as read from a .slip file); it is rotated with the body
every step like active_body_slip (multi_bodies_functions.py:123-140).

``solver="mixed"`` (no counterpart in the reference) solves each step with
RigidSuspension.solve_mixed: FP32-pair inner GMRES refined against FP64
residuals to the same FP64 tolerance.

Random numbers come from a torch generator on the device (the reference
draws z from numpy's global generator inside stochastic_forcing_lanczos);
``noise_source(kind, shape)`` with kind "z" replaces them, one call per
attempt, in the reference's order.
"""
from __future__ import annotations

import ctypes
import math
from typing import Callable, Optional, Sequence

import numpy as np
import torch

from . import _lib, forces
from . import _stream as _stream_mod
from .rigid import RigidSuspension

SCHEMES = ("deterministic_forward_euler", "deterministic_adams_bashforth", "stochastic_EM")


def _stream():
    return _stream_mod.current()


class RigidIntegrator:
    def __init__(self, locations, orientations, reference, scheme: str = "deterministic_forward_euler",
                 eta: float = 1.0, a: float = 1.0, g: float = 1.0, blob_mass: float = 1.0,
                 repulsion_strength_wall: float = 1.0, debye_length_wall: float = 1.0,
                 blob_blob_force_implementation: str = "None", repulsion_strength: float = 1.0,
                 debye_length: float = 1.0, periodic_length=None, tolerance: float = 1e-8,
                 precision: str = "fp64", slip_reference=None, domain: str = "single_wall", restart: int = 60,
                 update_PC: int = 1, solver: str = "fp64", kT: float = 0.0, seed: Optional[int] = None,
                 noise_source: Optional[Callable] = None, lanczos_precision: Optional[str] = None):
        if scheme not in SCHEMES:
            raise ValueError(f"scheme {scheme!r} is not provided; use one of {list(SCHEMES)}")
        if solver not in ("fp64", "mixed"):
            raise ValueError(f"solver must be 'fp64' or 'mixed', not {solver!r}")
        self.solver = solver
        if domain not in ("single_wall", "no_wall"):
            raise ValueError("rigid bodies: domain 'single_wall' or 'no_wall'")
        self.domain = domain
        self.scheme = scheme
        self.lib = _lib.load()
        self.device = torch.device("cuda", torch.cuda.current_device())
        self._centres = np.ascontiguousarray(np.asarray(locations, dtype=np.float64).reshape(-1, 3))
        self._quats = np.ascontiguousarray(np.asarray(orientations, dtype=np.float64).reshape(-1, 4))
        self.nbodies = self._centres.shape[0]
        self.reference = reference
        self.eta, self.a, self.g, self.blob_mass = float(eta), float(a), float(g), float(blob_mass)
        self.repulsion_strength_wall, self.debye_length_wall = float(repulsion_strength_wall), float(debye_length_wall)
        self.repulsion_strength, self.debye_length = float(repulsion_strength), float(debye_length)
        self.periodic_length = periodic_length
        self.tolerance, self.precision, self.restart = float(tolerance), precision, int(restart)
        self.blob_blob = None
        if blob_blob_force_implementation != "None":
            kw = forces.set_blob_blob_forces(blob_blob_force_implementation).keywords
            if blob_blob_force_implementation.startswith("radii_hip"):
                raise ValueError("use 'hip' or 'tree_hip' (one blob radius)")
            self.blob_blob = dict(precision=kw["precision"],
                                  cutoff=(2 * self.a + 30 * self.debye_length
                                          if blob_blob_force_implementation.startswith("tree_hip") else 0.0))
        self.slip_reference = slip_reference
        if int(update_PC) < 1:
            raise ValueError("update_PC must be >= 1")
        self.update_PC = int(update_PC)   # rebuild the preconditioner every update_PC steps
        self._pc_state = None
        self.pc_builds = 0
        self.step_count = 0
        self.first_step = True
        self.first_guess: Optional[torch.Tensor] = None   # scaled solution of the previous solve
        self.velocities_previous_step: Optional[torch.Tensor] = None
        self.invalid_configuration_count = 0
        self.det_iterations_count = 0
        self.stoch_iterations_count = 0
        self.kT = float(kT)
        # pair arithmetic of the Lanczos products (default: the solve's)
        self.lanczos_precision = lanczos_precision
        self.gen = torch.Generator(device=self.device)
        if seed is not None:
            self.gen.manual_seed(int(seed))
        else:
            self.gen.seed()
        self.noise_source = noise_source
        self._S: Optional[RigidSuspension] = None       # the current configuration
        self._S_next: Optional[RigidSuspension] = None  # buffers of the trial configuration
        self._slip_ref_dev: Optional[torch.Tensor] = None

    # host copies of the configuration, read back from the device on demand
    @property
    def centres(self) -> np.ndarray:
        if self._centres is None:
            self._centres = self._S.centres.cpu().numpy()
        return self._centres

    @centres.setter
    def centres(self, value):
        self._centres = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1, 3))
        if self._S is not None:
            self._S.set_configuration(self._centres, self._S.quats)

    @property
    def quats(self) -> np.ndarray:
        if self._quats is None:
            self._quats = self._S.quats.cpu().numpy()
        return self._quats

    @quats.setter
    def quats(self, value):
        self._quats = np.ascontiguousarray(np.asarray(value, dtype=np.float64).reshape(-1, 4))
        if self._S is not None:
            self._S.set_configuration(self._S.centres, self._quats)

    # ------------------------------------------------------------------
    def suspension(self, centres=None, quats=None) -> RigidSuspension:
        return RigidSuspension(self.centres if centres is None else centres, self.quats if quats is None else quats,
                               self.reference, self.eta, self.a, periodic_length=self.periodic_length,
                               precision=self.precision, domain=self.domain)

    def _current(self) -> RigidSuspension:
        if self._S is None:
            self._S = self.suspension()
        return self._S

    def blob_forces(self, S: RigidSuspension) -> torch.Tensor:
        F = torch.empty((S.Nblobs, 3), dtype=torch.float64, device=self.device)
        acc = 0
        if self.blob_blob is not None:
            forces.blob_forces(S.r_blobs, self.repulsion_strength, self.debye_length, self.a,
                               periodic_length=self.periodic_length, out=F, **self.blob_blob)
            acc = 1
        _lib.check(self.lib.rmw_rollers_one_blob_forces(S.r_blobs.data_ptr(), S.Nblobs, self.blob_mass, self.g,
                                                        self.a, self.repulsion_strength_wall, self.debye_length_wall,
                                                        acc, F.data_ptr(), _stream()))
        return F

    def force_torque(self, S: RigidSuspension) -> torch.Tensor:
        """(B, 6) [F, T] per body (force_torque_calculator_sort_by_bodies)."""
        F = self.blob_forces(S)
        ft = torch.empty((S.nbodies, 6), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.rmw_rigid_force_torque(S.r_rel.data_ptr(), S._offp(), S.nbodies, F.data_ptr(),
                                                   ft.data_ptr(), S.ws.data_ptr(), S.ws.numel(), _stream()))
        return ft

    def slip(self, S: RigidSuspension) -> torch.Tensor:
        if self.slip_reference is None:
            return torch.zeros((S.Nblobs, 3), dtype=torch.float64, device=self.device)
        if self._slip_ref_dev is None:
            ref = (np.vstack([np.asarray(s, dtype=np.float64).reshape(-1, 3) for s in self.slip_reference])
                   if isinstance(self.slip_reference, (list, tuple)) else
                   np.tile(np.asarray(self.slip_reference, dtype=np.float64).reshape(-1, 3), (self.nbodies, 1)))
            self._slip_ref_dev = torch.from_numpy(np.ascontiguousarray(ref)).to(self.device)
        sref = self._slip_ref_dev
        zero = torch.zeros_like(S.centres)
        rot = torch.empty((S.Nblobs, 3), dtype=torch.float64, device=self.device)
        dummy = torch.empty_like(rot)
        # the body rotation applied to the reference-frame slip (centres at 0)
        _lib.check(self.lib.rmw_body_blobs(zero.data_ptr(), S.quats.data_ptr(), sref.data_ptr(), S._offp(),
                                           S.nbodies, dummy.data_ptr(), rot.data_ptr(), S.ws.data_ptr(),
                                           S.ws.numel(), _stream()))
        return rot

    def _draw(self, kind: str, shape) -> torch.Tensor:
        if self.noise_source is not None:
            t = torch.as_tensor(self.noise_source(kind, shape))
            return t.to(self.device, torch.float64).reshape(shape).contiguous()
        return torch.randn(shape, dtype=torch.float64, device=self.device, generator=self.gen)

    def _preconditioner(self, S: RigidSuspension, step: int, stochastic: bool = False) -> None:
        """Rebuilt when step % update_PC == 0 or none exists, else the last
        one is reused (multi_bodies.py:766-819; the stochastic one with its
        Cholesky factors, :522-555)."""
        stale_kind = stochastic and self._pc_state is not None and self._pc_state[4] is None
        if self._pc_state is None or step % self.update_PC == 0 or stale_kind:
            if stochastic:
                S.build_stochastic_preconditioner()
            else:
                S.build_preconditioner()
            self.pc_builds += 1
            self._pc_state = S.preconditioner_state() if self.update_PC > 1 else None
            if self.update_PC > 1:
                S.adopt_preconditioner(self._pc_state)
        else:
            S.adopt_preconditioner(self._pc_state)

    def solve_mobility_problem(self, S: RigidSuspension, step: Optional[int] = None,
                               noise: Optional[torch.Tensor] = None, stochastic: bool = False):
        """Body velocities (B, 6) of the current configuration (:1441-1548);
        `noise` (3N) is subtracted from the slip of the right-hand side
        (:1490-1492)."""
        self._preconditioner(S, self.step_count if step is None else step, stochastic=stochastic)
        rhs = S.rhs(self.slip(S), self.force_torque(S))
        if noise is not None:
            rhs[: 3 * S.Nblobs] -= noise.reshape(-1)
        rnorm = S.norm(rhs)
        x0 = None if (self.first_guess is None or rnorm == 0.0) else self.first_guess * rnorm
        if self.solver == "mixed":
            res = S.solve_mixed(rhs=rhs, tol=self.tolerance, restart=self.restart, maxiter=1000, x0=x0,
                                rhs_norm=rnorm)
        else:
            res = S.solve_mobility_problem(rhs=rhs, tol=self.tolerance, restart=self.restart, maxiter=1000, x0=x0,
                                           rhs_norm=rnorm)
        if res.info != 0:
            raise RuntimeError(f"GMRES did not reach {self.tolerance} (residual {res.final_residual:.3e})")
        self.det_iterations_count += res.iterations
        self.first_guess = res.x / rnorm if rnorm > 0 else torch.zeros_like(res.x)
        return res.x[3 * S.Nblobs:].reshape(S.nbodies, 6).contiguous()

    def _trial(self, S: RigidSuspension, U: torch.Tensor, dt: float, ab: bool) -> bool:
        """Move the bodies with velocities U (and the previous step's, for
        Adams-Bashforth) into the trial configuration; True when every blob
        stays above the wall (check_positions, :1779-1829)."""
        if self._S_next is None:
            self._S_next = self.suspension()
            nb = ctypes.c_size_t()
            _lib.check(self.lib.rmw_count_below_workspace_size(S.Nblobs, ctypes.byref(nb)))
            self._count_ws = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=self.device)
            self._count = torch.empty(1, dtype=torch.int64, device=self.device)
        Sn = self._S_next
        # the new centres and quaternions go straight into the trial suspension
        _lib.check(self.lib.rmw_bodies_advance(S.centres.data_ptr(), S.quats.data_ptr(), U.data_ptr(),
                                               self.velocities_previous_step.data_ptr() if ab else None,
                                               S.nbodies, 1.5 if ab else 1.0, 0.5 if ab else 0.0, float(dt),
                                               Sn.centres.data_ptr(), Sn.quats.data_ptr(), _stream()))
        Sn.set_configuration(Sn.centres, Sn.quats)
        if self.domain != "single_wall":
            return True          # check_positions tests the wall only above one (:1786)
        self._count.zero_()
        _lib.check(self.lib.rmw_count_below(Sn.r_blobs.data_ptr(), Sn.Nblobs, 0.0, self._count.data_ptr(),
                                            self._count_ws.data_ptr(), self._count_ws.numel(), _stream()))
        return int(self._count.item()) == 0

    def _accept(self, S: RigidSuspension) -> None:
        self._S, self._S_next = self._S_next, S
        self._centres = None
        self._quats = None
        self.step_count += 1

    def _advance(self, dt: float, adams_bashforth: bool, step: Optional[int] = None) -> None:
        S = self._current()
        U = self.solve_mobility_problem(S, step)
        ab = adams_bashforth and not self.first_step
        if not self._trial(S, U, dt, ab):
            # check_positions fails; a deterministic step would be recomputed
            # unchanged forever by the reference
            self.invalid_configuration_count += 1
            raise RuntimeError("the deterministic step moves a blob below the wall")
        if adams_bashforth:
            self.first_step = False
            self.velocities_previous_step = U
        self._accept(S)

    def deterministic_forward_euler(self, dt: float, *args, **kwargs):
        self._advance(dt, adams_bashforth=False, step=kwargs.get("step"))

    def deterministic_adams_bashforth(self, dt: float, *args, **kwargs):
        self._advance(dt, adams_bashforth=True, step=kwargs.get("step"))

    def brownian_slip(self, S: RigidSuspension, dt: float) -> torch.Tensor:
        """sqrt(2 kT / dt) L (P^T M P)^1/2 z on the blobs (3N), Lanczos to the
        solver tolerance (:287-296)."""
        if self.kT == 0.0:
            return torch.zeros(3 * S.Nblobs, dtype=torch.float64, device=self.device)
        z = self._draw("z", (1, 3 * S.Nblobs))
        noise, its = S.brownian_slip(z, factor=math.sqrt(2.0 * self.kT / dt), tolerance=self.tolerance,
                                     precision=self.lanczos_precision)
        self.stoch_iterations_count += its
        return noise[0]

    def stochastic_EM(self, dt: float, *args, **kwargs):
        """Euler-Maruyama step (:262-300): build the stochastic preconditioner
        (every update_PC steps), draw the blob noise, solve the mobility
        problem with it, move the bodies with the solved velocities; redraw
        while a blob ends below the wall."""
        step = kwargs.get("step")
        while True:
            S = self._current()
            self._preconditioner(S, self.step_count if step is None else step, stochastic=True)
            noise = self.brownian_slip(S, dt)
            U = self.solve_mobility_problem(S, step, noise=noise, stochastic=True)
            if self._trial(S, U, dt, ab=False):
                self._accept(S)
                return
            self.invalid_configuration_count += 1

    def advance_time_step(self, dt: float, *args, **kwargs):
        """One step of the selected scheme (:68-72)."""
        return getattr(self, self.scheme)(dt, *args, **kwargs)

    def blob_positions(self) -> np.ndarray:
        return self._current().r_blobs.cpu().numpy()
