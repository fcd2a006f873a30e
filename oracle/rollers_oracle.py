"""FP64 CPU restatement of the rollers time step.

TEST INFRASTRUCTURE ONLY (imported by tests/ and the benchmark's CPU leg;
never by the product package).

Follows quaternion_integrator/quaternion_integrator_rollers.py:
  schemes                          :119-302
  deterministic velocity / torque  :905-983 (hydrodynamic), :985-1080 (uncorrelated)
  stochastic velocity              :1203-1261 (hydrodynamic), :1263-1313 (uncorrelated)
  free-kinematics torque           :1586-1597
and the forces bound by multi_bodies/multi_bodies.py:1325-1336:
  one-blob forces                  multi_bodies_functions.py:151-196
  blob-blob forces                 forces_numba.py:12-56 (the oracle's C restatement)
Products use the oracle's C restatement of the wall kernels with their
wrappers, the Lanczos square root and GMRES of oracle.py.  Random draws are
passed in (z and dx per attempt) so that the GPU integrator can be checked
on the same draws.
"""
from __future__ import annotations

import math
from typing import Callable, Iterator, Optional

import numpy as np

from . import oracle as O


def one_blob_forces(r, blob_mass, g, a, eps_wall, b_wall):
    """Gravity and wall repulsion per blob (multi_bodies_functions.py:151-196)."""
    r = np.asarray(r, dtype=np.float64).reshape(-1, 3)
    f = np.zeros_like(r)
    f[:, 2] = -g * blob_mass
    h = r[:, 2]
    lr = h > a
    f[lr, 2] += (eps_wall / b_wall) * np.exp(-(h[lr] - a) / b_wall)
    f[~lr, 2] += eps_wall / b_wall
    return f


def _self_coefs(z, eta, a):
    h_over_a = z / a
    h = max(h_over_a, 1.0)
    damping = 1.0
    if h_over_a < 0.0:
        damping = 0.0
    elif h_over_a <= 1.0:
        damping = h_over_a
    ftt = 1 / (6 * math.pi * eta * a)
    frt = 1 / (6 * math.pi * eta * a ** 2)
    mu_rt = frt * (3 / (32 * h ** 4)) * damping
    mu_perp = ftt * (1 - 9 / (8 * h) + 1 / (2 * h ** 3) - 1 / (8 * h ** 5)) * damping
    mu_para = ftt * (1 - 9 / (16 * h) + 2 / (16 * h ** 3) - 1 / (16 * h ** 5)) * damping
    return h, damping, mu_rt, mu_perp, mu_para


def uncorrelated_velocity(r, force, omega, free_kinematics, eta, a):
    """(velocity, torque), each (N, 3): quaternion_integrator_rollers.py:985-1080."""
    r = np.asarray(r, dtype=np.float64).reshape(-1, 3)
    f = np.asarray(force, dtype=np.float64).reshape(-1, 3)
    n = r.shape[0]
    v, t = np.empty((n, 3)), np.empty((n, 3))
    frr = 1 / (6 * math.pi * eta * a ** 3)
    for i in range(n):
        h, damping, mu_rt, mu_perp, mu_para = _self_coefs(r[i, 2], eta, a)
        if free_kinematics == "False":
            rr_perp_inv = 1 / (frr * (3 / 4 - 3 / (32 * h ** 3)) * damping)
            rr_para_inv = 1 / (frr * (3 / 4 - 15 / (64 * h ** 3)) * damping)
            t[i, 0] = rr_para_inv * (omega[0] + mu_rt * f[i, 1])
            t[i, 1] = rr_para_inv * (omega[1] - mu_rt * f[i, 0])
            t[i, 2] = rr_perp_inv * omega[2]
        else:
            t[i] = np.asarray(omega) * 8.0 * math.pi * eta * a ** 3
        v[i, 0] = mu_para * f[i, 0] + mu_rt * t[i, 1]
        v[i, 1] = mu_para * f[i, 1] - mu_rt * t[i, 0]
        v[i, 2] = mu_perp * f[i, 2]
    return v, t


def uncorrelated_noise(r, z, eta, a, kT, dt):
    """Stochastic velocity (N, 3): quaternion_integrator_rollers.py:1263-1313."""
    r = np.asarray(r, dtype=np.float64).reshape(-1, 3)
    z = np.asarray(z, dtype=np.float64).reshape(-1, 3)
    n = r.shape[0]
    v = np.empty((n, 3))
    ftt = 1 / (6 * math.pi * eta * a)
    fd = np.sqrt(2 * kT / dt)
    for k in range(n):
        h, damping, _, mu_perp, mu_para = _self_coefs(r[k, 2], eta, a)
        deriv = ftt * (9 / (8 * h ** 2) - 3 / (2 * h ** 4) + 5 / (8 * h ** 6)) * damping
        v[k, 0:2] = fd * np.sqrt(mu_para) * z[k, 0:2]
        v[k, 2] = fd * np.sqrt(mu_perp) * z[k, 2] + kT * deriv
    return v


class RollersStep:
    """The reference's rollers integrator state machine on the oracle."""

    def __init__(self, r, scheme, eta, a, kT=0.0, g=1.0, blob_mass=1.0, eps_wall=1.0, b_wall=1.0,
                 blob_blob=True, eps=1.0, b=1.0, periodic_length=None, omega=(0.0, 0.0, 0.0),
                 free_kinematics="True", hydro_interactions=1, tolerance=1e-8, rf_delta=1e-3, cutoff=0.0,
                 kind="strict", domain="single_wall"):
        self.r = np.asarray(r, dtype=np.float64).reshape(-1, 3).copy()
        self.kind = kind          # oracle build of the products ("fast": the CPU-baseline build)
        self.domain = domain      # quaternion_integrator_rollers.py:59-92
        self.B = {"single_wall": (O.UF, O.UT, O.WF, O.WT),
                  "no_wall": (O.UF_NOWALL, O.UT_NOWALL, O.WF_NOWALL, O.WT_NOWALL),
                  "in_plane": (O.UF_INPLANE, O.UT_INPLANE, None, None)}[domain]
        self.scheme = scheme.replace("_rollers", "")
        self.eta, self.a, self.kT, self.g, self.blob_mass = eta, a, kT, g, blob_mass
        self.eps_wall, self.b_wall, self.blob_blob, self.eps, self.b = eps_wall, b_wall, blob_blob, eps, b
        self.L = None if periodic_length is None else np.asarray(periodic_length, dtype=np.float64)
        self.omega = np.asarray(omega, dtype=np.float64)
        self.free_kinematics = free_kinematics
        self.hydro = hydro_interactions
        self.tol, self.rf_delta, self.cutoff = tolerance, rf_delta, cutoff
        self.first_step = True
        self.prev = None
        self.torque_prev = None
        self.invalid = 0
        self.wall_overlaps = 0

    def _p(self, b, r, v, eta, a, periodic_length=None, kind="strict"):
        return O.block_apply(self.B[b], r, v, eta, a, periodic_length, kind=kind)

    def forces(self, r):
        f = one_blob_forces(r, self.blob_mass, self.g, self.a, self.eps_wall, self.b_wall)
        if self.blob_blob:
            f = f + O.blob_forces(r, self.eps, self.b, self.a, periodic_length=self.L, cutoff=self.cutoff,
                                  kind=self.kind)
        return f

    def det_velocity(self, r):
        n = r.shape[0]
        F = self.forces(r).ravel()
        if self.hydro != 1:
            v, t = uncorrelated_velocity(r, F, self.omega, self.free_kinematics, self.eta, self.a)
            return v.ravel(), t.ravel()
        L = self.L
        if self.free_kinematics == "False":
            rhs = np.tile(self.omega, n) - self._p(2, r, F, self.eta, self.a, periodic_length=L, kind=self.kind)
            nrm = np.linalg.norm(rhs)
            if nrm > 0:
                rhs = rhs / nrm
            sol, _ = O.gmres_right(lambda x: self._p(3, r, x, self.eta, self.a, periodic_length=L, kind=self.kind),
                                   lambda x: x, rhs, tol=self.tol, restart=20, maxiter=1000, x0=self.torque_prev)
            self.torque_prev = sol
            T = sol * nrm if nrm > 0 else sol
        else:
            T = np.tile(self.omega * 8.0 * math.pi * self.eta * self.a ** 3, n)
        v = self._p(0, r, F, self.eta, self.a, periodic_length=L, kind=self.kind)
        if np.any(T):
            v = v + self._p(1, r, T, self.eta, self.a, periodic_length=L, kind=self.kind)
        return v, T

    def stoch_velocity(self, r, dt, z, dx):
        if self.hydro != 1:
            return uncorrelated_noise(r, z, self.eta, self.a, self.kT, dt).ravel()
        L = self.L
        noise, _ = O.lanczos_sqrt(lambda v: self._p(0, r, v, self.eta, self.a, periodic_length=L, kind=self.kind),
                                  np.asarray(z).ravel(), factor=np.sqrt(2 * self.kT / dt), tolerance=self.tol)
        if self.kT > 0.0 and self.domain != "no_wall":
            return noise + self.rfd_drift(r, dx)
        return noise

    def rfd_drift(self, r, dx):
        """kT/(delta a) [M_tt(r + delta a dx/2) - M_tt(r - delta a dx/2)] dx
        (quaternion_integrator_rollers.py:1242-1261)."""
        L = self.L
        dx = np.asarray(dx, dtype=np.float64).reshape(-1, 3)
        r1 = r + dx * (self.rf_delta * self.a * 0.5)
        div = self._p(0, r1, dx.ravel(), self.eta, self.a, periodic_length=L, kind=self.kind)
        r2 = r1 - dx * self.rf_delta * self.a
        div = div - self._p(0, r2, dx.ravel(), self.eta, self.a, periodic_length=L, kind=self.kind)
        return (self.kT / (self.rf_delta * self.a)) * div

    def advance(self, dt, draws: Optional[Iterator] = None):
        """One step; draws yields (z, dx) per attempt for the stochastic
        schemes.  Returns the number of attempts."""
        stochastic = self.scheme.startswith("stochastic")
        ab = self.scheme.endswith("adams_bashforth")
        attempts = 0
        while True:
            attempts += 1
            det, _ = self.det_velocity(self.r)
            stoch = 0.0
            if stochastic:
                z, dx = next(draws)
                stoch = self.stoch_velocity(self.r, dt, z, dx)
            if ab and not self.first_step:
                vel = 1.5 * det - 0.5 * self.prev + stoch
            else:
                vel = det + stoch
            new = self.r + dt * vel.reshape(-1, 3)
            if self.domain != "single_wall" or not (new[:, 2] < 0.0).any():
                if ab:
                    self.first_step = False
                    self.prev = det
                self.r = new
                if self.domain == "single_wall" or self.scheme == "stochastic_first_order":
                    self.wall_overlaps += int((new[:, 2] < self.a).sum())
                return attempts
            self.invalid += 1
            if not stochastic:
                raise RuntimeError("deterministic step leaves a roller below the wall")
