"""FP64 CPU restatement of the rigid-body time steps.

TEST INFRASTRUCTURE ONLY (imported by tests/; never by the product).

Follows quaternion_integrator/quaternion_integrator_multi_bodies.py:
  deterministic_forward_euler              :75-107
  deterministic_forward_euler_dense_algebra :110-139 (the dense pin)
  deterministic_adams_bashforth            :142-186
  stochastic_EM                            :262-300 (with the stochastic
      preconditioner of multi_bodies.py:480-623 and stochastic_forcing_lanczos
      with L_mult, stochastic_forcing.py:112-262)
  solve_mobility_problem                   :1441-1548
  check_positions                          :1779-1829
with quaternion.py:15-49 (from_rotation, product), the body forces of
multi_bodies_functions.py:151-196 and :411-445, and the oracle's rigid
system (operator, block-diagonal preconditioner, GMRES) of oracle.py.
"""
from __future__ import annotations

import numpy as np

from . import oracle as O
from . import rollers_oracle as RO


def quat_from_rotation(phi):
    """quaternion.py:15-27."""
    phi = np.asarray(phi, dtype=np.float64)
    n = np.linalg.norm(phi)
    s = np.cos(n / 2.0)
    p = np.sin(n / 2.0) * (phi / n) if n != 0 else np.zeros(3)
    return np.concatenate([[s], p])


def quat_mul(q1, q2):
    """quaternion.py:30-38 (q1 is the left factor)."""
    s1, p1, s2, p2 = q1[0], q1[1:], q2[0], q2[1:]
    return np.concatenate([[s1 * s2 - p1 @ p2], s1 * p2 + s2 * p1 + np.cross(p1, p2)])


class RigidStep:
    def __init__(self, locations, quats, refs, scheme, eta, a, g=1.0, blob_mass=1.0, eps_wall=1.0, b_wall=1.0,
                 blob_blob=False, eps=1.0, b=1.0, tol=1e-10, slip_refs=None):
        self.c = np.asarray(locations, dtype=np.float64).reshape(-1, 3).copy()
        self.q = np.asarray(quats, dtype=np.float64).reshape(-1, 4).copy()
        B = self.c.shape[0]
        self.refs = list(refs) if isinstance(refs, (list, tuple)) else [np.asarray(refs)] * B
        self.scheme = scheme
        self.eta, self.a, self.g, self.m = eta, a, g, blob_mass
        self.eps_wall, self.b_wall, self.blob_blob, self.eps, self.b = eps_wall, b_wall, blob_blob, eps, b
        self.tol = tol
        self.slip_refs = slip_refs
        self.first_step = True
        self.prev = None

    def system(self, c=None, q=None):
        c = self.c if c is None else c
        q = self.q if q is None else q
        return O.RigidSystem([(c[k], q[k], self.refs[k]) for k in range(len(c))], self.eta, self.a)

    def force_torque(self, sys):
        f = RO.one_blob_forces(sys.r, self.m, self.g, self.a, self.eps_wall, self.b_wall)
        if self.blob_blob:
            f = f + O.blob_forces(sys.r, self.eps, self.b, self.a)
        return sys.KT_dot(f.ravel()).reshape(-1, 6)

    def slip(self, sys, q=None):
        q = self.q if q is None else q
        if self.slip_refs is None:
            return np.zeros((sys.Nblobs, 3))
        return np.vstack([np.asarray(self.slip_refs[k]) @ O.rotation_matrix(q[k]).T for k in range(len(q))])

    def velocities(self, sys=None):
        sys = sys or self.system()
        rhs = sys.rhs(self.slip(sys), self.force_torque(sys))
        x, _ = sys.solve(rhs, tol=self.tol, restart=60)
        return x[3 * sys.Nblobs:].reshape(-1, 6)

    def velocities_dense(self, sys=None):
        """deterministic_forward_euler_dense_algebra: from M lambda - K U = slip
        and K^T lambda = F, U = N (F - K^T M^-1 slip) with the dense
        N = (K^T M^-1 K)^-1 of all bodies."""
        sys = sys or self.system()
        M = O.single_wall_fluid_mobility(sys.r, self.eta, self.a)
        nb3 = 3 * sys.Nblobs
        Kfull = np.zeros((nb3, 6 * sys.Nbodies))
        for k in range(sys.Nbodies):
            Kfull[3 * sys.offsets[k]:3 * sys.offsets[k + 1], 6 * k:6 * k + 6] = sys.K[k]
        Minv = np.linalg.inv(M)
        N = np.linalg.pinv(Kfull.T @ Minv @ Kfull)
        F = self.force_torque(sys).ravel()
        slip = self.slip(sys).ravel()
        return (N @ (F - Kfull.T @ Minv @ slip)).reshape(-1, 6)

    def advance(self, dt):
        U = self.velocities()
        ab = self.scheme == "deterministic_adams_bashforth" and not self.first_step
        V = 1.5 * U - 0.5 * self.prev if ab else U
        c_new = self.c + V[:, :3] * dt
        q_new = np.array([quat_mul(quat_from_rotation(V[k, 3:] * dt), self.q[k]) for k in range(len(self.c))])
        blobs = self.system(c_new, q_new).r
        if (blobs[:, 2] < 0.0).any():
            raise RuntimeError("deterministic step moves a blob below the wall")
        if self.scheme == "deterministic_adams_bashforth":
            self.first_step = False
            self.prev = U
        self.c, self.q = c_new, q_new


    def advance_stochastic(self, dt, kT, draws, lanczos_tol=None):
        """One stochastic_EM step; draws yields the Lanczos right-hand side z
        (3N) of each attempt.  Returns the number of attempts."""
        tol = self.tol if lanczos_tol is None else lanczos_tol
        attempts = 0
        while True:
            attempts += 1
            sys = self.system()
            sys.build_stochastic_pc()
            z = np.asarray(next(draws), dtype=np.float64).ravel()
            noise, _ = O.lanczos_sqrt(sys.mobility_pc, z, factor=np.sqrt(2.0 * kT / dt), tolerance=tol)
            noise = sys.P_inv_mult(noise)
            rhs = sys.rhs(self.slip(sys).ravel() - noise, self.force_torque(sys))
            x, _ = sys.solve(rhs, tol=self.tol, restart=60)
            U = x[3 * sys.Nblobs:].reshape(-1, 6)
            c_new = self.c + U[:, :3] * dt
            q_new = np.array([quat_mul(quat_from_rotation(U[k, 3:] * dt), self.q[k]) for k in range(len(self.c))])
            if (self.system(c_new, q_new).r[:, 2] >= 0.0).all():
                self.c, self.q = c_new, q_new
                return attempts
