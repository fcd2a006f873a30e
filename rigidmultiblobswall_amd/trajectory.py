"""Brownian trajectories of rigid multiblob spheres above a wall (C4 monolayer).

Used to generate (geometry, lambda) pairs for the learned GMRES warm start.

Forces follow the reference's blob-force model
(multi_bodies/multi_bodies_functions.py:153-190 and :297-320):
  gravity         f_z = -g m                               per blob
  wall repulsion  f_z = (eps_w / b_w) exp(-(h - a) / b_w)  (h > a), eps_w / b_w otherwise
  body repulsion  F = (eps / b) exp(-(r - d) / b) along the centre line
                  (r > d, d = 2 sphere radii), eps / b otherwise
and body force/torque are sums over blobs (force_torque_calculator_sort_by_
bodies, :411-445).

Time step (Euler-Maruyama): the deterministic velocity U is the rigid
mobility solve (GPU, FP64, warm-started from the previous step); the
Brownian displacement sqrt(2 kT dt) N_b^1/2 W uses each body's 6x6 mobility
N_b = pinv(K^T M_b^-1 K) from the block-diagonal preconditioner, i.e. it
omits the hydrodynamic correlation between bodies in the noise (the drift
is exact).  Quaternions are advanced by the rotation vector
w dt + dtheta (quaternion_integrator/quaternion.py:15-38).
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np
import torch

from .rigid import RigidSuspension
from .workloads import rotation_matrices


@dataclass
class ForceParams:
    g: float = 0.02            # per blob, blob mass 1
    eps_wall: float = 0.02
    b_wall: float = 0.1
    eps_body: float = 0.1
    b_body: float = 0.1
    contact: float = 2.0       # centre distance at contact (two unit spheres)


def body_forces(centres, quats, ref, a, fp: ForceParams) -> np.ndarray:
    """(B, 6) force and torque per body."""
    R = rotation_matrices(quats)
    rel = np.einsum("bij,kj->bki", R, ref)            # (B, nb, 3)
    r = rel + centres[:, None, :]
    h = r[..., 2]
    fz = -fp.g * np.ones_like(h)
    wall = np.where(h > a, (fp.eps_wall / fp.b_wall) * np.exp(-(h - a) / fp.b_wall), fp.eps_wall / fp.b_wall)
    fz = fz + wall
    f = np.zeros_like(r)
    f[..., 2] = fz
    F = f.sum(axis=1)
    T = np.cross(rel, f).sum(axis=1)
    # soft repulsion between sphere centres
    d = centres[:, None, :] - centres[None, :, :]
    dist = np.linalg.norm(d, axis=-1)
    np.fill_diagonal(dist, np.inf)
    mag = np.where(dist > fp.contact, (fp.eps_body / fp.b_body) * np.exp(-(dist - fp.contact) / fp.b_body),
                   fp.eps_body / fp.b_body)
    mag[~np.isfinite(dist)] = 0.0
    F += np.einsum("ij,ijk->ik", mag / np.where(np.isfinite(dist), dist, 1.0), d)
    return np.hstack([F, T])


def quat_from_rotation(phi: np.ndarray) -> np.ndarray:
    """(n, 3) rotation vectors -> (n, 4) quaternions (quaternion.py:15-26)."""
    n = np.linalg.norm(phi, axis=1)
    s = np.cos(n / 2.0)
    with np.errstate(invalid="ignore", divide="ignore"):
        p = np.where(n[:, None] > 0, np.sin(n / 2.0)[:, None] * phi / n[:, None], 0.0)
    return np.hstack([s[:, None], p])


def quat_mul(q1: np.ndarray, q2: np.ndarray) -> np.ndarray:
    """Batched q1 * q2 (quaternion.py:30-38)."""
    s1, p1 = q1[:, 0], q1[:, 1:]
    s2, p2 = q2[:, 0], q2[:, 1:]
    s = s1 * s2 - np.sum(p1 * p2, axis=1)
    p = s1[:, None] * p2 + s2[:, None] * p1 + np.cross(p1, p2)
    return np.hstack([s[:, None], p])


@dataclass
class Snapshot:
    centres: np.ndarray
    quats: np.ndarray
    force_torque: np.ndarray
    sol: torch.Tensor          # FP64 solution [lambda; U] on the device
    pc: torch.Tensor           # P^-1 b on the device
    iterations: int
    solve_s: float
    base: Optional[torch.Tensor] = None   # point-field block-Jacobi guess P^-1 (b - [u; 0])


@dataclass
class Trajectory:
    snapshots: List[Snapshot] = field(default_factory=list)
    params: dict = field(default_factory=dict)


def run_c4_trajectory(centres, quats, ref, a, eta, nsteps=200, dt=1e-3, kT=0.1, seed=4, tol=1e-8,
                      fp: Optional[ForceParams] = None, precision="fp64", log=None,
                      field_radius: float = 0.9) -> Trajectory:
    fp = fp or ForceParams()
    rng = np.random.default_rng(seed + 1000)
    centres = np.array(centres, dtype=np.float64)
    quats = np.array(quats, dtype=np.float64)
    traj = Trajectory(params=dict(nsteps=nsteps, dt=dt, kT=kT, eta=eta, a=a, tol=tol, forces=fp.__dict__,
                                  precision=precision))
    prev = None
    for step in range(nsteps):
        S = RigidSuspension(centres, quats, ref, eta, a, precision=precision)
        S.build_preconditioner()
        ft = body_forces(centres, quats, ref, a, fp)
        rhs = S.rhs(np.zeros((S.Nblobs, 3)), ft)
        pc = torch.empty_like(rhs)
        S.pc_apply(rhs, pc)
        base = S.point_field_guess(rhs, ft, radius=field_radius)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = S.solve_mobility_problem(rhs=rhs, tol=tol, x0=prev)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        if res.info != 0:
            raise RuntimeError(f"GMRES did not converge at step {step}")
        traj.snapshots.append(Snapshot(centres.copy(), quats.copy(), ft, res.x.clone(), pc, res.iterations, ts, base))
        if log:
            log(f"step {step}: GMRES iterations {res.iterations}, {ts * 1e3:.1f} ms")
        prev = res.x
        U = S.velocities(res.x)                       # (B, 6) [v, w]
        N = S.nbody.cpu().numpy()
        N = 0.5 * (N + np.transpose(N, (0, 2, 1)))
        L = np.linalg.cholesky(N + 1e-14 * np.eye(6))
        for _ in range(20):
            W = rng.standard_normal((len(centres), 6))
            dX = np.sqrt(2.0 * kT * dt) * np.einsum("bij,bj->bi", L, W)
            c_new = centres + U[:, :3] * dt + dX[:, :3]
            q_new = quat_mul(quat_from_rotation(U[:, 3:] * dt + dX[:, 3:]), quats)
            q_new /= np.linalg.norm(q_new, axis=1, keepdims=True)
            blobs_z = (np.einsum("bij,kj->bki", rotation_matrices(q_new), ref)[..., 2] + c_new[:, None, 2])
            if blobs_z.min() > 0.0:       # check_positions, qi_multi_bodies.py:1779
                break
        centres, quats = c_new, q_new
    return traj
