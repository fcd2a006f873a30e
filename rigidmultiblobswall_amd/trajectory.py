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

Time step (Euler-Maruyama, stochastic_EM of quaternion_integrator/
quaternion_integrator_multi_bodies.py:262-300): the Brownian blob noise
sqrt(2 kT / dt) L (P^T M P)^1/2 z comes from Lanczos on the stochastic
block-diagonal preconditioner (multi_bodies.py:480-623), so it carries the
hydrodynamic correlation between all blobs; the rigid solve of the
right-hand side [-noise; -F] gives the body velocities.  By linearity that
solve is split into the deterministic part [0; -F] (its solution is what a
snapshot stores: the (geometry, lambda) pair the warm start learns, solved
warm-started from the previous step) and the noise part [-noise; 0]; both
are solved to the same FP64 residual.  Quaternions are advanced by the
rotation vector w dt (quaternion.py:15-38); a step that moves a blob below
the wall is redrawn.
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
                      field_radius: float = 0.9, solver: str = "mixed", lanczos_tol: float = 1e-3,
                      lanczos_precision: str = "fp32_blocked") -> Trajectory:
    """Brownian trajectory of rigid bodies; solver 'mixed' (FP32-pair inner
    GMRES refined to the FP64 tolerance) or 'fp64' for both solves of a
    step."""
    fp = fp or ForceParams()
    dev = torch.device("cuda", torch.cuda.current_device())
    gen = torch.Generator(device=dev)
    gen.manual_seed(int(seed) + 1000)
    centres = np.array(centres, dtype=np.float64)
    quats = np.array(quats, dtype=np.float64)
    traj = Trajectory(params=dict(nsteps=nsteps, dt=dt, kT=kT, eta=eta, a=a, tol=tol, forces=fp.__dict__,
                                  precision=precision, solver=solver, lanczos_tol=lanczos_tol,
                                  lanczos_precision=lanczos_precision,
                                  noise="sqrt(2 kT/dt) L (P^T M P)^1/2 z (Lanczos, stochastic block-diagonal "
                                        "preconditioner)"))
    S = RigidSuspension(centres, quats, ref, eta, a, precision=precision)
    factor = float(np.sqrt(2.0 * kT / dt))

    def solve(rhs, x0):
        if solver == "mixed":
            res = S.solve_mixed(rhs=rhs, tol=tol, x0=x0)
        else:
            res = S.solve_mobility_problem(rhs=rhs, tol=tol, x0=x0)
        if res.final_residual is not None and res.final_residual > 1.5 * tol:
            raise RuntimeError(f"GMRES did not reach {tol} (residual {res.final_residual:.3e})")
        return res

    prev = None
    for step in range(nsteps):
        if step:
            S.set_configuration(centres, quats)
        S.build_stochastic_preconditioner()
        ft = body_forces(centres, quats, ref, a, fp)
        rhs = S.rhs(np.zeros((S.Nblobs, 3)), ft)
        pc = torch.empty_like(rhs)
        S.pc_apply(rhs, pc)
        base = S.point_field_guess(rhs, ft, radius=field_radius)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        res = solve(rhs, prev)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t0
        traj.snapshots.append(Snapshot(centres.copy(), quats.copy(), ft, res.x.clone(), pc, res.iterations, ts, base))
        prev = res.x
        U_det = S.velocities(res.x)                   # (B, 6) [v, w]
        attempts = 0
        while True:
            attempts += 1
            z = torch.randn((1, 3 * S.Nblobs), dtype=torch.float64, device=dev, generator=gen)
            noise, its = S.brownian_slip(z, factor=factor, tolerance=lanczos_tol, precision=lanczos_precision)
            rhs_n = torch.zeros_like(rhs)
            rhs_n[: 3 * S.Nblobs] = -noise[0]
            x0_n = torch.empty_like(rhs)
            S.pc_apply(rhs_n, x0_n)
            res_n = solve(rhs_n, x0_n)
            U = U_det + S.velocities(res_n.x)
            c_new = centres + U[:, :3] * dt
            q_new = quat_mul(quat_from_rotation(U[:, 3:] * dt), quats)
            q_new /= np.linalg.norm(q_new, axis=1, keepdims=True)
            blobs_z = (np.einsum("bij,kj->bki", rotation_matrices(q_new), ref)[..., 2] + c_new[:, None, 2])
            if blobs_z.min() >= 0.0 or attempts >= 20:   # check_positions, qi_multi_bodies.py:1779
                break
        if log:
            log(f"step {step}: GMRES iterations {res.iterations} (noise part {res_n.iterations}), Lanczos {its}, "
                f"{ts * 1e3:.1f} ms, {attempts} attempt(s)")
        centres, quats = c_new, q_new
    return traj
