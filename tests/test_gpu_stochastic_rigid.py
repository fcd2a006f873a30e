"""GPU parity of the stochastic rigid-body path: the stochastic block-diagonal
preconditioner (multi_bodies.py:480-623), the Brownian blob noise
L (P^T M P)^1/2 z by Lanczos with L_mult (stochastic_forcing.py:112-262) and
the Euler-Maruyama step (quaternion_integrator_multi_bodies.py:262-300),
against the oracle and the dense eigen-decomposition square root."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from oracle import rigid_step_oracle as RS
from rigidmultiblobswall_amd import lanczos, workloads
from rigidmultiblobswall_amd.rigid import RigidSuspension
from rigidmultiblobswall_amd.rigid_integrator import RigidIntegrator

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _system(nb, seed, nblob=42):
    w = workloads.c2_rigid_spheres(seed=seed, nbodies=nb)
    ref = workloads.icosphere_42(1.0)[:nblob]
    S = RigidSuspension(w["centres"], w["quats"], ref, w["eta"], w["a"], precision="fp64")
    Og = O.RigidSystem([(c, q, ref) for c, q in zip(w["centres"], w["quats"])], w["eta"], w["a"])
    return w, ref, S, Og


def _blocks(flat, offsets):
    out, o = [], 0
    for b in range(len(offsets) - 1):
        n = 3 * int(offsets[b + 1] - offsets[b])
        out.append(flat[o:o + n * n].reshape(n, n))
        o += n * n
    return out


@pytest.mark.parametrize("nblob", [42, 12])
def test_factors_vs_oracle(gpu, nblob):
    w, ref, S, Og = _system(5, 41, nblob)
    print(f"case stochastic PC factors, 5 bodies x {nblob} blobs", flush=True)
    S.build_stochastic_preconditioner()
    Og.build_stochastic_pc()
    L = _blocks(S.chol.cpu().numpy(), S.body_offsets)
    X = _blocks(S.chol_inv.cpu().numpy(), S.body_offsets)
    for k in range(5):
        assert np.abs(L[k] - Og.P_inv[k]).max() <= 1e-12 * np.abs(Og.P_inv[k]).max()
        assert np.abs(X[k] - Og.P[k].T).max() <= 1e-10 * np.abs(Og.P[k]).max()
    Minv = _blocks(S.minv.cpu().numpy(), S.body_offsets)
    for k in range(5):
        assert np.abs(Minv[k] - Og.Minv[k]).max() <= 1e-9 * np.abs(Og.Minv[k]).max()
        assert np.abs(S.nbody[k].cpu().numpy() - Og.N[k]).max() <= 1e-9 * np.abs(Og.N[k]).max()


def test_mobility_pc_and_P_inv_vs_oracle(gpu):
    w, ref, S, Og = _system(6, 42)
    print("case mobility_pc / P_inv_mult, 6 bodies x 42 blobs, 3 RHS", flush=True)
    S.build_stochastic_preconditioner()
    Og.build_stochastic_pc()
    rng = np.random.default_rng(3)
    n = 3 * S.Nblobs
    Xh = rng.standard_normal((3, n))
    X = torch.from_numpy(Xh).to(gpu)
    Y = torch.empty_like(X)
    S.mobility_pc(X, Y, precision="fp64")
    for r in range(3):
        assert rel(Y[r].cpu().numpy(), Og.mobility_pc(Xh[r])) <= 1e-12
        assert rel(S.P_inv_mult(X[r]).cpu().numpy(), Og.P_inv_mult(Xh[r])) <= 1e-13
    S.mobility_pc(X, Y, precision="fp32_blocked")
    assert rel(Y[0].cpu().numpy(), Og.mobility_pc(Xh[0])) <= 1e-6


def test_brownian_slip_matches_dense_square_root(gpu):
    """factor L (P^T M P)^1/2 z against the dense construction of the same
    square root, and the covariance identity."""
    w, ref, S, Og = _system(3, 43, nblob=12)
    print("case Brownian slip vs dense square root, 3 bodies x 12 blobs", flush=True)
    Og.build_stochastic_pc()
    n = 3 * S.Nblobs
    A = np.column_stack([Og.mobility_pc(e) for e in np.eye(n)])
    lam, V = np.linalg.eigh(0.5 * (A + A.T))
    Pinv = np.zeros((n, n))
    for k in range(3):
        s = slice(3 * Og.offsets[k], 3 * Og.offsets[k + 1])
        Pinv[s, s] = Og.P_inv[k]
    G = Pinv @ V @ np.diag(np.sqrt(np.maximum(lam, 0))) @ V.T
    Z = np.random.default_rng(5).standard_normal((2, n))
    noise, it = S.brownian_slip(torch.from_numpy(Z).to(gpu), factor=0.7, tolerance=1e-12, precision="fp64")
    for r in range(2):
        assert rel(noise[r].cpu().numpy(), 0.7 * G @ Z[r]) <= 1e-8
    print(f"  Lanczos iterations {it}")


def test_stochastic_em_steps_match_oracle(gpu):
    """Three Euler-Maruyama steps on shared draws: GPU integrator = oracle
    integrator (Lanczos and GMRES to 1e-10)."""
    w = workloads.c2_rigid_spheres(seed=44, nbodies=4)
    ref = w["ref"]
    draws = [np.random.default_rng(100 + i).standard_normal(3 * 4 * 42) for i in range(12)]
    it_gpu = iter(draws)
    kw = dict(eta=w["eta"], a=w["a"], g=0.05, repulsion_strength_wall=0.05, debye_length_wall=0.2)
    integ = RigidIntegrator(w["centres"], w["quats"], ref, scheme="stochastic_EM", tolerance=1e-10, kT=0.2,
                            noise_source=lambda kind, shape: next(it_gpu), **kw)
    ora = RS.RigidStep(w["centres"], w["quats"], ref, "stochastic_EM", w["eta"], w["a"], g=0.05, eps_wall=0.05,
                       b_wall=0.2, tol=1e-10)
    it_ora = iter(draws)
    dt = 1e-2
    print("case stochastic_EM 4 spheres x 42 blobs, 3 steps, kT 0.2", flush=True)
    c0 = w["centres"].copy()
    for step in range(3):
        integ.advance_time_step(dt)
        ora.advance_stochastic(dt, 0.2, it_ora)
    disp = np.linalg.norm(ora.c - c0)
    assert disp > 1e-3
    assert np.linalg.norm(integ.centres - ora.c) <= 1e-7 * disp
    qs = np.sign(np.sum(integ.quats * ora.q, axis=1))[:, None]
    assert np.abs(integ.quats - qs * ora.q).max() <= 1e-7
    assert integ.stoch_iterations_count > 0


def test_c4_brownian_slip_invariant(gpu):
    """Full C4 size (1024 spheres x 42 blobs): the Lanczos noise of the
    preconditioned mobility S = P^T M P satisfies |S^1/2 z|^2 = z^T S z,
    and the Brownian slip is finite."""
    w = workloads.c4_monolayer()
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    print("case C4 Brownian slip, 1024 spheres x 42 blobs", flush=True)
    S.build_stochastic_preconditioner()
    n = 3 * S.Nblobs
    Z = torch.from_numpy(np.random.default_rng(9).standard_normal((1, n))).to(gpu)
    pre, it = lanczos.lanczos_block(lambda X, Y: S.mobility_pc(X, Y, "fp64"), Z, tolerance=1e-6)
    SZ = torch.empty_like(Z)
    S.mobility_pc(Z, SZ, "fp64")
    zsz = float((Z * SZ).sum())
    nn = float((pre * pre).sum())
    print(f"  Lanczos iterations {it}: |S^1/2 z|^2 = {nn:.10e}, z^T S z = {zsz:.10e}")
    assert abs(nn - zsz) <= 1e-9 * abs(zsz)
    noise, _ = S.brownian_slip(Z, factor=1.0, tolerance=1e-3, precision="fp32_blocked")
    assert bool(torch.isfinite(noise).all())
