"""CPU pins of the oracle's rigid solve and Lanczos (no GPU).

* pair_active_rods (reference example, GMRES tol 1e-8): body velocities
  match the committed run_{low,mid,high}_res.velocity.dat.reference.
* the block-diagonal preconditioner is the exact inverse for one body.
* Lanczos M^1/2 z agrees with the eigen-decomposition square root
  (stochastic_forcing.py:44-82; stochastic_forcing/test.py:60-101) and is
  exact in one iteration for M = c I.
"""
import numpy as np
import pytest

from oracle import oracle as O
import golden_cases as G


def _oracle_system(case, matvec="dense"):
    bodies = [(l, q, case["ref"]) for l, q in zip(case["locations"], case["quats"])]
    S = O.RigidSystem(bodies, case["eta"], case["a"])
    if matvec == "dense":   # the example's 'python' implementation (dense A-E product)
        S.mvp = lambda r, v: O.single_wall_fluid_mobility(r, S.eta, S.a) @ v
    return S


@pytest.mark.parametrize("res,tol", [("low", 1e-8), ("mid", 1e-8), ("high", 5e-8)])
def test_pair_active_rods_golden(res, tol):
    case = G.load_case(res)
    S = _oracle_system(case, "kernel")   # numba-order kernel restatement as the matvec
    x, it = S.solve(S.rhs(case["slip"], case["force_torque"]), tol=case["tol"])
    vel = x[3 * S.Nblobs:].reshape(-1, 6)
    err = np.linalg.norm(vel - case["expected"]) / np.linalg.norm(case["expected"])
    assert err < tol, (res, err)


def test_block_pc_is_exact_for_one_body():
    case = G.load_case("mid")
    bodies = [(case["locations"][0], case["quats"][0], case["ref"])]
    S = O.RigidSystem(bodies, case["eta"], case["a"])
    S.mvp = lambda r, v: O.single_wall_fluid_mobility(r, S.eta, S.a) @ v
    S.build_pc()
    rng = np.random.default_rng(0)
    b = rng.standard_normal(S.size)
    y = S.apply(S.pc_apply(b))
    assert np.linalg.norm(y - b) < 1e-10 * np.linalg.norm(b)
    x, it = S.solve(b, tol=1e-10)
    assert it <= 2


def test_lanczos_matches_eigen_sqrt():
    rng = np.random.default_rng(1)
    n, a = 15, 0.3
    r = np.column_stack([rng.uniform(0, 2, n), rng.uniform(0, 2, n), rng.uniform(0.5, 1.5, n)])
    M = O.single_wall_fluid_mobility(r, 1.0, a)
    z = rng.standard_normal(3 * n)
    lam, V = np.linalg.eigh(M)
    ref = V @ (np.sqrt(np.maximum(lam, 0)) * (V.T @ z))
    noise, it = O.lanczos_sqrt(lambda v: M @ v, z, tolerance=1e-10)
    assert np.linalg.norm(noise - ref) < 1e-8 * np.linalg.norm(ref)
    noise1, it1 = O.lanczos_sqrt(lambda v: 3.0 * v, z, tolerance=1e-10)
    assert np.allclose(noise1, np.sqrt(3.0) * z, rtol=1e-13) and it1 <= 1


def _spheres(nb, seed, nblob=12):
    """nb small spheres (icosahedron vertices, radius 1) above the wall."""
    from rigidmultiblobswall_amd import workloads
    w = workloads.c2_rigid_spheres(seed=seed, nbodies=nb)
    ico = workloads.icosphere_42(1.0)[:nblob]
    bodies = [(c, q, ico) for c, q in zip(w["centres"], w["quats"])]
    return O.RigidSystem(bodies, 1.0, 0.35)


def _blockdiag(blocks):
    n = sum(b.shape[0] for b in blocks)
    D = np.zeros((n, n))
    o = 0
    for b in blocks:
        D[o:o + b.shape[0], o:o + b.shape[0]] = b
        o += b.shape[0]
    return D


def test_stochastic_pc_factors():
    """P_inv = U^T and P = U^-1 with M_b = U^T U (multi_bodies.py:520-534):
    P_inv P_inv^T = M_b, P_inv^T P = I; the preconditioned mobility is the
    identity on the body blocks (a single body: P^T M P = I exactly)."""
    S = _spheres(3, 5)
    S.mvp = lambda r, v: O.single_wall_fluid_mobility(r, S.eta, S.a) @ v
    S.build_stochastic_pc()
    for k in range(3):
        Mb = S.mblobs(S.r_bodies[k])
        assert np.abs(S.P_inv[k] @ S.P_inv[k].T - Mb).max() <= 1e-13 * np.abs(Mb).max()
        assert np.abs(S.P_inv[k].T @ S.P[k] - np.eye(Mb.shape[0])).max() <= 1e-11
        assert np.abs(np.tril(S.P_inv[k], -1) - S.P_inv[k] + np.diag(np.diag(S.P_inv[k]))).max() == 0.0
    one = O.RigidSystem([(S.locations[0], np.array([1.0, 0, 0, 0]), S.r_bodies[0] - S.locations[0])], 1.0, 0.35)
    one.mvp = S.mvp
    one.build_stochastic_pc()
    n = 3 * one.Nblobs
    D = np.column_stack([one.mobility_pc(e) for e in np.eye(n)])
    assert np.abs(D - np.eye(n)).max() <= 1e-11


def test_stochastic_noise_covariance_and_lanczos():
    """noise = P_inv (P^T M P)^1/2 z has covariance P_inv (P^T M P) P_inv^T =
    M (the dense matrices), and Lanczos on mobility_pc with L_mult gives
    that noise (stochastic_forcing.py:253-256)."""
    S = _spheres(4, 6)
    S.mvp = lambda r, v: O.single_wall_fluid_mobility(r, S.eta, S.a) @ v
    S.build_stochastic_pc()
    M = O.single_wall_fluid_mobility(S.r, S.eta, S.a)
    n = 3 * S.Nblobs
    A = np.column_stack([S.mobility_pc(e) for e in np.eye(n)])
    Pinv = _blockdiag(S.P_inv)
    assert np.abs(A - A.T).max() <= 1e-12 * np.abs(A).max()
    lam, V = np.linalg.eigh(A)
    Ah = V @ np.diag(np.sqrt(np.maximum(lam, 0))) @ V.T
    G = Pinv @ Ah
    assert np.abs(G @ G.T - M).max() <= 1e-12 * np.abs(M).max()
    z = np.random.default_rng(7).standard_normal(n)
    ref = G @ z
    noise, it = O.lanczos_sqrt(S.mobility_pc, z, factor=1.0, tolerance=1e-12)
    got = S.P_inv_mult(noise)
    assert np.linalg.norm(got - ref) <= 1e-9 * np.linalg.norm(ref)
    # the preconditioned operator is close to the identity: few iterations
    noise_m, it_m = O.lanczos_sqrt(lambda v: M @ v, z, tolerance=1e-12)
    assert it < it_m
