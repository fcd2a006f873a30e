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
