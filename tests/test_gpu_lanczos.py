"""GPU block Lanczos M^1/2 z (FP64 algebra, GPU tridiagonal QL) against the
oracle Lanczos, the eigen-decomposition square root, and size-independent
invariants at the rollers configuration (N = 32768, grand 6N, block of 4)."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import lanczos, workloads
from rigidmultiblobswall_amd.rollers import Rollers

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def test_dropin_dense_matches_eigh_and_oracle(gpu):
    rng = np.random.default_rng(0)
    n, a = 20, 0.3
    r = np.column_stack([rng.uniform(0, 2, n), rng.uniform(0, 2, n), rng.uniform(0.4, 1.5, n)])
    M = O.single_wall_fluid_mobility(r, 1.0, a)
    z = rng.standard_normal(3 * n)
    lam, V = np.linalg.eigh(M)
    exact = V @ (np.sqrt(np.maximum(lam, 0)) * (V.T @ z))
    print("case dense 60x60 Lanczos tol 1e-10", flush=True)
    got, it = lanczos.stochastic_forcing_lanczos(factor=1.0, tolerance=1e-10, mobility=M, z=z)
    ref, it_ref = O.lanczos_sqrt(lambda v: M @ v, z, tolerance=1e-10)
    assert rel(got, exact) < 1e-8
    assert rel(got, ref) < 1e-8
    assert abs(it - it_ref) <= 1
    g3, it3 = lanczos.stochastic_forcing_lanczos(factor=2.0, tolerance=1e-10, mobility_mult=lambda v: 3.0 * v, z=z)
    assert np.allclose(g3, 2.0 * np.sqrt(3.0) * z, rtol=1e-12) and it3 <= 1


@pytest.mark.parametrize("prec,tol", [("fp64", 1e-7), ("fp32_blocked", 1e-5)])
def test_grand_block_lanczos_vs_exact(gpu, prec, tol):
    rng = np.random.default_rng(1)
    n, a, eta = 120, 0.656, 1e-3
    r = np.column_stack([rng.uniform(0, 25, n), rng.uniform(0, 25, n), rng.uniform(0.8, 2.0, n)])
    G = np.block([[O.dense_block_matrix(O.UF, r, eta, a), O.dense_block_matrix(O.UT, r, eta, a)],
                  [O.dense_block_matrix(O.WF, r, eta, a), O.dense_block_matrix(O.WT, r, eta, a)]])
    lam, V = np.linalg.eigh(0.5 * (G + G.T))
    Z = rng.standard_normal((4, 6 * n))
    print(f"case grand N={n} k=4 {prec}", flush=True)
    R = Rollers(r, eta, a, precision=prec)
    noise, it = R.brownian_noise(torch.from_numpy(Z).to(gpu), factor=1.0, tolerance=1e-9)
    noise = noise.cpu().numpy()
    for j in range(4):
        exact = V @ (np.sqrt(np.maximum(lam, 0)) * (V.T @ Z[j]))
        e = rel(noise[j], exact)
        print(f"  rhs {j}: rel err vs eigh {e:.2e}")
        assert e < tol


def test_c3_rollers_block_lanczos_invariants(gpu):
    w = workloads.c3_rollers()
    n = w["r"].shape[0]
    Z = torch.from_numpy(w["z"]).to(gpu)          # (4, 6N)
    R = Rollers(w["r"], w["eta"], w["a"], precision="fp32_blocked")
    print("case C3 N=32768 grand, k=4, tol 1e-3, fp32_blocked", flush=True)
    noise, it = R.brownian_noise(Z, factor=1.0, tolerance=1e-3)
    print(f"  iterations {it}")
    assert 1 <= it < 200
    # ||M^1/2 z||^2 = z^T M z exactly for the Lanczos approximation (alpha_0)
    MZ = torch.empty_like(Z)
    R.grand_matvec(Z.contiguous(), MZ, precision="fp64")
    for j in range(4):
        zMz = float(Z[j] @ MZ[j])
        nn = float(noise[j] @ noise[j])
        assert abs(nn - zMz) <= 1e-6 * abs(zMz), (j, nn, zMz)
    # the same noise with FP64 pairs agrees to the Lanczos tolerance
    noise64, it64 = R.brownian_noise(Z, factor=1.0, tolerance=1e-3, precision="fp64")
    for j in range(4):
        assert rel(noise[j].cpu().numpy(), noise64[j].cpu().numpy()) < 1e-3
    # a single right-hand side reproduces its row of the block
    one, _ = R.brownian_noise(Z[:1].contiguous(), factor=1.0, tolerance=1e-3)
    assert rel(one[0].cpu().numpy(), noise[0].cpu().numpy()) < 2e-3


def test_unconverged_lanczos_is_reported(gpu):
    """max_iter steps that do not reach the tolerance: the last noise is
    returned with iterations = max_iter, as the reference does, plus a
    LanczosNotConverged warning; strict=True raises instead."""
    rng = np.random.default_rng(12)
    n, a = 20, 0.3
    r = np.column_stack([rng.uniform(0, 2, n), rng.uniform(0, 2, n), rng.uniform(0.4, 1.5, n)])
    M = torch.from_numpy(O.single_wall_fluid_mobility(r, 1.0, a)).to(gpu)
    Z = torch.from_numpy(rng.standard_normal((2, 3 * n))).to(gpu)
    mv = lambda X, Y: Y.copy_(X @ M)   # noqa: E731  (M symmetric)
    print("case unconverged Lanczos, max_iter 2, tol 1e-14", flush=True)
    with pytest.warns(lanczos.LanczosNotConverged):
        noise, it = lanczos.lanczos_block(mv, Z, tolerance=1e-14, max_iter=2)
    assert it == 2 and bool(torch.isfinite(noise).all())
    with pytest.raises(RuntimeError):
        lanczos.lanczos_block(mv, Z, tolerance=1e-14, max_iter=2, strict=True)
    import warnings
    with warnings.catch_warnings():
        warnings.simplefilter("error", lanczos.LanczosNotConverged)
        lanczos.lanczos_block(mv, Z, tolerance=1e-6, max_iter=200)     # converges: no warning
