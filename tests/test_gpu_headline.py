"""The benchmark's own launches against the FP64 oracle at full size.

The timed product of bench.py (UF block, N = 32768 rollers, FP32 pair terms
with blocked FP32 sub-sums, the launch shape the autotuner cache gives) and
the rigid-solve operator's product (256 spheres x 42 blobs, the symmetric
kernel with 64-source work items) are compared with the oracle
(mobility_numba.py:123-287 with the wrapper of mobility.py:1132-1163) on
target-row samples spread over the suspension.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import wall_rpy, workloads
from rigidmultiblobswall_amd._lib import RMW_UF

pytestmark = pytest.mark.gpu

TOL = {"fp32_blocked": 1e-6, "fp32": 1e-6, "fp32_kahan": 1e-6, "fp64": 1e-12}


def row_blocks(n, blocks=4, rows=128):
    starts = np.linspace(0, n - rows, blocks).astype(int)
    return [(int(s), int(s) + rows) for s in starts]


def sampled_error(u, r, f, eta, a, blocks):
    n = r.shape[0]
    got = u.reshape(n, 3).cpu().numpy()
    num = den = 0.0
    for t0, t1 in blocks:
        ref = O.uf(r, f, eta, a, rows=(t0, t1)).reshape(n, 3)[t0:t1]
        num += float(np.sum((got[t0:t1] - ref) ** 2))
        den += float(np.sum(ref ** 2))
    return float(np.sqrt(num / den))


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_c3_bench_product_vs_oracle(gpu, prec):
    """bench.py's step: the same inputs (C3 positions, f from seed 7), the
    same call (tuning from the autotuner cache), 512 sampled rows."""
    w = workloads.c3_rollers()
    r, eta, a = w["r"], w["eta"], w["a"]
    n = r.shape[0]
    f = np.random.default_rng(7).standard_normal(3 * n)
    plan = wall_rpy.describe(RMW_UF, n, 1, prec)
    print(f"case C3 bench product N={n} {prec}: {plan}", flush=True)
    assert plan["kernel"] == "symmetric"
    u = torch.empty((1, n, 3), dtype=torch.float64, device=gpu)
    rd = torch.from_numpy(r).to(gpu)
    fd = torch.from_numpy(f).to(gpu)
    for _ in range(2):      # a repeated launch into the same buffer gives the same product
        wall_rpy.apply(RMW_UF, rd, force=fd, eta=eta, a=a, precision=prec, out_velocity=u)
    torch.cuda.synchronize()
    e = sampled_error(u, r, f, eta, a, row_blocks(n))
    print(f"  rel L2 vs oracle (512 rows) = {e:.3e}")
    assert e <= TOL[prec]


@pytest.mark.parametrize("tuning", [{"symmetric": 1, "chunk_len": 64}, None], ids=["quarter_items", "cache"])
def test_c2_operator_shape_vs_oracle(gpu, tuning):
    """The rigid-solve product at full C2 size with the symmetric kernel on
    64-source work items (3612 items), and with whatever the cache holds."""
    w = workloads.c2_rigid_spheres()
    r, eta, a = w["r"], w["eta"], w["a"]
    n = r.shape[0]
    f = np.random.default_rng(8).standard_normal(3 * n)
    plan = wall_rpy.describe(RMW_UF, n, 1, "fp32_blocked", tuning=tuning)
    print(f"case C2 operator product N={n} fp32_blocked tuning={tuning}: {plan}", flush=True)
    if tuning is not None:
        assert plan["kernel"] == "symmetric" and plan["chunk_len"] == 64 and plan["work_items"] == 3612
    u, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=eta, a=a, precision="fp32_blocked", tuning=tuning)
    blocks = row_blocks(n, blocks=4, rows=96) + [(n - 64, n)]
    e = sampled_error(u, r, f, eta, a, blocks)
    print(f"  rel L2 vs oracle ({sum(t1 - t0 for t0, t1 in blocks)} rows) = {e:.3e}")
    assert e <= 1e-6


def test_c3_lanczos_block_product_vs_oracle(gpu):
    """The C3 Lanczos block's product at full size: the grand mobility on 4
    right-hand sides in one symmetric pass (explicit 3x3 blocks, the cache's
    launch shape), through the prepared strided product the Lanczos loop
    calls, against the oracle's four blocks on sampled rows of every
    right-hand side."""
    from rigidmultiblobswall_amd.rollers import GRAND, Rollers
    w = workloads.c3_rollers()
    r, eta, a = w["r"], w["eta"], w["a"]
    n = r.shape[0]
    Z = np.random.default_rng(9).standard_normal((4, 6 * n))
    plan = wall_rpy.describe(GRAND, n, 4, "fp32_blocked")
    print(f"case C3 grand k=4 N={n} fp32_blocked: {plan}", flush=True)
    assert plan["kernel"] == "symmetric" and plan["rhs_group"] == 4
    R = Rollers(r, eta, a, precision="fp32_blocked")
    X = torch.from_numpy(Z).to(gpu)
    Y = torch.empty_like(X)
    R.grand_matvec(X, Y)
    got = Y.cpu().numpy()
    blocks = [(0, 48), (n // 2, n // 2 + 48), (n - 32, n)]
    for q in range(4):
        f, t = Z[q, : 3 * n], Z[q, 3 * n:]
        num = den = 0.0
        for t0, t1 in blocks:
            ur, wr = O.grand(r, f, t, eta, a, rows=(t0, t1))
            for part, ref in ((got[q, : 3 * n], ur), (got[q, 3 * n:], wr)):
                g = part.reshape(n, 3)[t0:t1]
                rr = ref.reshape(n, 3)[t0:t1]
                num += float(np.sum((g - rr) ** 2))
                den += float(np.sum(rr ** 2))
        e = float(np.sqrt(num / den))
        print(f"  rhs {q}: rel L2 vs oracle (128 rows, u and w) = {e:.3e}")
        assert e <= 1e-6


def test_output_buffers_are_checked(gpu):
    """A caller's output buffer of the wrong size, type, device or layout is
    refused before any launch."""
    rng = np.random.default_rng(2)
    n = 300
    r = np.column_stack([rng.uniform(-5, 5, n), rng.uniform(-5, 5, n), rng.uniform(0.5, 2, n)])
    f = rng.standard_normal(3 * n)
    good = torch.empty((1, n, 3), dtype=torch.float64, device=gpu)
    wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.3, out_velocity=good)
    for bad in (torch.empty((1, n - 1, 3), dtype=torch.float64, device=gpu),
                torch.empty((1, n, 3), dtype=torch.float32, device=gpu),
                torch.empty((1, n, 3), dtype=torch.float64),
                torch.empty((1, 3, n), dtype=torch.float64, device=gpu).transpose(1, 2)):
        with pytest.raises(ValueError):
            wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.3, out_velocity=bad)
    with pytest.raises(ValueError):
        wall_rpy.apply_radii_self(r, f, np.full(n, 0.3), 1.0, out=torch.empty((n + 1, 3), dtype=torch.float64,
                                                                               device=gpu))
    X = torch.zeros((2, 6 * n), dtype=torch.float64, device=gpu)
    with pytest.raises(ValueError):
        wall_rpy.apply_strided(15, torch.from_numpy(r).to(gpu), n, X, X[:, : 6 * n - 3], 1.0, 0.3)
