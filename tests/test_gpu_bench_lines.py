"""The secondary bench lines' products at full size (N = 32768, the C3
geometry) against the oracle on sampled target rows: the UF product in
the unbounded, in-plane and free-surface domains (the symmetric kernel is
chosen automatically at this size) and the different-radii product above
the wall, in the free-surface form and unbounded (symmetric radii kernel,
sources = targets), with the inputs bench.py uses."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import wall_rpy, workloads
from rigidmultiblobswall_amd._lib import RMW_UF

pytestmark = pytest.mark.gpu

DOMAIN_BLOCK = {"no_wall": O.UF_NOWALL, "in_plane": O.UF_INPLANE, "free_surface": O.UF_FREE_SURFACE}
ROWS = [(0, 48), (16000, 16048), (32768 - 32, 32768)]


def sampled(got, ref_rows, n):
    num = den = 0.0
    g = np.asarray(got).reshape(n, 3)
    for (t0, t1), ref in zip(ROWS, ref_rows):
        r = ref.reshape(n, 3)[t0:t1]
        num += float(np.sum((g[t0:t1] - r) ** 2))
        den += float(np.sum(r ** 2))
    return float(np.sqrt(num / den))


@pytest.mark.parametrize("domain", ["no_wall", "in_plane", "free_surface"])
def test_domain_bench_product_vs_oracle(gpu, domain):
    w = workloads.c3_rollers()
    r, eta, a = w["r"], w["eta"], w["a"]
    n = r.shape[0]
    f = np.random.default_rng(7).standard_normal(3 * n)
    print(f"case C3 UF {domain} N={n} fp32_blocked", flush=True)
    u, _ = wall_rpy.apply(RMW_UF, torch.from_numpy(r).to(gpu), force=torch.from_numpy(f).to(gpu), eta=eta, a=a,
                          precision="fp32_blocked", domain=domain)
    refs = [O.block_apply(DOMAIN_BLOCK[domain], r, f, eta, a, rows=rw) for rw in ROWS]
    e = sampled(u.cpu().numpy(), refs, n)
    print(f"  rel L2 vs oracle (128 rows) = {e:.3e}")
    assert e <= 1e-6


@pytest.mark.parametrize("boundary", ["wall", "free_surface", "unbounded"])
def test_radii_bench_product_vs_oracle(gpu, boundary):
    w = workloads.c3_rollers()
    r, eta = w["r"], w["eta"]
    n = r.shape[0]
    rad = np.random.default_rng(33).uniform(0.5, 0.8, n)
    f = np.random.default_rng(7).standard_normal(3 * n)
    assert wall_rpy.describe_radii_self(n)["kernel"] == "symmetric"
    print(f"case C3 radii {boundary} N={n} fp32_blocked", flush=True)
    u = wall_rpy.apply_radii_self(torch.from_numpy(r).to(gpu), torch.from_numpy(f).to(gpu),
                                  torch.from_numpy(rad).to(gpu), eta, precision="fp32_blocked", boundary=boundary)
    refs = [O.radii_apply(r, r, f, rad, rad, eta, rows=rw, boundary=boundary) for rw in ROWS]
    e = sampled(u.cpu().numpy(), refs, n)
    print(f"  rel L2 vs oracle (128 rows) = {e:.3e}")
    assert e <= 1e-6
