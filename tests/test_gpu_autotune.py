"""Autotuner on the GPU: every timed launch shape gives the heuristic's
result, the chosen shape is used by wall_rpy when the caller passes no
tuning, and the RHS group size does not change results."""
import numpy as np
import pytest
import torch

from rigidmultiblobswall_amd import _lib, autotune, wall_rpy
from rigidmultiblobswall_amd._lib import RMW_UF

pytestmark = pytest.mark.gpu


def _geom(n, seed=0):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(0.2, 3.0, n)])


def test_tune_small_uf(gpu):
    r = _geom(3000)
    print("case autotune UF fp32_blocked N=3000", flush=True)
    res = autotune.tune(RMW_UF, "fp32_blocked", r, eta=1.0, a=0.13, reps=2, max_candidates=8)
    # 8 chunk lengths x symmetric on/off, then 3 grid caps (4, 8, 10 per CU) for the best 3
    assert len(res["table"]) == 2 * 8 + 3 * 3
    assert all(row["rel_diff"] <= 2e-6 for row in res["table"])
    best = res["best"]
    shapes = {(r["chunk_len"], r["max_wgs"], r["rhs_group"], r["symmetric"]) for r in res["table"] if r["ok"]}
    assert (best["chunk_len"], best["max_wgs"], best["rhs_group"], best["symmetric"]) in shapes
    confirmed = [r for r in res["table"] if "ms_confirmed" in r]
    assert len(confirmed) == 3 and best["ms"] == min(r["ms_confirmed"] for r in confirmed)
    assert best["chunk_len"] % 64 == 0


def test_rhs_group_and_chunk_do_not_change_results(gpu):
    r = _geom(1500, seed=1)
    rng = np.random.default_rng(2)
    f = rng.standard_normal((4, 1500, 3))
    t = rng.standard_normal((4, 1500, 3))
    print("case grand k=4 N=1500 rhs_group 1/2/4, chunks", flush=True)
    for prec, tol in (("fp64", 1e-13), ("fp32_blocked", 1e-6)):
        ref = wall_rpy.apply(15, r, force=f, torque=t, eta=1.0, a=0.13, precision=prec, k=4, tuning="heuristic")
        refv = torch.cat([ref[0].reshape(-1), ref[1].reshape(-1)])
        for g in (1, 2, 4):
            for ch in (256, 768, 1536):
                out = wall_rpy.apply(15, r, force=f, torque=t, eta=1.0, a=0.13, precision=prec, k=4,
                                     tuning=_lib.Tuning(ch, 0, g, 0, 0))
                v = torch.cat([out[0].reshape(-1), out[1].reshape(-1)])
                err = float(torch.linalg.vector_norm(v - refv) / torch.linalg.vector_norm(refv))
                assert err <= tol, (prec, g, ch, err)


def test_wall_rpy_uses_cached_shape(gpu, monkeypatch):
    r = _geom(2000, seed=3)
    f = np.random.default_rng(4).standard_normal((2000, 3))
    calls = []

    def fake_lookup(mask, precision, k, pairs, ragged):
        calls.append((mask, precision, k, pairs, ragged))
        return _lib.Tuning(512, 300, 0, 0, 0)
    monkeypatch.setattr(autotune, "lookup", fake_lookup)
    print("case cached shape is used", flush=True)
    u_auto, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.13, precision="fp64")
    assert calls == [(RMW_UF, "fp64", 1, 2000.0 ** 2, False)]
    u_expl, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.13, precision="fp64", tuning=_lib.Tuning(512, 300, 0, 0, 0))
    assert torch.equal(u_auto, u_expl)
    # "heuristic" bypasses the cache
    wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.13, precision="fp64", tuning="heuristic")
    assert len(calls) == 1
