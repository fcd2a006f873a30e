"""rmw_wall_rpy_apply_noalloc (no workspace argument, partial sums in the
library's fixed device array) against rmw_wall_rpy_apply, through the C ABI.

Calls come in pairs, the second on a smaller suspension, so the second reads
an array that still holds the first call's partial sums.  Where both entry
points run the same launch shape the results are bit-identical; where the
default shape needs more than the array holds (N = 32768) the no-allocation
call runs the ordered kernel with longer chunks and is compared to the
FP32 tolerance and against the oracle.
"""
import ctypes

import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, wall_rpy, workloads

pytestmark = pytest.mark.gpu

UF, UT, WF, WT = 1, 2, 4, 8
PREC = {"fp32_blocked": 3, "fp64": 2}


def _stream():
    return torch.cuda.current_stream().cuda_stream


def _L():
    return (ctypes.c_double * 3)(0.0, 0.0, 0.0)


def _rollers(n, seed):
    rng = np.random.default_rng(seed)
    side = 300.0 * np.sqrt(n / 32768.0)
    r = np.column_stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(0.5, 2.0, n)])
    return r


def _both(mask, prec, r, k, seed, dev):
    """(noalloc velocity, omega), (workspace velocity, omega) for k RHS."""
    lib = _lib.load()
    n = r.shape[0]
    rng = np.random.default_rng(seed)
    rd = torch.from_numpy(r).to(dev)
    f = torch.from_numpy(rng.standard_normal((k, n, 3))).to(dev)
    t = torch.from_numpy(rng.standard_normal((k, n, 3))).to(dev)
    outs = []
    for which in ("noalloc", "workspace"):
        u = torch.full((k, n, 3), np.nan, dtype=torch.float64, device=dev)
        w = torch.full((k, n, 3), np.nan, dtype=torch.float64, device=dev)
        fp = f.data_ptr() if mask & (UF | WF) else None
        tp = t.data_ptr() if mask & (UT | WT) else None
        up = u.data_ptr() if mask & (UF | UT) else None
        wp = w.data_ptr() if mask & (WF | WT) else None
        if which == "noalloc":
            _lib.check(lib.rmw_wall_rpy_apply_noalloc(mask, PREC[prec], rd.data_ptr(), fp, tp, up, wp, n, k, 0,
                                                      1e-3, 0.656, _L(), None, _stream()))
        else:
            nb = ctypes.c_size_t()
            _lib.check(lib.rmw_wall_rpy_workspace_size(mask, PREC[prec], n, k, None, ctypes.byref(nb)))
            ws = torch.empty(max(nb.value, 256), dtype=torch.uint8, device=dev)
            _lib.check(lib.rmw_wall_rpy_apply(mask, PREC[prec], rd.data_ptr(), fp, tp, up, wp, n, k, 0, 1e-3,
                                              0.656, _L(), None, ws.data_ptr(), ws.numel(), _stream()))
        torch.cuda.synchronize()
        outs.append((u.cpu().numpy() if up else None, w.cpu().numpy() if wp else None))
    return outs, f.cpu().numpy(), t.cpu().numpy()


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("first,second", [
    ((UF, "fp32_blocked", 8192, 1), (UF, "fp32_blocked", 1000, 1)),
    ((UF | UT | WF | WT, "fp64", 3000, 2), (UF | UT | WF | WT, "fp64", 700, 2)),
    ((UF | UT, "fp32_blocked", 2500, 3), (UF, "fp64", 333, 1)),
])
def test_noalloc_matches_workspace_entry(gpu, first, second):
    for (mask, prec, n, k), seed in ((first, 1), (second, 2)):
        r = _rollers(n, seed)
        print(f"case noalloc mask={mask} {prec} n={n} k={k}", flush=True)
        (na, ws), _, _ = _both(mask, prec, r, k, seed + 10, gpu)
        for a_, b_ in zip(na, ws):
            if a_ is None:
                continue
            assert np.isfinite(a_).all()
            assert np.array_equal(a_, b_)


def test_noalloc_beyond_default_workspace(gpu):
    """N = 32768: the default symmetric shape needs more partial-sum space
    than the array holds, so the ordered kernel runs with longer chunks;
    then a smaller call reuses the array."""
    w = workloads.c3_rollers()
    r = w["r"]
    n = r.shape[0]
    plan = wall_rpy.describe(UF, n, 1, "fp32_blocked", tuning="heuristic")
    print(f"case noalloc C3 n={n}: default plan {plan}", flush=True)
    assert plan["workspace_bytes"] > _lib.RMW_NOALLOC_SCRATCH_BYTES
    (na, ws), f, _ = _both(UF, "fp32_blocked", r, 1, 5, gpu)
    e = _rel(na[0], ws[0])
    print(f"  noalloc vs workspace entry: rel L2 {e:.2e}")
    assert e <= 1e-6
    rows = (20000, 20256)
    ref = O.uf(r, f[0].ravel(), 1e-3, 0.656, rows=rows).reshape(n, 3)[rows[0]:rows[1]]
    assert _rel(na[0][0, rows[0]:rows[1]], ref) <= 1e-6
    r2 = _rollers(2048, 6)
    print("case noalloc n=2048 after C3", flush=True)
    (na2, ws2), _, _ = _both(UF, "fp32_blocked", r2, 1, 7, gpu)
    assert np.array_equal(na2[0], ws2[0])


def test_noalloc_rejects_sizes_beyond_the_array(gpu):
    """A size whose partial sums cannot fit even in one ordered chunk is
    refused before any launch, with RMW_E_LIMIT."""
    lib = _lib.load()
    n = 2_900_000
    print(f"case noalloc limit n={n}", flush=True)
    r = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    r[:, 2] = 1.0
    f = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    u = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    rc = lib.rmw_wall_rpy_apply_noalloc(UF, 3, r.data_ptr(), f.data_ptr(), None, u.data_ptr(), None, n, 1, 0,
                                        1.0, 0.5, _L(), None, _stream())
    msg = lib.rmw_last_error().decode()
    print(f"  rc {rc}: {msg}")
    assert rc == _lib.RMW_E_LIMIT and "scratch" in msg
    torch.cuda.synchronize()
    assert float(u.abs().max().item()) == 0.0
