"""Symmetric evaluation (each unordered pair once: UF_ji = UF_ij^T, WT_ji =
WT_ij^T, UT_ji = WF_ij^T) of the UF block and of the grand mobility, on one
suspension and on ragged batches, against the ordered kernels and the FP64
oracle: tile edges, every precision mode, k right-hand sides, multi-tile
work items, persistent grids."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, wall_rpy, workloads
from rigidmultiblobswall_amd._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

pytestmark = pytest.mark.gpu

TOL = {"fp32_blocked": 2e-6, "fp32": 2e-6, "fp32_kahan": 2e-6, "fp64": 1e-13}


def _geom(n, seed):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(-8, 8, n), rng.uniform(-8, 8, n), rng.uniform(0.05, 3.0, n)])


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _uf(r, f, prec, k, sym, chunk=0, wgs=0):
    u, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.3, a=0.2, precision=prec, k=k,
                          tuning=_lib.Tuning(chunk, wgs, 0, sym, 0))
    return u.cpu().numpy()


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("n", [1, 2, 100, 255, 256, 257, 1000])
def test_symmetric_matches_ordered(gpu, prec, n):
    r = _geom(n, n)
    f = np.random.default_rng(n + 1).standard_normal((n, 3))
    print(f"case symmetric UF n={n} {prec}", flush=True)
    sym = _uf(r, f, prec, 1, 1)
    ordered = _uf(r, f, prec, 1, -1)
    assert rel(sym, ordered) <= TOL[prec]
    ref = O.block_apply(O.UF, r, f, 1.3, 0.2).reshape(n, 3)
    assert rel(sym, ref) <= (1e-12 if prec == "fp64" else 1e-6)


@pytest.mark.parametrize("prec,k", [("fp32_blocked", 4), ("fp32_blocked", 3), ("fp32", 2), ("fp64", 2),
                                    ("fp64", 3), ("fp32_kahan", 4)])
def test_symmetric_k_rhs(gpu, prec, k):
    n = 700
    r = _geom(n, 5)
    f = np.random.default_rng(6).standard_normal((k, n, 3))
    print(f"case symmetric UF k={k} {prec}", flush=True)
    sym = _uf(r, f, prec, k, 1)
    ordered = _uf(r, f, prec, k, -1)
    for j in range(k):
        assert rel(sym[j], ordered[j]) <= TOL[prec], j
        one = _uf(r, f[j], prec, 1, 1)
        assert rel(sym[j], one) <= TOL[prec], j


@pytest.mark.parametrize("chunk,wgs", [(64, 0), (128, 0), (192, 0), (256, 0), (512, 0), (768, 0), (1024, 0),
                                       (64, 5), (256, 7), (512, 3)])
def test_symmetric_work_split(gpu, chunk, wgs):
    n = 1900                                       # 8 tiles, last one ragged
    r = _geom(n, 9)
    f = np.random.default_rng(10).standard_normal((n, 3))
    print(f"case symmetric UF chunk={chunk} max_wgs={wgs}", flush=True)
    base = _uf(r, f, "fp64", 1, -1)
    got = _uf(r, f, "fp64", 1, 1, chunk, wgs)
    assert rel(got, base) <= 1e-13
    # deterministic: the same call gives the same bits
    assert np.array_equal(got, _uf(r, f, "fp64", 1, 1, chunk, wgs))


def test_symmetric_c1_and_overlaps(gpu):
    w = workloads.c1_uf_matvec()
    r, f = w["r"], w["f"].reshape(-1, 3)
    print("case symmetric UF C1 fp32_blocked / fp64", flush=True)
    ref = O.block_apply(O.UF, r, f, w["eta"], w["a"]).reshape(-1, 3)
    u32, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=w["eta"], a=w["a"], precision="fp32_blocked",
                            tuning=_lib.Tuning(0, 0, 0, 1, 0))
    u64, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=w["eta"], a=w["a"], precision="fp64",
                            tuning=_lib.Tuning(0, 0, 0, 1, 0))
    assert rel(u32.cpu().numpy(), ref) <= 1e-6
    assert rel(u64.cpu().numpy(), ref) <= 1e-12


def test_symmetric_c3_size_matches_ordered(gpu):
    w = workloads.c3_rollers()
    r = w["r"]
    f = np.random.default_rng(3).standard_normal((r.shape[0], 3))
    print("case symmetric UF N=32768 fp32_blocked", flush=True)
    sym = _uf(r, f, "fp32_blocked", 1, 1)
    ordered = _uf(r, f, "fp32_blocked", 1, -1)
    assert rel(sym, ordered) <= 1e-6


def test_symmetric_rejected_where_it_does_not_apply(gpu):
    r = _geom(50, 1)
    f = np.ones((50, 3))
    with pytest.raises(_lib.RmwError):      # not UT + WT (two outputs from one source vector)
        wall_rpy.apply(RMW_UT | RMW_WT, r, torque=f, eta=1.0, a=0.2, tuning=_lib.Tuning(0, 0, 0, 1, 0))
    # periodic images: the ordered kernel is used (auto), and agrees with the oracle
    L = np.array([6.0, 5.0, 0.0])
    u, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.2, periodic_length=L, precision="fp64")
    ref = O.block_apply(O.UF, r, f, 1.0, 0.2, periodic_length=L).reshape(-1, 3)
    assert rel(u.cpu().numpy(), ref) <= 1e-12


GRAND = 15


def _grand(r, f, t, prec, k, sym, offsets=None, chunk=0, wgs=0):
    tun = _lib.Tuning(chunk, wgs, 0, sym, 0)
    if offsets is None:
        u, w = wall_rpy.apply(GRAND, r, force=f, torque=t, eta=0.9, a=0.3, precision=prec, k=k, tuning=tun)
    else:
        u, w = wall_rpy.apply_ragged(GRAND, r, offsets, force=f, torque=t, eta=0.9, a=0.3, precision=prec, k=k,
                                     tuning=tun)
    return np.concatenate([u.cpu().numpy().reshape(k, -1), w.cpu().numpy().reshape(k, -1)], axis=1)


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("n,k", [(300, 1), (1000, 2), (513, 3), (700, 4)])
def test_symmetric_grand_matches_ordered(gpu, prec, n, k):
    r = _geom(n, 20 + n)
    rng = np.random.default_rng(n)
    f = rng.standard_normal((k, n, 3))
    t = rng.standard_normal((k, n, 3))
    print(f"case symmetric grand n={n} k={k} {prec}", flush=True)
    sym = _grand(r, f, t, prec, k, 1)
    ordered = _grand(r, f, t, prec, k, -1)
    for j in range(k):
        assert rel(sym[j], ordered[j]) <= TOL[prec], j
    if k == 4:
        # four right-hand sides in one symmetric pass (rhs_group 4)
        tun = _lib.Tuning(0, 0, 4, 1, 0)
        assert wall_rpy.describe(GRAND, n, 4, prec, tuning=tun)["rhs_group"] == 4
        u, w = wall_rpy.apply(GRAND, r, force=f, torque=t, eta=0.9, a=0.3, precision=prec, k=4, tuning=tun)
        one = np.concatenate([u.cpu().numpy().reshape(4, -1), w.cpu().numpy().reshape(4, -1)], axis=1)
        for j in range(4):
            assert rel(one[j], ordered[j]) <= TOL[prec], j
        uf4, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=0.9, a=0.3, precision=prec, k=4, tuning=tun)
        uf1, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=0.9, a=0.3, precision=prec, k=4,
                                tuning=_lib.Tuning(0, 0, 0, -1, 0))
        assert rel(uf4.cpu().numpy(), uf1.cpu().numpy()) <= TOL[prec]


def test_symmetric_grand_vs_oracle(gpu):
    n = 400
    r = _geom(n, 31)
    rng = np.random.default_rng(32)
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    print("case symmetric grand n=400 vs oracle", flush=True)
    ou = O.block_apply(O.UF, r, f, 0.9, 0.3) + O.block_apply(O.UT, r, t, 0.9, 0.3)
    ow = O.block_apply(O.WF, r, f, 0.9, 0.3) + O.block_apply(O.WT, r, t, 0.9, 0.3)
    ref = np.concatenate([ou, ow])
    for prec, tol in (("fp64", 1e-12), ("fp32_blocked", 1e-6), ("fp32", 1e-6)):
        got = _grand(r, f[None], t[None], prec, 1, 1)[0]
        assert rel(got, ref) <= tol, prec


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
def test_symmetric_uf_ut_vs_oracle_and_ordered(gpu, prec):
    """u = UF f + UT t (the fused product of mobility.py:384-410) through the
    symmetric kernel: target side UF and UT, source side UF^T f_i + WF^T t_i
    (UT_ji = WF_ij^T); 700 blobs (a ragged last tile), two right-hand sides."""
    n, k = 700, 2
    r = _geom(n, 33)
    rng = np.random.default_rng(34)
    f, t = rng.standard_normal((k, n, 3)), rng.standard_normal((k, n, 3))
    mask = RMW_UF | RMW_UT
    print(f"case symmetric UF+UT n={n} k={k} {prec}", flush=True)
    tol = 1e-12 if prec == "fp64" else 1e-6
    sym = _lib.Tuning(0, 0, 0, 1, 0)
    assert wall_rpy.describe(mask, n, k, prec, tuning=sym)["kernel"] == "symmetric"
    us, ws = wall_rpy.apply(mask, r, force=f, torque=t, eta=0.9, a=0.3, precision=prec, k=k, tuning=sym)
    assert ws is None or ws.numel() == 0 or not torch.any(ws)
    uo, _ = wall_rpy.apply(mask, r, force=f, torque=t, eta=0.9, a=0.3, precision=prec, k=k,
                           tuning=_lib.Tuning(0, 0, 0, -1, 0))
    us, uo = us.cpu().numpy().reshape(k, -1), uo.cpu().numpy().reshape(k, -1)
    for j in range(k):
        ref = O.block_apply(O.UF, r, f[j], 0.9, 0.3) + O.block_apply(O.UT, r, t[j], 0.9, 0.3)
        assert rel(us[j], ref) <= tol, j
        assert rel(uo[j], ref) <= tol, j


SIZES_R = [0, 300, 7, 0, 600, 1, 256, 257]


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
@pytest.mark.parametrize("mask", [RMW_UF, RMW_UF | RMW_UT, GRAND])
def test_symmetric_ragged_matches_ordered_and_oracle(gpu, prec, mask):
    off = np.concatenate([[0], np.cumsum(SIZES_R)]).astype(np.int64)
    n = int(off[-1])
    r = np.vstack([_geom(m, 40 + m) for m in SIZES_R if m > 0])
    rng = np.random.default_rng(41)
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    print(f"case symmetric ragged mask={mask} {prec}", flush=True)

    def run(sym, chunk=0, wgs=0):
        u, w = wall_rpy.apply_ragged(mask, r, off, force=f, torque=t if mask != RMW_UF else None, eta=0.9, a=0.3,
                                     precision=prec, tuning=_lib.Tuning(chunk, wgs, 0, sym, 0))
        parts = [u.cpu().numpy()] + ([w.cpu().numpy()] if w is not None else [])
        return parts
    sym = run(1)
    ordered = run(-1)
    for a_, b_ in zip(sym, ordered):
        assert rel(a_, b_) <= TOL[prec]
    split = run(1, 512, 5)                      # multi-tile items, small persistent grid
    for a_, b_ in zip(split, ordered):
        assert rel(a_, b_) <= TOL[prec]
    for chunk in (64, 128):                     # quarter- and half-tile items
        fine = run(1, chunk, 0)
        for a_, b_ in zip(fine, ordered):
            assert rel(a_, b_) <= TOL[prec], chunk
    # per system against the oracle
    for s in range(len(SIZES_R)):
        b0, b1 = int(off[s]), int(off[s + 1])
        if b1 == b0:
            continue
        rr = r[b0:b1]
        ou = O.block_apply(O.UF, rr, f[b0:b1], 0.9, 0.3)
        if mask & RMW_UT:
            ou = ou + O.block_apply(O.UT, rr, t[b0:b1], 0.9, 0.3)
        if mask == GRAND:
            ow = O.block_apply(O.WF, rr, f[b0:b1], 0.9, 0.3) + O.block_apply(O.WT, rr, t[b0:b1], 0.9, 0.3)
            assert rel(sym[1][b0:b1].ravel(), ow) <= (1e-12 if prec == "fp64" else 2e-6), s
        assert rel(sym[0][b0:b1].ravel(), ou) <= (1e-12 if prec == "fp64" else 2e-6), s


@pytest.mark.parametrize("chunk", [64, 128, 256])
def test_symmetric_sub_tile_items_c2(gpu, chunk):
    """The rigid-solve geometry (42 full tiles) with work items of a quarter,
    a half and a whole tile pair: the same product as the ordered kernel."""
    w = workloads.c2_rigid_spheres()
    r = w["r"]
    f = np.random.default_rng(61).standard_normal(r.shape)
    print(f"case symmetric C2 chunk={chunk}", flush=True)
    base = _uf(r, f, "fp64", 1, -1)
    got = _uf(r, f, "fp64", 1, 1, chunk, 0)
    assert rel(got, base) <= 1e-13
    g32 = _uf(r, f, "fp32_blocked", 1, 1, chunk, 0)
    assert rel(g32, base) <= 1e-6
    d = wall_rpy.describe(RMW_UF, r.shape[0], 1, "fp32_blocked", tuning=_lib.Tuning(chunk, 0, 0, 1, 0))
    assert d["kernel"] == "symmetric" and d["chunk_len"] == chunk
    assert d["work_items"] == 903 * 256 // chunk


def test_symmetric_c5_matches_ordered(gpu):
    w = workloads.c5_ragged_rollers()
    off, r = w["offsets"], w["r"]
    n = int(off[-1])
    rng = np.random.default_rng(51)
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    print("case symmetric C5 grand ragged fp32_blocked", flush=True)
    sym = _grand(r, f[None], t[None], "fp32_blocked", 1, 1, offsets=off)
    ordered = _grand(r, f[None], t[None], "fp32_blocked", 1, -1, offsets=off)
    assert rel(sym, ordered) <= 1e-6


def test_describe_reports_the_plan(gpu):
    d = wall_rpy.describe(RMW_UF, 32768, 1, "fp32_blocked", tuning=_lib.Tuning(0, 0, 0, 0, 0))
    assert d["kernel"] == "symmetric" and d["work_items"] > 0 and d["workspace_bytes"] > 0
    d = wall_rpy.describe(RMW_UF, 32768, 4, "fp32_blocked", periodic_length=[10.0, 0, 0],
                          tuning=_lib.Tuning(0, 0, 0, 0, 0))
    assert d["kernel"] == "symmetric" and d["rhs_group"] == 1      # images: one RHS per pass
    d = wall_rpy.describe(RMW_UF, 32768, 1, "fp32_blocked", tuning=_lib.Tuning(1024, 0, 0, -1, 0))
    assert d["kernel"] == "ordered" and d["chunk_len"] == 1024 and d["work_items"] == 128 * 32
    d = wall_rpy.describe(RMW_UF, 2048, 1, "fp64", tuning="heuristic")      # too few tile pairs
    assert d["kernel"] == "ordered"


PERIODIC_L = [(6.0, 5.0, 0.0), (8.0, 0.0, 0.0), (6.0, 5.0, 7.0)]


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("mask", [RMW_UF, RMW_UF | RMW_UT, GRAND])
def test_symmetric_periodic_vs_oracle_and_ordered(gpu, prec, mask):
    """Pseudo-periodic images through the symmetric kernel: each unordered
    pair's 3, 9 or 27 images on both sides, the self pair's outer images in
    the reduce; against the oracle and the ordered kernel, 700 blobs (two
    full tiles and a ragged one), boxes periodic in x and y, x only, and
    x, y and z, two right-hand sides (one per pass)."""
    n, k = 700, 2
    rng = np.random.default_rng(71)
    r = np.column_stack([rng.uniform(-4, 9, n), rng.uniform(0, 5, n), rng.uniform(0.25, 2.0, n)])
    f, t = rng.standard_normal((k, n, 3)), rng.standard_normal((k, n, 3))
    tol = 1e-12 if prec == "fp64" else 2e-6
    for L in PERIODIC_L:
        print(f"case symmetric periodic mask={mask} L={L} {prec}", flush=True)
        sym = _lib.Tuning(0, 0, 0, 1, 0)
        assert wall_rpy.describe(mask, n, k, prec, periodic_length=L, tuning=sym)["kernel"] == "symmetric"
        kw = dict(force=f, torque=t if mask != RMW_UF else None, eta=0.9, a=0.3, precision=prec, k=k,
                  periodic_length=np.array(L))
        us, ws = wall_rpy.apply(mask, r, tuning=sym, **kw)
        uo, wo = wall_rpy.apply(mask, r, tuning=_lib.Tuning(0, 0, 0, -1, 0), **kw)
        us, uo = us.cpu().numpy().reshape(k, -1), uo.cpu().numpy().reshape(k, -1)
        for j in range(k):
            ou = O.block_apply(O.UF, r, f[j], 0.9, 0.3, periodic_length=L)
            if mask & RMW_UT:
                ou = ou + O.block_apply(O.UT, r, t[j], 0.9, 0.3, periodic_length=L)
            assert rel(us[j], ou) <= tol, (L, j)
            assert rel(us[j], uo[j]) <= tol, (L, j)
            if mask == GRAND:
                ow = (O.block_apply(O.WF, r, f[j], 0.9, 0.3, periodic_length=L)
                      + O.block_apply(O.WT, r, t[j], 0.9, 0.3, periodic_length=L))
                wsj = ws.cpu().numpy().reshape(k, -1)[j]
                assert rel(wsj, ow) <= tol, (L, j)
                assert rel(wsj, wo.cpu().numpy().reshape(k, -1)[j]) <= tol, (L, j)


def test_symmetric_periodic_ragged_and_items(gpu):
    """Ragged systems and finer or coarser work items with periodic images."""
    off = np.concatenate([[0], np.cumsum(SIZES_R)]).astype(np.int64)
    n = int(off[-1])
    r = np.vstack([_geom(m, 40 + m) for m in SIZES_R if m > 0])
    rng = np.random.default_rng(72)
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    L = np.array([7.0, 6.0, 0.0])
    print("case symmetric periodic ragged", flush=True)

    def run(sym, chunk=0, wgs=0):
        u, w = wall_rpy.apply_ragged(GRAND, r, off, force=f, torque=t, eta=0.9, a=0.3, precision="fp32_blocked",
                                     periodic_length=L, tuning=_lib.Tuning(chunk, wgs, 0, sym, 0))
        return np.concatenate([u.cpu().numpy().ravel(), w.cpu().numpy().ravel()])
    ordered = run(-1)
    for chunk, wgs in ((0, 0), (64, 0), (512, 5)):
        assert rel(run(1, chunk, wgs), ordered) <= 2e-6, (chunk, wgs)


SINGLE_ROT = {RMW_UT: ("UT", "torque"), RMW_WF: ("WF", "force"), RMW_WT: ("WT", "torque")}


def _single(mask, r, v, prec, k, sym, L=None, offsets=None, chunk=0):
    kw = dict(eta=0.9, a=0.3, precision=prec, tuning=_lib.Tuning(chunk, 0, 0, sym, 0))
    kw[SINGLE_ROT[mask][1]] = v
    if L is not None:
        kw["periodic_length"] = np.array(L)
    if offsets is None:
        u, w = wall_rpy.apply(mask, r, k=k, **kw)
    else:
        u, w = wall_rpy.apply_ragged(mask, r, offsets, **kw)
    out = u if mask == RMW_UT else w
    return out.cpu().numpy().reshape(k, -1)


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("mask", [RMW_UT, RMW_WF, RMW_WT])
def test_symmetric_single_rotational_block(gpu, prec, mask):
    """One rotational block alone through the symmetric kernel: u = UT t
    (source side WF_ij^T t_i = UT_ji t_i), w = WF f (source side UT_ij^T f_i)
    or w = WT t (source side WT_ij^T t_i); against the oracle's block and
    the ordered kernel, 700 blobs (a ragged last tile), two right-hand sides,
    without and with pseudo-periodic images."""
    n, k = 700, 2
    rng = np.random.default_rng(90 + mask)
    r = np.column_stack([rng.uniform(-4, 9, n), rng.uniform(0, 5, n), rng.uniform(0.25, 2.0, n)])
    v = rng.standard_normal((k, n, 3))
    tol = 1e-12 if prec == "fp64" else 2e-6
    block = getattr(O, SINGLE_ROT[mask][0])
    for L in [None] + PERIODIC_L:
        print(f"case symmetric single block mask={mask} L={L} {prec}", flush=True)
        sym = _lib.Tuning(0, 0, 0, 1, 0)
        assert wall_rpy.describe(mask, n, k, prec, periodic_length=L, tuning=sym)["kernel"] == "symmetric"
        s_ = _single(mask, r, v, prec, k, 1, L)
        o_ = _single(mask, r, v, prec, k, -1, L)
        for j in range(k):
            ref = O.block_apply(block, r, v[j], 0.9, 0.3, periodic_length=L if L is not None else [0, 0, 0])
            assert rel(s_[j], ref) <= tol, (L, j)
            assert rel(o_[j], ref) <= tol, (L, j)


@pytest.mark.parametrize("mask", [RMW_UT, RMW_WF, RMW_WT])
def test_symmetric_single_rotational_ragged_and_items(gpu, mask):
    """Ragged batches (empty and one-blob systems included) and work items
    of one, two and eight 64-source blocks, against the ordered kernel."""
    off = np.concatenate([[0], np.cumsum(SIZES_R)]).astype(np.int64)
    n = int(off[-1])
    r = np.vstack([_geom(m, 40 + m) for m in SIZES_R if m > 0])
    v = np.random.default_rng(91).standard_normal((1, n, 3))
    print(f"case symmetric single block ragged mask={mask}", flush=True)
    ordered = _single(mask, r, v[0], "fp32_blocked", 1, -1, offsets=off)
    for chunk in (0, 64, 128, 512):
        assert rel(_single(mask, r, v[0], "fp32_blocked", 1, 1, offsets=off, chunk=chunk), ordered) <= 2e-6, chunk
