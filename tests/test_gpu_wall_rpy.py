"""GPU parity of the wall-RPY products (librmw_hip via its C ABI) against
the FP64 C oracle.

Tolerances (relative L2 over the whole output vector):
  fp32 / fp32_kahan  <= 1e-6   (FP32 pair terms)
  fp64               <= 1e-12
"""
import numpy as np
import pytest

from oracle import oracle as O
from rigidmultiblobswall_amd import mobility as mob
from rigidmultiblobswall_amd import wall_rpy, workloads
from rigidmultiblobswall_amd._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-6, "fp32_kahan": 1e-6, "fp32_blocked": 1e-6, "fp64": 1e-12}
BLOCKS = {"UF": (RMW_UF, O.UF), "UT": (RMW_UT, O.UT), "WF": (RMW_WF, O.WF), "WT": (RMW_WT, O.WT)}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _gpu_block(name, r, v, eta, a, L, prec):
    fn = {"UF": mob.single_wall_mobility_trans_times_force_hip,
          "UT": mob.single_wall_mobility_trans_times_torque_hip,
          "WF": mob.single_wall_mobility_rot_times_force_hip,
          "WT": mob.single_wall_mobility_rot_times_torque_hip}[name]
    return fn(r, v, eta, a, periodic_length=L, precision=prec)


@pytest.mark.parametrize("prec", ["fp32", "fp32_kahan", "fp32_blocked", "fp64"])
@pytest.mark.parametrize("block", ["UF", "UT", "WF", "WT"])
def test_block_parity_c1(gpu, block, prec):
    w = workloads.c1_uf_matvec()
    print(f"case C1 N=2048 block={block} prec={prec}", flush=True)
    ref = O.block_apply(BLOCKS[block][1], w["r"], w["f"], w["eta"], w["a"])
    got = _gpu_block(block, w["r"], w["f"], w["eta"], w["a"], w["L"], prec)
    assert got.shape == (3 * 2048,) and got.dtype == np.float64
    err = rel(got, ref)
    print(f"  rel L2 = {err:.3e}")
    assert err <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "fp32_blocked", "fp64"])
@pytest.mark.parametrize("block", ["UF", "UT", "WF", "WT"])
def test_block_parity_wall_overlap(gpu, block, prec):
    # mobility/test_blobs.py:31-34 stress layout: r = 5 a rand, blobs overlap
    # each other and the wall (B < 1, z_eff = a path).
    rng = np.random.default_rng(11)
    a, eta, n = 0.13, 7.0, 1000
    r = 5 * a * rng.random((n, 3))
    v = rng.standard_normal(3 * n)
    print(f"case overlap N={n} block={block} prec={prec}", flush=True)
    ref = O.block_apply(BLOCKS[block][1], r, v, eta, a)
    got = _gpu_block(block, r, v, eta, a, np.zeros(3), prec)
    err = rel(got, ref)
    print(f"  rel L2 = {err:.3e}")
    assert err <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_periodic_parity(gpu, prec):
    rng = np.random.default_rng(12)
    n, a, eta = 300, 0.3, 1.0
    L = np.array([6.0, 5.0, 0.0])
    r = np.column_stack([rng.uniform(-4, 9, n), rng.uniform(0, 5, n), rng.uniform(0.4, 2.0, n)])
    v = rng.standard_normal(3 * n)
    for name in ("UF", "UT", "WF", "WT"):
        print(f"case periodic N={n} L={L.tolist()} block={name} prec={prec}", flush=True)
        ref = O.block_apply(BLOCKS[name][1], r, v, eta, a, periodic_length=L)
        got = _gpu_block(name, r, v, eta, a, L, prec)
        err = rel(got, ref)
        print(f"  rel L2 = {err:.3e}")
        assert err <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32", "fp32_blocked", "fp64"])
@pytest.mark.parametrize("k", [1, 3, 4, 7])
def test_grand_fused_krhs(gpu, prec, k):
    rng = np.random.default_rng(13)
    n, a, eta = 700, 0.656, 1e-3
    r = np.column_stack([rng.uniform(0, 30, n), rng.uniform(0, 30, n), rng.uniform(0.5, 2.0, n)])
    f = rng.standard_normal((k, n, 3))
    t = rng.standard_normal((k, n, 3))
    print(f"case grand N={n} k={k} prec={prec}", flush=True)
    u, w = wall_rpy.apply(RMW_UF | RMW_UT | RMW_WF | RMW_WT, r, force=f, torque=t, eta=eta, a=a,
                          precision=prec, k=k)
    u = u.reshape(k, n, 3).cpu().numpy()
    w = w.reshape(k, n, 3).cpu().numpy()
    for j in range(k):
        ur, wr = O.grand(r, f[j], t[j], eta, a)
        eu, ew = rel(u[j].ravel(), ur), rel(w[j].ravel(), wr)
        print(f"  rhs {j}: rel u {eu:.3e} w {ew:.3e}")
        assert eu <= TOL[prec] and ew <= TOL[prec]


@pytest.mark.parametrize("mask", [RMW_UF | RMW_UT, RMW_WF | RMW_WT, RMW_UF | RMW_WF, RMW_UT | RMW_WT])
def test_two_block_masks(gpu, mask):
    rng = np.random.default_rng(14)
    n, a, eta = 513, 0.4, 2.0
    r = np.column_stack([rng.uniform(0, 20, n), rng.uniform(0, 20, n), rng.uniform(0.2, 2.0, n)])
    f = rng.standard_normal(3 * n)
    t = rng.standard_normal(3 * n)
    print(f"case mask={mask} N={n}", flush=True)
    u, w = wall_rpy.apply(mask, r, force=f, torque=t, eta=eta, a=a, precision="fp64")
    ur = np.zeros(3 * n)
    wr = np.zeros(3 * n)
    if mask & RMW_UF: ur += O.uf(r, f, eta, a)
    if mask & RMW_UT: ur += O.ut(r, t, eta, a)
    if mask & RMW_WF: wr += O.wf(r, f, eta, a)
    if mask & RMW_WT: wr += O.wt(r, t, eta, a)
    if u is not None:
        assert rel(u.cpu().numpy().ravel(), ur) <= 1e-12
    if w is not None:
        assert rel(w.cpu().numpy().ravel(), wr) <= 1e-12


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
@pytest.mark.parametrize("mask", [RMW_UF, RMW_UF | RMW_UT, 15])
def test_prepared_product(gpu, mask, prec):
    """wall_rpy.Product (shape, launch shape and workspace settled once)
    gives apply()'s result bit for bit, follows positions rewritten in place,
    matches the oracle, and refuses wrong buffers before any launch."""
    import torch
    rng = np.random.default_rng(16)
    n, a, eta = 1500, 0.4, 1.5
    r0 = np.column_stack([rng.uniform(0, 30, n), rng.uniform(0, 30, n), rng.uniform(0.2, 2.5, n)])
    r1 = r0 + rng.normal(0, 0.05, r0.shape)
    r1[:, 2] = np.abs(r1[:, 2])
    f, t = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    print(f"case Product mask={mask} prec={prec} N={n}", flush=True)
    dev = torch.device("cuda", 0)
    r = torch.from_numpy(r0).to(dev)
    fd, td = torch.from_numpy(f).to(dev), torch.from_numpy(t).to(dev)
    p = wall_rpy.Product(mask, r, eta=eta, a=a, precision=prec)
    u = torch.empty(3 * n, dtype=torch.float64, device=dev)
    w = torch.empty(3 * n, dtype=torch.float64, device=dev)
    has_w = bool(mask & (RMW_WF | RMW_WT))
    for rr in (r0, r1):
        r.copy_(torch.from_numpy(rr))
        p(force=fd, torque=td if mask & (RMW_UT | RMW_WT) else None, out_velocity=u,
          out_omega=w if has_w else None)
        ue, we = wall_rpy.apply(mask, r, force=fd, torque=td, eta=eta, a=a, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(u, ue.reshape(-1))
        if has_w:
            assert torch.equal(w, we.reshape(-1))
        ur = O.uf(rr, f, eta, a) + (O.ut(rr, t, eta, a) if mask & RMW_UT else 0.0)
        assert rel(u.cpu().numpy(), ur) <= TOL[prec]
        if has_w:
            assert rel(w.cpu().numpy(), O.wf(rr, f, eta, a) + O.wt(rr, t, eta, a)) <= TOL[prec]
    with pytest.raises(ValueError):
        p(force=fd[:-3], torque=td, out_velocity=u, out_omega=w)
    with pytest.raises(ValueError):
        p(force=fd, torque=td, out_velocity=u.float(), out_omega=w)
    with pytest.raises(ValueError):
        p(force=None, torque=td, out_velocity=u, out_omega=w)


def test_trans_times_force_torque_dropin(gpu):
    rng = np.random.default_rng(15)
    n, a, eta = 400, 0.5, 1.0
    r = np.column_stack([rng.uniform(0, 20, n), rng.uniform(0, 20, n), rng.uniform(0.3, 2.0, n)])
    f, t = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    got = mob.single_wall_mobility_trans_times_force_torque_hip(r, f, t, eta, a, precision="fp64")
    assert rel(got, O.uf(r, f, eta, a) + O.ut(r, t, eta, a)) <= 1e-12
    ft = np.concatenate([f, t])
    g = mob.single_wall_grand_mobility_hip(r, ft, eta, a, precision="fp64")
    ur, wr = O.grand(r, f, t, eta, a)
    assert rel(g, np.concatenate([ur, wr])) <= 1e-12


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_ragged_parity(gpu, prec):
    rng = np.random.default_rng(16)
    sizes = [0, 1, 100, 300, 257, 0, 513]
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(offsets[-1])
    a, eta = 0.656, 1e-3
    r = np.column_stack([rng.uniform(0, 40, n), rng.uniform(0, 40, n), rng.uniform(0.5, 2.0, n)])
    f = rng.standard_normal((2, n, 3))
    t = rng.standard_normal((2, n, 3))
    print(f"case ragged sizes={sizes} prec={prec}", flush=True)
    u, w = wall_rpy.apply_ragged(RMW_UF | RMW_UT | RMW_WF | RMW_WT, r, offsets, force=f, torque=t, eta=eta,
                                 a=a, precision=prec, k=2)
    u = u.cpu().numpy()
    w = w.cpu().numpy()
    for s in range(len(sizes)):
        b0, b1 = offsets[s], offsets[s + 1]
        if b1 == b0:
            continue
        for j in range(2):
            ur, wr = O.grand(r[b0:b1], f[j, b0:b1], t[j, b0:b1], eta, a)
            eu = rel(u[j, b0:b1].ravel(), ur)
            ew = rel(w[j, b0:b1].ravel(), wr)
            assert eu <= TOL[prec] and ew <= TOL[prec], (s, j, eu, ew)


def test_single_blob_and_empty(gpu):
    r = np.array([[0.3, -0.2, 0.9]])
    f = np.array([1.0, -2.0, 0.5])
    for name, ob in (("UF", O.UF), ("UT", O.UT), ("WF", O.WF), ("WT", O.WT)):
        got = _gpu_block(name, r, f, 1.0, 0.5, np.zeros(3), "fp64")
        ref = O.block_apply(ob, r, f, 1.0, 0.5)
        assert np.allclose(got, ref, rtol=1e-13, atol=0)
    u, _ = wall_rpy.apply(RMW_UF, np.zeros((0, 3)), force=np.zeros(0), eta=1.0, a=1.0)
    assert u.numel() == 0


def test_c3_size_properties(gpu):
    """N = 32768 rollers: fp32 vs fp64 on the GPU, oracle on a row sample,
    and symmetry of UF (x.My = y.Mx) at full size."""
    import torch
    w = workloads.c3_rollers()
    r, eta, a = w["r"], w["eta"], w["a"]
    n = r.shape[0]
    rng = np.random.default_rng(33)
    x = rng.standard_normal(3 * n)
    y = rng.standard_normal(3 * n)
    print("case C3 N=32768 UF fp32/fp64", flush=True)
    u32 = mob.single_wall_mobility_trans_times_force_hip(r, x, eta, a, precision="fp32")
    u64 = mob.single_wall_mobility_trans_times_force_hip(r, x, eta, a, precision="fp64")
    ub = mob.single_wall_mobility_trans_times_force_hip(r, x, eta, a, precision="fp32_blocked")
    e = rel(u32, u64)
    eb = rel(ub, u64)
    print(f"  fp32 vs fp64 rel L2 = {e:.3e}, fp32_blocked vs fp64 = {eb:.3e}")
    assert e <= 1e-6 and eb <= 1e-6
    rows = (1000, 1256)
    ref = O.uf(r, x, eta, a, rows=rows).reshape(n, 3)[rows[0]:rows[1]]
    assert rel(u64.reshape(n, 3)[rows[0]:rows[1]], ref) <= 1e-12
    assert rel(u32.reshape(n, 3)[rows[0]:rows[1]], ref) <= 1e-6
    My = mob.single_wall_mobility_trans_times_force_hip(r, y, eta, a, precision="fp64")
    s1, s2 = float(x @ My), float(y @ u64)
    assert abs(s1 - s2) <= 1e-11 * (abs(s1) + abs(s2))


@pytest.mark.parametrize("domain", ["no_wall", "in_plane", "free_surface"])
def test_prepared_product_domains(gpu, domain):
    """wall_rpy.Product in the other boundaries gives apply()'s result bit
    for bit, on the current stream and on a side stream."""
    import torch
    rng = np.random.default_rng(17)
    n, a, eta = 1200, 0.5, 1.0
    r = torch.from_numpy(np.column_stack([rng.uniform(0, 30, n), rng.uniform(0, 30, n),
                                          rng.uniform(0.6, 3.0, n)])).to(gpu)
    f = torch.from_numpy(rng.standard_normal(3 * n)).to(gpu)
    print(f"case Product domain={domain} N={n}", flush=True)
    p = wall_rpy.Product(RMW_UF, r, eta=eta, a=a, precision="fp32_blocked", domain=domain)
    u = torch.empty(3 * n, dtype=torch.float64, device=gpu)
    p(force=f, out_velocity=u)
    ue, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=eta, a=a, precision="fp32_blocked", domain=domain)
    torch.cuda.synchronize()
    assert torch.equal(u, ue.reshape(-1))
    side = torch.cuda.Stream()
    u2 = torch.empty_like(u)
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        p(force=f, out_velocity=u2)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    assert torch.equal(u2, u)


def test_pair_kernel_timer_launch_record(gpu):
    """PairKernelTimer reads the launch record launch by launch: one entry
    per pair-kernel launch in the wall-RPY family, positive times whose sum
    and median agree with its totals, nothing recorded outside the block."""
    import torch
    rng = np.random.default_rng(21)
    n = 4096
    r = torch.from_numpy(np.column_stack([rng.uniform(0, 60, n), rng.uniform(0, 60, n),
                                          rng.uniform(0.8, 2.0, n)])).to(gpu)
    f = torch.from_numpy(rng.standard_normal(3 * n)).to(gpu)
    print(f"case PairKernelTimer UF N={n} x5", flush=True)
    wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.5, precision="fp32_blocked")   # not recorded
    with wall_rpy.PairKernelTimer() as pk:
        for _ in range(5):
            wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=0.5, precision="fp32_blocked")
    d = pk.durations
    assert pk.launches == 5 and len(d["wall_rpy"]) == 5 and not d["radii"] and not d["forces"]
    assert all(t > 0.0 for t in d["wall_rpy"])
    assert abs(pk.total_ms - sum(d["wall_rpy"])) <= 1e-9 * pk.total_ms
    assert pk.by_kind["wall_rpy"][0] == 5 and min(d["wall_rpy"]) <= pk.median_ms <= max(d["wall_rpy"])
    with wall_rpy.PairKernelTimer() as empty:
        pass
    assert empty.launches == 0 and empty.median_ms == 0.0
