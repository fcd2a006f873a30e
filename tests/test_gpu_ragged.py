"""Ragged batches of independent suspensions: the segmented FP64 Krylov
kernels against numpy, the ragged Lanczos against per-system runs and the
eigen-decomposition square root, and the C5 batch (64 roller replicas of
128..4096 blobs) through one persistent grand launch per step, with
per-replica errors against FP64."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, lanczos, wall_rpy, workloads
from rigidmultiblobswall_amd.rollers import GRAND, RaggedRollers, Rollers

pytestmark = pytest.mark.gpu


def rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _stream():
    return torch.cuda.current_stream().cuda_stream


SIZES = [0, 130, 7, 0, 300, 1, 128, 129, 0]


def _seg(gpu, sizes=SIZES):
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    return lanczos.Segments(off, comp=3, nparts=2, device=gpu)


def test_segmented_dot_combination_scale(gpu):
    lib = _lib.load()
    seg = _seg(gpu)
    es = seg.element_system()
    rng = np.random.default_rng(0)
    nvec = 5
    Vh = rng.standard_normal((nvec, seg.dim))
    wh = rng.standard_normal(seg.dim)
    V = torch.from_numpy(Vh).to(gpu)
    w = torch.from_numpy(wh).to(gpu)
    print(f"case segmented kernels sizes={SIZES}", flush=True)
    ws = seg.workspace(nvec)
    out = torch.full((seg.nsys, 8), 7.0, dtype=torch.float64, device=gpu)
    _lib.check(lib.rmw_seg_dot_many(V.data_ptr(), seg.dim, nvec, w.data_ptr(), seg.ref, out.data_ptr(), 8, 0,
                                    ws.data_ptr(), ws.numel(), _stream()))
    got = out.cpu().numpy()
    for s in range(seg.nsys):
        ref = Vh[:, es == s] @ wh[es == s]
        assert np.allclose(got[s, :nvec], ref, rtol=1e-13, atol=1e-12)
        assert np.all(got[s, nvec:] == 7.0)                 # nothing beyond nvec written
    # deterministic: a second call gives the same bits
    out2 = out.clone()
    _lib.check(lib.rmw_seg_dot_many(V.data_ptr(), seg.dim, nvec, w.data_ptr(), seg.ref, out2.data_ptr(), 8, 0,
                                    ws.data_ptr(), ws.numel(), _stream()))
    assert torch.equal(out, out2)
    # combination with per-system coefficients, accumulate
    coef_h = rng.standard_normal((seg.nsys, nvec))
    y0 = rng.standard_normal(seg.dim)
    y = torch.from_numpy(y0.copy()).to(gpu)
    coef = torch.from_numpy(coef_h).to(gpu)
    _lib.check(lib.rmw_seg_combination(V.data_ptr(), seg.dim, nvec, coef.data_ptr(), nvec, y.data_ptr(), seg.ref, 1,
                                       _stream()))
    ref = y0 + np.einsum("ij,ji->i", Vh.T, coef_h[es].T)
    assert np.allclose(y.cpu().numpy(), ref, rtol=1e-13, atol=1e-12)
    # scale by a per-system divisor; zero divisor -> e_0 of the system
    div_h = rng.uniform(0.5, 2.0, seg.nsys)
    div_h[4] = 0.0
    div = torch.from_numpy(div_h).to(gpu)
    yy = torch.full((seg.dim,), 9.0, dtype=torch.float64, device=gpu)
    _lib.check(lib.rmw_seg_scale(w.data_ptr(), yy.data_ptr(), div.data_ptr(), 1, seg.ref, _stream()))
    g = yy.cpu().numpy()
    for s in range(seg.nsys):
        sel = es == s
        if s == 4:
            e0 = np.zeros(sel.sum())
            e0[0] = 1.0
            assert np.array_equal(g[sel], e0)
        else:
            assert np.allclose(g[sel], wh[sel] / div_h[s], rtol=1e-15)


def test_segmented_arnoldi_orthonormal_per_system(gpu):
    lib = _lib.load()
    seg = _seg(gpu)
    es = seg.element_system()
    rng = np.random.default_rng(1)
    m = 6
    dim = seg.dim
    V = torch.zeros((m + 1, dim), dtype=torch.float64, device=gpu)
    z = torch.from_numpy(rng.standard_normal(dim)).to(gpu)
    nrm = seg.dot(z, z, sqrt=True)
    _lib.check(lib.rmw_seg_scale(z.data_ptr(), V[0].data_ptr(), nrm.data_ptr(), 1, seg.ref, _stream()))
    H = torch.zeros((seg.nsys, m, m + 1), dtype=torch.float64, device=gpu)
    hs = torch.empty(seg.nsys * (m + 1), dtype=torch.float64, device=gpu)
    ws = seg.workspace(m + 1)
    A = rng.standard_normal((dim, dim)) * 0.1
    print("case segmented CGS2 Arnoldi, 6 steps", flush=True)
    for j in range(m):
        # block-diagonal operator: each system only sees itself
        v = V[j].cpu().numpy()
        wv = np.zeros(dim)
        for s in range(seg.nsys):
            sel = es == s
            wv[sel] = A[np.ix_(sel, sel)] @ v[sel]
        w = torch.from_numpy(wv).to(gpu)
        _lib.check(lib.rmw_seg_arnoldi_cgs2(V.data_ptr(), dim, j, w.data_ptr(), seg.ref, H[0, j].data_ptr(),
                                            H.stride(0), hs.data_ptr(), ws.data_ptr(), ws.numel(), _stream()))
    Vh = V.cpu().numpy()
    Hh = H.cpu().numpy()
    sizes = np.diff(seg.offsets)
    for s in range(seg.nsys):
        if sizes[s] == 0:
            continue
        sel = es == s
        Q = Vh[:, sel]
        d = sel.sum()
        mm = min(m + 1, d)
        G = Q[:mm] @ Q[:mm].T
        assert np.allclose(G, np.eye(mm), atol=1e-12), s
        # Arnoldi relation A Q_j = sum_i H[j, i] Q_i
        As = A[np.ix_(sel, sel)]
        for j in range(min(m, d - 1)):
            lhs = As @ Q[j]
            rhs = Hh[s, j, :j + 2] @ Q[:j + 2]
            assert np.allclose(lhs, rhs, atol=1e-12), (s, j)


def _dense_grand(r, eta, a):
    return np.block([[O.dense_block_matrix(O.UF, r, eta, a), O.dense_block_matrix(O.UT, r, eta, a)],
                     [O.dense_block_matrix(O.WF, r, eta, a), O.dense_block_matrix(O.WT, r, eta, a)]])


def test_ragged_lanczos_matches_per_system_runs(gpu):
    rng = np.random.default_rng(3)
    sizes = [9, 0, 40, 23, 2]
    eta, a = 1e-3, 0.656
    rs = [np.column_stack([rng.uniform(0, 8, n), rng.uniform(0, 8, n), rng.uniform(0.8, 2.0, n)]) for n in sizes]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    r = np.vstack(rs)
    RR = RaggedRollers(r, off, eta, a, precision="fp64", device=gpu)
    Z = rng.standard_normal((2, 6 * off[-1]))
    print(f"case ragged Lanczos sizes={sizes} k=2 fp64 tol 1e-10", flush=True)
    noise, its = RR.brownian_noise(torch.from_numpy(Z).to(gpu), tolerance=1e-10)
    noise = noise.cpu().numpy()
    assert its.shape == (2, len(sizes))
    for s, n in enumerate(sizes):
        if n == 0:
            assert its[0, s] == 0
            continue
        idx = RR.system_slices(s)
        G = _dense_grand(rs[s], eta, a)
        lam, Q = np.linalg.eigh(0.5 * (G + G.T))
        single = Rollers(rs[s], eta, a, precision="fp64")
        for row in range(2):
            zs = Z[row, idx]
            exact = Q @ (np.sqrt(np.maximum(lam, 0)) * (Q.T @ zs))
            got = noise[row, idx]
            e = rel(got, exact)
            one, it1 = single.brownian_noise(torch.from_numpy(zs[None]).to(gpu), tolerance=1e-10)
            ref, it_ref = O.lanczos_sqrt(lambda v: G @ v, zs, tolerance=1e-10)
            print(f"  system {s} (n={n}) row {row}: it {its[row, s]} (single {it1}, oracle {it_ref}), "
                  f"rel vs eigh {e:.2e}")
            assert e < 1e-7
            assert rel(got, one[0].cpu().numpy()) < 1e-8
            assert abs(its[row, s] - it1) <= 1
            assert abs(its[row, s] - it_ref) <= 1
            assert rel(got, ref) < 1e-8


@pytest.fixture(scope="module")
def c5():
    return workloads.c5_ragged_rollers()


def test_c5_ragged_grand_matvec_per_replica(gpu, c5):
    r, off, eta, a = c5["r"], c5["offsets"], c5["eta"], c5["a"]
    n = int(off[-1])
    rng = np.random.default_rng(8)
    f = rng.standard_normal((n, 3))
    t = rng.standard_normal((n, 3))
    print(f"case C5 ragged grand matvec: {len(off) - 1} replicas, {n} blobs, fp32_blocked", flush=True)
    u, w = wall_rpy.apply_ragged(GRAND, r, off, force=f, torque=t, eta=eta, a=a, precision="fp32_blocked")
    u, w = u.cpu().numpy(), w.cpu().numpy()
    errs = []
    rows = 48
    for s in range(len(off) - 1):
        b0, b1 = int(off[s]), int(off[s + 1])
        rr, fs, ts = r[b0:b1], f[b0:b1], t[b0:b1]
        ou = (O.block_apply(O.UF, rr, fs, eta, a, rows=(0, rows)) + O.block_apply(O.UT, rr, ts, eta, a, rows=(0, rows)))
        ow = (O.block_apply(O.WF, rr, fs, eta, a, rows=(0, rows)) + O.block_apply(O.WT, rr, ts, eta, a, rows=(0, rows)))
        gu = u[b0:b0 + rows].reshape(-1)
        gw = w[b0:b0 + rows].reshape(-1)
        ref = np.concatenate([ou[:3 * rows], ow[:3 * rows]])
        errs.append(rel(np.concatenate([gu, gw]), ref))
    errs = np.array(errs)
    print(f"  per-replica rel L2 vs FP64 oracle (first {rows} rows): median {np.median(errs):.2e} max {errs.max():.2e}")
    assert errs.max() <= 1e-6
    # With the same kernel (ordered or symmetric) and the same work split, a
    # replica of the batch gives the bits of the replica alone.
    for sym in (-1, 1):
        tun = _lib.Tuning(1024, 0, 0, sym, 0)
        for prec in ("fp64", "fp32_blocked"):
            ub, wb = wall_rpy.apply_ragged(GRAND, r, off, force=f, torque=t, eta=eta, a=a, precision=prec,
                                           tuning=tun)
            for s in (0, 17, 63):
                b0, b1 = int(off[s]), int(off[s + 1])
                us, wsv = wall_rpy.apply(GRAND, r[b0:b1], force=f[b0:b1], torque=t[b0:b1], eta=eta, a=a,
                                         precision=prec, tuning=tun)
                assert torch.equal(ub[b0:b1], us) and torch.equal(wb[b0:b1], wsv), (sym, prec, s)


def test_c5_ragged_lanczos_per_replica(gpu, c5):
    r, off, eta, a = c5["r"], c5["offsets"], c5["eta"], c5["a"]
    n = int(off[-1])
    nsys = len(off) - 1
    rng = np.random.default_rng(9)
    Z = torch.from_numpy(rng.standard_normal((1, 6 * n))).to(gpu)
    RR = RaggedRollers(r, off, eta, a, precision="fp32_blocked", device=gpu)
    print(f"case C5 ragged Lanczos: {nsys} replicas, {n} blobs, tol 1e-3, fp32_blocked", flush=True)
    noise, its = RR.brownian_noise(Z, tolerance=1e-3)
    print(f"  iterations per replica: min {its.min()} max {its.max()}")
    assert its.min() >= 1 and its.max() < 200
    # ||noise_s||^2 = z_s^T M_s z_s (FP64 pairs)
    MZ = torch.empty_like(Z)
    RR.grand_matvec(Z.contiguous(), MZ, precision="fp64")
    zMz = RR.seg.dot(Z[0], MZ[0]).cpu().numpy()
    nn = RR.seg.dot(noise[0], noise[0]).cpu().numpy()
    e_inv = np.abs(nn - zMz) / np.abs(zMz)
    print(f"  | ||noise||^2 - z^T M z | / z^T M z: max {e_inv.max():.2e}")
    assert e_inv.max() <= 1e-6
    # per-replica error against the same batch with FP64 pairs
    noise64, its64 = RR.brownian_noise(Z, tolerance=1e-3, precision="fp64")
    nz, n64 = noise[0].cpu().numpy(), noise64[0].cpu().numpy()
    errs = np.array([rel(nz[RR.system_slices(s)], n64[RR.system_slices(s)]) for s in range(nsys)])
    print(f"  per-replica rel L2 vs FP64 pairs: median {np.median(errs):.2e} max {errs.max():.2e}")
    assert errs.max() < 2e-3
    # FP64 batch == FP64 per-replica runs
    for s in (3, 40):
        idx = RR.system_slices(s)
        b0, b1 = int(off[s]), int(off[s + 1])
        one, it1 = Rollers(r[b0:b1], eta, a, precision="fp64").brownian_noise(
            Z[:, torch.from_numpy(idx).to(gpu)].contiguous(), tolerance=1e-3)
        assert it1 == its64[0, s]
        assert rel(n64[idx], one[0].cpu().numpy()) < 1e-9, s


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_prepared_strided_products(gpu, prec):
    """wall_rpy.StridedProduct (launch shape and workspace settled once per
    row count) equals apply_strided / apply_ragged_strided bit for bit, for
    one suspension and for a ragged batch, at two row counts, and refuses
    rows that do not hold the grand vectors."""
    rng = np.random.default_rng(21)
    sizes = [130, 7, 300, 1, 129]
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    n = int(off[-1])
    r = np.column_stack([rng.uniform(0, 25, n), rng.uniform(0, 25, n), rng.uniform(0.7, 3.0, n)])
    rd = torch.from_numpy(r).to(gpu)
    eta, a = 1.0, 0.45
    print(f"case StridedProduct prec={prec} N={n} systems={len(sizes)}", flush=True)
    single = wall_rpy.StridedProduct(GRAND, rd, eta, a, precision=prec)
    ragged = wall_rpy.StridedProduct(GRAND, rd, eta, a, precision=prec, offsets=off)
    for k in (3, 1, 3):
        X = torch.from_numpy(rng.standard_normal((k, 6 * n))).to(gpu)
        Y1, Y2 = torch.empty_like(X), torch.empty_like(X)
        single(X, Y1)
        wall_rpy.apply_strided(GRAND, rd, n, X, Y2, eta, a, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(Y1, Y2)
        ragged(X, Y1)
        wall_rpy.apply_ragged_strided(GRAND, rd, off, X, Y2, eta, a, precision=prec)
        torch.cuda.synchronize()
        assert torch.equal(Y1, Y2)
    # every system's block of the ragged product is that system alone
    b0, b1 = int(off[2]), int(off[3])
    rs = torch.from_numpy(np.ascontiguousarray(r[b0:b1])).to(gpu)
    Xs = torch.cat([X[:, 3 * b0:3 * b1], X[:, 3 * n + 3 * b0:3 * n + 3 * b1]], dim=1).contiguous()
    Ys = torch.empty_like(Xs)
    wall_rpy.StridedProduct(GRAND, rs, eta, a, precision=prec)(Xs, Ys)
    got = torch.cat([Y1[:, 3 * b0:3 * b1], Y1[:, 3 * n + 3 * b0:3 * n + 3 * b1]], dim=1)
    assert rel(got.cpu().numpy(), Ys.cpu().numpy()) <= (1e-13 if prec == "fp64" else 2e-6)
    with pytest.raises(ValueError):
        single(X[:, : 6 * n - 3], Y1)
    with pytest.raises(ValueError):
        wall_rpy.StridedProduct(GRAND, rd, eta, a, precision=prec, offsets=off[:-1])
