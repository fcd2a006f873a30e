"""GPU checks of the learned warm start: BF16 MFMA GEMM, network forward and
gradients against the FP64 CPU oracle, graph features, and an end-to-end
warm start on a small monolayer."""
import ctypes

import numpy as np
import pytest
import torch

from oracle import gnn_oracle as GO
from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, gnn, workloads

pytestmark = pytest.mark.gpu


def _bf16_tensor(x, dev):
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16).astype(np.uint16)
    return torch.from_numpy(u.view(np.int16).copy()).to(dev)


@pytest.mark.parametrize("tA,tB", [(0, 0), (0, 1), (1, 0), (1, 1)])
@pytest.mark.parametrize("M,N,K", [(77, 45, 37), (128, 64, 16), (5, 132, 300)])
def test_gemm_bf16_matches_fp64(gpu, tA, tB, M, N, K):
    lib = _lib.load()
    rng = np.random.default_rng(M * 7 + N + K)
    Am = GO.bf16_round(rng.standard_normal((M, K)))
    Bm = GO.bf16_round(rng.standard_normal((K, N)))
    bias = rng.standard_normal(N).astype(np.float32)
    As = Am.T.copy() if tA else Am
    Bs = Bm.T.copy() if tB else Bm
    A = _bf16_tensor(As, gpu)
    B = _bf16_tensor(Bs, gpu)
    C = torch.empty((M, N), dtype=torch.float32, device=gpu)
    H = torch.empty((M, N), dtype=torch.int16, device=gpu)
    bd = torch.from_numpy(bias).to(gpu)
    print(f"case gemm tA={tA} tB={tB} M={M} N={N} K={K}", flush=True)
    _lib.check(lib.rmw_gemm_bf16(tA, tB, M, N, K, A.data_ptr(), As.shape[1], B.data_ptr(), Bs.shape[1],
                                 C.data_ptr(), N, 0.0, H.data_ptr(), N, bd.data_ptr(), 1,
                                 torch.cuda.current_stream().cuda_stream))
    ref = Am @ Bm + bias
    got = C.cpu().numpy()
    assert np.abs(got - ref).max() <= 1e-5 * np.abs(ref).max()
    hs = (H.cpu().numpy().view(np.uint16).astype(np.uint32) << 16).view(np.float32)
    silu = ref / (1 + np.exp(-ref))
    assert np.abs(hs - silu).max() <= 1e-2 * np.abs(silu).max()


@pytest.mark.parametrize("M,N,K,splits", [(64, 16, 16384, 128), (128, 80, 2048, 16), (132, 128, 2048, 3),
                                          (64, 64, 1000, 7), (5, 9, 40, 4)])
def test_gemm_bf16_splitk_matches_fp64(gpu, M, N, K, splits):
    """Weight-gradient form C = A^T B with K split into slices (summed in
    slice order), with beta accumulation; K not a multiple of the 32-wide
    step and more slices than 32-wide steps exercised."""
    lib = _lib.load()
    rng = np.random.default_rng(M + N + K + splits)
    Am = GO.bf16_round(rng.standard_normal((M, K)))
    Bm = GO.bf16_round(rng.standard_normal((K, N)))
    A = _bf16_tensor(Am.T.copy(), gpu)          # stored K x M
    B = _bf16_tensor(Bm, gpu)                   # stored K x N
    C0 = rng.standard_normal((M, N)).astype(np.float32)
    C = torch.from_numpy(C0.copy()).to(gpu)
    nb = ctypes.c_size_t()
    _lib.check(lib.rmw_gemm_bf16_splitk_workspace(M, N, K, splits, ctypes.byref(nb)))
    part = torch.empty(max(1, nb.value // 4), dtype=torch.float32, device=gpu)
    print(f"case gemm split-K M={M} N={N} K={K} splits={splits}", flush=True)
    _lib.check(lib.rmw_gemm_bf16_splitk(1, 0, M, N, K, A.data_ptr(), M, B.data_ptr(), N, C.data_ptr(), N, 0.5,
                                        splits, part.data_ptr(), part.numel() * 4,
                                        torch.cuda.current_stream().cuda_stream))
    ref = Am @ Bm + 0.5 * C0
    assert np.abs(C.cpu().numpy() - ref).max() <= 1e-5 * np.abs(ref).max() * max(1.0, np.sqrt(K / 2048))
    # the unsplit kernel agrees
    C1 = torch.empty((M, N), dtype=torch.float32, device=gpu)
    _lib.check(lib.rmw_gemm_bf16(1, 0, M, N, K, A.data_ptr(), M, B.data_ptr(), N, C1.data_ptr(), N, 0.0, None, 0,
                                 None, 0, torch.cuda.current_stream().cuda_stream))
    assert np.abs(C1.cpu().numpy() - Am @ Bm).max() <= 1e-5 * np.abs(Am @ Bm).max() * max(1.0, np.sqrt(K / 2048))


@pytest.mark.parametrize("rows,cols,ldx", [(16384, 64, 64), (2048, 132, 132), (2048, 16, 80), (3, 200, 201),
                                           (1, 1, 1)])
def test_colsum_bf16_matches_fp64(gpu, rows, cols, ldx):
    lib = _lib.load()
    rng = np.random.default_rng(rows + cols)
    X = GO.bf16_round(rng.standard_normal((rows, ldx)))
    Xd = _bf16_tensor(X, gpu)
    out0 = rng.standard_normal(cols).astype(np.float32)
    out = torch.from_numpy(out0.copy()).to(gpu)
    print(f"case colsum rows={rows} cols={cols} ldx={ldx}", flush=True)
    _lib.check(lib.rmw_colsum_bf16(Xd.data_ptr(), ldx, rows, cols, out.data_ptr(), 2.0,
                                   torch.cuda.current_stream().cuda_stream))
    ref = X[:, :cols].sum(axis=0) + 2.0 * out0
    assert np.abs(out.cpu().numpy() - ref).max() <= 1e-5 * max(1.0, np.abs(ref).max()) * max(1.0, np.sqrt(rows / 256))


def _net_and_batch(gpu, B=37, seed=3):
    net = gnn.GraphNet(blobs_per_body=42, device=gpu, seed=seed)
    rng = np.random.default_rng(seed)
    Xn = rng.standard_normal((B, gnn.NIN)).astype(np.float32)
    Xe = rng.standard_normal((B * gnn.DEG, gnn.EIN)).astype(np.float32)
    T = rng.standard_normal((B, net.nout)).astype(np.float32)
    batch = gnn.Batch(Xn=torch.from_numpy(Xn).to(gpu), Xe=torch.from_numpy(Xe).to(gpu), T=torch.from_numpy(T).to(gpu))
    P = {k: GO.bf16_round(v).astype(np.float64) for k, v in net.params_numpy().items()}
    for b in ("be1", "be2", "bn1", "bn2", "bn3"):   # biases stay FP32 on the GPU
        P[b] = net.params_numpy()[b].astype(np.float64)
    return net, batch, P, GO.bf16_round(Xn), GO.bf16_round(Xe), T.astype(np.float64)


def test_forward_and_gradients_match_oracle(gpu):
    net, batch, P, Xn, Xe, T = _net_and_batch(gpu)
    # non-trivial biases
    rng = np.random.default_rng(9)
    for b in ("be1", "be2", "bn1", "bn2", "bn3"):
        v = (0.1 * rng.standard_normal(net.W(b).shape)).astype(np.float32)
        net.W(b).copy_(torch.from_numpy(v))
        P[b] = v.astype(np.float64)
    print("case gnn forward/backward B=37", flush=True)
    loss = net.loss_and_grad(batch)
    Y = net.forward(batch).cpu().numpy()
    Yr, _ = GO.forward(P, Xn.astype(np.float64), Xe.astype(np.float64))
    assert np.linalg.norm(Y - Yr) <= 3e-2 * np.linalg.norm(Yr)
    Lr, gr = GO.loss_and_grad(P, Xn.astype(np.float64), Xe.astype(np.float64), T)
    assert abs(loss * net.nout - Lr) <= 5e-2 * abs(Lr)
    for name in net.names:
        g = net.G(name).cpu().numpy().astype(np.float64)
        r = gr[name]
        cos = float((g * r).sum() / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-300))
        err = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-300)
        print(f"  {name}: rel {err:.2e} cos {cos:.6f}")
        # measured on MI355X: rel <= 4e-3, cos >= 0.99999 (BF16 operands)
        assert cos > 0.9999 and err < 1e-2, name


def test_training_reduces_loss(gpu):
    net, batch, P, Xn, Xe, T = _net_and_batch(gpu, B=512, seed=5)
    # a smooth target of the inputs, learnable by the architecture
    xn = batch.Xn.cpu().numpy()
    xe = batch.Xe.cpu().numpy().reshape(512, gnn.DEG, gnn.EIN)
    feat = np.concatenate([xn, np.tanh(xe).sum(axis=1)], axis=1)
    W = np.random.default_rng(1).standard_normal((feat.shape[1], net.nout)) / np.sqrt(feat.shape[1])
    tgt = np.tanh(feat @ W).astype(np.float32)
    batch.T = torch.from_numpy(tgt).to(gpu)
    l0 = net.loss_and_grad(batch)
    for it in range(300):
        net.loss_and_grad(batch)
        net.adam_step(3e-3)
    l1 = net.loss_and_grad(batch)
    print(f"case training: loss {l0:.3e} -> {l1:.3e}")
    assert l1 < 0.1 * l0


@pytest.mark.parametrize("nb,deg,lattice", [(9, 8, False), (300, 8, False), (1024, 8, False), (400, 16, True),
                                            (700, 3, True)])
def test_knn_matches_sorted_distances(gpu, nb, deg, lattice):
    """Neighbour lists: ascending squared distance, ties by index, the body
    itself excluded; a square lattice gives exact ties."""
    rng = np.random.default_rng(nb + deg)
    if lattice:
        side = int(np.ceil(np.sqrt(nb)))
        g = np.stack(np.meshgrid(np.arange(side), np.arange(side), indexing="ij"), -1).reshape(-1, 2)[:nb]
        c = np.column_stack([g * 2.0, np.full(nb, 1.3)])
    else:
        c = np.column_stack([rng.uniform(0, 60, nb), rng.uniform(0, 60, nb), rng.uniform(1.2, 1.4, nb)])
    cd = torch.from_numpy(np.ascontiguousarray(c)).to(gpu)
    nbr = torch.empty(nb * deg, dtype=torch.int32, device=gpu)
    lib = _lib.load()
    print(f"case knn nb={nb} deg={deg} lattice={lattice}", flush=True)
    _lib.check(lib.rmw_gnn_knn(cd.data_ptr(), nb, deg, nbr.data_ptr(), torch.cuda.current_stream().cuda_stream))
    got = nbr.cpu().numpy().reshape(nb, deg)
    d2 = ((c[:, None, :] - c[None, :, :]) ** 2).sum(-1)
    idx = np.arange(nb)
    for i in range(nb):
        keys = np.lexsort((idx, d2[i]))          # by distance, then index
        keys = keys[keys != i][:deg]
        assert np.array_equal(got[i], keys), i


def test_features_and_body_frame_round_trip(gpu):
    w = workloads.c4_monolayer(nbodies=64, seed=7)
    net = gnn.GraphNet(42, gpu)
    feats = gnn.GraphFeatures(net)
    rng = np.random.default_rng(2)
    c = torch.from_numpy(w["centres"]).to(gpu)
    q = torch.from_numpy(w["quats"]).to(gpu)
    ft = torch.from_numpy(rng.standard_normal((64, 6))).to(gpu)
    b = feats.batch(c, q, ft)
    # neighbours: 8 nearest centres
    d = np.linalg.norm(w["centres"][:, None] - w["centres"][None], axis=-1)
    np.fill_diagonal(d, np.inf)
    nn = np.argsort(d, axis=1)[:, :8]
    Xe = b.Xe.cpu().numpy().reshape(64, 8, 16)
    R = workloads.rotation_matrices(w["quats"])
    for i in (0, 17, 63):
        rel = w["centres"][nn[i]] - w["centres"][i]
        body = rel @ R[i]                                   # R^T d for each neighbour
        assert np.allclose(Xe[i, :, 0:3], body, atol=1e-5)
        assert np.allclose(Xe[i, :, 3], np.linalg.norm(rel, axis=1), atol=1e-5)
    zb = b.Xn.cpu().numpy()[:, 0:3]
    assert np.allclose(zb, R[:, 2, :], atol=1e-6)           # R^T z-hat = third row of R
    # targets -> initial guess reproduces the solution
    nb, nbb = 64, 42
    offs = torch.arange(0, (nb + 1) * nbb, nbb, dtype=torch.int64, device=gpu)
    n = 3 * nb * nbb + 6 * nb
    sol = torch.from_numpy(rng.standard_normal(n)).to(gpu)
    pc = torch.from_numpy(rng.standard_normal(n)).to(gpu)
    net.fscale = 2.5
    T = feats.targets(q, offs, nb * nbb, sol, pc)
    x0 = feats.initial_guess(q, offs, nb * nbb, T, pc)
    assert torch.allclose(x0, sol, rtol=0, atol=1e-5 * float(sol.abs().max()))


def test_small_warm_start_end_to_end(gpu):
    """64 bodies, 16 Brownian steps: the trained network's x0 beats the
    preconditioner guess and its own point-field base on held-out steps, and
    GMRES from it converges to the same FP64 tolerance in no more
    iterations."""
    from rigidmultiblobswall_amd.trajectory import run_c4_trajectory
    from rigidmultiblobswall_amd.warmstart import WarmStart, evaluate
    w = workloads.c4_monolayer(nbodies=64, seed=11)
    print("case warm start 64 bodies x 16 steps", flush=True)
    traj = run_c4_trajectory(w["centres"], w["quats"], w["ref"], w["a"], w["eta"], nsteps=16, kT=0.1)
    ws = WarmStart(w["ref"], w["a"], w["eta"], device=gpu)
    fit = ws.fit(traj.snapshots[:12], epochs=60, batch=256)
    assert fit["loss_history"][-1] < 0.5 * fit["loss_history"][0]
    rows = evaluate(ws, traj.snapshots[13:15], traj.snapshots[12:14])
    for row in rows:
        print("  ", {k: (v["x0_rel_err"], v["iterations"]) for k, v in row.items() if isinstance(v, dict)})
        assert row["gnn/fp64"]["converged"] and row["zero/fp64"]["converged"] and row["gnn/mixed"]["converged"]
        assert row["gnn/fp64"]["x0_rel_err"] < row["pc/fp64"]["x0_rel_err"]
        assert row["gnn/fp64"]["iterations"] <= row["pc/fp64"]["iterations"]
        # the point-field base beats P^-1 b, and the network improves on its base
        assert row["point/fp64"]["x0_rel_err"] < row["pc/fp64"]["x0_rel_err"]
        assert row["gnn/fp64"]["x0_rel_err"] < row["point/fp64"]["x0_rel_err"]


def test_motion_features(gpu):
    """Motion features of the previous-step network: the first 10 node
    features equal the static ones; a configuration compared with itself has
    zero motion; a known displacement and rotation of every body appear in
    the body frame (rotation vector, displacement, neighbour-offset change)."""
    from rigidmultiblobswall_amd.trajectory import quat_from_rotation, quat_mul
    w = workloads.c4_monolayer(nbodies=64, seed=8)
    net = gnn.GraphNet(42, gpu, nin=gnn.NIN_MOTION)
    net.fscale, net.dscale, net.rscale = 2.0, 0.01, 0.02
    feats = gnn.GraphFeatures(net)
    rng = np.random.default_rng(3)
    c0 = w["centres"]
    q0 = w["quats"]
    f0 = rng.standard_normal((64, 6))
    dev = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(gpu)   # noqa: E731
    print("case motion features, 64 bodies", flush=True)
    same = feats.batch(dev(c0), dev(q0), dev(f0), previous=(dev(c0), dev(q0), dev(f0)))
    Xn = same.Xn.cpu().numpy()
    # q q^-1 is the identity up to rounding: rotation vectors of ~1e-16 / rscale
    assert Xn.shape == (64, 22) and np.abs(Xn[:, 10:]).max() < 1e-12
    assert np.abs(same.Xe.cpu().numpy().reshape(64, 8, 16)[:, :, 5:9]).max() == 0.0
    net10 = gnn.GraphNet(42, gpu)
    net10.fscale = 2.0
    static = gnn.GraphFeatures(net10).batch(dev(c0), dev(q0), dev(f0))
    assert np.array_equal(Xn[:, :10], static.Xn.cpu().numpy())
    # move every body
    dc = 0.01 * rng.standard_normal((64, 3))
    th = 0.02 * rng.standard_normal((64, 3))
    c1 = c0 + dc
    q1 = quat_mul(quat_from_rotation(th), q0)
    f1 = f0 + 0.1 * rng.standard_normal((64, 6))
    b = feats.batch(dev(c1), dev(q1), dev(f1), previous=(dev(c0), dev(q0), dev(f0)))
    Xn = b.Xn.cpu().numpy().astype(np.float64)
    R1 = workloads.rotation_matrices(q1)
    body = lambda v: np.einsum("bji,bj->bi", R1, v)   # noqa: E731  R^T v per body
    assert np.allclose(Xn[:, 10:13], body(dc) / 0.01, atol=1e-4)
    assert np.allclose(Xn[:, 13:16], body(th) / 0.02, atol=1e-4)
    assert np.allclose(Xn[:, 16:19], body(f1[:, :3] - f0[:, :3]) / 2.0, atol=1e-5)
    assert np.allclose(Xn[:, 19:22], body(f1[:, 3:] - f0[:, 3:]) / 2.0, atol=1e-5)
    d = np.linalg.norm(c1[:, None] - c1[None], axis=-1)
    np.fill_diagonal(d, np.inf)
    nn = np.argsort(d, axis=1)[:, :8]
    Xe = b.Xe.cpu().numpy().reshape(64, 8, 16).astype(np.float64)
    for i in (0, 31, 63):
        dd = (c1[nn[i]] - c1[i]) - (c0[nn[i]] - c0[i])
        assert np.allclose(Xe[i, :, 5:8], (dd @ R1[i]) / 0.01, atol=1e-4)
        dl = np.linalg.norm(c1[nn[i]] - c1[i], axis=1) - np.linalg.norm(c0[nn[i]] - c0[i], axis=1)
        assert np.allclose(Xe[i, :, 8], dl / 0.01, atol=1e-4)


def test_motion_network_gradients_match_oracle(gpu):
    """The 22-input network (node width 96): GPU forward and gradients =
    the FP64 oracle on BF16-rounded operands."""
    rng = np.random.default_rng(13)
    B = 41
    net = gnn.GraphNet(12, gpu, seed=4, nin=gnn.NIN_MOTION)
    Xn = rng.standard_normal((B, gnn.NIN_MOTION)).astype(np.float32)
    Xe = rng.standard_normal((B * gnn.DEG, gnn.EIN)).astype(np.float32)
    T = rng.standard_normal((B, net.nout)).astype(np.float32)
    batch = gnn.Batch(Xn=torch.from_numpy(Xn).to(gpu), Xe=torch.from_numpy(Xe).to(gpu), T=torch.from_numpy(T).to(gpu))
    P = {k: GO.bf16_round(v).astype(np.float64) for k, v in net.params_numpy().items()}
    for bname in ("be1", "be2", "bn1", "bn2", "bn3"):
        v = (0.1 * rng.standard_normal(net.W(bname).shape)).astype(np.float32)
        net.W(bname).copy_(torch.from_numpy(v))
        P[bname] = v.astype(np.float64)
    print("case motion network forward/backward B=41, nin 22", flush=True)
    loss = net.loss_and_grad(batch)
    Y = net.forward(batch).cpu().numpy()
    Xn64, Xe64 = GO.bf16_round(Xn), GO.bf16_round(Xe)
    Yr, _ = GO.forward(P, Xn64, Xe64)
    assert np.linalg.norm(Y - Yr) <= 3e-2 * np.linalg.norm(Yr)
    Lr, gr = GO.loss_and_grad(P, Xn64, Xe64, T.astype(np.float64))
    assert abs(loss * net.nout - Lr) <= 5e-2 * abs(Lr)
    for name in net.names:
        g = net.G(name).cpu().numpy().astype(np.float64)
        r = gr[name]
        cos = float((g * r).sum() / (np.linalg.norm(g) * np.linalg.norm(r) + 1e-300))
        err = np.linalg.norm(g - r) / (np.linalg.norm(r) + 1e-300)
        print(f"  {name}: rel {err:.2e} cos {cos:.6f}")
        assert cos > 0.9999 and err < 1e-2, name
