"""The learned-warm-start configuration at its full size: 1024 spheres x 42
blobs (43008 blobs, system size 135168) in a monolayer of area fraction 0.4.

Checks the pieces the warm-start study runs on at this size: the FP64 rigid
solve to 1e-8 (true residual of the returned solution, judged with the FP64
operator), the FP32-pair UF product against the oracle on sampled rows, the
point-field product against the oracle's different-radii product on sampled
bodies, and a finite network guess.
"""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import wall_rpy, workloads
from rigidmultiblobswall_amd._lib import RMW_UF
from rigidmultiblobswall_amd.rigid import RigidSuspension
from rigidmultiblobswall_amd.trajectory import ForceParams, Snapshot, body_forces

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def c4(gpu):
    w = workloads.c4_monolayer()
    assert w["centres"].shape[0] == 1024 and w["r"].shape[0] == 43008
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    ft = body_forces(w["centres"], w["quats"], w["ref"], w["a"], ForceParams())
    return w, S, ft


def test_c4_fp64_solve_reaches_tolerance(c4):
    w, S, ft = c4
    print("case C4 1024 spheres x 42 blobs: FP64 GMRES to 1e-8 from zero", flush=True)
    rhs = S.rhs(np.zeros((S.Nblobs, 3)), ft)
    out = S.solve_mobility_problem(rhs=rhs, tol=1e-8)
    r = torch.empty_like(rhs)
    S.operator(out.x, r, precision="fp64")
    true_res = float(torch.linalg.norm(r - rhs) / torch.linalg.norm(rhs))
    print(f"  iterations {out.iterations}, true residual {true_res:.2e}")
    assert out.info == 0 and true_res <= 1.5e-8
    # the mixed-precision solve meets the same FP64 tolerance
    mixed = S.solve_mixed(rhs=rhs, tol=1e-8)
    S.operator(mixed.x, r, precision="fp64")
    mres = float(torch.linalg.norm(r - rhs) / torch.linalg.norm(rhs))
    print(f"  mixed: inner iterations {mixed.iterations}, FP64 residual {mres:.2e}")
    assert mres <= 1.5e-8


def test_c4_uf_product_vs_oracle(c4):
    w, S, _ = c4
    n = S.Nblobs
    r = S.r_blobs.cpu().numpy()
    assert np.abs(r - w["r"]).max() < 1e-12
    f = np.random.default_rng(44).standard_normal(3 * n)
    # blobs below z = a: the damping path is live at this size
    assert (r[:, 2] < w["a"]).any()
    print(f"case C4 UF product N={n}, {int((r[:, 2] < w['a']).sum())} blobs below a", flush=True)
    for prec, tol in (("fp32_blocked", 1e-6), ("fp64", 1e-12)):
        u, _ = wall_rpy.apply(RMW_UF, S.r_blobs, force=torch.from_numpy(f).to(S.device), eta=w["eta"], a=w["a"],
                              precision=prec)
        got = u.cpu().numpy()
        low = np.nonzero(r[:, 2] < w["a"])[0][:64]
        rows = [(0, 128), (n // 2, n // 2 + 128), (n - 128, n)] + [(int(i), int(i) + 1) for i in low]
        num = den = 0.0
        for t0, t1 in rows:
            ref = O.uf(r, f, w["eta"], w["a"], rows=(t0, t1)).reshape(n, 3)[t0:t1]
            num += float(np.sum((got[t0:t1] - ref) ** 2))
            den += float(np.sum(ref ** 2))
        e = float(np.sqrt(num / den))
        print(f"  {prec}: rel L2 vs oracle = {e:.3e}")
        assert e <= tol


def test_c4_point_field_vs_oracle(c4):
    w, S, ft = c4
    print("case C4 point field on sampled bodies", flush=True)
    u = S.point_field(ft, radius=0.9).cpu().numpy()
    rb = S.r_blobs.cpu().numpy()
    c = w["centres"]
    nb = c.shape[0]
    worst = 0.0
    for b in (0, 137, 511, 1023):
        others = np.arange(nb) != b
        ref = O.radii_apply(c[others], rb[42 * b: 42 * (b + 1)], ft[others, :3], np.full(nb - 1, 0.9),
                            np.full(42, w["a"]), w["eta"], wall=True).reshape(42, 3)
        worst = max(worst, float(np.abs(u[42 * b: 42 * (b + 1)] - ref).max() / np.abs(ref).max()))
    print(f"  max rel deviation {worst:.2e}")
    assert worst <= 1e-12


def test_c4_network_guess_is_finite(c4):
    from rigidmultiblobswall_amd.warmstart import WarmStart
    w, S, ft = c4
    print("case C4 network guess (untrained weights)", flush=True)
    rhs = S.rhs(np.zeros((S.Nblobs, 3)), ft)
    S.build_preconditioner()
    pc = torch.empty_like(rhs)
    S.pc_apply(rhs, pc)
    base = S.point_field_guess(rhs, ft, radius=0.9)
    ws = WarmStart(w["ref"], w["a"], w["eta"])
    ws.net.fscale = float(np.sqrt(np.mean(np.sum(ft ** 2, axis=1))))
    snap = Snapshot(w["centres"], w["quats"], ft, torch.zeros_like(rhs), pc, 0, 0.0, base)
    x0 = ws.predict_x0(snap)
    assert x0 is not None and x0.shape == rhs.shape and bool(torch.isfinite(x0).all())
