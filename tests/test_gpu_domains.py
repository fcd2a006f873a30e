"""The no-wall, in-plane and free-surface products on the GPU against the
oracle (every block combination the library serves for each, both
precision families, periodic images, blobs overlapping the wall), several
right-hand sides, the drop-in signatures, the rollers integrator in the
other domains, and the argument checks."""
import numpy as np
import pytest

from oracle import oracle as O
from oracle import rollers_oracle as R
from rigidmultiblobswall_amd import mobility, multi_bodies, wall_rpy
from rigidmultiblobswall_amd._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

pytestmark = pytest.mark.gpu

TOL = {"fp32_blocked": 1e-6, "fp32": 1e-6, "fp64": 1e-12}
BLOCKS = {"no_wall": (O.UF_NOWALL, O.UT_NOWALL, O.WF_NOWALL, O.WT_NOWALL),
          "in_plane": (O.UF_INPLANE, O.UT_INPLANE, None, None),
          "free_surface": (O.UF_FREE_SURFACE, None, None, None)}
MASKS = {"no_wall": (1, 2, 4, 8, 3, 12, 5, 10, 15), "in_plane": (1, 2, 3), "free_surface": (1,)}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _ref(domain, mask, r, f, t, eta, a, L):
    uf, ut, wf, wt = BLOCKS[domain]
    u = np.zeros(r.size)
    w = np.zeros(r.size)
    if mask & RMW_UF:
        u += O.block_apply(uf, r, f, eta, a, L)
    if mask & RMW_UT:
        u += O.block_apply(ut, r, t, eta, a, L)
    if mask & RMW_WF:
        w += O.block_apply(wf, r, f, eta, a, L)
    if mask & RMW_WT:
        w += O.block_apply(wt, r, t, eta, a, L)
    return u, w


@pytest.mark.parametrize("domain", ["no_wall", "in_plane", "free_surface"])
@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_domain_products_match_oracle(gpu, domain, prec):
    rng = np.random.default_rng(1)
    n, a, eta = 640, 0.3, 0.8
    r = np.column_stack([rng.uniform(0, 12, n), rng.uniform(0, 12, n), rng.uniform(0.1, 3.0, n)])
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    for L in (None, np.array([7.0, 6.0, 0.0])):
        for mask in MASKS[domain]:
            print(f"case domain {domain} mask {mask} {prec} L={L}", flush=True)
            u, w = wall_rpy.apply(mask, r, force=f, torque=t, eta=eta, a=a, periodic_length=L, precision=prec,
                                  domain=domain)
            ur, wr = _ref(domain, mask, r, f, t, eta, a, L)
            if mask & (RMW_UF | RMW_UT):
                assert rel(u.cpu().numpy().ravel(), ur) <= TOL[prec], (mask, L)
            if mask & (RMW_WF | RMW_WT):
                assert rel(w.cpu().numpy().ravel(), wr) <= TOL[prec], (mask, L)


def test_several_right_hand_sides(gpu):
    rng = np.random.default_rng(2)
    n, k = 500, 3
    r = np.column_stack([rng.uniform(0, 10, n), rng.uniform(0, 10, n), rng.uniform(0.5, 3.0, n)])
    f, t = rng.standard_normal((k, n, 3)), rng.standard_normal((k, n, 3))
    print("case domain several rhs", flush=True)
    for domain in ("no_wall", "in_plane"):
        u, _ = wall_rpy.apply(RMW_UF | RMW_UT, r, force=f, torque=t, eta=1.0, a=0.3, precision="fp64", k=k,
                              domain=domain)
        for q in range(k):
            ur, _ = _ref(domain, 3, r, f[q], t[q], 1.0, 0.3, None)
            assert rel(u[q].cpu().numpy().ravel(), ur) <= 1e-12, (domain, q)


def test_in_plane_outputs_no_normal_velocity(gpu):
    rng = np.random.default_rng(3)
    n = 300
    r = np.column_stack([rng.uniform(0, 8, n), rng.uniform(0, 8, n), rng.uniform(0.1, 2.0, n)])
    u, _ = wall_rpy.apply(RMW_UF | RMW_UT, r, force=rng.standard_normal((n, 3)), torque=rng.standard_normal((n, 3)),
                          eta=1.0, a=0.3, domain="in_plane")
    assert not u[:, 2].any()


def test_domain_dropins(gpu):
    rng = np.random.default_rng(4)
    n, a, eta = 260, 0.25, 1.3
    r = np.column_stack([rng.uniform(0, 7, n), rng.uniform(0, 7, n), rng.uniform(0.2, 2.0, n)])
    f, t = rng.standard_normal(3 * n), rng.standard_normal(3 * n)
    L = np.array([5.0, 0.0, 0.0])
    kw = dict(periodic_length=L, precision="fp64")
    print("case domain drop-ins", flush=True)
    cases = [
        (mobility.no_wall_mobility_trans_times_force_hip(r, f, eta, a, **kw), O.block_apply(O.UF_NOWALL, r, f, eta, a, L)),
        (mobility.no_wall_mobility_trans_times_torque_hip(r, t, eta, a, **kw), O.block_apply(O.UT_NOWALL, r, t, eta, a, L)),
        (mobility.no_wall_mobility_rot_times_force_hip(r, f, eta, a, **kw), O.block_apply(O.WF_NOWALL, r, f, eta, a, L)),
        (mobility.no_wall_mobility_rot_times_torque_hip(r, t, eta, a, **kw), O.block_apply(O.WT_NOWALL, r, t, eta, a, L)),
        (mobility.no_wall_mobility_trans_times_force_torque_hip(r, f, t, eta, a, **kw),
         O.block_apply(O.UF_NOWALL, r, f, eta, a, L) + O.block_apply(O.UT_NOWALL, r, t, eta, a, L)),
        (mobility.in_plane_mobility_trans_times_force_hip(r, f, eta, a, **kw), O.block_apply(O.UF_INPLANE, r, f, eta, a, L)),
        (mobility.in_plane_mobility_trans_times_torque_hip(r, t, eta, a, **kw), O.block_apply(O.UT_INPLANE, r, t, eta, a, L)),
        (mobility.free_surface_mobility_trans_times_force_hip(r, f, eta, a, **kw),
         O.block_apply(O.UF_FREE_SURFACE, r, f, eta, a, L)),
    ]
    for got, ref in cases:
        assert isinstance(got, np.ndarray) and got.shape == (3 * n,) and rel(got, ref) <= 1e-12
    fn = multi_bodies.set_mobility_vector_prod("hip_no_wall_fp64")
    assert rel(fn(r, f, eta, a), O.block_apply(O.UF_NOWALL, r, f, eta, a)) <= 1e-12
    fn = multi_bodies.set_mobility_vector_prod("hip_free_surface")
    assert rel(fn(r, f, eta, a), O.block_apply(O.UF_FREE_SURFACE, r, f, eta, a)) <= 1e-6
    blocks = multi_bodies.set_rollers_blocks("hip_fp64", domain="in_plane")
    assert blocks[2] is None and rel(blocks[0](r, f, eta, a), O.block_apply(O.UF_INPLANE, r, f, eta, a)) <= 1e-12


def test_domain_argument_checks(gpu):
    r = np.random.default_rng(5).uniform(0.5, 2.0, (20, 3))
    v = np.ones((20, 3))
    with pytest.raises(RuntimeError, match="in-plane"):
        wall_rpy.apply(RMW_WF, r, force=v, domain="in_plane")
    with pytest.raises(RuntimeError, match="free-surface"):
        wall_rpy.apply(RMW_UT, r, torque=v, domain="free_surface")
    with pytest.raises(ValueError, match="domain"):
        wall_rpy.apply(RMW_UF, r, force=v, domain="two_walls")


@pytest.mark.parametrize("domain,free", [("no_wall", "True"), ("no_wall", "False"), ("in_plane", "True")])
def test_rollers_in_other_domains(gpu, domain, free):
    from rigidmultiblobswall_amd.rollers_integrator import RollersIntegrator
    rng = np.random.default_rng(6)
    n, steps = 120, 2
    r0 = np.column_stack([rng.uniform(0, 25, n), rng.uniform(0, 25, n), rng.uniform(0.8, 3.0, n)])
    zs, dxs = rng.standard_normal((steps, 3 * n)), rng.standard_normal((steps, n, 3))
    it_z, it_dx = iter(zs), iter(dxs)
    src = lambda kind, shape: next(it_z) if kind == "z" else next(it_dx)  # noqa: E731
    kw = dict(eta=1e-3, a=0.656, kT=0.0041419464, g=0.0024892)
    g = RollersIntegrator(r0, scheme="stochastic_adams_bashforth_rollers", repulsion_strength_wall=0.0165677856,
                          debye_length_wall=0.0656, omega_one_roller=(0.0, 62.8, 0.0), free_kinematics=free,
                          tolerance=1e-10, precision="fp64", domain=domain, noise_source=src, **kw)
    o = R.RollersStep(r0, "stochastic_adams_bashforth_rollers", eps_wall=0.0165677856, b_wall=0.0656,
                      blob_blob=False, omega=(0.0, 62.8, 0.0), free_kinematics=free, tolerance=1e-10,
                      domain=domain, **kw)
    draws = iter(zip(zs, dxs))
    print(f"case rollers domain {domain} free={free}", flush=True)
    for _ in range(steps):
        g.advance_time_step(0.016)
        o.advance(0.016, draws)
    assert np.linalg.norm(g.locations() - o.r) <= 1e-8 * np.linalg.norm(o.r - r0)


@pytest.mark.parametrize("domain,mask", [("no_wall", 1), ("no_wall", 15), ("no_wall", 3), ("no_wall", 2),
                                         ("no_wall", 4), ("no_wall", 8), ("in_plane", 1), ("in_plane", 3),
                                         ("in_plane", 2), ("free_surface", 1)])
@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_symmetric_kernel_in_other_domains(gpu, domain, mask, prec):
    """The symmetric kernel (each unordered pair once) serves the UF block of
    every boundary, UF + UT and the single UT, WF and WT blocks where the
    boundary has them (unbounded; in plane: UT) and the unbounded grand
    mobility."""
    from rigidmultiblobswall_amd import _lib
    rng = np.random.default_rng(7)
    n, a, eta = 777, 0.3, 0.8
    r = np.column_stack([rng.uniform(0, 12, n), rng.uniform(0, 12, n), rng.uniform(0.1, 3.0, n)])
    f, t = rng.standard_normal((n, 3)), rng.standard_normal((n, 3))
    print(f"case symmetric domain {domain} mask {mask} {prec}", flush=True)
    for chunk in (256, 768):
        u, w = wall_rpy.apply(mask, r, force=f, torque=t, eta=eta, a=a, precision=prec, domain=domain,
                              tuning=_lib.Tuning(chunk, 0, 0, 1, 0))
        ur, wr = _ref(domain, mask, r, f, t, eta, a, None)
        if mask & (RMW_UF | RMW_UT):
            assert rel(u.cpu().numpy().ravel(), ur) <= TOL[prec], chunk
        if mask & (RMW_WF | RMW_WT):
            assert rel(w.cpu().numpy().ravel(), wr) <= TOL[prec], chunk
