"""Blobs of different radii on the GPU against the FP64 oracle (source ->
target, wall, unbounded and free surface, overlaps with the wall, periodic
images, every precision mode), the equal-radii case against the
single-radius kernels, and the drop-in signatures."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, mobility, multi_bodies, wall_rpy
from rigidmultiblobswall_amd._lib import RMW_UF

pytestmark = pytest.mark.gpu

TOL = {"fp32_blocked": 1e-6, "fp32": 1e-6, "fp32_kahan": 1e-6, "fp64": 1e-12}


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _blobs(n, seed, side=12.0, zlo=0.05):
    rng = np.random.default_rng(seed)
    r = np.column_stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(zlo, 3.0, n)])
    return r, rng.uniform(0.1, 0.6, n), rng


BOUNDARIES = ["wall", "unbounded", "free_surface"]


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("bnd", BOUNDARIES)
def test_radii_self_interaction_matches_oracle(gpu, prec, bnd):
    n = 700
    r, rad, rng = _blobs(n, 1)
    f = rng.standard_normal((n, 3))
    print(f"case radii n={n} {bnd} {prec}", flush=True)
    u = wall_rpy.apply_radii(r, r, f, rad, rad, 0.9, boundary=bnd, precision=prec).cpu().numpy().ravel()
    ref = O.radii_apply(r, r, f, rad, rad, 0.9, boundary=bnd)
    assert rel(u, ref) <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_radii_source_target_sets_and_periodic(gpu, prec):
    src, rs, rng = _blobs(333, 2)
    trg, rt, _ = _blobs(517, 3)
    f = rng.standard_normal((333, 3))
    for L in (None, np.array([7.0, 6.0, 0.0]), np.array([7.0, 6.0, 9.0])):
        print(f"case radii source/target 333 -> 517 L={L} {prec}", flush=True)
        for bnd in BOUNDARIES:
            u = wall_rpy.apply_radii(src, trg, f, rs, rt, 1.2, periodic_length=L, boundary=bnd,
                                     precision=prec).cpu().numpy().ravel()
            ref = O.radii_apply(src, trg, f, rs, rt, 1.2, periodic_length=L, boundary=bnd)
            assert rel(u, ref) <= TOL[prec], (L, bnd)


def test_radii_work_split_and_edges(gpu):
    r, rad, rng = _blobs(1100, 4)
    f = rng.standard_normal((1100, 3))
    ref = O.radii_apply(r, r, f, rad, rad, 1.0)
    print("case radii chunks / grid caps / empty sources", flush=True)
    for chunk, wgs in ((64, 0), (192, 0), (1024, 0), (256, 5)):
        u = wall_rpy.apply_radii(r, r, f, rad, rad, 1.0, precision="fp64",
                                 tuning=_lib.Tuning(chunk, wgs, 0, 0, 0)).cpu().numpy().ravel()
        assert rel(u, ref) <= 1e-12, (chunk, wgs)
    # no sources: zero velocities; one blob: the self mobility with its image
    z = wall_rpy.apply_radii(np.zeros((0, 3)), r[:5], np.zeros((0, 3)), np.zeros(0), rad[:5], 1.0)
    assert torch.count_nonzero(z) == 0
    one = wall_rpy.apply_radii(r[:1], r[:1], f[:1], rad[:1], rad[:1], 1.0, precision="fp64").cpu().numpy().ravel()
    assert rel(one, O.radii_apply(r[:1], r[:1], f[:1], rad[:1], rad[:1], 1.0)) <= 1e-13


def test_equal_radii_matches_single_radius_kernel(gpu):
    r, _, rng = _blobs(900, 5)
    a = 0.3
    f = rng.standard_normal((900, 3))
    rad = np.full(900, a)
    print("case radii equal radii vs UF kernel", flush=True)
    u_r = wall_rpy.apply_radii(r, r, f, rad, rad, 1.0, precision="fp64").cpu().numpy().ravel()
    u_uf, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=a, precision="fp64")
    assert rel(u_r, u_uf.cpu().numpy().ravel()) <= 1e-12
    u_r = wall_rpy.apply_radii(r, r, f, rad, rad, 1.0, boundary="free_surface", precision="fp64")
    u_fs, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1.0, a=a, precision="fp64", domain="free_surface")
    assert rel(u_r.cpu().numpy().ravel(), u_fs.cpu().numpy().ravel()) <= 1e-12


def test_radii_dropin_signatures(gpu):
    r, rad, rng = _blobs(300, 6)
    f = rng.standard_normal(900)
    ref = O.radii_apply(r, r, f, rad, rad, 0.7)
    u = mobility.single_wall_mobility_trans_times_force_source_target_hip(r, r, f, rad, rad, 0.7, precision="fp64")
    assert isinstance(u, np.ndarray) and u.shape == (900,) and rel(u, ref) <= 1e-12
    fn = multi_bodies.set_mobility_vector_prod("radii_hip_fp64", radius_blobs=rad)
    assert rel(fn(r, f, 0.7, 0.0), ref) <= 1e-12
    fn0 = multi_bodies.set_mobility_vector_prod("radii_hip_no_wall", radius_blobs=rad)
    assert rel(fn0(r, f, 0.7, 0.0), O.radii_apply(r, r, f, rad, rad, 0.7, wall=False)) <= 1e-6
    ref_fs = O.radii_apply(r, r, f, rad, rad, 0.7, boundary="free_surface")
    u = mobility.free_surface_mobility_trans_times_force_source_target_hip(r, r, f, rad, rad, 0.7, precision="fp64")
    assert isinstance(u, np.ndarray) and u.shape == (900,) and rel(u, ref_fs) <= 1e-12
    for name, tol in (("radii_hip_free_surface", 1e-6), ("radii_hip_free_surface_fp64", 1e-12)):
        fn = multi_bodies.set_mobility_vector_prod(name, radius_blobs=rad)
        assert rel(fn(r, f, 0.7, 0.0), ref_fs) <= tol, name
    # an unknown boundary is refused by the library itself
    lib = _lib.load()
    with pytest.raises(ValueError):
        wall_rpy.apply_radii(r, r, f, rad, rad, 0.7, boundary="two_walls")
    assert lib.rmw_radii_apply(0, 3, None, None, None, None, None, None, 0, 0, 1.0, None, None, None, 0,
                               None) == -1  # RMW_E_ARG


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("bnd", BOUNDARIES)
def test_radii_symmetric_kernel_matches_oracle(gpu, prec, bnd):
    """Sources = targets through the symmetric kernel (forced on at this
    size): 1100 blobs = 4 full tiles and a ragged one, radii 0.1-0.6 so the
    far, overlapping and engulfed regimes all occur, blobs below their
    radius (wall damping)."""
    n = 1100
    r, rad, rng = _blobs(n, 11, side=8.0)
    f = rng.standard_normal((n, 3))
    ref = O.radii_apply(r, r, f, rad, rad, 0.9, boundary=bnd)
    print(f"case radii symmetric n={n} {bnd} {prec}", flush=True)
    for chunk in (0, 512):
        t = _lib.Tuning(chunk, 0, 0, 1, 0)
        assert wall_rpy.describe_radii_self(n, prec, tuning=t)["kernel"] == "symmetric"
        u = wall_rpy.apply_radii_self(r, f, rad, 0.9, boundary=bnd, precision=prec, tuning=t).cpu().numpy().ravel()
        assert rel(u, ref) <= TOL[prec], chunk
    # the ordered kernel through the same entry point
    t = _lib.Tuning(0, 0, 0, -1, 0)
    assert wall_rpy.describe_radii_self(n, prec, tuning=t)["kernel"] == "ordered"
    u = wall_rpy.apply_radii_self(r, f, rad, 0.9, boundary=bnd, precision=prec, tuning=t).cpu().numpy().ravel()
    assert rel(u, ref) <= TOL[prec]


def test_radii_symmetric_edges_and_choice(gpu):
    r, rad, rng = _blobs(300, 12)
    f = rng.standard_normal((300, 3))
    print("case radii symmetric edges: grid caps, tiny n, periodic, default choice", flush=True)
    ref = O.radii_apply(r, r, f, rad, rad, 1.0)
    for wgs in (1, 3):
        u = wall_rpy.apply_radii_self(r, f, rad, 1.0, precision="fp64", tuning=_lib.Tuning(0, wgs, 0, 1, 0))
        assert rel(u.cpu().numpy().ravel(), ref) <= 1e-12, wgs
    for n in (1, 2, 65):
        u = wall_rpy.apply_radii_self(r[:n], f[:n], rad[:n], 1.0, precision="fp64", tuning=_lib.Tuning(0, 0, 0, 1, 0))
        assert rel(u.cpu().numpy().ravel(), O.radii_apply(r[:n], r[:n], f[:n], rad[:n], rad[:n], 1.0)) <= 1e-12, n
    # periodic boxes take the ordered kernel; forcing the symmetric one is an error
    L = np.array([7.0, 6.0, 0.0])
    assert wall_rpy.describe_radii_self(300, "fp64", periodic_length=L)["kernel"] == "ordered"
    u = wall_rpy.apply_radii_self(r, f, rad, 1.0, periodic_length=L, precision="fp64")
    assert rel(u.cpu().numpy().ravel(), O.radii_apply(r, r, f, rad, rad, 1.0, periodic_length=L)) <= 1e-12
    with pytest.raises(RuntimeError):
        wall_rpy.apply_radii_self(r, f, rad, 1.0, periodic_length=L, tuning=_lib.Tuning(0, 0, 0, 1, 0))
    # default: ordered below one tile pair per CU, symmetric from there on
    assert wall_rpy.describe_radii_self(300)["kernel"] == "ordered"
    assert wall_rpy.describe_radii_self(32768)["kernel"] == "symmetric"
    # the drop-in routes sources = targets through this entry point
    fn = multi_bodies.set_mobility_vector_prod("radii_hip_fp64", radius_blobs=rad)
    assert rel(fn(r, f.ravel(), 1.0, 0.0), ref) <= 1e-12


def test_radii_symmetric_large_matches_ordered(gpu):
    """At the benchmark's size (32768 blobs): symmetric and ordered kernels
    agree to the FP32 tolerance and the FP64 symmetric product is
    reproducible bit for bit."""
    from rigidmultiblobswall_amd import workloads
    w = workloads.c3_rollers()
    r = w["r"]
    n = r.shape[0]
    rad = np.random.default_rng(33).uniform(0.5, 0.8, n)
    f = np.random.default_rng(13).standard_normal((n, 3))
    print("case radii symmetric vs ordered, 32768 blobs", flush=True)
    us = wall_rpy.apply_radii_self(r, f, rad, 1e-3, tuning=_lib.Tuning(0, 0, 0, 1, 0)).cpu().numpy().ravel()
    uo = wall_rpy.apply_radii_self(r, f, rad, 1e-3, tuning=_lib.Tuning(0, 0, 0, -1, 0)).cpu().numpy().ravel()
    assert rel(us, uo) <= 1e-6
    a1 = wall_rpy.apply_radii_self(r, f, rad, 1e-3, precision="fp64", tuning=_lib.Tuning(0, 0, 0, 1, 0))
    a2 = wall_rpy.apply_radii_self(r, f, rad, 1e-3, precision="fp64", tuning=_lib.Tuning(0, 0, 0, 1, 0))
    assert torch.equal(a1, a2)
    assert rel(a1.cpu().numpy().ravel(), us) <= 1e-6
