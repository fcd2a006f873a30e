"""Blobs of different radii on the GPU against the FP64 oracle (source ->
target, wall and unbounded, overlaps with the wall, periodic images, every
precision mode), the equal-radii case against the single-radius kernels,
and the drop-in signatures."""
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


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp32", "fp32_kahan", "fp64"])
@pytest.mark.parametrize("wall", [True, False])
def test_radii_self_interaction_matches_oracle(gpu, prec, wall):
    n = 700
    r, rad, rng = _blobs(n, 1)
    f = rng.standard_normal((n, 3))
    print(f"case radii n={n} wall={wall} {prec}", flush=True)
    u = wall_rpy.apply_radii(r, r, f, rad, rad, 0.9, wall=wall, precision=prec).cpu().numpy().ravel()
    ref = O.radii_apply(r, r, f, rad, rad, 0.9, wall=wall)
    assert rel(u, ref) <= TOL[prec]


@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_radii_source_target_sets_and_periodic(gpu, prec):
    src, rs, rng = _blobs(333, 2)
    trg, rt, _ = _blobs(517, 3)
    f = rng.standard_normal((333, 3))
    for L in (None, np.array([7.0, 6.0, 0.0])):
        print(f"case radii source/target 333 -> 517 L={L} {prec}", flush=True)
        for wall in (True, False):
            u = wall_rpy.apply_radii(src, trg, f, rs, rt, 1.2, periodic_length=L, wall=wall,
                                     precision=prec).cpu().numpy().ravel()
            ref = O.radii_apply(src, trg, f, rs, rt, 1.2, periodic_length=L, wall=wall)
            assert rel(u, ref) <= TOL[prec], (L, wall)


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
