"""Blob-blob soft repulsion on the GPU against the FP64 oracle (open and
periodic boxes, one radius and different radii, the neighbour-list cutoff,
both precision modes), edge cases, Newton's third law at the rollers size,
and the drop-in signatures."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from rigidmultiblobswall_amd import _lib, forces

pytestmark = pytest.mark.gpu

TOL = {"fp32": 1e-6, "fp64": 1e-12}
ROLLERS = dict(repulsion_strength=0.0165677856, debye_length=0.0656, blob_radius=0.656)


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _cloud(n, seed, side=14.0):
    rng = np.random.default_rng(seed)
    r = np.column_stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(0.5, 3.0, n)])
    return r, rng


def _sym(s):
    return _lib.Tuning(0, 0, 0, s, 0)


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
@pytest.mark.parametrize("L", [None, (9.0, 8.0, 0.0), (9.0, 8.0, 3.5)])
@pytest.mark.parametrize("sym", [1, -1])
def test_forces_match_oracle(gpu, prec, L, sym):
    n = 1500                                   # 6 tiles, the last one partial
    r, _ = _cloud(n, 1)
    eps, b, a = 0.3, 0.08, 0.3
    print(f"case forces n={n} L={L} {prec} symmetric={sym}", flush=True)
    F = forces.blob_forces(r, eps, b, a, periodic_length=L, precision=prec, tuning=_sym(sym)).cpu().numpy()
    ref = O.blob_forces(r, eps, b, a, periodic_length=L)
    assert F.shape == (n, 3)
    assert rel(F, ref) <= TOL[prec]


def test_symmetric_kernel_choice_and_runs(gpu):
    """Automatic choice: symmetric once the tile pairs fill the device;
    tile-run lengths from 1 to beyond the tile count; full and ragged last
    tiles."""
    assert forces.describe(32768)["kernel"] == "symmetric"
    assert forces.describe(300)["kernel"] == "ordered"
    assert forces.describe(32768, tuning=_sym(-1))["kernel"] == "ordered"
    for n in (512, 777):
        r, _ = _cloud(n, 6, side=7.0)
        ref = O.blob_forces(r, 0.3, 0.1, 0.25)
        print(f"case forces symmetric runs n={n}", flush=True)
        for chunk in (256, 512, 768, 4096):
            F = forces.blob_forces(r, 0.3, 0.1, 0.25, precision="fp64",
                                   tuning=_lib.Tuning(chunk, 0, 0, 1, 0)).cpu().numpy()
            assert rel(F, ref) <= 1e-12, (n, chunk)
        F = forces.blob_forces(r, 0.3, 0.1, 0.25, precision="fp64", tuning=_lib.Tuning(256, 3, 0, 1, 0))
        assert rel(F.cpu().numpy(), ref) <= 1e-12


@pytest.mark.parametrize("sym", [1, -1])
def test_fp32_overlap_branch_is_exact(gpu, sym):
    """Inside the overlap the FP32-exponential mode takes 2^0 = 1: the force
    magnitude is eps/b to FP64 rounding."""
    eps, b, a = 0.4, 0.05, 0.5
    r = np.array([[0.0, 0.0, 1.0], [0.3, 0.4, 1.0]])
    F = forces.blob_forces(r, eps, b, a, tuning=_sym(sym)).cpu().numpy()
    assert abs(np.linalg.norm(F[0]) - eps / b) <= 1e-14 * eps / b
    assert np.array_equal(F[0], -F[1])


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_forces_radii_and_cutoff(gpu, prec):
    n = 1100
    r, rng = _cloud(n, 2)
    rad = rng.uniform(0.1, 0.5, n)
    print(f"case forces radii / cutoff n={n} {prec}", flush=True)
    for L in (None, (9.0, 8.0, 0.0)):
        F = forces.blob_forces(r, 0.2, 0.1, periodic_length=L, radius_blobs=rad, precision=prec).cpu().numpy()
        assert rel(F, O.blob_forces(r, 0.2, 0.1, 0.0, periodic_length=L, radii=rad)) <= TOL[prec], L
    d_max = 2 * 0.3 + 30 * 0.05
    Fc = forces.blob_forces(r, 0.2, 0.05, 0.3, cutoff=d_max, precision=prec).cpu().numpy()
    assert rel(Fc, O.blob_forces(r, 0.2, 0.05, 0.3, cutoff=d_max)) <= TOL[prec]
    # a tight cutoff really drops pairs
    Ft = forces.blob_forces(r, 0.2, 0.5, 0.3, cutoff=0.9, precision=prec).cpu().numpy()
    ref_t = O.blob_forces(r, 0.2, 0.5, 0.3, cutoff=0.9)
    assert rel(Ft, ref_t) <= TOL[prec]
    assert rel(ref_t, O.blob_forces(r, 0.2, 0.5, 0.3)) > 1e-3


def test_forces_work_split_and_edges(gpu):
    r, _ = _cloud(1300, 3, side=10.0)
    ref = O.blob_forces(r, 0.3, 0.1, 0.25)
    print("case forces chunks / grid caps / edge cases", flush=True)
    for chunk, wgs in ((64, 0), (192, 0), (2048, 0), (256, 5)):
        F = forces.blob_forces(r, 0.3, 0.1, 0.25, precision="fp64",
                               tuning=_lib.Tuning(chunk, wgs, 0, 0, 0)).cpu().numpy()
        assert rel(F, ref) <= 1e-12, (chunk, wgs)
    # no blobs, one blob, coincident blobs, zero strength
    assert forces.blob_forces(np.zeros((0, 3)), 0.3, 0.1, 0.25).numel() == 0
    assert torch.count_nonzero(forces.blob_forces(r[:1], 0.3, 0.1, 0.25)) == 0
    two = np.array([[1.0, 2.0, 3.0], [1.0, 2.0, 3.0], [1.5, 2.0, 3.0]])
    F2 = forces.blob_forces(two, 0.3, 0.1, 0.25, precision="fp64").cpu().numpy()
    assert rel(F2, O.blob_forces(two, 0.3, 0.1, 0.25)) <= 1e-14 and np.isfinite(F2).all()
    assert torch.count_nonzero(forces.blob_forces(r, 0.0, 0.1, 0.25)) == 0
    # 2a/b far beyond the FP32 exponent range: overlapping pairs stay finite
    Fo = forces.blob_forces(r, 0.3, 0.002, 0.6).cpu().numpy()
    assert np.isfinite(Fo).all()
    assert rel(Fo, O.blob_forces(r, 0.3, 0.002, 0.6)) <= 1e-6


def test_forces_argument_errors(gpu):
    r, _ = _cloud(10, 4)
    with pytest.raises(RuntimeError, match="debye_length"):
        forces.blob_forces(r, 0.3, 0.0, 0.25)
    with pytest.raises(RuntimeError, match="cutoff"):
        forces.blob_forces(r, 0.3, 0.1, 0.25, cutoff=-1.0)
    with pytest.raises(ValueError, match="radius_blobs"):
        forces.blob_forces(r, 0.3, 0.1, radius_blobs=np.ones(9))


@pytest.mark.parametrize("prec", ["fp32", "fp64"])
def test_rollers_size_third_law_and_sampled_rows(gpu, prec):
    """32768 rollers with the reference's rollers parameters: the forces sum
    to zero and sampled rows match the oracle."""
    from rigidmultiblobswall_amd.workloads import c3_rollers
    r = c3_rollers()["r"]
    n = r.shape[0]
    print(f"case forces rollers n={n} {prec}", flush=True)
    F = forces.blob_forces(r, precision=prec, **ROLLERS).cpu().numpy()
    assert np.abs(F.sum(axis=0)).max() <= 1e-12 * np.abs(F).sum(axis=0).max()
    for rows in ((0, 300), (n - 300, n), (16000, 16300)):
        ref = O.blob_forces(r, ROLLERS["repulsion_strength"], ROLLERS["debye_length"], ROLLERS["blob_radius"],
                            rows=rows)[rows[0]:rows[1]]
        assert rel(F[rows[0]:rows[1]], ref) <= TOL[prec], rows


class _Body:
    def __init__(self, radii):
        self.blobs_radius = radii


def test_dropins_and_dispatch(gpu):
    r, rng = _cloud(400, 5, side=8.0)
    kw = dict(periodic_length=np.array([6.0, 7.0, 0.0]), repulsion_strength=0.3, debye_length=0.1,
              blob_radius=0.25)
    ref = O.blob_forces(r, 0.3, 0.1, 0.25, periodic_length=kw["periodic_length"])
    F = forces.calc_blob_blob_forces_hip(r, **kw)
    assert isinstance(F, np.ndarray) and F.shape == (400, 3) and rel(F, ref) <= 1e-6
    assert rel(forces.set_blob_blob_forces("hip_fp64")(r.ravel(), **kw), ref) <= 1e-12
    d_max = 2 * 0.25 + 30 * 0.1
    ref_c = O.blob_forces(r, 0.3, 0.1, 0.25, periodic_length=kw["periodic_length"], cutoff=d_max)
    assert rel(forces.set_blob_blob_forces("tree_hip_fp64")(r, **kw), ref_c) <= 1e-12
    rad = rng.uniform(0.1, 0.4, 400)
    fn = forces.set_blob_blob_forces("radii_hip_fp64", bodies=[_Body(rad[:150]), _Body(rad[150:])])
    ref_r = O.blob_forces(r, 0.3, 0.1, 0.0, periodic_length=kw["periodic_length"], radii=rad)
    assert rel(fn(r, **kw), ref_r) <= 1e-12
