"""Pins of the oracle for blobs of different radii (source -> target wall
RPY, mobility_numba.py:1479-1658 with the wrapper mobility.py:551-591).

* equal radii: the independent single-radius UF oracle (a4, incl. overlaps,
  damping and periodic images) and its unbounded form;
* reciprocity M_ij(a, b) = M_ji(b, a)^T in every RPY regime;
* the point-particle limit a, b -> 0: the Blake tensor written from the
  mathematics;
* continuity of the RPY regimes at r = a + b and r = |a - b|;
* positive definiteness of the dense matrix.

The free-surface form (mobility_numba.py:1941-2093 behind
mobility.py:1409-1429) is pinned by the method of images: it equals the
unbounded product over the sources and their mirror images (x, y, -z)
carrying D f, D = diag(1, 1, -1); by the equal-radii free-surface oracle;
by reciprocity and by positive definiteness.
"""
import numpy as np
import pytest

from oracle import oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _config(n, seed, zlo=0.1):
    rng = np.random.default_rng(seed)
    return np.column_stack([rng.uniform(0, 4, n), rng.uniform(0, 4, n), rng.uniform(zlo, 2.0, n)]), rng


@pytest.mark.parametrize("L", [None, (5.0, 4.0, 0.0)])
def test_equal_radii_is_the_single_radius_oracle(L):
    r, rng = _config(60, 1)                      # heights from 0.1: overlaps and damping live
    f = rng.standard_normal((60, 3))
    a = 0.3
    rad = np.full(60, a)
    u = O.radii_apply(r, r, f, rad, rad, 1.1, periodic_length=L)
    ref = O.block_apply(O.UF, r, f, 1.1, a, periodic_length=L)
    assert rel(u, ref) < 1e-13
    u0 = O.radii_apply(r, r, f, rad, rad, 1.1, periodic_length=L, wall=False)
    ref0 = O.block_apply(O.UF_NOWALL, r, f, 1.1, a, periodic_length=L)
    assert rel(u0, ref0) < 1e-13


def test_reciprocity_all_regimes():
    rng = np.random.default_rng(2)
    worst = 0.0
    regimes = set()
    for t in range(3000):
        ri = np.array([*rng.uniform(-1.5, 1.5, 2), rng.uniform(0.5, 2.5)])
        rj = np.array([*rng.uniform(-1.5, 1.5, 2), rng.uniform(0.5, 2.5)])
        a, b = rng.uniform(0.05, 0.8, 2)
        d = np.linalg.norm(ri - rj)
        regimes.add(0 if d > a + b else (1 if d > abs(a - b) else 2))
        for bnd in ("wall", "unbounded", "free_surface"):
            A = O.radii_pair_matrix(ri, rj, a, b, 1.0, boundary=bnd)
            B = O.radii_pair_matrix(rj, ri, b, a, 1.0, boundary=bnd)
            worst = max(worst, np.abs(A - B.T).max() / np.abs(A).max())
    assert regimes == {0, 1, 2}
    assert worst < 1e-13


def test_point_particle_limit_is_blake():
    rng = np.random.default_rng(3)
    for t in range(50):
        x = np.array([*rng.uniform(-2, 2, 2), rng.uniform(0.5, 2.5)])
        y = np.array([*rng.uniform(-2, 2, 2), rng.uniform(0.5, 2.5)])
        if np.linalg.norm(x - y) < 0.3:
            continue
        M = O.radii_pair_matrix(x, y, 1e-5, 1e-5, 1.0)
        G = O.blake_tensor(x, y, 1.0)
        assert np.abs(M - G).max() <= 1e-7 * np.abs(G).max()
        # the finite-size terms are O(a^2): visible at a = 0.1
        M1 = O.radii_pair_matrix(x, y, 0.1, 0.1, 1.0)
        assert np.abs(M1 - G).max() > 1e-6 * np.abs(G).max()


def test_regime_boundaries_are_continuous():
    rng = np.random.default_rng(4)
    for t in range(200):
        a, b = rng.uniform(0.1, 0.8, 2)
        e = rng.standard_normal(3)
        e /= np.linalg.norm(e)
        y = np.array([0.0, 0.0, 3.0])
        for rb in (a + b, abs(a - b)):
            if rb < 1e-3:
                continue
            lo = O.radii_pair_matrix(y + e * rb * (1 - 1e-9), y, a, b, 1.0, False)
            hi = O.radii_pair_matrix(y + e * rb * (1 + 1e-9), y, a, b, 1.0, False)
            assert np.abs(lo - hi).max() <= 1e-7 * np.abs(hi).max()


def test_dense_matrix_is_spd():
    r, rng = _config(40, 5, zlo=0.9)
    rad = rng.uniform(0.1, 0.45, 40)
    n = 40
    for bnd in ("wall", "unbounded", "free_surface"):
        M = np.zeros((3 * n, 3 * n))
        for j in range(n):
            for c in range(3):
                e = np.zeros((n, 3))
                e[j, c] = 1.0
                M[:, 3 * j + c] = O.radii_apply(r, r, e, rad, rad, 1.0, boundary=bnd)
        assert np.abs(M - M.T).max() <= 1e-13 * np.abs(M).max()
        assert np.linalg.eigvalsh(0.5 * (M + M.T)).min() > 0


def test_source_target_sets_and_rows():
    rng = np.random.default_rng(6)
    src = np.column_stack([rng.uniform(0, 3, 30), rng.uniform(0, 3, 30), rng.uniform(0.2, 2, 30)])
    trg = np.column_stack([rng.uniform(0, 3, 17), rng.uniform(0, 3, 17), rng.uniform(0.2, 2, 17)])
    rs, rt = rng.uniform(0.1, 0.4, 30), rng.uniform(0.1, 0.4, 17)
    f = rng.standard_normal((30, 3))
    u = O.radii_apply(src, trg, f, rs, rt, 0.8)
    # brute force from the pair blocks with the wrapper (shift + damping)
    ref = np.zeros((17, 3))
    for i in range(17):
        zi = max(trg[i, 2], rt[i])
        Bt = trg[i, 2] / rt[i] if trg[i, 2] < rt[i] else 1.0
        for j in range(30):
            zj = max(src[j, 2], rs[j])
            Bs = src[j, 2] / rs[j] if src[j, 2] < rs[j] else 1.0
            M = O.radii_pair_matrix([*trg[i, :2], zi], [*src[j, :2], zj], rt[i], rs[j], 0.8)
            ref[i] += Bt * (M @ (Bs * f[j]))
    assert rel(u, ref.ravel()) < 1e-14
    part = O.radii_apply(src, trg, f, rs, rt, 0.8, rows=(5, 9)).reshape(17, 3)
    assert np.allclose(part[5:9], u.reshape(17, 3)[5:9], rtol=0, atol=0)
    assert not part[:5].any() and not part[9:].any()


def _mirror_failures(kind="strict"):
    """Free surface = unbounded product over sources + mirrored sources."""
    fails = []
    rng = np.random.default_rng(8)
    for L in (None, (5.0, 4.0, 0.0)):
        src = np.column_stack([rng.uniform(0, 4, 40), rng.uniform(0, 4, 40), rng.uniform(0.05, 1.5, 40)])
        trg = np.column_stack([rng.uniform(0, 4, 25), rng.uniform(0, 4, 25), rng.uniform(0.05, 1.5, 25)])
        rs, rt = rng.uniform(0.05, 0.6, 40), rng.uniform(0.05, 0.6, 25)
        f = rng.standard_normal((40, 3))
        u = O.radii_apply(src, trg, f, rs, rt, 0.7, periodic_length=L, boundary="free_surface", kind=kind)
        img = src * [1.0, 1.0, -1.0]
        ref = O.radii_apply(np.vstack([src, img]), trg, np.vstack([f, f * [1.0, 1.0, -1.0]]),
                            np.concatenate([rs, rs]), rt, 0.7, periodic_length=L, wall=False, kind=kind)
        if rel(u, ref) > 1e-13:
            fails.append(f"mirror L={L}")
    return fails


def test_free_surface_is_the_mirrored_unbounded_product():
    assert _mirror_failures() == []


def test_free_surface_planted_mistake_is_caught():
    O.set_mutation(9)
    try:
        assert _mirror_failures("mutant")
    finally:
        O.set_mutation(0)
    assert _mirror_failures("mutant") == []


@pytest.mark.parametrize("L", [None, (5.0, 4.0, 0.0), (5.0, 4.0, 6.0)])
def test_free_surface_equal_radii_is_the_single_radius_oracle(L):
    r, rng = _config(50, 9, zlo=0.05)
    f = rng.standard_normal((50, 3))
    a = 0.3
    rad = np.full(50, a)
    u = O.radii_apply(r, r, f, rad, rad, 1.1, periodic_length=L, boundary="free_surface")
    ref = O.block_apply(O.UF_FREE_SURFACE, r, f, 1.1, a, periodic_length=L)
    assert rel(u, ref) < 1e-13


def test_free_surface_point_particles_at_the_surface():
    """Point particles: no normal velocity on the surface and twice the
    unbounded tangential velocity there."""
    rng = np.random.default_rng(10)
    for _ in range(20):
        y = np.array([*rng.uniform(-1, 1, 2), rng.uniform(0.5, 2)])
        x = np.array([*rng.uniform(-1, 1, 2), 0.0])
        M = O.radii_pair_matrix(x, y, 1e-4, 2e-4, 1.0, boundary="free_surface")
        M0 = O.radii_pair_matrix(x, y, 1e-4, 2e-4, 1.0, boundary="unbounded")
        assert np.abs(M[2]).max() <= 1e-12 * np.abs(M).max()
        assert np.abs(M[:2] - 2 * M0[:2]).max() <= 1e-7 * np.abs(M0).max()
