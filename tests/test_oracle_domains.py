"""Pins of the oracle's no-wall, in-plane and free-surface blocks
(mobility_numba.py:12-120, 290-437, 440-545, 690-830, 832-936, 1077-1186,
1770-1938):

* the unbounded grand mobility [UF UT; WF WT] is symmetric and positive
  definite (overlapping pairs included);
* far apart (r > 2a) its rotational blocks are the curls of the Oseen
  tensor written from the mathematics (1/2 curl for WF, 1/4 curl curl for WT);
* the single-wall blocks tend to the unbounded ones far above the wall;
* the in-plane blocks are the x-y part of the single-wall blocks;
* the free-surface UF tends to the unbounded one far below the surface, is
  symmetric positive definite, and in the point-particle limit leaves the
  surface z = 0 with no normal velocity and no shear (u_z = 0,
  du_x/dz = du_y/dz = 0);
* a planted sign mistake in the free-surface image is caught.
"""
import numpy as np
import pytest

from oracle import oracle as O

EPS = np.zeros((3, 3, 3))
for i, j, k in ((0, 1, 2), (1, 2, 0), (2, 0, 1)):
    EPS[i, j, k], EPS[j, i, k] = 1.0, -1.0


def oseen(x, y, eta=1.0):
    d = np.asarray(x, dtype=float) - np.asarray(y, dtype=float)
    r = np.linalg.norm(d)
    return (np.eye(3) / r + np.outer(d, d) / r ** 3) / (8 * np.pi * eta)


def _d(F, x, l, h):
    e = np.zeros(3)
    e[l] = h
    return (F(x + e) - F(x - e)) / (2 * h)


def _half_curl(F, x, h=1e-4):
    return 0.5 * sum(np.einsum("im,mn->in", EPS[:, l, :], _d(F, x, l, h)) for l in range(3))


def _grand(r, eta, a, blocks):
    n = r.shape[0]
    G = np.zeros((6 * n, 6 * n))
    for (bi, bj), blk in blocks.items():
        G[3 * n * bi:3 * n * (bi + 1), 3 * n * bj:3 * n * (bj + 1)] = O.dense_block_matrix(blk, r, eta, a)
    return G


def test_unbounded_grand_mobility_is_spd():
    rng = np.random.default_rng(1)
    n, a = 30, 0.4
    r = rng.uniform(0, 3.0, (n, 3))                     # many pairs closer than 2a
    G = _grand(r, 1.0, a, {(0, 0): O.UF_NOWALL, (0, 1): O.UT_NOWALL, (1, 0): O.WF_NOWALL, (1, 1): O.WT_NOWALL})
    assert np.abs(G - G.T).max() <= 1e-14 * np.abs(G).max()
    assert np.linalg.eigvalsh(0.5 * (G + G.T)).min() > 0


def test_unbounded_rotational_blocks_are_oseen_curls():
    rng = np.random.default_rng(2)
    a = 1e-3
    for _ in range(20):
        x, y = rng.uniform(-1, 1, 3), rng.uniform(-1, 1, 3)
        if np.linalg.norm(x - y) < 0.3:
            continue
        wf = _half_curl(lambda xx: oseen(xx, y), x)
        m = O.pair_matrix(O.WF_NOWALL, x, y, False, 1.0, a)
        assert np.abs(m - wf).max() <= 1e-7 * np.abs(wf).max()
        ut = O.pair_matrix(O.UT_NOWALL, x, y, False, 1.0, a)
        assert np.abs(ut - _half_curl(lambda yy: oseen(x, yy), y).T).max() <= 1e-7 * np.abs(wf).max()
        wt = _half_curl(lambda xx: _half_curl(lambda yy: oseen(yy, xx), y, 1e-3).T, x, 1e-3)
        m = O.pair_matrix(O.WT_NOWALL, x, y, False, 1.0, a)
        assert np.abs(m - wt).max() <= 5e-5 * np.abs(wt).max()


def test_wall_blocks_tend_to_unbounded_far_from_the_wall():
    rng = np.random.default_rng(3)
    for _ in range(20):
        x = np.array([*rng.uniform(-2, 2, 2), 1e6])
        y = x + rng.uniform(-2, 2, 3)
        for wall, free in ((O.UF, O.UF_NOWALL), (O.UT, O.UT_NOWALL), (O.WF, O.WF_NOWALL), (O.WT, O.WT_NOWALL)):
            mw = O.pair_matrix(wall, x, y, False, 1.0, 0.3)
            mf = O.pair_matrix(free, x, y, False, 1.0, 0.3)
            assert np.abs(mw - mf).max() <= 1e-5 * np.abs(mf).max(), wall
            assert np.abs(mf).max() > 0


def test_in_plane_is_the_xy_part_of_the_wall():
    rng = np.random.default_rng(4)
    n, a = 25, 0.3
    r = np.column_stack([rng.uniform(0, 4, n), rng.uniform(0, 4, n), rng.uniform(0.1, 2, n)])
    v = rng.standard_normal((n, 3))
    P = np.array([1.0, 1.0, 0.0])
    for ip, w in ((O.UF_INPLANE, O.UF), (O.UT_INPLANE, O.UT)):
        got = O.block_apply(ip, r, v, 1.0, a).reshape(n, 3)
        ref = O.block_apply(w, r, v * P, 1.0, a).reshape(n, 3) * P
        assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


def _fs_failures(kind="strict"):
    fails = []
    rng = np.random.default_rng(5)
    # far from the surface: the unbounded RPY
    x = np.array([0.1, 0.2, 1e6])
    y = x + np.array([0.7, -0.4, 0.3])
    if not np.allclose(O.pair_matrix(O.UF_FREE_SURFACE, x, y, False, 1.0, 0.3, kind),
                       O.pair_matrix(O.UF_NOWALL, x, y, False, 1.0, 0.3, kind), rtol=0, atol=1e-6):
        fails.append("far")
    # point particles: no normal velocity and no shear on the surface
    a = 1e-4
    for _ in range(10):
        y = np.array([*rng.uniform(-1, 1, 2), rng.uniform(0.5, 2)])
        x = np.array([*rng.uniform(-1, 1, 2), 0.0])
        M = O.pair_matrix(O.UF_FREE_SURFACE, x, y, False, 1.0, a, kind)
        if np.abs(M[2]).max() > 1e-12 * np.abs(M).max():
            fails.append("normal")
        h = 1e-5
        dM = (O.pair_matrix(O.UF_FREE_SURFACE, x + [0, 0, h], y, False, 1.0, a, kind)
              - O.pair_matrix(O.UF_FREE_SURFACE, x - [0, 0, h], y, False, 1.0, a, kind)) / (2 * h)
        if np.abs(dM[:2]).max() > 1e-6 * np.abs(M).max():
            fails.append("shear")
    # symmetric positive definite above the surface
    n = 30
    r = np.column_stack([rng.uniform(0, 4, n), rng.uniform(0, 4, n), rng.uniform(0.35, 2, n)])
    D = O.dense_block_matrix(O.UF_FREE_SURFACE, r, 1.0, 0.3, kind)
    if np.abs(D - D.T).max() > 1e-14 * np.abs(D).max():
        fails.append("symmetric")
    if np.linalg.eigvalsh(0.5 * (D + D.T)).min() <= 0:
        fails.append("spd")
    return fails


def test_free_surface_pins():
    assert _fs_failures() == []


def test_free_surface_planted_mistake_is_caught():
    O.set_mutation(8)
    try:
        assert _fs_failures("mutant")
    finally:
        O.set_mutation(0)
    assert _fs_failures("mutant") == []
