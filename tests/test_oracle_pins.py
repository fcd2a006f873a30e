"""CPU pins of the FP64 oracle (no GPU).

Every pair block of the oracle is tied to something the reference does not
itself supply, so that a dropped term, a wrong sign or a transposed index
fails at least one check:

* UF: the second, independent algebraic statement of the same mobility
  (dense Swan-Brady A-E form, mobility/mobility.py:1018-1116), including
  the wall-overlap regularisation B M(z_eff) B.
* all blocks: the Blake (1971) point-force image system written from the
  mathematics (pinned here by no-slip, reciprocity and incompressibility).
  As a -> 0, UF -> G, WF -> 1/2 curl_x G, UT -> 1/2 curl_y G, WT -> 1/4
  curl_x curl_y G; the O(a^2) terms of UF, WF, UT are the Faxen corrections
  (1 + a^2/6 lap) on the translational side(s).
* self blocks: closed forms (mobility/mobility.py:739-772;
  quaternion_integrator/quaternion_integrator_rollers.py:1017-1060) and the
  coincidence limit of the pair wall terms.
* invariants: M_ij = M_ji^T (UF, WT), UT_ij = WF_ji^T, SPD of the grand
  mobility (mobility/mobility_test.py:116-130), zero mobility at the wall
  (mobility_test.py:132-144), periodic images by brute force.
* each planted mistake of the mutant build is caught.
"""
import numpy as np
import pytest

from oracle import oracle as O

EPS = np.zeros((3, 3, 3))
EPS[0, 1, 2] = EPS[1, 2, 0] = EPS[2, 0, 1] = 1.0
EPS[0, 2, 1] = EPS[2, 1, 0] = EPS[1, 0, 2] = -1.0


def _rel(a, b):
    return float(np.abs(np.asarray(a) - np.asarray(b)).max() / np.abs(np.asarray(b)).max())


def _d(F, x, l, h):
    e = np.zeros(3)
    e[l] = h
    return (F(x + e) - F(x - e)) / (2 * h)


def _lap(F, x, h):
    return sum((F(x + np.eye(3)[l] * h) - 2 * F(x) + F(x - np.eye(3)[l] * h)) / h ** 2 for l in range(3))


def _half_curl(F, x, h=1e-4):
    """(1/2) eps_ilm d_l F_mn for a matrix field F(x)."""
    return 0.5 * sum(np.einsum("im,mn->in", EPS[:, l, :], _d(F, x, l, h)) for l in range(3))


def _pairs(seed, count, zlo=0.5, zhi=2.0):
    rng = np.random.default_rng(seed)
    for _ in range(count):
        x = np.array([*rng.uniform(-1, 1, 2), rng.uniform(zlo, zhi)])
        y = np.array([*rng.uniform(-1, 1, 2), rng.uniform(zlo, zhi)])
        yield x, y


# ---------------------------------------------------------------------------
# Blake tensor itself
# ---------------------------------------------------------------------------
def test_blake_tensor_defining_properties():
    for x, y in _pairs(1, 4):
        xw = x.copy()
        xw[2] = 0.0
        assert np.abs(O.blake_tensor(xw, y)).max() < 1e-15            # no slip on the wall
        assert np.abs(O.blake_tensor(x, y) - O.blake_tensor(y, x).T).max() < 1e-15  # reciprocity
        h = 1e-5
        div = sum(_d(lambda xx: O.blake_tensor(xx, y), x, i, h)[i] for i in range(3))
        assert np.abs(div).max() < 1e-9                                # incompressible
        # far from the wall it tends to the free Stokeslet
        far = x + np.array([0, 0, 1e4])
        yf = y + np.array([0, 0, 1e4])
        S = O.blake_tensor(far, yf) * 8 * np.pi
        r = far - yf
        Sfree = np.eye(3) / np.linalg.norm(r) + np.outer(r, r) / np.linalg.norm(r) ** 3
        assert _rel(S, Sfree) < 1e-3


# ---------------------------------------------------------------------------
# UF
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("zlo", [0.05, 0.5])
def test_uf_matches_dense_swan_brady_form(zlo):
    """Numba-order restatement == A-E dense form, incl. z < a (B damping)."""
    rng = np.random.default_rng(2)
    n, a, eta = 40, 0.13, 0.9
    r = np.column_stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(zlo, 0.8, n)])
    f = rng.standard_normal(3 * n)
    M = O.single_wall_fluid_mobility(r, eta, a)
    assert np.linalg.norm(O.uf(r, f, eta, a) - M @ f) <= 1e-13 * np.linalg.norm(M @ f)


def test_no_wall_matches_dense_rpy():
    rng = np.random.default_rng(3)
    n, a, eta = 40, 0.2, 1.3
    r = rng.uniform(-1, 1, (n, 3))
    f = rng.standard_normal(3 * n)
    M = O.rotne_prager_tensor(r, eta, a)
    u = O.block_apply(O.UF_NOWALL, r, f, eta, a)
    assert np.linalg.norm(u - M @ f) <= 1e-13 * np.linalg.norm(M @ f)


def test_uf_point_limit_is_blake_with_order_a2():
    for x, y in _pairs(4, 3):
        errs = [_rel(O.pair_matrix(O.UF, x, y, False, 1.0, a), O.blake_tensor(x, y)) for a in (1e-2, 1e-3)]
        assert errs[1] < 1e-5
        assert 80 < errs[0] / errs[1] < 120   # O(a^2) convergence


def test_uf_order_a2_is_faxen():
    a, h = 0.05, 2e-3
    for x, y in _pairs(5, 3):
        G = O.blake_tensor(x, y)
        pred = G + a * a / 6 * (_lap(lambda xx: O.blake_tensor(xx, y), x, h)
                                + _lap(lambda yy: O.blake_tensor(x, yy), y, h))
        M = O.pair_matrix(O.UF, x, y, False, 1.0, a)
        assert _rel(M, G) > 1e-4            # the a^2 terms are visible ...
        assert _rel(M, pred) < 1e-5         # ... and are the Faxen terms


# ---------------------------------------------------------------------------
# WF, UT, WT
# ---------------------------------------------------------------------------
def test_wf_ut_point_limit_is_half_curl_of_blake():
    a = 1e-4
    for x, y in _pairs(6, 3):
        wf = _half_curl(lambda xx: O.blake_tensor(xx, y), x)
        assert _rel(O.pair_matrix(O.WF, x, y, False, 1.0, a), wf) < 1e-6
        ut = _half_curl(lambda yy: O.blake_tensor(yy, x), y).T
        assert _rel(O.pair_matrix(O.UT, x, y, False, 1.0, a), ut) < 1e-6


def test_wf_ut_order_a2_is_faxen():
    a, h = 0.05, 2e-3
    for x, y in _pairs(7, 3):
        wf = _half_curl(lambda xx: O.blake_tensor(xx, y)
                        + a * a / 6 * _lap(lambda yy: O.blake_tensor(xx, yy), y, h), x)
        assert _rel(O.pair_matrix(O.WF, x, y, False, 1.0, a), wf) < 1e-6
        ut = _half_curl(lambda yy: O.blake_tensor(yy, x)
                        + a * a / 6 * _lap(lambda xx: O.blake_tensor(yy, xx), x, h), y).T
        assert _rel(O.pair_matrix(O.UT, x, y, False, 1.0, a), ut) < 1e-6


def test_wt_is_quarter_curl_curl_of_blake():
    # WT has no a-dependence beyond its normalisation (no Faxen term).
    for x, y in _pairs(8, 3):
        wt = _half_curl(lambda xx: _half_curl(lambda yy: O.blake_tensor(yy, xx), y, 1e-3).T, x, 1e-3)
        for a in (1e-3, 0.05):
            assert _rel(O.pair_matrix(O.WT, x, y, False, 1.0, a), wt) < 5e-5


# ---------------------------------------------------------------------------
# self terms
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("h_over_a", [1.0, 1.3, 2.0, 5.0, 40.0])
def test_self_blocks_closed_forms(h_over_a):
    eta, a = 0.7, 0.3
    p = np.array([0.2, -0.1, h_over_a * a])
    C = O.self_mobility_with_rotation(p, eta, a)
    uf = O.pair_matrix(O.UF, p, p, True, eta, a)
    wt = O.pair_matrix(O.WT, p, p, True, eta, a)
    ut = O.pair_matrix(O.UT, p, p, True, eta, a)
    wf = O.pair_matrix(O.WF, p, p, True, eta, a)
    np.testing.assert_allclose(uf, C[:3, :3], rtol=1e-13, atol=1e-15)
    np.testing.assert_allclose(wt, C[3:, 3:], rtol=1e-13, atol=1e-15)
    # coupling: same magnitude, right-handed sign (torque +y rolls the blob
    # in +x near the wall; the closed form uses the opposite convention).
    np.testing.assert_allclose(np.abs(ut), np.abs(C[:3, 3:]), rtol=1e-13, atol=1e-15)
    assert ut[0, 1] > 0 and ut[1, 0] < 0
    np.testing.assert_allclose(wf, ut.T, rtol=1e-13, atol=1e-15)
    # rollers' closed forms (quaternion_integrator_rollers.py:1017-1060)
    h = h_over_a
    ftt, frr, frt = 1 / (6 * np.pi * eta * a), 1 / (6 * np.pi * eta * a ** 3), 1 / (6 * np.pi * eta * a ** 2)
    assert np.isclose(uf[0, 0], ftt * (1 - 9 / (16 * h) + 2 / (16 * h ** 3) - 1 / (16 * h ** 5)), rtol=1e-13)
    assert np.isclose(uf[2, 2], ftt * (1 - 9 / (8 * h) + 1 / (2 * h ** 3) - 1 / (8 * h ** 5)), rtol=1e-13)
    assert np.isclose(wt[0, 0], frr * (3 / 4 - 15 / (64 * h ** 3)), rtol=1e-13)
    assert np.isclose(wt[2, 2], frr * (3 / 4 - 3 / (32 * h ** 3)), rtol=1e-13)
    assert np.isclose(ut[0, 1], frt * 3 / (32 * h ** 4), rtol=1e-13)


@pytest.mark.parametrize("block", [O.UF, O.UT, O.WF, O.WT])
def test_pair_wall_term_tends_to_self_wall_term(block):
    """As j -> i the pair's wall part equals the self wall part (R -> 2 h z).
    The unbounded part is removed by differencing against the same pair
    lifted by `up` (same d, weaker image), compared with the self blocks at
    the two heights."""
    eta, a = 1.0, 0.4
    p = np.array([0.0, 0.0, 0.9])
    q = p + 1e-8 * np.array([1.0, -2.0, 1.5])
    up = np.array([0.0, 0.0, 50.0])
    pair = O.pair_matrix(block, p, q, False, eta, a) - O.pair_matrix(block, p + up, q + up, False, eta, a)
    self_ = O.pair_matrix(block, p, p, True, eta, a) - O.pair_matrix(block, p + up, p + up, True, eta, a)
    np.testing.assert_allclose(pair, self_, rtol=1e-6, atol=1e-6 * np.abs(self_).max())


# ---------------------------------------------------------------------------
# invariants
# ---------------------------------------------------------------------------
def _dense(block, r, eta, a, kind="strict"):
    return O.dense_block_matrix(block, r, eta, a, kind=kind)


def test_symmetries_and_spd():
    rng = np.random.default_rng(9)
    n, eta, a = 6, 1.0, 1.0
    r = rng.normal(10.0, 1.0, (n, 3))  # mobility_test.py:116-130 layout
    uf, wt = _dense(O.UF, r, eta, a), _dense(O.WT, r, eta, a)
    ut, wf = _dense(O.UT, r, eta, a), _dense(O.WF, r, eta, a)
    assert np.abs(uf - uf.T).max() < 1e-15
    assert np.abs(wt - wt.T).max() < 1e-15
    assert np.abs(ut - wf.T).max() < 1e-15
    grand = np.block([[uf, ut], [wf, wt]])
    assert np.linalg.eigvalsh(grand).min() > 0


def test_grand_spd_near_wall_with_overlaps():
    rng = np.random.default_rng(10)
    n, eta, a = 12, 1.0, 0.3
    r = np.column_stack([rng.uniform(0, 2, n), rng.uniform(0, 2, n), rng.uniform(0.35, 1.0, n)])
    grand = np.block([[_dense(O.UF, r, eta, a), _dense(O.UT, r, eta, a)],
                      [_dense(O.WF, r, eta, a), _dense(O.WT, r, eta, a)]])
    assert np.abs(grand - grand.T).max() < 1e-12 * np.abs(grand).max()
    assert np.linalg.eigvalsh(grand).min() > 0


def test_wrapper_equals_dense_with_damping():
    rng = np.random.default_rng(11)
    n, eta, a = 30, 1.0, 0.13
    r = 5 * a * rng.random((n, 3))  # mobility/test_blobs.py:31-34 stress layout
    v = rng.standard_normal(3 * n)
    for b in (O.UF, O.UT, O.WF, O.WT):
        M = O.dense_block_matrix(b, r, eta, a, wrapper=True)
        assert np.linalg.norm(O.block_apply(b, r, v, eta, a) - M @ v) <= 1e-13 * np.linalg.norm(M @ v)


def test_mobility_vanishes_for_tiny_blob_at_wall():
    a = 1e-4
    r = np.array([[0.0, 0.0, a], [1.0, 1.0, 8.0]])
    M = O.single_wall_fluid_mobility(r, 1.0, a)
    assert np.abs(M[0:3, 3:6]).max() < 1e-6     # mobility_test.py:132-144
    Mc = _dense(O.UF, r, 1.0, a)
    assert np.abs(Mc[0:3, 3:6]).max() < 1e-6


def test_periodic_images_brute_force():
    rng = np.random.default_rng(12)
    n, eta, a = 7, 1.0, 0.3
    L = np.array([3.0, 2.5, 0.0])
    r = np.column_stack([rng.uniform(-2, 5, n), rng.uniform(0, 2.5, n), rng.uniform(0.4, 1.5, n)])
    v = rng.standard_normal((n, 3))
    for block in (O.UF, O.UT, O.WF, O.WT):
        u = np.zeros((n, 3))
        for i in range(n):
            for j in range(n):
                d = r[i] - r[j]
                for c in range(2):
                    d[c] -= np.trunc(d[c] / L[c] + 0.5 * np.sign(d[c])) * L[c]
                for bx in (-1, 0, 1):
                    for by in (-1, 0, 1):
                        shift = np.array([bx * L[0], by * L[1], 0.0])
                        self_term = i == j and bx == 0 and by == 0
                        rj = r[i] - d - shift   # source image position
                        u[i] += O.pair_matrix(block, r[i], rj, self_term, eta, a) @ v[j]
        got = O.block_apply(block, r, v, eta, a, periodic_length=L, wrapper=False)
        assert np.linalg.norm(got - u.ravel()) <= 1e-12 * np.linalg.norm(u)


# ---------------------------------------------------------------------------
# planted mistakes are caught
# ---------------------------------------------------------------------------
def _pin_failures(kind):
    """Run a cheap subset of the pins against one oracle build; return the
    names of the failing pins."""
    fails = []
    rng = np.random.default_rng(20)
    n, a, eta = 30, 0.13, 1.0
    r = np.column_stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0.2, 0.8, n)])
    f = rng.standard_normal(3 * n)
    M = O.single_wall_fluid_mobility(r, eta, a)
    if not np.linalg.norm(O.block_apply(O.UF, r, f, eta, a, kind=kind) - M @ f) <= 1e-12 * np.linalg.norm(M @ f):
        fails.append("uf_dense")  # (NaN-safe comparisons throughout)
    p = np.array([0.1, 0.2, 1.4 * a])
    C = O.self_mobility_with_rotation(p, eta, a)
    if not np.allclose(O.pair_matrix(O.UF, p, p, True, eta, a, kind), C[:3, :3], rtol=1e-12):
        fails.append("uf_self")
    x, y = np.array([0.3, -0.2, 0.9]), np.array([-0.4, 0.5, 1.3])
    wf = _half_curl(lambda xx: O.blake_tensor(xx, y), x)
    if not _rel(O.pair_matrix(O.WF, x, y, False, 1.0, 1e-4, kind), wf) <= 1e-6:
        fails.append("wf_curl")
    ut_t = O.pair_matrix(O.UT, y, x, False, 1.0, 0.2, kind).T
    if not _rel(O.pair_matrix(O.WF, x, y, False, 1.0, 0.2, kind), ut_t) <= 1e-13:
        fails.append("ut_wf_transpose")
    wt = _half_curl(lambda xx: _half_curl(lambda yy: O.blake_tensor(yy, xx), y, 1e-3).T, x, 1e-3)
    if not _rel(O.pair_matrix(O.WT, x, y, False, 1.0, 1e-3, kind), wt) <= 5e-5:
        fails.append("wt_curlcurl")
    return fails


def test_strict_oracle_passes_pin_subset():
    assert _pin_failures("strict") == []


@pytest.mark.parametrize("mutation", [1, 2, 3, 4, 5, 6, 7])
def test_planted_mistake_is_caught(mutation):
    O.set_mutation(mutation)
    try:
        assert _pin_failures("mutant"), f"mutation {mutation} was not caught"
    finally:
        O.set_mutation(0)
    assert _pin_failures("mutant") == []
