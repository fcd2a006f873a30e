"""Pins of the oracle for the blob-blob soft repulsion
(multi_bodies/forces_numba.py:12-56, radii :75-122, neighbour list
:141-201; the same force as forces.cpp:11-42 and the Python pair loop
multi_bodies_functions.py:297-345).

* the forces are minus the gradient of the potential written in the
  reference's docstring, U(r) = eps (1 + (2a - r)/b) for r < 2a and
  eps exp(-(r - 2a)/b) otherwise (central differences of the total energy,
  open and periodic boxes, one radius and a = (a_i + a_j)/2);
* Newton's third law (the forces sum to zero);
* inside the overlap the force magnitude is eps/b whatever the distance;
* the cutoff version equals a neighbour list built by scipy's cKDTree with
  the reference's cutoff 2a + 30b, and the dense sum to the e^-30 tail;
* each planted mistake of the mutant build fails one of these.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import oracle as O


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _min_image(d, L):
    if L is None:
        return d
    L = np.asarray(L, dtype=float)
    out = d.copy()
    for k in range(3):
        if L[k] > 0:
            out[..., k] -= L[k] * np.round(d[..., k] / L[k])
    return out


def energy(r, eps, b, a, L=None, radii=None):
    """Total potential sum_{i<j} U(r_ij), from the docstring's potential."""
    n = r.shape[0]
    i, j = np.triu_indices(n, 1)
    d = np.linalg.norm(_min_image(r[j] - r[i], L), axis=1)
    aa = a if radii is None else 0.5 * (radii[i] + radii[j])
    U = np.where(d < 2 * aa, eps * (1 + (2 * aa - d) / b), eps * np.exp(-(d - 2 * aa) / b))
    return U.sum()


def fd_forces(r, eps, b, a, L=None, radii=None, h=1e-6):
    F = np.zeros_like(r)
    for p in range(r.shape[0]):
        for k in range(3):
            rp, rm = r.copy(), r.copy()
            rp[p, k] += h
            rm[p, k] -= h
            F[p, k] = -(energy(rp, eps, b, a, L, radii) - energy(rm, eps, b, a, L, radii)) / (2 * h)
    return F


def _cloud(n, seed, side=3.0):
    rng = np.random.default_rng(seed)
    return rng.uniform(0, side, (n, 3)), rng


def _check_all(kind="strict"):
    """Cheap subset of the pins; returns the names of the failing ones."""
    fails = []
    eps, b, a = 0.7, 0.3, 0.25
    r, rng = _cloud(24, 1)
    F = O.blob_forces(r, eps, b, a, kind=kind)
    d = np.linalg.norm(r[:, None] - r[None], axis=-1)[np.triu_indices(24, 1)]
    assert (d < 2 * a).any() and (d > 2 * a).any()          # both branches live
    if not rel(F, fd_forces(r, eps, b, a)) < 1e-7:
        fails.append("gradient")
    L = (2.5, 2.0, 0.0)
    Fp = O.blob_forces(r, eps, b, a, periodic_length=L, kind=kind)
    if not rel(Fp, fd_forces(r, eps, b, a, L)) < 1e-7:
        fails.append("gradient_periodic")
    rad = rng.uniform(0.1, 0.4, 24)
    Fr = O.blob_forces(r, eps, b, a, radii=rad, kind=kind)
    if not rel(Fr, fd_forces(r, eps, b, a, radii=rad)) < 1e-7:
        fails.append("gradient_radii")
    return fails


def test_forces_are_minus_the_energy_gradient():
    assert _check_all() == []


@pytest.mark.parametrize("L", [None, (2.5, 2.0, 0.0), (3.0, 3.0, 3.0)])
def test_newton_third_law(L):
    r, rng = _cloud(200, 2)
    rad = rng.uniform(0.1, 0.4, 200)
    for radii in (None, rad):
        F = O.blob_forces(r, 0.3, 0.1, 0.2, periodic_length=L, radii=radii)
        assert np.abs(F.sum(axis=0)).max() <= 1e-13 * np.abs(F).max() * 200


def test_overlap_force_magnitude_is_constant():
    eps, b, a = 0.4, 0.05, 0.5
    for d in (1e-3, 0.2, 0.7, 0.999):
        r = np.array([[0.0, 0.0, 1.0], [d, 0.0, 1.0]])
        F = O.blob_forces(r, eps, b, a)
        assert abs(np.linalg.norm(F[0]) - eps / b) <= 1e-14 * eps / b
        assert F[0, 0] < 0 < F[1, 0]                        # repulsive
    # coincident blobs (and a blob with itself) exert nothing
    r = np.array([[0.3, 0.2, 1.0], [0.3, 0.2, 1.0]])
    assert not O.blob_forces(r, eps, b, a).any()


def test_periodic_images_and_large_box():
    r, _ = _cloud(60, 3)
    F0 = O.blob_forces(r, 0.3, 0.2, 0.2)
    # a box far larger than the cloud changes nothing
    assert rel(O.blob_forces(r, 0.3, 0.2, 0.2, periodic_length=(1e4, 1e4, 1e4)), F0) < 1e-14
    # shifting one blob by whole box lengths changes nothing
    L = np.array([2.0, 2.5, 0.0])
    Fp = O.blob_forces(r, 0.3, 0.2, 0.2, periodic_length=L)
    r2 = r.copy()
    r2[7] += [2.0, -5.0, 0.0]
    assert rel(O.blob_forces(r2, 0.3, 0.2, 0.2, periodic_length=L), Fp) < 1e-12


def test_cutoff_matches_kdtree_neighbour_list():
    eps, b, a = 0.2, 0.05, 0.3
    r, _ = _cloud(300, 4, side=6.0)
    d_max = 2 * a + 30 * b
    pairs = cKDTree(r).query_ball_tree(cKDTree(r), d_max)
    ref = np.zeros_like(r)
    for i, nb in enumerate(pairs):
        for j in nb:
            if i == j:
                continue
            dr = r[j] - r[i]
            dn = np.linalg.norm(dr)
            f0 = -(eps / b) * np.exp(-(dn - 2 * a) / b) / dn if dn > 2 * a else -(eps / b) / dn
            ref[i] += f0 * dr
    F = O.blob_forces(r, eps, b, a, cutoff=d_max)
    assert rel(F, ref) < 1e-14
    assert rel(O.blob_forces(r, eps, b, a), F) < 1e-12      # only the e^-30 tail differs


def test_rows():
    r, rng = _cloud(50, 5)
    rad = rng.uniform(0.1, 0.3, 50)
    full = O.blob_forces(r, 0.3, 0.1, 0.2, radii=rad)
    part = O.blob_forces(r, 0.3, 0.1, 0.2, radii=rad, rows=(10, 33))
    assert np.array_equal(part[10:33], full[10:33])
    assert not part[:10].any() and not part[33:].any()


@pytest.mark.parametrize("mutation", [11, 12, 13, 14])
def test_planted_force_mistake_is_caught(mutation):
    O.set_mutation(mutation)
    try:
        assert _check_all("mutant"), f"mutation {mutation} was not caught"
    finally:
        O.set_mutation(0)
    assert _check_all("mutant") == []
