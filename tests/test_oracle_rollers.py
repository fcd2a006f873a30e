"""Pins of the rollers time-step oracle (oracle/rollers_oracle.py):

* one-blob forces are minus the derivative of gravity plus the wall
  potential of the reference's docstring (central differences);
* the uncorrelated velocity / torque formulas equal the oracle's wall
  kernels applied to one roller (self terms), and the constrained torque
  makes that roller spin at omega;
* the uncorrelated drift is the height derivative of the perpendicular self
  mobility (central differences of the independent closed form);
* the random finite difference of the hydrodynamic step approaches
  kT d(mu_zz)/dz for one roller;
* one deterministic Euler step moves every roller by dt times the oracle's
  velocity, Adams-Bashforth uses 1.5 / -0.5 weights, and a stochastic
  attempt that crosses the wall is retried with the next draws.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle import rollers_oracle as R

ETA, A = 1e-3, 0.656
P = dict(g=0.0024892, eps_wall=0.0165677856, b_wall=0.0656)


def test_one_blob_forces_are_the_potential_gradient():
    def U(z):
        e, b = P["eps_wall"], P["b_wall"]
        wall = e * np.exp(-(z - A) / b) if z >= A else e * (1 + (A - z) / b)
        return P["g"] * 1.0 * z + wall
    for z in (0.3, 0.6, 0.7, 1.0, 2.5):
        f = R.one_blob_forces(np.array([[0.1, 0.2, z]]), 1.0, P["g"], A, P["eps_wall"], P["b_wall"])
        h = 1e-6
        fd = -(U(z + h) - U(z - h)) / (2 * h)
        assert abs(f[0, 2] - fd) <= 1e-6 * max(1.0, abs(fd)) and f[0, 0] == 0 and f[0, 1] == 0


@pytest.mark.parametrize("z", [0.9, 1.5, 4.0])
def test_uncorrelated_equals_wall_self_terms(z):
    r = np.array([[0.3, 0.2, z]])
    F = np.array([0.2, -0.4, 0.7])
    om = np.array([0.5, 62.8, -3.0])
    v, T = R.uncorrelated_velocity(r, F, om, "True", ETA, A)
    assert np.allclose(T[0], om * 8 * math.pi * ETA * A ** 3, rtol=1e-15)
    ref = O.uf(r, F, ETA, A) + O.ut(r, T.ravel(), ETA, A)
    assert np.abs(v.ravel() - ref).max() <= 1e-13 * np.abs(ref).max()
    v2, T2 = R.uncorrelated_velocity(r, F, om, "False", ETA, A)
    w = O.wf(r, F, ETA, A) + O.wt(r, T2.ravel(), ETA, A)
    assert np.abs(w - om).max() <= 1e-13 * np.abs(om).max()
    ref2 = O.uf(r, F, ETA, A) + O.ut(r, T2.ravel(), ETA, A)
    assert np.abs(v2.ravel() - ref2).max() <= 1e-13 * np.abs(ref2).max()


def test_uncorrelated_drift_is_dmu_perp_dh():
    kT, dt = 0.0041419464, 0.016
    for z in (0.9, 1.3, 3.0):
        v = R.uncorrelated_noise(np.array([[0.0, 0.0, z]]), np.zeros(3), ETA, A, kT, dt)
        mu = lambda zz: O.self_mobility_with_rotation(np.array([0, 0, zz]), ETA, A)[2, 2]  # noqa: E731
        h = 1e-6
        dmu_dh = A * (mu(z + h) - mu(z - h)) / (2 * h)     # derivative in units of the radius
        assert abs(v[0, 2] - kT * dmu_dh) <= 1e-6 * abs(kT * dmu_dh)
        assert v[0, 0] == 0 and v[0, 1] == 0
    # the noise part: sqrt(2 kT/dt) sqrt(mu) z
    z3 = np.array([0.3, -1.2, 0.0])
    v = R.uncorrelated_noise(np.array([[0.0, 0.0, 2.0]]), z3, ETA, A, kT, dt)
    M = O.self_mobility_with_rotation(np.array([0, 0, 2.0]), ETA, A)
    assert np.allclose(v[0, :2], math.sqrt(2 * kT / dt) * np.sqrt(M[0, 0]) * z3[:2], rtol=1e-13)


def test_rfd_drift_approaches_the_divergence():
    kT, dt = 0.0041419464, 0.016
    for z in (1.0, 2.0):
        st = R.RollersStep(np.array([[0.0, 0.0, z]]), "stochastic_first_order", ETA, A, kT=kT, blob_blob=False,
                           rf_delta=1e-4)
        drift = st.rfd_drift(st.r, np.array([[0.0, 0.0, 1.0]]))
        mu = lambda zz: O.uf(np.array([[0.0, 0.0, zz]]), np.array([0, 0, 1.0]), ETA, A)[2]  # noqa: E731
        h = 1e-5
        assert abs(drift[2] - kT * (mu(z + h) - mu(z - h)) / (2 * h)) <= 1e-6 * abs(drift[2])


def test_scheme_updates_and_retry():
    rng = np.random.default_rng(3)
    n = 12
    r = np.column_stack([rng.uniform(0, 6, n), rng.uniform(0, 6, n), rng.uniform(1.0, 2.5, n)])
    kw = dict(kT=0.0041419464, g=P["g"], eps_wall=P["eps_wall"], b_wall=P["b_wall"], eps=0.0165677856, b=0.0656,
              omega=(0.0, 62.8, 0.0))
    # deterministic Euler: r + dt v
    st = R.RollersStep(r, "deterministic_forward_euler_rollers", ETA, A, **kw)
    v, _ = st.det_velocity(st.r)
    st.advance(0.016)
    assert np.allclose(st.r, r + 0.016 * v.reshape(-1, 3), rtol=0, atol=1e-15)
    # Adams-Bashforth: second step uses 1.5 v1 - 0.5 v0
    st = R.RollersStep(r, "deterministic_adams_bashforth_rollers", ETA, A, **kw)
    v0, _ = st.det_velocity(st.r)
    st.advance(0.016)
    r1 = st.r.copy()
    v1, _ = st.det_velocity(r1)
    st.advance(0.016)
    assert np.allclose(st.r, r1 + 0.016 * (1.5 * v1 - 0.5 * v0).reshape(-1, 3), rtol=0, atol=1e-15)
    # a stochastic attempt that puts a roller below the wall is redrawn
    st = R.RollersStep(r, "stochastic_adams_bashforth_rollers", ETA, A, **kw)
    bad = np.zeros((n, 3))
    bad[0, 2] = -1e4
    good = rng.standard_normal((n, 3))
    dxs = rng.standard_normal((2, n, 3))
    draws = iter([(bad.ravel(), dxs[0]), (good.ravel(), dxs[1])])
    assert st.advance(0.016, draws) == 2 and st.invalid == 1
    assert (st.r[:, 2] > 0).all()
