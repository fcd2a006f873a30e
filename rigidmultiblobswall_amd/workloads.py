"""Synthetic workloads of the benchmark configurations.

Each generator draws from numpy.random.default_rng(seed) in the order
positions -> orientations -> forces/torques, and returns plain float64
numpy arrays.

  C1  single-wall UF matvec, N = 2048 blobs, a = 0.13, eta = 1,
      x, y ~ U[-20, 20], z ~ U[0.2, 3.0], f ~ N(0, 1).
  C2  256 rigid spheres, each a 42-blob shell (once-subdivided icosahedron,
      radius 1, blob a = 0.243), centres x, y ~ U[-30, 30], z ~ U[1.3, 4.0]
      with centre distance >= 2 (1 + a), random unit quaternions, body
      force/torque ~ N(0, 1).
  C3  32768 single-blob rollers, a = 0.656, eta = 1e-3, x, y ~ U[0, 300],
      z ~ U[0.8, 2.0]; grand (6N) noise z ~ N(0, 1), block of 4.
  C4  1024 42-blob spheres in a monolayer of area fraction 0.4
      (box side sqrt(1024 pi / 0.4)), z_c = 1.3 + N(0, 0.05^2).
  C5  64 independent roller suspensions, N_r ~ U{128..4096}, same areal
      density as C3, f ~ N(0, 1).
  C3 strip (time-step workload): 32768 rollers in the footprint of the
      rollers example's clones file, x ~ U[0, 39.36], y ~ U[0, 3224.3],
      z = 0.66 + Exp(mean 2.25), centres at least 1.38 apart (the example's
      smallest neighbour distances), with the example's parameters.
"""
from __future__ import annotations

import itertools

import numpy as np


def icosphere_42(radius: float = 1.0) -> np.ndarray:
    """42 vertices of the once-subdivided icosahedron on a sphere (the
    geometry of multi_bodies/Structures/shell_N_42_Rg_1_Rh_1_1220.vertex, up
    to a rotation)."""
    phi = (1.0 + 5.0 ** 0.5) / 2.0
    base = []
    for s1, s2 in itertools.product((-1.0, 1.0), repeat=2):
        base += [(0.0, s1, s2 * phi), (s1, s2 * phi, 0.0), (s2 * phi, 0.0, s1)]
    V = np.array(base)
    V /= np.linalg.norm(V, axis=1, keepdims=True)
    # icosahedron edges join vertices at the minimum distance
    d = np.linalg.norm(V[:, None] - V[None], axis=-1)
    emin = d[d > 1e-9].min()
    edges = [(i, j) for i in range(12) for j in range(i + 1, 12) if abs(d[i, j] - emin) < 1e-9]
    mids = np.array([(V[i] + V[j]) / 2.0 for i, j in edges])
    mids /= np.linalg.norm(mids, axis=1, keepdims=True)
    P = np.vstack([V, mids])
    assert P.shape == (42, 3)
    return radius * P


def random_quaternions(rng: np.random.Generator, n: int) -> np.ndarray:
    """Normalised N(0, I_4) draws (quaternion_integrator/quaternion.py:77-84)."""
    q = rng.standard_normal((n, 4))
    return q / np.linalg.norm(q, axis=1, keepdims=True)


def rotation_matrices(q: np.ndarray) -> np.ndarray:
    """Rotation matrices of unit quaternions (s, p) as in
    quaternion_integrator/quaternion.py:42-52; q (n, 4) -> (n, 3, 3)."""
    s, p0, p1, p2 = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
    diag = s * s - 0.5
    R = np.empty((q.shape[0], 3, 3))
    R[:, 0, 0] = p0 * p0 + diag
    R[:, 0, 1] = p0 * p1 - s * p2
    R[:, 0, 2] = p0 * p2 + s * p1
    R[:, 1, 0] = p1 * p0 + s * p2
    R[:, 1, 1] = p1 * p1 + diag
    R[:, 1, 2] = p1 * p2 - s * p0
    R[:, 2, 0] = p2 * p0 - s * p1
    R[:, 2, 1] = p2 * p1 + s * p0
    R[:, 2, 2] = p2 * p2 + diag
    return 2.0 * R


def _rsa_centres(rng, n, lo_xy, hi_xy, z_sampler, min_dist, max_tries=10_000_000):
    """Random sequential placement with a minimum centre distance."""
    C = np.empty((n, 3))
    k = 0
    tries = 0
    while k < n:
        tries += 1
        if tries > max_tries:
            raise RuntimeError("could not place bodies at this density")
        c = np.array([rng.uniform(lo_xy, hi_xy), rng.uniform(lo_xy, hi_xy), z_sampler()])
        if k and np.min(np.sum((C[:k] - c) ** 2, axis=1)) < min_dist * min_dist:
            continue
        C[k] = c
        k += 1
    return C


def c1_uf_matvec(seed: int = 1, n: int = 2048):
    rng = np.random.default_rng(seed)
    r = np.column_stack([rng.uniform(-20, 20, n), rng.uniform(-20, 20, n), rng.uniform(0.2, 3.0, n)])
    f = rng.standard_normal(3 * n)
    return {"r": r, "f": f, "eta": 1.0, "a": 0.13, "L": np.zeros(3)}


def bodies_blobs(centres, quats, ref):
    R = rotation_matrices(quats)
    return (np.einsum("bij,kj->bki", R, ref) + centres[:, None, :]).reshape(-1, 3)


def c2_rigid_spheres(seed: int = 2, nbodies: int = 256):
    rng = np.random.default_rng(seed)
    a = 0.243
    centres = _rsa_centres(rng, nbodies, -30.0, 30.0, lambda: rng.uniform(1.3, 4.0), 2.0 * (1.0 + a))
    quats = random_quaternions(rng, nbodies)
    ft = rng.standard_normal((nbodies, 6))
    ref = icosphere_42(1.0)
    return {"centres": centres, "quats": quats, "ref": ref, "force_torque": ft, "eta": 1.0, "a": a,
            "r": bodies_blobs(centres, quats, ref), "L": np.zeros(3)}


def c3_rollers(seed: int = 3, n: int = 32768, k: int = 4):
    rng = np.random.default_rng(seed)
    r = np.column_stack([rng.uniform(0, 300, n), rng.uniform(0, 300, n), rng.uniform(0.8, 2.0, n)])
    z = rng.standard_normal((k, 6 * n))
    return {"r": r, "z": z, "eta": 1e-3, "a": 0.656, "L": np.zeros(3)}


def c4_monolayer(seed: int = 4, nbodies: int = 1024, area_fraction: float = 0.4):
    rng = np.random.default_rng(seed)
    a = 0.243
    side = float(np.sqrt(nbodies * np.pi / area_fraction))
    centres = _rsa_centres(rng, nbodies, 0.0, side, lambda: 1.3 + 0.05 * rng.standard_normal(), 2.1)
    quats = random_quaternions(rng, nbodies)
    ref = icosphere_42(1.0)
    return {"centres": centres, "quats": quats, "ref": ref, "eta": 1.0, "a": a, "side": side,
            "r": bodies_blobs(centres, quats, ref), "L": np.zeros(3)}


def c5_ragged_rollers(seed: int = 5, nrep: int = 64, nmin: int = 128, nmax: int = 4096):
    rng = np.random.default_rng(seed)
    sizes = rng.integers(nmin, nmax + 1, size=nrep)
    rs = []
    for nr in sizes:
        side = 300.0 * np.sqrt(nr / 32768.0)
        rs.append(np.column_stack([rng.uniform(0, side, nr), rng.uniform(0, side, nr), rng.uniform(0.8, 2.0, nr)]))
    offsets = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    r = np.vstack(rs)
    f = rng.standard_normal(3 * int(offsets[-1]))
    return {"r": r, "f": f, "offsets": offsets, "eta": 1e-3, "a": 0.656, "L": np.zeros(3)}


# Parameters of the rollers example (multi_bodies/examples/rollers/inputfile_rollers.dat).
ROLLERS_EXAMPLE = dict(scheme="stochastic_adams_bashforth_rollers", dt=0.016, solver_tolerance=1e-3, eta=1.0e-3,
                       g=0.0024892, blob_radius=0.656, kT=0.0041419464, omega_one_roller=(0.0, 62.8, 0.0),
                       free_kinematics="True", repulsion_strength=0.0165677856, debye_length=0.0656,
                       repulsion_strength_wall=0.0165677856, debye_length_wall=0.0656)


def c3_rollers_strip(seed: int = 6, n: int = 32768, min_dist: float = 1.38):
    """Rollers placed like the example's suspension (see the module notes),
    by batched rejection with a k-d tree."""
    from scipy.spatial import cKDTree
    rng = np.random.default_rng(seed)
    Lx, Ly = 39.36, 3224.3
    pts = np.empty((0, 3))
    while pts.shape[0] < n:
        m = n - pts.shape[0]
        cand = np.column_stack([rng.uniform(0, Lx, 2 * m), rng.uniform(0, Ly, 2 * m),
                                0.66 + rng.exponential(2.25, 2 * m)])
        keep = np.ones(cand.shape[0], dtype=bool)
        if pts.shape[0]:
            d, _ = cKDTree(pts).query(cand, k=1)
            keep &= d >= min_dist
        for i, j in sorted(cKDTree(cand).query_pairs(min_dist)):
            if keep[i] and keep[j]:
                keep[j] = False
        pts = np.vstack([pts, cand[keep][:m]])
    return {"r": pts[:n], "eta": ROLLERS_EXAMPLE["eta"], "a": ROLLERS_EXAMPLE["blob_radius"], "L": np.zeros(3),
            "params": dict(ROLLERS_EXAMPLE)}
