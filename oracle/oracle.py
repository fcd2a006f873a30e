"""FP64 CPU oracle for the wall-RPY mobility path.

TEST INFRASTRUCTURE ONLY.  Only the test suite, the smoke check and the
CPU-baseline leg of the benchmark import this module; the product package
never does.  It is the checker, never the thing measured or shipped.

Contents
--------
* ctypes front end of ``wall_rpy_oracle.c`` -- the C restatement of the four
  single-wall numba kernels (mobility/mobility_numba.py:123-287, 547-686,
  937-1073, 1188-1328) and their B M(z_eff) B wrappers (mobility/mobility.py:
  52-84, 1132-1163, 1213-1232, 1273-1298, 1316-1342).
* numpy restatements of the *second* algebraic statement of the wall
  mobility, the dense Swan-Brady A-E form (mobility/mobility.py:1018-1116),
  of the dense RPY tensor (mobility.py:967-1012), of the Blake point-particle
  image system (mobility.py:122-149, 174-182) and of the closed-form 6x6
  self mobility (mobility.py:739-772).  These pin the C restatement.
* rigid-body pieces: quaternion rotation (quaternion_integrator/quaternion.py:
  42-52), K = [J, rot] (body/body.py:81-113), the rigid linear operator
  (multi_bodies/multi_bodies.py:424-477), the block-diagonal preconditioner
  (multi_bodies/multi_bodies.py:752-907) and right-preconditioned GMRES
  (general_application_utils.py:514-635, with the warm start honoured).
* the Lanczos M^1/2 z iteration (stochastic_forcing/stochastic_forcing.py:
  112-262).
"""
from __future__ import annotations

import ctypes
import os
from typing import Callable, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

UF, UT, WF, WT, UF_NOWALL = 0, 1, 2, 3, 4
UT_NOWALL, WF_NOWALL, WT_NOWALL, UF_INPLANE, UT_INPLANE, UF_FREE_SURFACE = 5, 6, 7, 8, 9, 10
BLOCK_POWER = {UF: 1, UT: 2, WF: 2, WT: 3, UF_NOWALL: 1, UT_NOWALL: 2, WF_NOWALL: 2, WT_NOWALL: 3,
               UF_INPLANE: 1, UT_INPLANE: 2, UF_FREE_SURFACE: 1}
# blocks whose wrapper applies the wall height shift and damping
WALL_BLOCKS = (UF, UT, WF, WT, UF_INPLANE, UT_INPLANE)

_LIBS: dict = {}


def _lib(kind: str = "strict") -> ctypes.CDLL:
    """Load one of the oracle builds ('strict', 'fast', 'fast_avx512',
    'mutant'); 'fast_avx512' only on a CPU with the x86-64-v4 features."""
    if kind in _LIBS:
        return _LIBS[kind]
    if kind == "fast_avx512" and not host_has_avx512():
        raise RuntimeError("this CPU lacks the AVX-512 features of the fast_avx512 oracle build")
    name = {"strict": "liboracle_wall_rpy.so",
            "fast": "liboracle_wall_rpy_fast.so",
            "fast_avx512": "liboracle_wall_rpy_fast_avx512.so",
            "mutant": "liboracle_wall_rpy_mutant.so"}[kind]
    path = os.path.join(_HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library {path} is not built; run `make -C oracle`")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    lib.orc_block_kernel.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                     ctypes.c_int64, ctypes.c_double, ctypes.c_double, dp]
    lib.orc_block_kernel.restype = ctypes.c_int
    lib.orc_block_apply.argtypes = [ctypes.c_int, dp, dp, dp, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_double, ctypes.c_double, dp, dp, dp]
    lib.orc_block_apply.restype = ctypes.c_int
    lib.orc_block_pair_matrix.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_double,
                                          ctypes.c_double, dp]
    lib.orc_block_pair_matrix.restype = ctypes.c_int
    lib.orc_num_threads.restype = ctypes.c_int
    lib.orc_radii_apply.argtypes = [dp, dp, dp, dp, dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                    ctypes.c_int64, ctypes.c_double, dp, ctypes.c_int, dp]
    lib.orc_radii_apply.restype = ctypes.c_int
    lib.orc_radii_pair_matrix.argtypes = [dp, dp, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                          dp]
    lib.orc_radii_pair_matrix.restype = ctypes.c_int
    lib.orc_blob_forces.argtypes = [dp, dp, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_double, ctypes.c_double, dp, dp]
    lib.orc_blob_forces.restype = ctypes.c_int
    if kind == "mutant":
        lib.orc_set_mutation.argtypes = [ctypes.c_int]
    _LIBS[kind] = lib
    return lib


def host_has_avx512() -> bool:
    """Whether the CPU reports the AVX-512 subsets x86-64-v4 requires."""
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    return {"avx512f", "avx512bw", "avx512cd", "avx512dq", "avx512vl"} <= flags
    except OSError:
        pass
    return False


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def set_mutation(m: int) -> None:
    """Select a planted mistake in the mutant build (0 = none)."""
    _lib("mutant").orc_set_mutation(int(m))


def num_threads(kind: str = "strict") -> int:
    return int(_lib(kind).orc_num_threads())


def set_num_threads(n: int, kind: str = "strict") -> None:
    _lib(kind).orc_set_num_threads(int(n))


def block_apply(block: int, r_vectors, vec, eta: float, a: float, periodic_length=None,
                rows: Optional[tuple] = None, kind: str = "strict", wrapper: bool = True) -> np.ndarray:
    """out = B M(z_eff) B vec for one block (wrapper=True), or the bare
    kernel product M(r) vec (wrapper=False).  Returns (3N,) float64; with
    ``rows=(t0, t1)`` only those targets are computed (others are zero)."""
    r = np.ascontiguousarray(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3))
    n = r.shape[0]
    v = np.ascontiguousarray(np.asarray(vec, dtype=np.float64).reshape(n, 3))
    out = np.zeros((n, 3))
    L = np.zeros(3) if periodic_length is None else np.ascontiguousarray(
        np.asarray(periodic_length, dtype=np.float64).reshape(3))
    t0, t1 = (0, n) if rows is None else rows
    lib = _lib(kind)
    if wrapper:
        work = np.empty(4 * n)
        vwork = np.empty(3 * n)
        rc = lib.orc_block_apply(block, _ptr(r), _ptr(v), _ptr(out), n, t0, t1, eta, a, _ptr(L),
                                 _ptr(work), _ptr(vwork))
    else:
        rc = lib.orc_block_kernel(block, _ptr(r), _ptr(v), _ptr(out), n, t0, t1, eta, a, _ptr(L))
    if rc != 0:
        raise ValueError(f"oracle rejected arguments (rc={rc})")
    return out.reshape(-1)


def uf(r, f, eta, a, periodic_length=None, **kw):
    return block_apply(UF, r, f, eta, a, periodic_length, **kw)


def ut(r, t, eta, a, periodic_length=None, **kw):
    return block_apply(UT, r, t, eta, a, periodic_length, **kw)


def wf(r, f, eta, a, periodic_length=None, **kw):
    return block_apply(WF, r, f, eta, a, periodic_length, **kw)


def wt(r, t, eta, a, periodic_length=None, **kw):
    return block_apply(WT, r, t, eta, a, periodic_length, **kw)


def grand(r, f, t, eta, a, periodic_length=None, **kw):
    """[UF UT; WF WT] [f; t] as the rollers' grand mobility closure
    (quaternion_integrator/quaternion_integrator_rollers.py:1114-1122)."""
    u = uf(r, f, eta, a, periodic_length, **kw) + ut(r, t, eta, a, periodic_length, **kw)
    w = wf(r, f, eta, a, periodic_length, **kw) + wt(r, t, eta, a, periodic_length, **kw)
    return u, w


def pair_matrix(block: int, ri, rj, self_term: bool, eta: float, a: float, kind="strict") -> np.ndarray:
    m = np.empty(9)
    ri = np.ascontiguousarray(ri, dtype=np.float64)
    rj = np.ascontiguousarray(rj, dtype=np.float64)
    rc = _lib(kind).orc_block_pair_matrix(block, _ptr(ri), _ptr(rj), int(self_term), eta, a, _ptr(m))
    if rc:
        raise ValueError("bad block")
    return m.reshape(3, 3)


RADII_BOUNDARIES = {"unbounded": 0, "wall": 1, "free_surface": 2}


def _radii_boundary(wall, boundary) -> int:
    if boundary is None:
        return 1 if wall else 0
    return RADII_BOUNDARIES[boundary]


def radii_apply(source, target, force, radius_source, radius_target, eta: float, periodic_length=None,
                wall: bool = True, rows: Optional[tuple] = None, kind: str = "strict",
                boundary: Optional[str] = None) -> np.ndarray:
    """u = B_t M B_s f for blobs of different radii, source -> target
    (mobility_numba.py:1479-1658 with the wrapper mobility.py:551-591;
    wall=False: the unbounded kernel of mobility.py:593-615;
    boundary='free_surface': mobility_numba.py:1941-2093 behind
    mobility.py:1409-1429, heights as given).  (3 N_t,)."""
    src = np.ascontiguousarray(np.asarray(source, dtype=np.float64).reshape(-1, 3))
    trg = np.ascontiguousarray(np.asarray(target, dtype=np.float64).reshape(-1, 3))
    f = np.ascontiguousarray(np.asarray(force, dtype=np.float64).reshape(-1, 3))
    rs = np.ascontiguousarray(np.asarray(radius_source, dtype=np.float64).reshape(-1))
    rt = np.ascontiguousarray(np.asarray(radius_target, dtype=np.float64).reshape(-1))
    ns, nt = src.shape[0], trg.shape[0]
    if f.shape[0] != ns or rs.size != ns or rt.size != nt:
        raise ValueError("size mismatch")
    L = np.zeros(3) if periodic_length is None else np.ascontiguousarray(
        np.asarray(periodic_length, dtype=np.float64).reshape(3))
    out = np.zeros((nt, 3))
    t0, t1 = (0, nt) if rows is None else rows
    rc = _lib(kind).orc_radii_apply(_ptr(src), _ptr(trg), _ptr(f), _ptr(rs), _ptr(rt), ns, nt, t0, t1, float(eta),
                                    _ptr(L), _radii_boundary(wall, boundary), _ptr(out))
    if rc:
        raise ValueError("bad row range")
    return out.reshape(-1)


def blob_forces(r_vectors, eps: float, debye: float, a: float, periodic_length=None, radii=None,
                cutoff: float = 0.0, rows: Optional[tuple] = None, kind: str = "strict") -> np.ndarray:
    """Blob-blob soft repulsion forces (N, 3): multi_bodies/forces_numba.py
    :12-56 (radii given: :75-122, a = (a_i + a_j)/2; cutoff > 0: only pairs
    with r <= cutoff, the neighbour-list version :141-201)."""
    r = np.ascontiguousarray(np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3))
    n = r.shape[0]
    L = np.zeros(3) if periodic_length is None else np.ascontiguousarray(
        np.asarray(periodic_length, dtype=np.float64).reshape(3))
    rad = None if radii is None else np.ascontiguousarray(np.asarray(radii, dtype=np.float64).reshape(n))
    out = np.zeros((n, 3))
    t0, t1 = (0, n) if rows is None else rows
    rc = _lib(kind).orc_blob_forces(_ptr(r), _ptr(rad) if rad is not None else None, n, t0, t1, float(eps),
                                    float(debye), float(a), float(cutoff), _ptr(L), _ptr(out))
    if rc:
        raise ValueError("bad row range")
    return out


def radii_pair_matrix(ri, rj, a: float, b: float, eta: float, wall: bool = True, kind="strict",
                      boundary: Optional[str] = None) -> np.ndarray:
    """3x3 block of one (target radius a, source radius b) pair, no wrapper."""
    m = np.empty(9)
    ri = np.ascontiguousarray(ri, dtype=np.float64)
    rj = np.ascontiguousarray(rj, dtype=np.float64)
    _lib(kind).orc_radii_pair_matrix(_ptr(ri), _ptr(rj), float(a), float(b), float(eta),
                                     _radii_boundary(wall, boundary), _ptr(m))
    return m.reshape(3, 3)


def dense_block_matrix(block: int, r_vectors, eta: float, a: float, kind="strict",
                       wrapper: bool = True) -> np.ndarray:
    """Dense 3N x 3N matrix of one block assembled from the C pair function
    (small N only).  With wrapper=True the shift/damping regularisation
    B M(z_eff) B is applied like the product wrappers."""
    r = np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
    n = r.shape[0]
    if wrapper and block in WALL_BLOCKS:
        r_eff, Bd, _ = shift_and_damping(r, a)
    else:
        r_eff, Bd = r, np.ones(n)
    M = np.zeros((3 * n, 3 * n))
    for i in range(n):
        for j in range(n):
            M[3 * i:3 * i + 3, 3 * j:3 * j + 3] = pair_matrix(block, r_eff[i], r_eff[j], i == j, eta, a, kind)
    b3 = np.repeat(Bd, 3)
    return (b3[:, None] * M) * b3[None, :]


# ---------------------------------------------------------------------------
# numpy restatements (independent algebra)
# ---------------------------------------------------------------------------

def shift_and_damping(r_vectors, a: float):
    """shift_heights (mobility.py:52-64) and damping_matrix_B (mobility.py:67-84)."""
    r = np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
    r_eff = r.copy()
    r_eff[r[:, 2] <= a, 2] = a
    B = np.ones(r.shape[0])
    low = r[:, 2] < a
    B[low] = r[low, 2] / a
    return r_eff, B, bool(low.any())


def rotne_prager_tensor(r_vectors, eta: float, a: float) -> np.ndarray:
    """Dense unbounded RPY tensor, mobility.py:967-1012 (vectorised form)."""
    r = np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
    d = r[None, :, :] - r[:, None, :]            # d[i,j] = r_j - r_i (sign irrelevant: even)
    dr = np.sqrt((d ** 2).sum(-1))
    fac = 1.0 / (6.0 * np.pi * eta)
    fr = np.zeros_like(dr)
    gr = np.zeros_like(dr)
    far = dr > 2.0 * a
    near = ~far
    zero = dr == 0.0
    near[zero] = False
    fr[far] = fac * (0.75 / dr[far] + a ** 2 / (2.0 * dr[far] ** 3))
    gr[far] = fac * (0.75 / dr[far] ** 3 - 1.5 * a ** 2 / dr[far] ** 5)
    fr[zero] = fac / a
    fr[near] = fac * (1.0 / a - 0.28125 * dr[near] / a ** 2)
    gr[near] = fac * (3.0 / (32.0 * a ** 2 * dr[near]))
    n = r.shape[0]
    M = np.zeros((3 * n, 3 * n))
    for p in range(3):
        for q in range(3):
            M[p::3, q::3] = (fr if p == q else 0.0) + gr * d[:, :, p] * d[:, :, q]
    return M


def single_wall_fluid_mobility(r_vectors, eta: float, a: float) -> np.ndarray:
    """Dense wall mobility in the Swan-Brady A-E form, mobility.py:1018-1116:
    M = RPY(r_eff) + (6 pi eta a)^-1 [A I + B e e^T + C^T e d3^T + D^T d3 e^T
    + E^T d3 d3^T] (plus the self terms), then B_damp M B_damp."""
    r = np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
    r_eff, Bd, overlap = shift_and_damping(r, a)
    n = r.shape[0]
    x, y, z = r_eff[:, 0], r_eff[:, 1], r_eff[:, 2]
    dx = (x[:, None] - x) / a
    dy = (y[:, None] - y) / a
    dz = (z[:, None] + z) / a
    dr = np.sqrt(dx ** 2 + dy ** 2 + dz ** 2)
    h_hat = z[:, None] / (a * dz)
    h = z / a
    ex, ey, ez = dx / dr, dy / dr, dz / dr
    fac = 1.0 / (6.0 * np.pi * eta * a)
    sel = dr > 1e-12 * a
    sel[np.diag_indices(n)] = False
    A = np.zeros_like(dr); Bm = np.zeros_like(dr); C = np.zeros_like(dr)
    D = np.zeros_like(dr); E = np.zeros_like(dr)
    s = sel
    A[s] = -0.25 * (3.0 * (1.0 + 2 * h_hat[s] * (1.0 - h_hat[s]) * ez[s] ** 2) / dr[s]
                    + 2.0 * (1 - 3.0 * ez[s] ** 2) / dr[s] ** 3 - 2.0 * (1 - 5.0 * ez[s] ** 2) / dr[s] ** 5)
    Bm[s] = -0.25 * (3.0 * (1.0 - 6.0 * h_hat[s] * (1.0 - h_hat[s]) * ez[s] ** 2) / dr[s]
                     - 6.0 * (1.0 - 5.0 * ez[s] ** 2) / dr[s] ** 3 + 10.0 * (1.0 - 7.0 * ez[s] ** 2) / dr[s] ** 5)
    C[s] = 0.5 * ez[s] * (3.0 * h_hat[s] * (1.0 - 6.0 * (1.0 - h_hat[s]) * ez[s] ** 2) / dr[s]
                          - 6.0 * (1.0 - 5.0 * ez[s] ** 2) / dr[s] ** 3 + 10.0 * (2.0 - 7.0 * ez[s] ** 2) / dr[s] ** 5)
    D[s] = 0.5 * ez[s] * (3.0 * h_hat[s] / dr[s] - 10.0 / dr[s] ** 5)
    E[s] = -(3.0 * h_hat[s] ** 2 * ez[s] ** 2 / dr[s] + 3.0 * ez[s] ** 2 / dr[s] ** 3
             + (2.0 - 15.0 * ez[s] ** 2) / dr[s] ** 5)
    M = np.zeros((3 * n, 3 * n))
    A_vec = -0.0625 * (9.0 / h - 2.0 / h ** 3 + 1.0 / h ** 5)
    E_vec = -A_vec - 0.125 * (9.0 / h - 4.0 / h ** 3 + 1.0 / h ** 5)
    di = np.arange(n)
    M[3 * di, 3 * di] += A_vec
    M[3 * di + 1, 3 * di + 1] += A_vec
    M[3 * di + 2, 3 * di + 2] += A_vec + E_vec
    M[0::3, 0::3] += A + Bm * ex * ex
    M[0::3, 1::3] += Bm * ex * ey
    M[0::3, 2::3] += Bm * ex * ez + C.T * ex
    M[1::3, 0::3] += Bm * ey * ex
    M[1::3, 1::3] += A + Bm * ey * ey
    M[1::3, 2::3] += Bm * ey * ez + C.T * ey
    M[2::3, 0::3] += Bm * ez * ex + D.T * ex
    M[2::3, 1::3] += Bm * ez * ey + D.T * ey
    M[2::3, 2::3] += A + Bm * ez * ez + C.T * ez + D.T * ez + E.T
    M *= fac
    M += rotne_prager_tensor(r_eff, eta, a)
    if overlap:
        b3 = np.repeat(Bd, 3)
        M = (b3[:, None] * M) * b3[None, :]
    return M


def image_singular_stokeslet(r_vectors, eta: float, a: float) -> np.ndarray:
    """Blake point-particle image system, mobility.py:122-149 with the
    doublet/dipole pieces of mobility.py:174-182."""
    r = np.asarray(r_vectors, dtype=np.float64).reshape(-1, 3)
    n = r.shape[0]
    M = np.zeros((3 * n, 3 * n))
    e3 = np.array([0.0, 0.0, 1.0])
    for j in range(n):
        for k in range(n):
            if j == k:
                M[3 * j:3 * j + 3, 3 * k:3 * k + 3] = np.eye(3) / (6 * np.pi * eta * a)
                continue
            rp = r[j] - r[k]
            rn = np.linalg.norm(rp)
            h = r[k][2]
            rr = r[j] - (r[k] - 2.0 * np.array([0.0, 0.0, h]))
            rrn = np.linalg.norm(rr)
            blk = (np.eye(3) / rn + np.outer(rp, rp) / rn ** 3
                   - (np.eye(3) / rrn + np.outer(rr, rr) / rrn ** 3)) / (8.0 * np.pi)
            # doublet_and_dipole (mobility.py:174-182)
            dd = 2.0 * h * (np.outer(rr, e3) - np.outer(e3, rr)) / (8.0 * np.pi * rrn ** 3)
            dd[:, 0:2] = -dd[:, 0:2]
            M[3 * j:3 * j + 3, 3 * k:3 * k + 3] = blk + dd
    return M


def blake_tensor(x, y, eta: float = 1.0) -> np.ndarray:
    """Velocity at x due to a unit point force at y above a no-slip wall at
    z = 0: the Blake (1971) image system, written from the mathematics (not
    from the reference):

        8 pi eta G = S(r) - S(R) + 2 h P_jk d/dR_k [h R_i/|R|^3 - (d_i3/|R| + R_i R_3/|R|^3)]

    with S(v) = I/|v| + v v^T/|v|^3, r = x - y, R = x - y* (y* the mirror
    image of y), h = y_3 and P = diag(1, 1, -1).  It is pinned by its
    defining properties (zero velocity on the wall, reciprocity
    G(x,y) = G(y,x)^T, incompressibility) in the tests.
    """
    x = np.asarray(x, dtype=np.float64)
    y = np.asarray(y, dtype=np.float64)
    h = y[2]
    r = x - y
    R = x - np.array([y[0], y[1], -h])
    rn, Rn = np.linalg.norm(r), np.linalg.norm(R)
    I = np.eye(3)
    G = I / rn + np.outer(r, r) / rn ** 3 - I / Rn - np.outer(R, R) / Rn ** 3
    P = np.array([1.0, 1.0, -1.0])
    d3 = np.array([0.0, 0.0, 1.0])
    RR = np.outer(R, R)
    # dPhi[i, k] = d/dR_k [h R_i/|R|^3 - d_i3/|R| - R_i R_3/|R|^3]
    dPhi = (h * (I / Rn ** 3 - 3.0 * RR / Rn ** 5)
            + np.outer(d3, R) / Rn ** 3
            - (I * R[2] + np.outer(R, d3)) / Rn ** 3
            + 3.0 * R[2] * RR / Rn ** 5)
    G += 2.0 * h * dPhi * P[None, :]
    return G / (8.0 * np.pi * eta)


def self_mobility_with_rotation(location, eta: float, a: float) -> np.ndarray:
    """Closed-form 6x6 wall self mobility, mobility.py:739-772.  NB: its
    coupling block uses the opposite (left-handed) sign convention to the
    numba UT/WF kernels; compare magnitudes there."""
    h = location[2] / a
    M = np.zeros((6, 6))
    M[:3, :3] = np.eye(3) / (6.0 * np.pi * eta * a)
    for l in range(3):
        M[l, l] += (1.0 / (6.0 * np.pi * eta * a)) * (
            (l != 2) * (-1.0 / 16.0) * (9.0 / h - 2.0 / h ** 3 + 1.0 / h ** 5)
            + (l == 2) * (-1.0 / 8.0) * (9.0 / h - 4.0 / h ** 3 + 1.0 / h ** 5))

    def eps(i, j, k):
        if j == (i + 1) % 3 and k == (j + 1) % 3:
            return 1.0
        if i == (j + 1) % 3 and j == (k + 1) % 3:
            return -1.0
        return 0.0
    for l in range(3):
        for m in range(3):
            M[3 + l, m] += (1.0 / (6.0 * np.pi * eta * a * a)) * (3.0 / 32.0) * h ** -4 * eps(2, l, m)
            M[m, 3 + l] += M[3 + l, m]
    for l in range(3):
        for m in range(3):
            M[3 + l, 3 + m] += ((1.0 / (8.0 * np.pi * eta * a ** 3)) * (l == m)
                                - (1.0 / (6.0 * np.pi * eta * a ** 3)) * (
                                    (15.0 / 64.0) * h ** -3 * (l == m) * (l != 2)
                                    + (3.0 / 32.0) * h ** -3 * (m == 2) * (l == 2)))
    return M


# ---------------------------------------------------------------------------
# rigid bodies
# ---------------------------------------------------------------------------

def rotation_matrix(q) -> np.ndarray:
    """quaternion_integrator/quaternion.py:42-52, q = (s, p1, p2, p3)."""
    s, p = q[0], np.asarray(q[1:4])
    diag = s ** 2 - 0.5
    return 2.0 * np.array([
        [p[0] ** 2 + diag, p[0] * p[1] - s * p[2], p[0] * p[2] + s * p[1]],
        [p[1] * p[0] + s * p[2], p[1] ** 2 + diag, p[1] * p[2] - s * p[0]],
        [p[2] * p[0] - s * p[1], p[2] * p[1] + s * p[0], p[2] ** 2 + diag]])


def body_r_vectors(location, q, ref) -> np.ndarray:
    """body/body.py:64-77."""
    return np.asarray(ref) @ rotation_matrix(q).T + np.asarray(location)


def K_matrix(r_body_relative) -> np.ndarray:
    """K = [J, rot] with (rot_i x) = -(r_i x x), body/body.py:81-113."""
    rv = np.asarray(r_body_relative).reshape(-1, 3)
    nb = rv.shape[0]
    K = np.zeros((3 * nb, 6))
    K[0::3, 0] = 1.0
    K[1::3, 1] = 1.0
    K[2::3, 2] = 1.0
    K[0::3, 4] = rv[:, 2]
    K[0::3, 5] = -rv[:, 1]
    K[1::3, 3] = -rv[:, 2]
    K[1::3, 5] = rv[:, 0]
    K[2::3, 3] = rv[:, 1]
    K[2::3, 4] = -rv[:, 0]
    return K


class RigidSystem:
    """Oracle rigid mobility problem for bodies sharing one reference shape
    list.  ``bodies`` is a list of (location, quaternion, ref_config)."""

    def __init__(self, bodies, eta: float, a: float, periodic_length=None,
                 mobility_vector_prod: Optional[Callable] = None,
                 mobility_blobs: Optional[Callable] = None):
        self.eta, self.a = eta, a
        self.L = np.zeros(3) if periodic_length is None else np.asarray(periodic_length, float)
        self.locations = [np.asarray(b[0], float) for b in bodies]
        self.r_bodies = [body_r_vectors(b[0], b[1], b[2]) for b in bodies]
        self.nb = [rb.shape[0] for rb in self.r_bodies]
        self.offsets = np.concatenate([[0], np.cumsum(self.nb)])
        self.r = np.concatenate(self.r_bodies, axis=0)
        self.Nblobs = int(self.offsets[-1])
        self.Nbodies = len(bodies)
        self.K = [K_matrix(rb - loc) for rb, loc in zip(self.r_bodies, self.locations)]
        self.mvp = mobility_vector_prod or (lambda r, v: uf(r, v, eta, a, self.L))
        self.mblobs = mobility_blobs or (lambda r: single_wall_fluid_mobility(r, eta, a))
        self.size = 3 * self.Nblobs + 6 * self.Nbodies

    def K_dot(self, U):
        U = U.reshape(self.Nbodies, 6)
        return np.concatenate([self.K[k] @ U[k] for k in range(self.Nbodies)])

    def KT_dot(self, lam):
        return np.concatenate([self.K[k].T @ lam[3 * self.offsets[k]:3 * self.offsets[k + 1]]
                               for k in range(self.Nbodies)])

    def apply(self, x):
        """[M lam - K U; -K^T lam], multi_bodies.py:424-477 (no constraints)."""
        nb3 = 3 * self.Nblobs
        lam, U = x[:nb3], x[nb3:]
        out = np.empty_like(x)
        out[:nb3] = self.mvp(self.r, lam) - self.K_dot(U)
        out[nb3:] = -self.KT_dot(lam)
        return out

    def build_pc(self):
        """Per body M^-1 (via Cholesky) and N = pinv(K^T M^-1 K),
        multi_bodies.py:766-785."""
        import scipy.linalg as sla
        self.Minv, self.N = [], []
        for k in range(self.Nbodies):
            M = self.mblobs(self.r_bodies[k])
            c = sla.cho_factor(M)
            Minv = sla.cho_solve(c, np.eye(M.shape[0]))
            self.Minv.append(Minv)
            self.N.append(np.linalg.pinv(self.K[k].T @ sla.cho_solve(c, self.K[k])))

    def build_stochastic_pc(self):
        """build_block_diagonal_preconditioners_det_stoch, multi_bodies.py
        :480-560: besides the deterministic factors, per body scipy's upper
        Cholesky factor U (M_b = U^T U), P_b = U^-1 (its
        M_factorization_blobs_inv) and P_inv_b = U^T (its
        M_factorization_blobs)."""
        import scipy.linalg as sla
        self.build_pc()
        self.P, self.P_inv = [], []
        for k in range(self.Nbodies):
            M = self.mblobs(self.r_bodies[k])
            U = np.triu(sla.cho_factor(M)[0])
            self.P_inv.append(U.T)
            self.P.append(sla.solve_triangular(U, np.eye(M.shape[0])))

    def mobility_pc(self, w):
        """P^T M P w with P = blockdiag(P_b): the mobility_pc closure,
        multi_bodies.py:591-606."""
        res = np.empty_like(w)
        for k in range(self.Nbodies):
            s = slice(3 * self.offsets[k], 3 * self.offsets[k + 1])
            res[s] = self.P[k] @ w[s]
        res2 = self.mvp(self.r, res)
        for k in range(self.Nbodies):
            s = slice(3 * self.offsets[k], 3 * self.offsets[k + 1])
            res[s] = self.P[k].T @ res2[s]
        return res

    def P_inv_mult(self, w):
        """blockdiag(P_inv_b) w: the P_inv_mult closure, multi_bodies.py:608-615."""
        out = np.empty_like(w)
        for k in range(self.Nbodies):
            s = slice(3 * self.offsets[k], 3 * self.offsets[k + 1])
            out[s] = self.P_inv[k] @ w[s]
        return out

    def pc_apply(self, v):
        """multi_bodies.py:825-849."""
        nb3 = 3 * self.Nblobs
        out = np.empty_like(v)
        for k in range(self.Nbodies):
            s = slice(3 * self.offsets[k], 3 * self.offsets[k + 1])
            slip = v[s]
            lt = self.Minv[k] @ slip
            F = v[nb3 + 6 * k: nb3 + 6 * k + 6]
            Y = self.N[k] @ (-F - self.K[k].T @ lt)
            out[s] = self.Minv[k] @ (slip + self.K[k] @ Y)
            out[nb3 + 6 * k: nb3 + 6 * k + 6] = Y
        return out

    def rhs(self, slip, force_torque):
        return np.concatenate([np.asarray(slip, float).ravel(), -np.asarray(force_torque, float).ravel()])

    def solve(self, rhs, tol=1e-8, restart=60, maxiter=1000, x0=None):
        """Normalised-RHS right-preconditioned GMRES, as solve_mobility_problem
        (quaternion_integrator_multi_bodies.py:1515-1536) + utils.gmres (right PC)."""
        if not hasattr(self, "Minv"):
            self.build_pc()
        bnorm = np.linalg.norm(rhs)
        if bnorm == 0:
            return np.zeros_like(rhs), 0
        b = rhs / bnorm
        xw = None if x0 is None else x0 / bnorm
        x, it = gmres_right(self.apply, self.pc_apply, b, tol=tol, restart=restart,
                            maxiter=maxiter, x0=xw)
        return x * bnorm, it


def gmres_right(A: Callable, Pinv: Callable, b: np.ndarray, tol=1e-8, restart=60, maxiter=1000,
                x0: Optional[np.ndarray] = None):
    """Restarted GMRES on A P^-1 y = r0 with x = x0 + P^-1 y; stop when
    ||b - A x|| <= tol ||b|| (general_application_utils.py:514-635 semantics,
    scipy gmres rtol=tol atol=0), with the warm start applied by residual
    correction.  Returns (x, total inner iterations)."""
    x = np.zeros_like(b) if x0 is None else x0.copy()
    bnorm = np.linalg.norm(b)
    total = 0
    for _outer in range(maxiter):
        r = b - A(x)
        beta = np.linalg.norm(r)
        if beta <= tol * bnorm:
            return x, total
        n = b.size
        m = restart
        V = np.zeros((m + 1, n))
        H = np.zeros((m + 1, m))
        V[0] = r / beta
        g = np.zeros(m + 1)
        g[0] = beta
        cs, sn = np.zeros(m), np.zeros(m)
        j_used = 0
        for j in range(m):
            w = A(Pinv(V[j]))
            for i in range(j + 1):          # modified Gram-Schmidt
                H[i, j] = V[i] @ w
                w = w - H[i, j] * V[i]
            H[j + 1, j] = np.linalg.norm(w)
            if H[j + 1, j] > 0:
                V[j + 1] = w / H[j + 1, j]
            for i in range(j):
                t = cs[i] * H[i, j] + sn[i] * H[i + 1, j]
                H[i + 1, j] = -sn[i] * H[i, j] + cs[i] * H[i + 1, j]
                H[i, j] = t
            den = np.hypot(H[j, j], H[j + 1, j])
            cs[j], sn[j] = H[j, j] / den, H[j + 1, j] / den
            H[j, j] = den
            H[j + 1, j] = 0.0
            g[j + 1] = -sn[j] * g[j]
            g[j] = cs[j] * g[j]
            total += 1
            j_used = j + 1
            if abs(g[j + 1]) <= tol * bnorm:
                break
        yk = np.linalg.solve(np.triu(H[:j_used, :j_used]), g[:j_used])
        x = x + Pinv(V[:j_used].T @ yk)
        if total >= maxiter:
            break
    return x, total


def lanczos_sqrt(mobility_mult: Callable, z: np.ndarray, factor: float = 1.0,
                 tolerance: float = 1e-6, max_iter: int = 1000):
    """factor M^1/2 z by Lanczos with full re-orthogonalisation and the
    noise-difference stop rule, stochastic_forcing.py:112-262."""
    dim = z.size
    if factor == 0.0:
        return np.zeros(dim), 0
    v = [z.astype(np.float64) / np.linalg.norm(z)]
    vnorm = np.linalg.norm(z)
    h_sup, h_diag = [], []
    noise_old = np.zeros(dim)
    noise = noise_old
    for i in range(max_iter + 1):
        w = np.asarray(mobility_mult(v[i]), dtype=np.float64).reshape(dim)
        if i > 0:
            w = w - h_sup[i - 1] * v[i - 1]
        h_diag.append(w @ v[i])
        w = w - h_diag[i] * v[i]
        h_sup.append(np.linalg.norm(w))
        if h_sup[i] > 0:
            w = w / h_sup[i]
        else:
            w = np.zeros(dim)
            w[0] = 1.0
        m = len(h_diag)
        H = np.diag(h_diag) + np.diag(h_sup[:m - 1], -1) + np.diag(h_sup[:m - 1], 1)
        lam, Q = np.linalg.eigh(H)
        sq = np.where(lam > 0, np.sqrt(np.maximum(lam, 0.0)), 0.0)
        coef = Q @ (vnorm * factor * sq * Q[0, :])
        noise = np.asarray(v).T @ coef
        for row in v:
            w = w - (row @ w) * row
        v.append(w)
        if i > 0:
            on = np.linalg.norm(noise_old)
            dn = np.linalg.norm(noise - noise_old)
            if dn / max(on, np.finfo(float).eps) < tolerance:
                return noise, i
        noise_old = noise.copy()
    return noise, max_iter
