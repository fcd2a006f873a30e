"""CPU checks of the boundary and host logic (no GPU, no compute calls):

* librmw_hip.so loads and exports every function include/rmw_hip.h declares,
  and the ctypes prototypes cover exactly that set;
* the product's pair-term header, compiled for the host, reproduces the
  oracle's 3x3 blocks (FP64 to 1e-13, FP32 to the FP32 budget);
* workload generators (42-blob shell vs the reference's vertex file).
"""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

from oracle import oracle as O

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(REPO, "include", "rmw_hip.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(rmw_[a-z0-9_]+)\s*\(", text)))


def test_header_symbols_exported():
    from rigidmultiblobswall_amd import _lib
    syms = _declared_symbols()
    assert len(syms) >= 8
    lib = _lib.load()
    for s in syms:
        assert hasattr(lib, s), f"{s} declared in include/rmw_hip.h but not exported"
    assert sorted(_lib.PROTOTYPES) == syms


def test_library_has_no_oracle_dependency():
    """The product library must not link or name the CPU oracle."""
    from rigidmultiblobswall_amd import _lib
    data = open(_lib.library_path(), "rb").read()
    assert b"orc_block" not in data and b"liboracle" not in data


def test_error_path_reports_without_gpu():
    """Argument validation runs before any device call and reports a code
    and a message through rmw_last_error (size query with a bad mask)."""
    from rigidmultiblobswall_amd import _lib
    lib = _lib.load()
    nbytes = ctypes.c_size_t()
    rc = lib.rmw_wall_rpy_workspace_size(0, 0, 10, 1, None, ctypes.byref(nbytes))
    assert rc == -1
    assert b"mask" in lib.rmw_last_error()
    rc = lib.rmw_wall_rpy_workspace_size(1, 7, 10, 1, None, ctypes.byref(nbytes))
    assert rc == -1 and b"precision" in lib.rmw_last_error()


def _host_lib():
    path = os.path.join(REPO, "tests", "helpers", "libpair_terms_host.so")
    if not os.path.exists(path):
        pytest.fail("tests/helpers/libpair_terms_host.so not built (run __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    dp = ctypes.POINTER(ctypes.c_double)
    for fn in (lib.pt_pair_matrix_f64, lib.pt_pair_matrix_f32):
        fn.argtypes = [ctypes.c_int, dp, dp, ctypes.c_int, ctypes.c_double, ctypes.c_double, dp]
    return lib


@pytest.mark.parametrize("block", [O.UF, O.UT, O.WF, O.WT])
def test_pair_term_header_matches_oracle(block):
    lib = _host_lib()
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rng = np.random.default_rng(block)
    a, eta = 0.3, 0.7
    worst64 = worst32 = 0.0
    for t in range(3000):
        ri = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        rj = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        self_ = int(t % 97 == 0)
        if self_:
            rj = ri.copy()
        m64, m32 = np.empty(9), np.empty(9)
        assert lib.pt_pair_matrix_f64(block, P(ri), P(rj), self_, eta, a, P(m64)) == 0
        assert lib.pt_pair_matrix_f32(block, P(ri), P(rj), self_, eta, a, P(m32)) == 0
        ref = O.pair_matrix(block, ri, rj, bool(self_), eta, a).ravel()
        sc = np.abs(ref).max()
        worst64 = max(worst64, np.abs(m64 - ref).max() / sc)
        worst32 = max(worst32, np.abs(m32 - ref).max() / sc)
    assert worst64 < 1e-13
    # FP32 of the pair formula itself (d exact): bounded by the far-field
    # Stokeslet/image cancellation, ~eps_32 * r^2 / (z_i z_j).
    assert worst32 < 2e-5


# explicit matrix forms (blocks 8-11) -> oracle block of the same pair, and
# oracle block of the swapped pair that their transposes must equal
EXPLICIT = {8: (O.UF, O.UF), 9: (O.UT, O.WF), 10: (O.WF, O.UT), 11: (O.WT, O.WT)}


@pytest.mark.parametrize("eblock", sorted(EXPLICIT))
def test_explicit_block_matrices_match_oracle(eblock):
    """mat_uf / mat_ut / mat_wf / mat_wt (the blocks the symmetric kernels
    form once per pair when several right-hand sides share it) equal the
    oracle's block of (i, j), and read by columns its block of (j, i)."""
    lib = _host_lib()
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rng = np.random.default_rng(31 + eblock)
    a, eta = 0.3, 0.7
    block, swapped = EXPLICIT[eblock]
    worst = worst_t = worst32 = 0.0
    for t in range(2000):
        ri = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        rj = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        if t % 5 == 0:   # overlapping pairs (near RPY branch)
            rj = ri + rng.uniform(-0.25, 0.25, 3)
            rj[2] = max(rj[2], a)
        m, m32 = np.empty(9), np.empty(9)
        assert lib.pt_pair_matrix_f64(eblock, P(ri), P(rj), 0, eta, a, P(m)) == 0
        assert lib.pt_pair_matrix_f32(eblock, P(ri), P(rj), 0, eta, a, P(m32)) == 0
        ref = O.pair_matrix(block, ri, rj, False, eta, a).ravel()
        ref_t = O.pair_matrix(swapped, rj, ri, False, eta, a).T.ravel()
        sc = np.abs(ref).max()
        worst = max(worst, np.abs(m - ref).max() / sc)
        worst_t = max(worst_t, np.abs(m - ref_t).max() / sc)
        worst32 = max(worst32, np.abs(m32 - ref).max() / sc)
    assert worst < 1e-13 and worst_t < 1e-13
    assert worst32 < 2e-5


# transposed block of the pair (i, j) -> (oracle block of the pair (j, i),
# untransposed block of (i, j) that it must differ from)
TRANSPOSES = {4: (O.UF, 0), 5: (O.WF, 1), 6: (O.UT, 2), 7: (O.WT, 3)}


@pytest.mark.parametrize("tblock", sorted(TRANSPOSES))
def test_transposed_blocks_are_the_swapped_pair(tblock):
    """apply_*_t on the pair (i, j) is the oracle's block of (j, i) (UF_ji =
    UF_ij^T, UT_ji = WF_ij^T, WF_ji = UT_ij^T, WT_ji = WT_ij^T): the
    identities the symmetric kernels use to serve both blobs of a pair."""
    lib = _host_lib()
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    rng = np.random.default_rng(11 + tblock)
    a, eta = 0.3, 0.7
    oblock, plain = TRANSPOSES[tblock]
    worst, differs = 0.0, 0.0
    for t in range(2000):
        ri = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        rj = np.array([*rng.uniform(-3, 3, 2), rng.uniform(a, 3)])
        m, m0 = np.empty(9), np.empty(9)
        assert lib.pt_pair_matrix_f64(tblock, P(ri), P(rj), 0, eta, a, P(m)) == 0
        ref = O.pair_matrix(oblock, rj, ri, False, eta, a).ravel()
        worst = max(worst, np.abs(m - ref).max() / np.abs(ref).max())
        lib.pt_pair_matrix_f64(plain, P(ri), P(rj), 0, eta, a, P(m0))
        differs = max(differs, np.abs(m0 - m).max() / np.abs(m).max())
    assert worst < 1e-13
    assert differs > 1e-3          # the transpose is not the block itself


def test_icosphere_matches_reference_vertex_file():
    from rigidmultiblobswall_amd.workloads import icosphere_42
    ref = np.loadtxt(os.path.join(REPO, "tests", "golden", "shell_N_42_Rg_1_Rh_1_1220.vertex"), skiprows=1)
    P = icosphere_42()
    d1 = np.sort(np.linalg.norm(P[:, None] - P[None], axis=-1).ravel())
    d2 = np.sort(np.linalg.norm(ref[:, None] - ref[None], axis=-1).ravel())
    assert ref.shape == (42, 3)
    assert np.abs(d1 - d2).max() < 1e-12
    assert abs(d1[42] - 0.5465330578253432) < 1e-12   # header: min spacing


def test_workload_shapes_and_ranges():
    from rigidmultiblobswall_amd import workloads as W
    c1 = W.c1_uf_matvec()
    assert c1["r"].shape == (2048, 3) and c1["r"][:, 2].min() >= 0.2
    c2 = W.c2_rigid_spheres()
    assert c2["r"].shape == (10752, 3)
    d = np.linalg.norm(c2["centres"][:, None] - c2["centres"][None], axis=-1)
    assert d[np.triu_indices(256, 1)].min() >= 2 * (1 + 0.243)
    c5 = W.c5_ragged_rollers()
    sizes = np.diff(c5["offsets"])
    assert sizes.size == 64 and sizes.min() >= 128 and sizes.max() <= 4096


def test_output_buffer_check_host():
    """The host-side check of caller-supplied output buffers (size, dtype,
    device, layout) refuses a wrong buffer before any launch."""
    import torch
    from rigidmultiblobswall_amd.wall_rpy import _check_out
    cpu = torch.device("cpu")
    ok = torch.empty((1, 10, 3), dtype=torch.float64)
    assert _check_out(ok, 30, cpu, "out") is ok
    for bad, count in ((torch.empty((1, 9, 3), dtype=torch.float64), 30),
                       (torch.empty((1, 10, 3), dtype=torch.float32), 30),
                       (torch.empty((1, 3, 10), dtype=torch.float64).transpose(1, 2), 30),
                       (np.empty((10, 3)), 30)):
        with pytest.raises(ValueError):
            _check_out(bad, count, cpu, "out")


def test_tuning_flags_layout_and_merge():
    """rmw_tuning's flags field sits where include/rmw_hip.h puts it, and
    the finite-check flag is merged into any launch shape a call resolves
    (none, a dict, a cached Tuning) without changing the shape."""
    import ctypes
    from rigidmultiblobswall_amd import _lib
    assert ctypes.sizeof(_lib.Tuning) == 24 and _lib.Tuning.flags.offset == 20
    hdr = open(os.path.join(os.path.dirname(os.path.dirname(__file__)), "include", "rmw_hip.h")).read()
    assert re.search(r"#define RMW_FLAG_CHECK_FINITE 1\b", hdr) and re.search(r"#define RMW_E_NOTFINITE \(-4\)", hdr)
    t = _lib.with_flags(None, _lib.RMW_FLAG_CHECK_FINITE)
    assert (t.chunk_len, t.max_wgs, t.rhs_group, t.symmetric, t.flags) == (0, 0, 0, 0, 1)
    t = _lib.with_flags({"chunk_len": 128, "symmetric": 1}, 1)
    assert (t.chunk_len, t.symmetric, t.flags) == (128, 1, 1)
    base = _lib.Tuning(704, 0, 4, 1, 0)
    t = _lib.with_flags(base, 1)
    assert (t.chunk_len, t.rhs_group, t.symmetric, t.flags) == (704, 4, 1, 1) and base.flags == 0


def test_masked_geometry_gives_exact_zeros():
    """A pair the symmetric kernels do not take (the self pair, the lower
    half of a diagonal tile, padding beyond the suspension) is masked
    through its geometry (make_geom_on: 1/|d| = 1/|R| = 0, far-field
    branches): every block of every boundary, dyadic and explicit, both
    orientations, is then exactly zero -- also at d = 0 and for near pairs --
    and a taken pair is unchanged."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(REPO, "tests", "helpers", "libpair_terms_host.so"))
    dp = ctypes.POINTER(ctypes.c_double)
    P = lambda x: x.ctypes.data_as(dp)   # noqa: E731
    for fn in (lib.pt_masked_blocks_f32, lib.pt_masked_blocks_f64):
        fn.argtypes = [dp, dp, ctypes.c_int, dp]
    for fn32, fn in ((True, lib.pt_masked_blocks_f32), (False, lib.pt_masked_blocks_f64)):
        for ri, rj in ((np.array([1.0, 2.0, 1.5]), np.array([1.0, 2.0, 1.5])),      # d = 0
                       (np.array([0.3, 0.1, 0.9]), np.array([0.0, -0.4, 1.2])),     # overlapping
                       (np.array([5.0, 2.0, 3.5]), np.array([-1.0, 0.4, 0.8]))):    # far
            off = np.full(324, np.nan)
            on = np.full(324, np.nan)
            assert fn(P(ri), P(rj), 0, P(off)) == 324
            assert np.all(off == 0.0), (fn32, ri, rj)
            if np.any(ri != rj):
                assert fn(P(ri), P(rj), 1, P(on)) == 324
                assert np.all(np.isfinite(on)) and np.abs(on).max() > 0.0


def test_work_item_schedule_covers_every_tile_pair_once():
    """The work-item decode of the symmetric kernels (rmw_tile_pairs.h,
    host build): items of C 64-source blocks (C = 1, 2 or a multiple of 4)
    and of C whole tiles cover every unordered (target tile, source block /
    tile >= target) pair exactly once and nothing else, for every size up to
    130 tiles; the closed-form run count equals the direct sum."""
    import ctypes
    lib = ctypes.CDLL(os.path.join(REPO, "tests", "helpers", "libpair_terms_host.so"))
    for fn in (lib.tp_block_cover_errors, lib.tp_tile_cover_errors, lib.tp_ceil_sum):
        fn.argtypes = [ctypes.c_int64, ctypes.c_int64]
        fn.restype = ctypes.c_int64
    for C in (1, 2, 4, 8, 12, 44):
        bad = [n_t for n_t in range(1, 131) if lib.tp_block_cover_errors(n_t, C)]
        assert not bad, (C, bad[:5])
    for C in range(1, 21):
        bad = [n_t for n_t in range(1, 101) if lib.tp_tile_cover_errors(n_t, C)]
        assert not bad, (C, bad[:5])
    for C in range(1, 41):
        for M in range(0, 301, 7):
            assert lib.tp_ceil_sum(M, C) == sum(-(-m // C) for m in range(1, M + 1))


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with status 2 and a message,
    before touching a GPU."""
    import subprocess
    out = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                         text=True, timeout=120, cwd=REPO)
    assert out.returncode == 2
    assert "--gpus 2 is not supported" in out.stderr
