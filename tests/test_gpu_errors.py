"""Error reporting of the wall-RPY products.

Two distinct blobs at the same position make the RPY pair tensor 0/0: the
reference returns the NaN silently (mobility/mobility_numba.py:88, a TODO
there).  The products here do the same by default; with the finite check
(RMW_FLAG_CHECK_FINITE, `check_finite=True`) the call reports the
non-finite outputs as RMW_E_NOTFINITE instead.
"""
import numpy as np
import pytest
import torch

from rigidmultiblobswall_amd import _lib, mobility, wall_rpy
from rigidmultiblobswall_amd._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT

pytestmark = pytest.mark.gpu


def _suspension(n, seed, coincident=True):
    rng = np.random.default_rng(seed)
    side = 300.0 * np.sqrt(n / 32768.0)
    r = np.column_stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(0.8, 2.0, n)])
    if coincident:
        r[n // 3] = r[n // 2]
    return r


@pytest.mark.parametrize("n", [500, 8192], ids=["ordered", "symmetric"])
@pytest.mark.parametrize("prec", ["fp32_blocked", "fp64"])
def test_coincident_blobs_reported(gpu, n, prec):
    r = _suspension(n, 3)
    f = np.random.default_rng(4).standard_normal(3 * n)
    print(f"case coincident blobs n={n} {prec}: {wall_rpy.describe(RMW_UF, n, 1, prec)['kernel']}", flush=True)
    # default: silent NaN, as the reference
    u, _ = wall_rpy.apply(RMW_UF, r, force=f, eta=1e-3, a=0.656, precision=prec)
    bad = ~torch.isfinite(u).all(dim=1).cpu().numpy()
    assert bad[n // 3] and bad[n // 2]
    with pytest.raises(_lib.RmwError) as ei:
        wall_rpy.apply(RMW_UF, r, force=f, eta=1e-3, a=0.656, precision=prec, check_finite=True)
    assert ei.value.code == _lib.RMW_E_NOTFINITE
    with pytest.raises(_lib.RmwError) as ei:
        mobility.single_wall_mobility_trans_times_force_hip(r, f, 1e-3, 0.656, precision=prec, check_finite=True)
    assert ei.value.code == _lib.RMW_E_NOTFINITE


def test_grand_mobility_coincident_blobs_reported(gpu):
    n = 600
    r = _suspension(n, 5)
    ft = np.random.default_rng(6).standard_normal(6 * n)
    print("case coincident blobs, grand mobility", flush=True)
    with pytest.raises(_lib.RmwError) as ei:
        mobility.single_wall_grand_mobility_hip(r, ft, 1e-3, 0.656, precision="fp32_blocked", check_finite=True)
    assert ei.value.code == _lib.RMW_E_NOTFINITE


def test_finite_check_passes_valid_products(gpu):
    """With distinct positions the checked call returns the same product."""
    for n in (500, 8192):
        r = _suspension(n, 7, coincident=False)
        f = np.random.default_rng(8).standard_normal((2, n, 3))
        t = np.random.default_rng(9).standard_normal((2, n, 3))
        print(f"case finite check on a valid suspension n={n}", flush=True)
        mask = RMW_UF | RMW_UT | RMW_WF | RMW_WT
        u0, w0 = wall_rpy.apply(mask, r, force=f, torque=t, eta=1e-3, a=0.656, k=2, precision="fp32_blocked")
        u1, w1 = wall_rpy.apply(mask, r, force=f, torque=t, eta=1e-3, a=0.656, k=2, precision="fp32_blocked",
                                check_finite=True)
        assert torch.equal(u0, u1) and torch.equal(w0, w1)


def test_noalloc_refuses_the_finite_check(gpu):
    import ctypes
    lib = _lib.load()
    n = 64
    r = torch.from_numpy(_suspension(n, 1, coincident=False)).to(gpu)
    f = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    u = torch.zeros((n, 3), dtype=torch.float64, device=gpu)
    t = _lib.Tuning(0, 0, 0, 0, _lib.RMW_FLAG_CHECK_FINITE)
    rc = lib.rmw_wall_rpy_apply_noalloc(RMW_UF, 3, r.data_ptr(), f.data_ptr(), None, u.data_ptr(), None, n, 1, 0,
                                        1.0, 0.5, (ctypes.c_double * 3)(0, 0, 0), ctypes.byref(t),
                                        torch.cuda.current_stream().cuda_stream)
    assert rc == _lib.RMW_E_ARG
