"""GPU checks of the FP64 Krylov vector kernels (dot products against many
vectors, CGS2 Arnoldi step) against torch FP64 on the same data."""
import ctypes

import numpy as np
import pytest
import torch

from rigidmultiblobswall_amd import _lib, krylov

pytestmark = pytest.mark.gpu


def _stream():
    return torch.cuda.current_stream().cuda_stream


# lengths that take each launch form: one element per thread, 2, 4 and 8
# per thread with the vector slice held in registers, and the slice-walking
# kernel beyond 512 blocks of 2048 elements
@pytest.mark.parametrize("n,nvec", [(1, 1), (100, 5), (4096, 64), (135174, 65), (300000, 130), (600000, 9),
                                    (1000000, 4), (2500000, 3)])
def test_dot_many_matches_torch(gpu, n, nvec):
    print(f"case dot_many n={n} nvec={nvec}", flush=True)
    g = torch.Generator(device="cpu").manual_seed(n + nvec)
    V = torch.randn((nvec, n), generator=g, dtype=torch.float64).to(gpu)
    w = torch.randn(n, generator=g, dtype=torch.float64).to(gpu)
    ops = krylov.KrylovOps(n, nvec + 1, gpu)
    out = torch.empty(nvec, dtype=torch.float64, device=gpu)
    ops.dot_many(V, nvec, w, out)
    ref = V @ w
    scale = (V.abs() @ w.abs()).max().item()
    assert (out - ref).abs().max().item() <= 1e-14 * scale * max(1.0, np.sqrt(n) / 64)
    again = torch.empty_like(out)
    ops.dot_many(V, nvec, w, again)
    assert torch.equal(out, again)          # fixed summation order
    assert abs(ops.norm(w) - float(torch.linalg.norm(w))) <= 1e-13 * float(torch.linalg.norm(w))


@pytest.mark.parametrize("n", [1000, 135174])
def test_arnoldi_cgs2_builds_orthonormal_basis(gpu, n):
    print(f"case arnoldi n={n}", flush=True)
    m = 24
    g = torch.Generator(device="cpu").manual_seed(7)
    A = torch.randn((n,), generator=g, dtype=torch.float64).to(gpu)
    ops = krylov.KrylovOps(n, m + 2, gpu)
    V = torch.zeros((m + 1, n), dtype=torch.float64, device=gpu)
    V[0] = torch.randn(n, generator=g, dtype=torch.float64).to(gpu)
    V[0] /= torch.linalg.norm(V[0])
    h = torch.empty(m + 2, dtype=torch.float64, device=gpu)
    hs = torch.empty(m + 2, dtype=torch.float64, device=gpu)
    H = torch.zeros((m + 1, m), dtype=torch.float64, device=gpu)
    for j in range(m):
        w = A * V[j] + 0.1 * torch.roll(V[j], 1)      # a non-symmetric operator
        w_in = w.clone()
        ops.arnoldi(V, j, w, h, hs)
        H[: j + 2, j] = h[: j + 2]
        # w_in = sum_i h_i V_i + h_{j+1} V_{j+1}
        rec = V[: j + 2].T @ h[: j + 2]
        assert float(torch.linalg.norm(rec - w_in)) <= 1e-12 * float(torch.linalg.norm(w_in))
    G = V @ V.T
    assert float((G - torch.eye(m + 1, dtype=torch.float64, device=gpu)).abs().max()) <= 1e-13
