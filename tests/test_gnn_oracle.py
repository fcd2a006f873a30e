"""CPU pins of the warm-start network oracle (FP64): the analytic backward
equals central finite differences of the loss for every parameter tensor."""
import numpy as np
import pytest

from oracle import gnn_oracle as G


def _params(rng, nout, nin=G.NIN):
    P = {}
    nc = (nin + G.NMSG + 15) // 16 * 16
    shapes = [("We1", (G.HE, G.EIN)), ("be1", (G.HE,)), ("We2", (G.NMSG, G.HE)), ("be2", (G.NMSG,)),
              ("Wn1", (G.HN, nc)), ("bn1", (G.HN,)), ("Wn2", (G.HN, G.HN)), ("bn2", (G.HN,)),
              ("Wn3", (nout, G.HN)), ("bn3", (nout,))]
    for n, s in shapes:
        P[n] = rng.standard_normal(s) * (1.0 / np.sqrt(s[1]) if len(s) == 2 else 0.1)
    P["Wn1"][:, nin + G.NMSG:] = 0.0
    return P


@pytest.mark.parametrize("nin", [10, 22])
def test_backward_matches_finite_differences(nin):
    """The 10-input network and the 22-input one of the motion features
    (node width 80 and 96)."""
    rng = np.random.default_rng(0)
    B, nout = 3, 12
    P = _params(rng, nout, nin)
    nc = P["Wn1"].shape[1]
    Xn = rng.standard_normal((B, nin))
    Xe = rng.standard_normal((B * G.DEG, G.EIN))
    T = rng.standard_normal((B, nout))
    L, g = G.loss_and_grad(P, Xn, Xe, T)
    h = 1e-6
    for name, W in P.items():
        flat = W.reshape(-1)
        picks = rng.choice(flat.size, size=min(6, flat.size), replace=False)
        for i in picks:
            if name == "Wn1" and (i % nc) >= nin + G.NMSG:
                continue
            old = flat[i]
            flat[i] = old + h
            lp = G.loss(P, Xn, Xe, T)
            flat[i] = old - h
            lm = G.loss(P, Xn, Xe, T)
            flat[i] = old
            fd = (lp - lm) / (2 * h)
            an = g[name].reshape(-1)[i]
            assert abs(fd - an) <= 1e-6 * max(1e-2, abs(fd)), (name, i, fd, an)


def test_bf16_round():
    x = np.array([1.0, 1.00390625, 1.01171875, -3.14159, 1e-3])
    r = G.bf16_round(x)
    # 1 + 2^-8 and 1 + 3 2^-8 are ties: round to the even mantissa
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.015625
    assert abs(r[3] + 3.140625) < 1e-12
