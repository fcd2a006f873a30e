"""FP64 CPU oracle of the warm-start graph network (test infrastructure only).

Forward and analytic backward of the same architecture as the GPU
implementation (edge MLP 16 -> 64 -> 64, sum over 8 edges, node MLP
(nin + 64, padded to 16) -> 128 -> 128 -> out, SiLU), in numpy float64.  The tests pin the
analytic backward with central finite differences and compare the GPU's
BF16 MFMA forward and gradients with it.  The GPU rounds operands to BF16;
`bf16_round` lets a test apply the same rounding to inputs and weights.
"""
import numpy as np

DEG, NIN, EIN, HE, NMSG, HN, NC = 8, 10, 16, 64, 64, 128, 80


def bf16_round(x):
    """Round-to-nearest-even to bfloat16, returned as float64."""
    a = np.asarray(x, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return (u.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(P, Xn, Xe):
    """P: dict of float64 parameters (W [out][in]).  Returns (Y, cache)."""
    B, nin = Xn.shape
    Ze1 = Xe @ P["We1"].T + P["be1"]
    He1 = silu(Ze1)
    Msg = He1 @ P["We2"].T + P["be2"]
    Xc = np.zeros((B, P["Wn1"].shape[1]))     # node inputs, message, zero padding
    Xc[:, :nin] = Xn
    Xc[:, nin:nin + NMSG] = Msg.reshape(B, DEG, NMSG).sum(axis=1)
    Zn1 = Xc @ P["Wn1"].T + P["bn1"]
    Hn1 = silu(Zn1)
    Zn2 = Hn1 @ P["Wn2"].T + P["bn2"]
    Hn2 = silu(Zn2)
    Y = Hn2 @ P["Wn3"].T + P["bn3"]
    return Y, dict(Xe=Xe, Ze1=Ze1, He1=He1, Xc=Xc, Zn1=Zn1, Hn1=Hn1, Zn2=Zn2, Hn2=Hn2, nin=nin)


def loss(P, Xn, Xe, T):
    Y, _ = forward(P, Xn, Xe)
    return float(((Y - T) ** 2).sum() / Xn.shape[0])


def backward(P, cache, dY):
    B = dY.shape[0]
    g = {}
    g["Wn3"] = dY.T @ cache["Hn2"]
    g["bn3"] = dY.sum(0)
    dH2 = dY @ P["Wn3"]
    dZ2 = dH2 * dsilu(cache["Zn2"])
    g["Wn2"] = dZ2.T @ cache["Hn1"]
    g["bn2"] = dZ2.sum(0)
    dH1 = dZ2 @ P["Wn2"]
    dZ1 = dH1 * dsilu(cache["Zn1"])
    nin = cache["nin"]
    g["Wn1"] = dZ1.T @ cache["Xc"]
    g["Wn1"][:, nin + NMSG:] = 0.0
    g["bn1"] = dZ1.sum(0)
    dXc = dZ1 @ P["Wn1"]
    dM = np.repeat(dXc[:, nin:nin + NMSG], DEG, axis=0)
    g["We2"] = dM.T @ cache["He1"]
    g["be2"] = dM.sum(0)
    dHe = dM @ P["We2"]
    dZe = dHe * dsilu(cache["Ze1"])
    g["We1"] = dZe.T @ cache["Xe"]
    g["be1"] = dZe.sum(0)
    return g


def loss_and_grad(P, Xn, Xe, T):
    Y, cache = forward(P, Xn, Xe)
    B = Xn.shape[0]
    dY = 2.0 * (Y - T) / B
    return float(((Y - T) ** 2).sum() / B), backward(P, cache, dY)
