"""Learned GMRES warm start: a per-body graph network (BF16 MFMA kernels).

For every rigid body b the network predicts the part of the solution that
a base guess misses, in the body's own frame:

    x0 = base + R_b (network output) * fscale

(the base is P^-1 b, the exact solution of the isolated-body problem, or
the point-field block-Jacobi step P^-1 (b - [u; 0]) that adds the flow of
the other bodies' net forces; the network adds the correction due to the
8 nearest neighbours.)  The prediction is
only the GMRES initial guess: GMRES to the same tolerance remains the judge,
and a non-finite prediction is discarded (x0 = 0).

Architecture (~48k parameters, W stored [out][in]):
  edge MLP   e = W_e2 SiLU(W_e1 x_e + b_e1) + b_e2          16 -> 64 -> 64
  message    m_b = sum of e over the 8 edges of b
  node MLP   y = W_n3 SiLU(W_n2 SiLU(W_n1 [x_b, m_b] + b_n1) + b_n2) + b_n3
                                                     (10 + 64, padded 80) -> 128 -> 128 -> 3 nb + 6
Every layer and its backward pass is one call of the BF16 MFMA GEMM
(FP32 accumulation); Adam keeps FP32 master weights and refreshes the BF16
copies.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import Dict, List, Optional, Tuple

import numpy as np
import torch

from . import _lib
from . import _stream as _stream_mod

DEG, NIN, EIN, HE, NMSG, HN = 8, 10, 16, 64, 64, 128
NC = 80  # node input width: 10 + 64 padded to a multiple of 16
NIN_MOTION = 22   # node inputs of the network that corrects the previous step's solution


def node_width(nin: int) -> int:
    """Node-layer input width: nin features + the message, padded to 16."""
    return (nin + NMSG + 15) // 16 * 16


def _stream():
    return _stream_mod.current()


def param_shapes(nout: int, nin: int = NIN) -> List[Tuple[str, Tuple[int, ...]]]:
    return [("We1", (HE, EIN)), ("be1", (HE,)), ("We2", (NMSG, HE)), ("be2", (NMSG,)),
            ("Wn1", (HN, node_width(nin))), ("bn1", (HN,)), ("Wn2", (HN, HN)), ("bn2", (HN,)),
            ("Wn3", (nout, HN)), ("bn3", (nout,))]


def init_params(nout: int, seed: int = 0, nin: int = NIN) -> Dict[str, np.ndarray]:
    """Glorot-uniform weights, zero biases (FP32)."""
    rng = np.random.default_rng(seed)
    out = {}
    for name, shp in param_shapes(nout, nin):
        if len(shp) == 2:
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            w = rng.uniform(-lim, lim, shp).astype(np.float32)
            if name == "Wn1":
                w[:, nin + NMSG:] = 0.0   # padding columns multiply zero inputs
            if name == "Wn3":
                w *= 0.1                  # small initial correction
            out[name] = w
        else:
            out[name] = np.zeros(shp, np.float32)
    return out


@dataclass
class Batch:
    Xn: torch.Tensor   # (B, nin) f32
    Xe: torch.Tensor   # (8B, 16) f32
    T: Optional[torch.Tensor] = None  # (B, nout) f32 standardised targets


class GraphNet:
    """Device-resident network: FP32 master parameters in one flat buffer,
    BF16 copies, gradients and Adam moments alongside."""

    def __init__(self, blobs_per_body: int, device=None, seed: int = 0, nin: int = NIN):
        self.lib = _lib.load()
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.nbb = blobs_per_body
        self.nout = 3 * blobs_per_body + 6
        self.nin, self.nc = int(nin), node_width(nin)
        shapes = param_shapes(self.nout, self.nin)
        self.sizes = [int(np.prod(s)) for _, s in shapes]
        self.offsets = np.concatenate([[0], np.cumsum(self.sizes)]).astype(np.int64)
        self.nparam = int(self.offsets[-1])
        init = init_params(self.nout, seed, self.nin)
        flat = np.concatenate([init[n].ravel() for n, _ in shapes]).astype(np.float32)
        self.w = torch.from_numpy(flat).to(self.dev)
        self.wb = torch.empty(self.nparam, dtype=torch.int16, device=self.dev)
        self.g = torch.zeros(self.nparam, dtype=torch.float32, device=self.dev)
        self.m = torch.zeros_like(self.g)
        self.v = torch.zeros_like(self.g)
        self.step = 0
        _lib.check(self.lib.rmw_f32_to_bf16(self.w.data_ptr(), self.wb.data_ptr(), self.nparam, _stream()))
        self.names = [n for n, _ in shapes]
        self.shapes = dict(shapes)
        self.mean = torch.zeros(self.nout, dtype=torch.float32, device=self.dev)
        self.sd = torch.ones(self.nout, dtype=torch.float32, device=self.dev)
        self.fscale = 1.0
        self.dscale = 1.0     # motion features: displacement and rotation scales
        self.rscale = 1.0
        self._act = {}

    # parameter views
    def _view(self, buf, name):
        i = self.names.index(name)
        return buf[self.offsets[i]: self.offsets[i + 1]].view(self.shapes[name])

    def W(self, name):      # FP32 master
        return self._view(self.w, name)

    def Wb(self, name):     # BF16 copy
        return self._view(self.wb, name)

    def G(self, name):      # gradient
        return self._view(self.g, name)

    def params_numpy(self) -> Dict[str, np.ndarray]:
        return {n: self.W(n).cpu().numpy().copy() for n in self.names}

    def load_params(self, params: Dict[str, np.ndarray]):
        flat = np.concatenate([np.asarray(params[n], np.float32).ravel() for n in self.names])
        self.w.copy_(torch.from_numpy(flat))
        _lib.check(self.lib.rmw_f32_to_bf16(self.w.data_ptr(), self.wb.data_ptr(), self.nparam, _stream()))

    # -- kernels ------------------------------------------------------------------
    def _gemm(self, tA, tB, M, N, K, A, lda, B, ldb, C=None, ldc=0, beta=0.0, H=None, ldh=0, bias=None, act=0):
        _lib.check(self.lib.rmw_gemm_bf16(tA, tB, M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb,
                                          C.data_ptr() if C is not None else None, ldc, float(beta),
                                          H.data_ptr() if H is not None else None, ldh,
                                          bias.data_ptr() if bias is not None else None, act, _stream()))

    def _gemm_wgrad(self, M, N, K, A, lda, B, ldb, C, ldc):
        """C = A^T B for a weight gradient (K = batch rows): K split so that
        about 256 workgroups share it."""
        tiles = ((M + 63) // 64) * ((N + 63) // 64)
        splits = max(1, min(256 // tiles, K // 128))
        if splits == 1:
            self._gemm(1, 0, M, N, K, A, lda, B, ldb, C=C, ldc=ldc)
            return
        nbytes = ctypes.c_size_t()
        _lib.check(self.lib.rmw_gemm_bf16_splitk_workspace(M, N, K, splits, ctypes.byref(nbytes)))
        part = torch.empty(max(1, nbytes.value // 4), dtype=torch.float32, device=self.dev)
        _lib.check(self.lib.rmw_gemm_bf16_splitk(1, 0, M, N, K, A.data_ptr(), lda, B.data_ptr(), ldb, C.data_ptr(),
                                                 ldc, 0.0, splits, part.data_ptr(), part.numel() * 4, _stream()))

    def _bf16(self, x: torch.Tensor) -> torch.Tensor:
        y = torch.empty(x.shape, dtype=torch.int16, device=self.dev)
        _lib.check(self.lib.rmw_f32_to_bf16(x.data_ptr(), y.data_ptr(), x.numel(), _stream()))
        return y

    def _empty(self, shape, dtype=torch.float32):
        return torch.empty(shape, dtype=dtype, device=self.dev)

    # -- forward / backward ---------------------------------------------------------
    def forward(self, batch: Batch, keep: bool = False) -> torch.Tensor:
        B = batch.Xn.shape[0]
        E = B * DEG
        Xe_b = self._bf16(batch.Xe.contiguous())
        Ze1 = self._empty((E, HE))
        He1 = self._empty((E, HE), torch.int16)
        self._gemm(0, 1, E, HE, EIN, Xe_b, EIN, self.Wb("We1"), EIN, C=Ze1, ldc=HE, H=He1, ldh=HE,
                   bias=self.W("be1"), act=1)
        Msg = self._empty((E, NMSG))
        self._gemm(0, 1, E, NMSG, HE, He1, HE, self.Wb("We2"), HE, C=Msg, ldc=NMSG, bias=self.W("be2"))
        nc = self.nc
        if batch.Xn.shape[1] != self.nin:
            raise ValueError(f"the network takes {self.nin} node features, the batch has {batch.Xn.shape[1]}")
        Xc = self._empty((B, nc), torch.int16)
        _lib.check(self.lib.rmw_gnn_gather(batch.Xn.data_ptr(), self.nin, Msg.data_ptr(), NMSG, DEG, B, Xc.data_ptr(),
                                           nc, _stream()))
        Zn1 = self._empty((B, HN))
        Hn1 = self._empty((B, HN), torch.int16)
        self._gemm(0, 1, B, HN, nc, Xc, nc, self.Wb("Wn1"), nc, C=Zn1, ldc=HN, H=Hn1, ldh=HN, bias=self.W("bn1"),
                   act=1)
        Zn2 = self._empty((B, HN))
        Hn2 = self._empty((B, HN), torch.int16)
        self._gemm(0, 1, B, HN, HN, Hn1, HN, self.Wb("Wn2"), HN, C=Zn2, ldc=HN, H=Hn2, ldh=HN, bias=self.W("bn2"),
                   act=1)
        Y = self._empty((B, self.nout))
        self._gemm(0, 1, B, self.nout, HN, Hn2, HN, self.Wb("Wn3"), HN, C=Y, ldc=self.nout, bias=self.W("bn3"))
        if keep:
            self._act = dict(B=B, E=E, Xe_b=Xe_b, Ze1=Ze1, He1=He1, Xc=Xc, Zn1=Zn1, Hn1=Hn1, Zn2=Zn2, Hn2=Hn2)
        return Y

    def backward(self, dY_b: torch.Tensor):
        """Gradients of the loss with respect to every parameter, given
        dL/dY in BF16 (written into self.g)."""
        a = self._act
        B, E = a["B"], a["E"]
        no = self.nout
        # node layer 3
        self._gemm_wgrad(no, HN, B, dY_b, no, a["Hn2"], HN, self.G("Wn3"), HN)
        _lib.check(self.lib.rmw_colsum_bf16(dY_b.data_ptr(), no, B, no, self.G("bn3").data_ptr(), 0.0, _stream()))
        dH = self._empty((B, HN))
        self._gemm(0, 0, B, HN, no, dY_b, no, self.Wb("Wn3"), HN, C=dH, ldc=HN)
        # node layer 2
        dZ = self._empty((B, HN), torch.int16)
        _lib.check(self.lib.rmw_silu_backward(dH.data_ptr(), a["Zn2"].data_ptr(), dZ.data_ptr(), B * HN, _stream()))
        self._gemm_wgrad(HN, HN, B, dZ, HN, a["Hn1"], HN, self.G("Wn2"), HN)
        _lib.check(self.lib.rmw_colsum_bf16(dZ.data_ptr(), HN, B, HN, self.G("bn2").data_ptr(), 0.0, _stream()))
        dH1 = self._empty((B, HN))
        self._gemm(0, 0, B, HN, HN, dZ, HN, self.Wb("Wn2"), HN, C=dH1, ldc=HN)
        # node layer 1
        dZ1 = self._empty((B, HN), torch.int16)
        _lib.check(self.lib.rmw_silu_backward(dH1.data_ptr(), a["Zn1"].data_ptr(), dZ1.data_ptr(), B * HN,
                                              _stream()))
        nc = self.nc
        self._gemm_wgrad(HN, nc, B, dZ1, HN, a["Xc"], nc, self.G("Wn1"), nc)
        _lib.check(self.lib.rmw_colsum_bf16(dZ1.data_ptr(), HN, B, HN, self.G("bn1").data_ptr(), 0.0, _stream()))
        dXc = self._empty((B, nc))
        self._gemm(0, 0, B, nc, HN, dZ1, HN, self.Wb("Wn1"), nc, C=dXc, ldc=nc)
        # messages -> edge layer 2
        dM = self._empty((E, NMSG), torch.int16)
        _lib.check(self.lib.rmw_gnn_scatter(dXc.data_ptr(), nc, self.nin, NMSG, DEG, B, dM.data_ptr(), _stream()))
        self._gemm_wgrad(NMSG, HE, E, dM, NMSG, a["He1"], HE, self.G("We2"), HE)
        _lib.check(self.lib.rmw_colsum_bf16(dM.data_ptr(), NMSG, E, NMSG, self.G("be2").data_ptr(), 0.0, _stream()))
        dHe = self._empty((E, HE))
        self._gemm(0, 0, E, HE, NMSG, dM, NMSG, self.Wb("We2"), HE, C=dHe, ldc=HE)
        # edge layer 1
        dZe = self._empty((E, HE), torch.int16)
        _lib.check(self.lib.rmw_silu_backward(dHe.data_ptr(), a["Ze1"].data_ptr(), dZe.data_ptr(), E * HE,
                                              _stream()))
        self._gemm_wgrad(HE, EIN, E, dZe, HE, a["Xe_b"], EIN, self.G("We1"), EIN)
        _lib.check(self.lib.rmw_colsum_bf16(dZe.data_ptr(), HE, E, HE, self.G("be1").data_ptr(), 0.0, _stream()))
        # the padding columns of Wn1 stay zero
        self.G("Wn1")[:, self.nin + NMSG:].zero_()

    def loss_and_grad(self, batch: Batch, weights: Optional[torch.Tensor] = None) -> float:
        Y = self.forward(batch, keep=True)
        B = batch.Xn.shape[0]
        dY = self._empty((B, self.nout), torch.int16)
        nparts = 64
        part = self._empty(nparts)
        _lib.check(self.lib.rmw_mse(Y.data_ptr(), batch.T.data_ptr(), weights.data_ptr() if weights is not None
                                    else None, B, self.nout, dY.data_ptr(), part.data_ptr(), nparts, _stream()))
        self.backward(dY)
        return float(part.cpu().numpy().astype(np.float64).sum()) / self.nout

    def adam_step(self, lr: float, b1: float = 0.9, b2: float = 0.999, eps: float = 1e-8, wd: float = 0.0):
        self.step += 1
        _lib.check(self.lib.rmw_adam(self.w.data_ptr(), self.g.data_ptr(), self.m.data_ptr(), self.v.data_ptr(),
                                     self.wb.data_ptr(), self.nparam, lr, b1, b2, eps, self.step, wd, _stream()))


class GraphFeatures:
    """Builds network inputs and targets from body geometry on the device."""

    def __init__(self, net: GraphNet):
        self.net = net
        self.lib = net.lib
        self.dev = net.dev

    def batch(self, centres: torch.Tensor, quats: torch.Tensor, force_torque: torch.Tensor, previous=None) -> Batch:
        """Inputs of the network; previous = (centres0, quats0, force_torque0)
        of the earlier configuration for the motion features (nin 22)."""
        nb = centres.shape[0]
        nbr = torch.empty(nb * DEG, dtype=torch.int32, device=self.dev)
        _lib.check(self.lib.rmw_gnn_knn(centres.data_ptr(), nb, DEG, nbr.data_ptr(), _stream()))
        Xe = torch.empty((nb * DEG, EIN), dtype=torch.float32, device=self.dev)
        if previous is None:
            if self.net.nin != NIN:
                raise ValueError("this network needs the previous configuration")
            Xn = torch.empty((nb, NIN), dtype=torch.float32, device=self.dev)
            _lib.check(self.lib.rmw_gnn_features(centres.data_ptr(), quats.data_ptr(), force_torque.data_ptr(),
                                                 nbr.data_ptr(), nb, DEG, float(self.net.fscale), Xn.data_ptr(),
                                                 Xe.data_ptr(), _stream()))
        else:
            if self.net.nin != NIN_MOTION:
                raise ValueError(f"motion features need a network of {NIN_MOTION} node inputs")
            c0, q0, f0 = previous
            Xn = torch.empty((nb, NIN_MOTION), dtype=torch.float32, device=self.dev)
            _lib.check(self.lib.rmw_gnn_motion_features(centres.data_ptr(), quats.data_ptr(), force_torque.data_ptr(),
                                                        c0.data_ptr(), q0.data_ptr(), f0.data_ptr(), nbr.data_ptr(),
                                                        nb, DEG, float(self.net.fscale), float(self.net.dscale),
                                                        float(self.net.rscale), Xn.data_ptr(), Xe.data_ptr(),
                                                        _stream()))
        return Batch(Xn=Xn, Xe=Xe)

    def targets(self, quats, offs_dev, nblobs, sol, pc, raw: bool = False) -> torch.Tensor:
        """Body-frame (sol - base)/fscale (base passed as `pc`), standardised with the net's mean/sd
        (raw=True: mean 0, sd 1)."""
        nb = quats.shape[0]
        T = torch.empty((nb, self.net.nout), dtype=torch.float32, device=self.dev)
        mean = torch.zeros_like(self.net.mean) if raw else self.net.mean
        sd = torch.ones_like(self.net.sd) if raw else self.net.sd
        _lib.check(self.lib.rmw_gnn_body_frame(quats.data_ptr(), offs_dev.data_ptr(), nb, nblobs, self.net.nbb,
                                               sol.data_ptr(), pc.data_ptr(), T.data_ptr(), mean.data_ptr(),
                                               sd.data_ptr(), float(self.net.fscale), 0, _stream()))
        return T

    def initial_guess(self, quats, offs_dev, nblobs, pred: torch.Tensor, pc: torch.Tensor) -> torch.Tensor:
        """x0 = base + R_b (pred sd + mean) fscale (base passed as `pc`)."""
        nb = quats.shape[0]
        x0 = torch.empty_like(pc)
        _lib.check(self.lib.rmw_gnn_body_frame(quats.data_ptr(), offs_dev.data_ptr(), nb, nblobs, self.net.nbb,
                                               x0.data_ptr(), pc.data_ptr(), pred.data_ptr(), self.net.mean.data_ptr(),
                                               self.net.sd.data_ptr(), float(self.net.fscale), 1, _stream()))
        return x0
