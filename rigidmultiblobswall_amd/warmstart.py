"""Learned GMRES warm start on the C4 monolayer: data, training, evaluation.

  python -m rigidmultiblobswall_amd.warmstart [--steps 200] [--out results.json]

1. Brownian trajectory of 1024 42-blob spheres (C4, trajectory.py); every
   step's FP64 GMRES solution, the preconditioner guess P^-1 b and the
   point-field guess P^-1 (b - [u; 0]) (u: velocity at each blob from the
   other bodies' net forces) are kept.
2. The graph network is trained (BF16 MFMA kernels, Adam) on the body-frame
   correction (x - point-field guess) of the first `train_frac` of the
   steps.
   A second network (mode 'previous') learns how the solution changed since
   the previous step from the motion in between, on consecutive pairs.
3. On held-out steps (and on a fresh configuration) the rigid solve to 1e-8
   is timed for x0 in {0, previous step, P^-1 b, point field, network,
   previous step + learned change},
   each with FP64 GMRES and with mixed precision (FP32-pair inner GMRES,
   FP64 residual refinement); the guesses' own cost is in their times.
   GMRES to the same FP64 tolerance judges every variant.
"""
from __future__ import annotations

import argparse
import json
import math
import time
from typing import Dict, List, Optional

import numpy as np
import torch

from . import gnn, workloads
from .rigid import RigidSuspension
from .trajectory import ForceParams, Snapshot, body_forces, run_c4_trajectory


def _sync():
    torch.cuda.synchronize()


def _rotation_vectors(q, q0) -> np.ndarray:
    """(n, 3) rotation vectors of q q0^-1 (the rotation carrying q0 to q)."""
    from .trajectory import quat_mul
    c0 = np.asarray(q0, dtype=np.float64) * np.array([1.0, -1.0, -1.0, -1.0])
    d = quat_mul(np.asarray(q, dtype=np.float64), c0)
    d *= np.where(d[:, :1] < 0, -1.0, 1.0)
    pn = np.linalg.norm(d[:, 1:], axis=1)
    k = np.where(pn > 0, 2.0 * np.arctan2(pn, d[:, 0]) / np.where(pn > 0, pn, 1.0), 2.0)
    return d[:, 1:] * k[:, None]


class WarmStart:
    """The network's x0 = base + R_b (prediction) fscale.

    mode 'point': the base is the point-field block-Jacobi guess
    P^-1 (b - [u; 0]) (u: velocity at each blob from the other bodies' net
    forces, spheres of `field_radius`) and the network sees the current
    geometry.  mode 'previous': the base is the previous step's solution and
    the network also sees the motion since that step (body-frame
    displacement, rotation and force changes of each body, the change of
    each neighbour offset): it predicts how the solution changed."""

    def __init__(self, ref, a, eta, device=None, seed=0, field_radius: float = 0.9, mode: str = "point"):
        if mode not in ("point", "previous"):
            raise ValueError("mode must be 'point' or 'previous'")
        self.mode = mode
        self.ref, self.a, self.eta = ref, a, eta
        self.field_radius = float(field_radius)
        self.dev = device or torch.device("cuda", torch.cuda.current_device())
        self.nbb = ref.shape[0]
        self.net = gnn.GraphNet(self.nbb, self.dev, seed=seed,
                                nin=gnn.NIN_MOTION if mode == "previous" else gnn.NIN)
        self.feats = gnn.GraphFeatures(self.net)
        self.offs_cache = {}

    def _offs(self, nb):
        if nb not in self.offs_cache:
            self.offs_cache[nb] = torch.arange(0, (nb + 1) * self.nbb, self.nbb, dtype=torch.int64, device=self.dev)
        return self.offs_cache[nb]

    def _dev(self, x):
        return torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(self.dev)

    def inputs(self, snap: Snapshot, prev: Optional[Snapshot] = None) -> gnn.Batch:
        previous = None
        if self.mode == "previous":
            if prev is None:
                raise ValueError("the 'previous' network needs the previous step's snapshot")
            previous = (self._dev(prev.centres), self._dev(prev.quats), self._dev(prev.force_torque))
        return self.feats.batch(self._dev(snap.centres), self._dev(snap.quats), self._dev(snap.force_torque),
                                previous=previous)

    def raw_targets(self, snap: Snapshot, prev: Optional[Snapshot] = None) -> torch.Tensor:
        nb = snap.centres.shape[0]
        return self.feats.targets(self._dev(snap.quats), self._offs(nb), nb * self.nbb, snap.sol,
                                  self._base(snap, prev), raw=True)

    def _base(self, snap: Snapshot, prev: Optional[Snapshot] = None) -> torch.Tensor:
        if self.mode == "previous":
            return prev.sol
        return snap.base if snap.base is not None else snap.pc

    # -- training -----------------------------------------------------------------
    def fit(self, snaps: List[Snapshot], epochs=40, batch=2048, lr=2e-3, lr_min=2e-5, seed=0, log=None) -> Dict:
        """Train on a trajectory's snapshots (mode 'previous': on the pairs of
        consecutive snapshots)."""
        fts = np.concatenate([s.force_torque for s in snaps])
        self.net.fscale = float(np.sqrt(np.mean(np.sum(fts ** 2, axis=1))))
        pairs = [(s, None) for s in snaps] if self.mode == "point" else list(zip(snaps[1:], snaps[:-1]))
        if self.mode == "previous":
            dc = np.concatenate([s.centres - p.centres for s, p in pairs])
            th = np.concatenate([_rotation_vectors(s.quats, p.quats) for s, p in pairs])
            self.net.dscale = float(np.sqrt(np.mean(np.sum(dc ** 2, axis=1)))) or 1.0
            self.net.rscale = float(np.sqrt(np.mean(np.sum(th ** 2, axis=1)))) or 1.0
        Xn, Xe, T = [], [], []
        for s, p in pairs:
            b = self.inputs(s, p)
            Xn.append(b.Xn)
            Xe.append(b.Xe)
            T.append(self.raw_targets(s, p))
        Xn = torch.cat(Xn)
        Xe = torch.cat(Xe)
        T = torch.cat(T)
        Th = T.cpu().numpy().astype(np.float64)
        mean = Th.mean(axis=0)
        sd = Th.std(axis=0)
        sd = np.where(sd > 1e-12, sd, 1.0)
        self.net.mean.copy_(torch.from_numpy(mean.astype(np.float32)))
        self.net.sd.copy_(torch.from_numpy(sd.astype(np.float32)))
        T = torch.from_numpy(((Th - mean) / sd).astype(np.float32)).to(self.dev)
        n = Xn.shape[0]
        rng = np.random.default_rng(seed)
        steps_per_epoch = max(1, n // batch)
        total = epochs * steps_per_epoch
        hist = []
        t0 = time.perf_counter()
        it = 0
        for ep in range(epochs):
            perm = torch.from_numpy(rng.permutation(n)).to(self.dev)
            ep_loss = 0.0
            for s in range(steps_per_epoch):
                idx = perm[s * batch:(s + 1) * batch]
                eidx = (idx[:, None] * gnn.DEG + torch.arange(gnn.DEG, device=self.dev)[None, :]).reshape(-1)
                bt = gnn.Batch(Xn=Xn.index_select(0, idx).contiguous(), Xe=Xe.index_select(0, eidx).contiguous(),
                               T=T.index_select(0, idx).contiguous())
                loss = self.net.loss_and_grad(bt)
                frac = it / max(1, total - 1)
                cur = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * frac))
                self.net.adam_step(cur)
                ep_loss += loss
                it += 1
            hist.append(ep_loss / steps_per_epoch)
            if log:
                log(f"epoch {ep}: loss {hist[-1]:.4e}")
        _sync()
        return {"loss_history": hist, "train_seconds": time.perf_counter() - t0, "samples": int(n),
                "steps": it, "parameters": self.net.nparam}

    # -- prediction ---------------------------------------------------------------
    def predict_x0(self, snap: Snapshot, base: Optional[torch.Tensor] = None,
                   prev: Optional[Snapshot] = None) -> torch.Tensor:
        nb = snap.centres.shape[0]
        pred = self.net.forward(self.inputs(snap, prev))
        x0 = self.feats.initial_guess(self._dev(snap.quats), self._offs(nb), nb * self.nbb, pred,
                                      self._base(snap, prev) if base is None else base)
        if not bool(torch.isfinite(x0).all().item()):
            return None   # discarded: GMRES starts from zero
        return x0

    def heldout_loss(self, snaps: List[Snapshot]) -> float:
        """Mean square standardised error on snapshots not trained on (mode
        'previous': on their consecutive pairs)."""
        tot, cnt = 0.0, 0
        pairs = [(s, None) for s in snaps] if self.mode == "point" else list(zip(snaps[1:], snaps[:-1]))
        for s, p in pairs:
            pred = self.net.forward(self.inputs(s, p)).cpu().numpy().astype(np.float64)
            raw = self.raw_targets(s, p).cpu().numpy().astype(np.float64)
            t = (raw - self.net.mean.cpu().numpy()) / self.net.sd.cpu().numpy()
            tot += float(((pred - t) ** 2).sum())
            cnt += pred.size
        return tot / cnt


def _rel(a, b):
    return float(torch.linalg.vector_norm(a - b).item() / torch.linalg.vector_norm(b).item())


def evaluate(ws: WarmStart, snaps: List[Snapshot], prev_snaps: List, tol=1e-8, repeats=1, log=None,
             ws_prev: Optional[WarmStart] = None) -> List[Dict]:
    """Solve each snapshot's problem to tol from every guess; ws_prev (a
    mode 'previous' network) adds the guess 'gnn_prev' = previous solution +
    learned change, where a previous snapshot exists."""
    rows = []
    if snaps:
        ws.predict_x0(snaps[0])    # untimed: first-call set-up (buffers, kernel loading)
        if ws_prev is not None and prev_snaps and prev_snaps[0] is not None:
            ws_prev.predict_x0(snaps[0], prev=prev_snaps[0])
        _sync()
    for snap, prev in zip(snaps, prev_snaps):
        S = RigidSuspension(snap.centres, snap.quats, ws.ref, ws.eta, ws.a, precision="fp64")
        _sync()
        t0 = time.perf_counter()
        S.build_preconditioner()
        _sync()
        t_pc = time.perf_counter() - t0
        rhs = S.rhs(np.zeros((S.Nblobs, 3)), snap.force_torque)
        xref = snap.sol
        _sync()
        t0 = time.perf_counter()
        x_pt = S.point_field_guess(rhs, snap.force_torque, radius=ws.field_radius)
        _sync()
        t_pt = time.perf_counter() - t0
        t0 = time.perf_counter()
        x_gnn = ws.predict_x0(snap, base=x_pt)
        _sync()
        t_gnn = time.perf_counter() - t0
        guesses = {"zero": (None, 0.0), "pc": (snap.pc, 0.0), "point": (x_pt, t_pt), "gnn": (x_gnn, t_pt + t_gnn)}
        t_gp = 0.0
        if prev is not None:
            guesses["previous"] = (prev.sol, 0.0)
            if ws_prev is not None:
                _sync()
                t0 = time.perf_counter()
                x_gp = ws_prev.predict_x0(snap, prev=prev)
                _sync()
                t_gp = time.perf_counter() - t0
                guesses["gnn_prev"] = (x_gp, t_gp)
        row = {"pc_build_ms": t_pc * 1e3, "point_field_ms": t_pt * 1e3, "gnn_inference_ms": t_gnn * 1e3,
               "gnn_prev_inference_ms": t_gp * 1e3}
        for name, (x0, t_extra) in guesses.items():
            err = _rel(x0, xref) if x0 is not None else 1.0
            for mode in ("fp64", "mixed"):
                best = None
                for _ in range(repeats):
                    _sync()
                    t0 = time.perf_counter()
                    if mode == "fp64":
                        res = S.solve_mobility_problem(rhs=rhs, tol=tol, x0=x0)
                    else:
                        res = S.solve_mixed(rhs=rhs, tol=tol, x0=x0)
                    _sync()
                    dt = time.perf_counter() - t0
                    best = dt if best is None else min(best, dt)
                # the judge: FP64 residual of the returned solution
                r = torch.empty_like(rhs)
                S.operator(res.x, r, precision="fp64")
                true_res = _rel(r, rhs)
                row[f"{name}/{mode}"] = {"x0_rel_err": err, "iterations": res.iterations, "matvecs": res.matvecs,
                                         "solve_ms": best * 1e3, "total_ms": (best + t_extra) * 1e3,
                                         "true_residual": true_res, "converged": bool(true_res <= 1.5 * tol)}
        rows.append(row)
        if log:
            log(json.dumps({k: (v if not isinstance(v, dict) else {kk: round(vv, 5) if isinstance(vv, float) else vv
                                                                 for kk, vv in v.items()}) for k, v in row.items()}))
    return rows


def summarise(rows: List[Dict]) -> Dict:
    keys = [k for k in rows[0] if "/" in k]
    out = {}
    for k in keys:
        vals = [r[k] for r in rows if k in r]
        out[k] = {"iterations": float(np.mean([v["iterations"] for v in vals])),
                  "total_ms": float(np.mean([v["total_ms"] for v in vals])),
                  "x0_rel_err": float(np.mean([v["x0_rel_err"] for v in vals])),
                  "all_converged": all(v["converged"] for v in vals)}
    base = out["zero/fp64"]["total_ms"]
    for k in out:
        out[k]["speedup_vs_plain_fp64"] = base / out[k]["total_ms"]
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--bodies", type=int, default=1024)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--train-frac", type=float, default=0.75)
    ap.add_argument("--epochs", type=int, default=120)
    ap.add_argument("--eval", type=int, default=5, help="held-out steps evaluated")
    ap.add_argument("--out", default=None)
    args = ap.parse_args(argv)
    torch.cuda.set_device(0)
    log = lambda m: print(m, flush=True)  # noqa: E731
    w = workloads.c4_monolayer(nbodies=args.bodies)
    t0 = time.perf_counter()
    traj = run_c4_trajectory(w["centres"], w["quats"], w["ref"], w["a"], w["eta"], nsteps=args.steps,
                             log=lambda m: log(m) if m.startswith("step ") and int(m.split()[1][:-1]) % 20 == 0
                             else None)
    t_traj = time.perf_counter() - t0
    log(f"trajectory: {args.steps} steps in {t_traj:.1f} s, mean GMRES iterations "
        f"{np.mean([s.iterations for s in traj.snapshots]):.1f}")
    ntr = int(args.train_frac * args.steps)
    ws = WarmStart(w["ref"], w["a"], w["eta"])
    fit = ws.fit(traj.snapshots[:ntr], epochs=args.epochs, log=log)
    held = traj.snapshots[ntr:]
    fit["heldout_loss"] = ws.heldout_loss(held)
    fit["train_loss_final"] = fit["loss_history"][-1]
    log(f"training: {fit['steps']} Adam steps in {fit['train_seconds']:.1f} s, final loss "
        f"{fit['train_loss_final']:.3e}, held-out loss {fit['heldout_loss']:.3e}")
    ws_prev = WarmStart(w["ref"], w["a"], w["eta"], mode="previous")
    fit_prev = ws_prev.fit(traj.snapshots[:ntr], epochs=args.epochs)
    fit_prev["heldout_loss"] = ws_prev.heldout_loss(held)
    log(f"previous-step network: {fit_prev['steps']} Adam steps in {fit_prev['train_seconds']:.1f} s, held-out loss "
        f"{fit_prev['heldout_loss']:.3e}")
    picks = np.linspace(0, len(held) - 1, min(args.eval, len(held))).astype(int)
    ev = [held[i] for i in picks]
    prev = [traj.snapshots[ntr + i - 1] for i in picks]
    rows = evaluate(ws, ev, prev, log=log, ws_prev=ws_prev)
    # a fresh configuration (different seed, never seen in training)
    w2 = workloads.c4_monolayer(seed=40, nbodies=args.bodies)
    fresh = run_c4_trajectory(w2["centres"], w2["quats"], w2["ref"], w2["a"], w2["eta"], nsteps=1)
    rows_fresh = evaluate(ws, fresh.snapshots, [None], log=log)
    res = {"config": {"bodies": args.bodies, "blobs_per_body": int(w["ref"].shape[0]), "steps": args.steps,
                      "train_steps": ntr, "trajectory": traj.params, "trajectory_seconds": t_traj},
           "training": fit, "training_previous": fit_prev, "heldout": summarise(rows),
           "fresh_configuration": summarise(rows_fresh),
           "rows": rows, "rows_fresh": rows_fresh}
    if args.out:
        with open(args.out, "w") as fh:
            json.dump(res, fh, indent=1)
    log(json.dumps({"heldout": res["heldout"], "fresh": res["fresh_configuration"]}))
    return res


if __name__ == "__main__":
    main()
