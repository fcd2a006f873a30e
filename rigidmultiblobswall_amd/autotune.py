"""Autotuner for the launch shape of the wall-RPY pair kernels.

A pair-kernel work item is a tile of 256 targets against one chunk of
sources.  Three launch parameters are tuned (``rmw_tuning``):

  chunk_len  sources per work item (multiples of 64): short chunks give
             more workgroups and more FP64 partial sums to reduce, long
             chunks fewer of both;
  max_wgs    cap on the resident grid of the persistent loop (0 = one
             workgroup per work item, ragged batches default to 8 per CU);
  rhs_group  right-hand sides per pass (1, 2 or 4): more share the
             geometry of a pair, but need more registers;
  symmetric  for the blocks that have it (UF, UF + UT, UT, WF or WT alone,
             all four): each unordered pair once for both blobs (on) or the
             ordered kernel (off).

`tune` times candidates with HIP events on the launch stream (whole
product: pair kernel + reduce) in three stages -- every chunk length (with
the symmetric kernel on and off where it applies), then grid caps and RHS
groups for the three best, then the three fastest again interleaved with
more repetitions -- keeps only candidates whose result agrees with the
built-in rule's to the precision's tolerance, and returns the fastest.  Results live in a JSON cache keyed by device architecture, block
mask, precision, right-hand sides, single/ragged and the problem size
(half-octave bucket of the pair count); `wall_rpy` consults the cache
whenever the caller passes no tuning.

  python -m rigidmultiblobswall_amd.autotune [--workloads c1,c2,c3,c4,c5] [--out FILE]

writes the cache for the benchmark workloads (the package ships
``tuning_gfx950.json`` made this way on an MI355X).  RMW_TUNING_FILE
selects another cache file; RMW_AUTOTUNE=off ignores every cache.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import threading
import time
from typing import Dict, Iterable, List, Optional

import numpy as np

from . import _lib

HERE = os.path.dirname(os.path.abspath(__file__))
BLOCK = 256


def default_cache_path(arch: str = "gfx950") -> str:
    return os.environ.get("RMW_TUNING_FILE") or os.path.join(HERE, f"tuning_{arch}.json")


def pairs_bucket(pairs: float) -> float:
    """Half-octave bucket of the pair count."""
    return round(2.0 * math.log2(max(pairs, 1.0))) / 2.0


def cache_key(arch: str, mask: int, precision: str, k: int, pairs: float, ragged: bool) -> str:
    prec = "fp32" if precision == "fp32_f64acc" else precision
    return f"{arch.split(':')[0]}|mask{mask}|{prec}|k{k}|{'ragged' if ragged else 'single'}|p{pairs_bucket(pairs):g}"


def _key_fields(key: str):
    arch, mask, prec, k, kind, p = key.split("|")
    return arch, mask, prec, k, kind, float(p[1:])


class TuningCache:
    """JSON-backed map key -> {"chunk_len", "max_wgs", "rhs_group", ...}."""

    def __init__(self, path: Optional[str] = None):
        self.path = path
        self.entries: Dict[str, dict] = {}
        if path and os.path.exists(path):
            with open(path) as fh:
                data = json.load(fh)
            self.entries = data.get("entries", {})

    def lookup(self, key: str, max_octaves: float = 1.0) -> Optional[dict]:
        """Exact key, else the nearest size bucket of the same kind within
        max_octaves."""
        if key in self.entries:
            return self.entries[key]
        arch, mask, prec, k, kind, p = _key_fields(key)
        best, dist = None, None
        for other, e in self.entries.items():
            a2, m2, pr2, k2, kd2, p2 = _key_fields(other)
            if (a2, m2, pr2, k2, kd2) != (arch, mask, prec, k, kind):
                continue
            d = abs(p2 - p) / 2.0          # bucket unit = log2 of pairs
            if d <= max_octaves and (dist is None or d < dist):
                best, dist = e, d
        return best

    def store(self, key: str, entry: dict) -> None:
        self.entries[key] = entry

    def save(self, path: Optional[str] = None) -> str:
        path = path or self.path
        with open(path, "w") as fh:
            json.dump({"format": 1, "entries": dict(sorted(self.entries.items()))}, fh, indent=1)
            fh.write("\n")
        return path


_cache_lock = threading.Lock()
_cache: Dict[str, TuningCache] = {}


def _shared_cache(arch: str) -> TuningCache:
    path = default_cache_path(arch.split(":")[0])
    with _cache_lock:
        if path not in _cache:
            _cache[path] = TuningCache(path)
        return _cache[path]


def lookup(mask: int, precision: str, k: int, pairs: float, ragged: bool) -> Optional[_lib.Tuning]:
    """The cached launch shape for this problem, or None (built-in heuristic)."""
    if os.environ.get("RMW_AUTOTUNE", "").lower() in ("off", "0", "no"):
        return None
    from .wall_rpy import device_info
    arch = device_info()["arch"]
    e = _shared_cache(arch).lookup(cache_key(arch, mask, precision, k, pairs, ragged))
    if e is None:
        return None
    return _lib.Tuning(int(e["chunk_len"]), int(e["max_wgs"]), int(e.get("rhs_group", 0)),
                       int(e.get("symmetric", 0)), 0)


def chunk_candidates(nmax: int, max_chunks: int = 48) -> List[int]:
    """Chunk lengths to try (explicit, so a cached entry does not depend on
    the library's built-in rule): 64 and 128 (for the symmetric kernel a
    quarter or a half of a tile pair per work item), powers of two from 256
    and every distinct length ceil(n / c) rounded to 64 for c = 1..max_chunks
    chunks."""
    q = 64
    top = max(q, (nmax + q - 1) // q * q)
    out = {c for c in (64, 128) if c < top}
    c = 256
    while c < top:
        out.add(c)
        c *= 2
    for nch in range(1, max_chunks + 1):
        ln = ((nmax + nch - 1) // nch + q - 1) // q * q
        out.add(max(q, min(ln, top)))
    return sorted(out)


def candidates(nmax: int, k: int, cu: int, ragged: bool, symmetric_possible: bool = False) -> List[dict]:
    """Second-stage shapes for one chunk length: grid caps (none, 4, 8 and
    10 per CU, plus 16 per CU for ragged batches) x RHS groups x (where the
    symmetric UF kernel applies) symmetric on / off."""
    wgs = [0, 4 * cu, 8 * cu, 10 * cu] + ([16 * cu] if ragged else [])
    groups = ([0] + [g for g in (1, 2) if g < min(k, 4)] + ([4] if k >= 4 else [])) if k > 1 else [0]
    syms = [0, -1] if symmetric_possible else [0]
    return [{"max_wgs": w, "rhs_group": g, "symmetric": sy} for w in wgs for g in groups for sy in syms]


def _tolerance(precision: str) -> float:
    return 1e-12 if precision == "fp64" else 2e-6


def tune(mask: int, precision: str, r: np.ndarray, offsets: Optional[np.ndarray] = None, k: int = 1,
         eta: float = 1.0, a: float = 1.0, periodic_length=None, reps: int = 5, seed: int = 0,
         log=None, max_candidates: Optional[int] = None) -> dict:
    """Time every candidate launch shape on this geometry and return
    {"best": entry, "default_ms": ..., "table": [...]} (entry has the
    three parameters, its time and the heuristic's time)."""
    import torch
    from . import wall_rpy
    from ._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT
    dev = torch.device("cuda", torch.cuda.current_device())
    n = r.shape[0]
    ragged = offsets is not None
    sizes = np.diff(offsets) if ragged else np.array([n])
    pairs = float(np.sum(sizes.astype(np.float64) ** 2))
    rng = np.random.default_rng(seed)
    rd = torch.from_numpy(np.ascontiguousarray(r, dtype=np.float64)).to(dev)
    f = torch.from_numpy(rng.standard_normal((k, n, 3))).to(dev) if mask & (RMW_UF | RMW_WF) else None
    t = torch.from_numpy(rng.standard_normal((k, n, 3))).to(dev) if mask & (RMW_UT | RMW_WT) else None

    def run(tun):
        if ragged:
            return wall_rpy.apply_ragged(mask, rd, offsets, force=f, torque=t, eta=eta, a=a,
                                         periodic_length=periodic_length, precision=precision, k=k, tuning=tun)
        return wall_rpy.apply(mask, rd, force=f, torque=t, eta=eta, a=a, periodic_length=periodic_length,
                              precision=precision, k=k, tuning=tun)

    def timed(tun, n_reps=None):
        out = run(tun)                      # warm-up (and the result to compare)
        torch.cuda.synchronize()
        ts = []
        for _ in range(n_reps or reps):
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            run(tun)
            e1.record()
            e1.synchronize()
            ts.append(e0.elapsed_time(e1))
        return float(np.median(ts)), out

    heuristic = "heuristic"
    base_ms, base = timed(heuristic)
    ref = torch.cat([x.reshape(-1) for x in base if x is not None])
    refn = float(torch.linalg.vector_norm(ref).item())
    tol = _tolerance(precision)
    cu = wall_rpy.device_info()["cu_count"]
    nmax = int(sizes.max()) if sizes.size else 0
    table = []

    def trial(c):
        tun = _lib.Tuning(c["chunk_len"], c["max_wgs"], c["rhs_group"], c.get("symmetric", 0), 0)
        ms, out = timed(tun)
        got = torch.cat([x.reshape(-1) for x in out if x is not None])
        err = float(torch.linalg.vector_norm(got - ref).item()) / max(refn, 1e-300)
        row = dict(c, ms=ms, rel_diff=err, ok=err <= tol)
        table.append(row)
        return row

    # stage 1: chunk length at the default grid and RHS grouping
    chunks = chunk_candidates(nmax)
    if max_candidates:
        chunks = chunks[:max_candidates]
    sym_possible = mask in (RMW_UF, RMW_UF | RMW_UT, RMW_UF | RMW_UT | RMW_WF | RMW_WT) and periodic_length is None
    stage1 = [trial({"chunk_len": c, "max_wgs": 0, "rhs_group": 0, "symmetric": 0}) for c in chunks]
    if sym_possible:
        stage1 += [trial({"chunk_len": c, "max_wgs": 0, "rhs_group": 0, "symmetric": -1}) for c in chunks]
    good = sorted((r for r in stage1 if r["ok"]), key=lambda r: r["ms"])[:3]
    # stage 2: grid cap and RHS group for the three best (chunk length, kernel)
    for r0 in good:
        for c in candidates(nmax, k, cu, ragged):
            if c["max_wgs"] == 0 and c["rhs_group"] == 0:
                continue
            trial(dict(c, chunk_len=r0["chunk_len"], symmetric=r0["symmetric"]))
    cands = table
    # stage 3: the three fastest again, interleaved, with four times the
    # repetitions (single timings of near-equal shapes are noise-limited)
    finals = sorted((r for r in table if r["ok"]), key=lambda r: r["ms"])[:3]
    rounds = {id(r): [] for r in finals}
    for _ in range(3):
        for r0 in finals:
            tun = _lib.Tuning(r0["chunk_len"], r0["max_wgs"], r0["rhs_group"], r0.get("symmetric", 0), 0)
            ms, _ = timed(tun, 4 * reps)
            rounds[id(r0)].append(ms)
    for r0 in finals:
        r0["ms_confirmed"] = float(np.median(rounds[id(r0)]))
    best = min(finals, key=lambda r: r["ms_confirmed"])
    best = dict(best, ms=best["ms_confirmed"])
    if log:
        log(f"  {len(cands)} shapes: heuristic {base_ms:.4f} ms, best {best['ms']:.4f} ms "
            f"(chunk {best['chunk_len']}, max_wgs {best['max_wgs']}, rhs_group {best['rhs_group']}, "
            f"symmetric {best.get('symmetric', 0)})")
    entry = {"chunk_len": best["chunk_len"], "max_wgs": best["max_wgs"], "rhs_group": best["rhs_group"],
             "symmetric": best.get("symmetric", 0),
             "ms": best["ms"], "heuristic_ms": base_ms, "pairs": pairs, "n_max": int(sizes.max()),
             "systems": int(sizes.size)}
    return {"best": entry, "default_ms": base_ms, "table": table, "pairs": pairs, "ragged": ragged}


def _jobs(names: Iterable[str]):
    """(label, mask, precision, k, geometry) of the benchmark workloads."""
    from . import workloads
    from ._lib import RMW_UF, RMW_UT
    grand = 15
    for name in names:
        if name == "c1":
            w = workloads.c1_uf_matvec()
            for prec in ("fp32_blocked", "fp32", "fp64"):
                yield "c1", RMW_UF, prec, 1, w["r"], None, w["eta"], w["a"]
        elif name == "c2":
            w = workloads.c2_rigid_spheres()
            for prec in ("fp32_blocked", "fp32", "fp32_kahan", "fp64"):
                yield "c2", RMW_UF, prec, 1, w["r"], None, w["eta"], w["a"]
        elif name == "c3":
            w = workloads.c3_rollers()
            yield "c3", RMW_UF, "fp32_blocked", 1, w["r"], None, w["eta"], w["a"]
            # the rollers' fused u = UF f + UT t (constrained kinematics, RFD)
            yield "c3", RMW_UF | RMW_UT, "fp32_blocked", 1, w["r"], None, w["eta"], w["a"]
            for prec in ("fp32_blocked", "fp64"):
                yield "c3", grand, prec, 4, w["r"], None, w["eta"], w["a"]
        elif name == "c4":
            w = workloads.c4_monolayer()
            from .rigid import RigidSuspension
            import torch
            S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"])
            r = S.r_blobs.cpu().numpy()
            for prec in ("fp32_blocked", "fp64"):
                yield "c4", RMW_UF, prec, 1, r, None, w["eta"], w["a"]
            del S
            torch.cuda.empty_cache()
        elif name == "c5":
            w = workloads.c5_ragged_rollers()
            for prec in ("fp32_blocked", "fp64"):
                yield "c5", grand, prec, 1, w["r"], w["offsets"], w["eta"], w["a"]
        else:
            raise SystemExit(f"unknown workload {name}")


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--workloads", default="c1,c2,c3,c4,c5")
    ap.add_argument("--out", default=None)
    ap.add_argument("--reps", type=int, default=5)
    args = ap.parse_args(argv)
    import torch
    from .wall_rpy import device_info
    torch.cuda.set_device(0)
    log = lambda m: print(m, flush=True)  # noqa: E731
    arch = device_info()["arch"]
    path = args.out or default_cache_path(arch.split(":")[0])
    cache = TuningCache(path)
    t0 = time.perf_counter()
    for label, mask, prec, k, r, off, eta, a in _jobs(args.workloads.split(",")):
        log(f"{label}: mask {mask} {prec} k={k} {'ragged' if off is not None else 'single'} n={r.shape[0]}")
        res = tune(mask, prec, r, offsets=off, k=k, eta=eta, a=a, reps=args.reps, log=log)
        key = cache_key(arch, mask, prec, k, res["pairs"], res["ragged"])
        entry = dict(res["best"], workload=label)
        cache.store(key, entry)
        cache.save(path)
    log(f"wrote {path} ({len(cache.entries)} entries) in {time.perf_counter() - t0:.1f} s")
    return 0


if __name__ == "__main__":
    sys.exit(main())
