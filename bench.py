"""Benchmark: wall-RPY blob mobility matvec on one MI355X.

One step = one wall-corrected RPY translational product u = B M_UF(z_eff) B f
over N = 32768 single-blob rollers (a = 0.656, eta = 1e-3, x, y ~ U[0, 300],
z ~ U[0.8, 2.0], seed 3; f ~ N(0, 1), seed 7), FP32 pair terms with FP32
sub-sums folded into FP64 per-target sums (--precision), inputs resident in
HBM: the N = 32768 point of the 50x-over-the-host-CPU target.

Prints ONE JSON line (rank 0):
  value        blob-pairs/s = N^2 / (time per step)
  oracle_error relative L2 error of the timed product against the FP64 C
               oracle (strict IEEE build) on a sample of 512 target rows
  fp64         the same product with FP64 pair terms: time, pairs/s, fraction
               of the FP64 vector peak and its oracle error (run before the
               FP32 warm-up: its sustained load brings the GPU to its
               sustained clock, which the first ~30 ms of load from idle
               do not reach; lines without this leg repeat their product,
               untimed, for 40 ms before the warm-up instead)
  roofline     pair kernel: algorithmic FLOP per launch (164 FLOP/pair x N^2)
               / mean pair-kernel duration (HIP events on the launch stream),
               against the FP32 vector peak; traffic from committed rocprof
               PMC runs when present
  cpu_baseline FP64 C oracle (-O3 -ffast-math, OpenMP over all usable host
               cores; the x86-64-v3 build and, on AVX-512 hosts, the x86-64-v4
               build, the faster reported) on a bounded sample of target rows
               of the same workload
  gmres        rigid solve to 1e-8 on the 1024-sphere monolayer: plain FP64
               GMRES from x0 = 0 against the learned warm start (network
               trained on the GPU from a 200-step Brownian trajectory), per-solve
               times

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
                       [--workload c3|c2|c1|c3r|c3f|c3step|c2step|c3lz|c5lz|prodtiming]
       (c2: the rigid-solve suspension, 256 spheres x 42 blobs;
       c3r: the different-radii product on the C3 geometry, above the wall,
       or with --domain no_wall / free_surface; c3f: the
       blob-blob soft-repulsion forces on it; c3step: whole rollers time
       steps, steps/s; c2step: rigid-body time steps of the C2 suspension;
       c3lz / c5lz: Lanczos Brownian-slip solves of the C3 block of 4 RHS
       and of the C5 ragged batch, solves/s; prodtiming: the reference's
       four-block timing harness, --blobs-per-body 1|12|42|162)
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

METRIC = "wall-RPY matvec blob-pairs/s; GMRES time-to-1e-8 residual (learned x0 vs plain)"
FLOP_PER_PAIR = {"UF": 164, "UT": 95, "WF": 95, "WT": 101, "grand": 386,
                 "radii": 154,   # radii: counted on the dyadic form of the kernel (FMA = 2)
                 "forces": 22,   # difference 3, r^2 5, sqrt + reciprocal 2, exponent 2, exp 1,
                                 # prefactor 2, radius sum 1, accumulation 6
                 # UF of the other boundaries, same convention: RPY 36 + 3x3 product 18
                 # (no wall); + the image RPY 36 (free surface); the wall terms as UF (in plane)
                 "UF_no_wall": 54, "UF_free_surface": 90, "UF_in_plane": 164,
                 # different radii, same convention as the equal-radii domains (the
                 # radius products and regime branches of the radii form not counted)
                 "radii_no_wall": 54, "radii_free_surface": 90}
RADII_BOUNDARY = {"single_wall": "wall", "no_wall": "unbounded", "free_surface": "free_surface"}
PEAK_FP32_TFLOPS = 157.3
PRECISION_NOTE = {"fp32": " (FP32 pair terms, FP64 per-target sums)",
                  "fp32_blocked": " (FP32 pair terms, FP32 sub-sums of <= 16 sources folded into FP64 per-target sums)",
                  "fp32_kahan": " (FP32 pair terms, compensated FP32 per-target sums)",
                  "fp64": " (FP64 throughout)"}
PEAK_FP64_TFLOPS = 78.6


def load_workload(name: str):
    from rigidmultiblobswall_amd import workloads
    if name == "c2":
        w = workloads.c2_rigid_spheres()
        desc = "C2 rigid-solve matvec: 256 spheres x 42 blobs = 10752 blobs, a=0.243, eta=1, UF block"
    elif name == "c1":
        w = workloads.c1_uf_matvec()
        desc = "C1: N=2048 blobs, a=0.13, eta=1, UF block"
    elif name == "c3":
        w = workloads.c3_rollers()
        desc = ("C3 matvec: N=32768 single-blob rollers above a wall, a=0.656, eta=1e-3, x,y ~ U[0,300], "
                "z ~ U[0.8,2.0], f ~ N(0,1), UF block")
    elif name == "c3f":
        w = workloads.c3_rollers()
        w = dict(w, forces=dict(repulsion_strength=0.0165677856, debye_length=0.0656, blob_radius=0.656))
        desc = ("C3 geometry, blob-blob soft repulsion with the rollers example's parameters "
                "(eps=0.0165677856, b=0.0656, a=0.656): N=32768, every pair, open box")
    elif name == "c3r":
        w = workloads.c3_rollers()
        w = dict(w, radii=np.random.default_rng(33).uniform(0.5, 0.8, w["r"].shape[0]))
        desc = ("C3 geometry with blob radii U[0.5, 0.8] (seed 33): N=32768, eta=1e-3, different-radii wall "
                "RPY (radii kernel, sources = targets)")
    else:
        raise SystemExit(f"unknown workload {name}")
    return w, desc


def _cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(r, f, eta, a, target_seconds=16.0, radii=None, force_params=None, boundary="wall"):
    """Time the FP64 oracle (fast-math build, all usable cores) on a
    bounded sample of the same matvec (whole products repeated, or a row
    block when one product is longer than the budget); returns pairs/s,
    with the CPU model, core count and a single-thread rate on a small
    row block."""
    from oracle import oracle as O
    kinds = ["fast"] + (["fast_avx512"] if O.host_has_avx512() else [])
    runs = [_cpu_baseline_kind(O, kind, r, f, eta, a, target_seconds / len(kinds), radii, force_params, boundary)
            for kind in kinds]
    best = max(runs, key=lambda x: x["value"])
    best["builds_timed"] = {x["build"]: x["value"] for x in runs}
    return best


def _cpu_baseline_kind(O, kind, r, f, eta, a, target_seconds, radii, force_params, boundary):
    n = r.shape[0]
    threads = O.num_threads(kind)
    if force_params is not None:
        fp = force_params
        def run(rows):
            O.blob_forces(r, fp["repulsion_strength"], fp["debye_length"], fp["blob_radius"], rows=rows, kind=kind)
    elif radii is not None:
        # different radii: the oracle's radii product on a block of rows
        def run(rows):
            O.radii_apply(r, r, f, radii, radii, eta, rows=rows, kind=kind, boundary=boundary)
    else:
        def run(rows):
            O.block_apply(O.UF, r, f, eta, a, rows=rows, kind=kind)
    rows0 = max(1, min(n, 256))
    run((0, rows0))          # warm-up
    t = time.perf_counter()
    run((0, rows0))
    dt = max(time.perf_counter() - t, 1e-6)
    per_pair = dt / (rows0 * n)
    rows = int(min(n, max(rows0, target_seconds / (per_pair * n))))
    reps = max(1, int(target_seconds / (per_pair * rows * n)))
    t = time.perf_counter()
    for _ in range(reps):
        run((0, rows))
    dt = time.perf_counter() - t
    # one thread on a small block, for the per-core rate
    O.set_num_threads(1, kind)
    rows1 = max(1, min(n, 32))
    t = time.perf_counter()
    run((0, rows1))
    single = rows1 * n / (time.perf_counter() - t)
    O.set_num_threads(threads, kind)
    isa = "x86-64-v4 (AVX-512)" if kind == "fast_avx512" else "x86-64-v3 (AVX2)"
    return {"value": reps * rows * n / dt, "unit": "blob-pairs/s", "cores": threads, "kind": "port",
            "sample": f"{reps} x {'forces' if force_params else 'radii' if radii is not None else 'UF'} rows 0..{rows} "
                      f"of {n} (x {n} sources) "
                      f"of the same workload; "
                      f"FP64 C oracle -O3 -ffast-math -march={isa.split()[0]}, OpenMP {threads} threads; {dt:.1f} s",
            "build": kind, "isa": isa,
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "single_thread_pairs_per_s": single}


def oracle_rows(n: int, blocks: int = 4, rows: int = 128):
    """`blocks` row blocks of `rows` targets spread over [0, n)."""
    if n <= blocks * rows:
        return [(0, n)]
    starts = np.linspace(0, n - rows, blocks).astype(int)
    return [(int(s0), int(s0) + rows) for s0 in starts]


def oracle_error(u, r_h, f_h, eta, a, domain="single_wall"):
    """Relative L2 error of a UF product u (device (n, 3)) against the FP64
    C oracle (strict build) on a sample of target rows."""
    from oracle import oracle as O
    block = {"single_wall": O.UF}.get(domain)
    if block is None:
        return None
    n = r_h.shape[0]
    got = u.reshape(n, 3).cpu().numpy()
    num = den = 0.0
    sample = oracle_rows(n)
    for t0, t1 in sample:
        ref = O.block_apply(block, r_h, f_h, eta, a, rows=(t0, t1)).reshape(n, 3)[t0:t1]
        num += float(np.sum((got[t0:t1] - ref) ** 2))
        den += float(np.sum(ref ** 2))
    return {"rel_l2": float(np.sqrt(num / den)), "rows": int(sum(t1 - t0 for t0, t1 in sample)),
            "oracle": "FP64 C restatement of mobility_numba.py:123-287 with the wrapper's height shift and damping "
                      "(strict IEEE build)"}


def fp64_leg(r, f, u, r_h, f_h, eta, a, steps, warmup, check: bool = True):
    """The same product with FP64 pair terms and sums (the 1e-12 mode);
    check=False leaves out the oracle comparison (the caller makes it later
    from u)."""
    import torch
    from rigidmultiblobswall_amd import wall_rpy
    from rigidmultiblobswall_amd._lib import RMW_UF
    prod = wall_rpy.Product(RMW_UF, r, eta=eta, a=a, precision="fp64")
    step = lambda: prod(force=f, out_velocity=u)  # noqa: E731
    settle(step)
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with wall_rpy.PairKernelTimer() as pk:
        torch.cuda.synchronize()
        start.record()
        for _ in range(steps):
            step()
        stop.record()
        torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / steps
    n = r_h.shape[0]
    pairs = float(n) * n
    achieved = FLOP_PER_PAIR["UF"] * pairs / (pk.mean_ms * 1e-3) / 1e12
    launch = wall_rpy.describe(RMW_UF, n, 1, "fp64")
    return {"ms_per_step": ms, "pairs_per_s": pairs / (ms * 1e-3), "kernel_mean_ms": pk.mean_ms,
            "kernel_median_ms": pk.median_ms,
            "kernel": f"{'mutual_kernel<UF,K=1,fp64>' if launch['kernel'] == 'symmetric' else 'pair_kernel_f64<UF,K=1>'}",
            "achieved_tflops": achieved, "peak_tflops": PEAK_FP64_TFLOPS, "frac": achieved / PEAK_FP64_TFLOPS,
            **({"oracle_error": oracle_error(u, r_h, f_h, eta, a)} if check else {})}


def end_to_end(r_h, f_h, eta, a, precision, reps=20):
    """Median time of the numpy drop-in signature (host arrays in, host
    array out: H2D copies, the product, D2H copy), the path a caller of the
    reference's mobility functions sees."""
    import torch
    from rigidmultiblobswall_amd import mobility
    for _ in range(3):
        mobility.single_wall_mobility_trans_times_force_hip(r_h, f_h, eta, a, precision=precision)
    ts = []
    for _ in range(reps):
        torch.cuda.synchronize()
        t = time.perf_counter()
        mobility.single_wall_mobility_trans_times_force_hip(r_h, f_h, eta, a, precision=precision)
        ts.append(time.perf_counter() - t)
    ms = 1e3 * float(np.median(ts))
    n = r_h.shape[0]
    return {"ms": ms, "pairs_per_s": float(n) * n / (ms * 1e-3),
            "path": "mobility.single_wall_mobility_trans_times_force_hip(numpy r, numpy f) -> numpy u"}


def gmres_leg(steps: int, epochs: int, evals: int, log):
    """Rigid-body solve to 1e-8 on the learned-warm-start monolayer (1024
    spheres x 42 blobs): a short Brownian trajectory generates the training
    solves, the graph network is trained on the GPU, and held-out steps are
    solved from x0 = 0 with plain FP64 GMRES and from the network's x0 with
    FP64 GMRES and with mixed precision (FP32-pair inner GMRES, FP64
    residual refinement).  The network corrects the point-field guess
    P^-1 (b - [u; 0]), which is also solved from on its own.  Every solve's
    returned x is checked against the FP64 operator; times include the
    guesses' own cost (point field, network inference)."""
    from rigidmultiblobswall_amd import warmstart, workloads
    from rigidmultiblobswall_amd.trajectory import run_c4_trajectory
    w = workloads.c4_monolayer()
    t0 = time.perf_counter()
    traj = run_c4_trajectory(w["centres"], w["quats"], w["ref"], w["a"], w["eta"], nsteps=steps,
                             log=lambda m: log(m) if m.startswith("step ") and int(m.split()[1][:-1]) % 25 == 0
                             else None)
    t_traj = time.perf_counter() - t0
    log(f"gmres leg: trajectory {steps} steps in {t_traj:.1f} s")
    ntr = int(0.75 * steps)
    ws = warmstart.WarmStart(w["ref"], w["a"], w["eta"])
    fit = ws.fit(traj.snapshots[:ntr], epochs=epochs)
    log(f"gmres leg: trained {fit['steps']} Adam steps in {fit['train_seconds']:.1f} s, loss {fit['loss_history'][-1]:.3e}")
    ws_prev = warmstart.WarmStart(w["ref"], w["a"], w["eta"], mode="previous")
    fit_prev = ws_prev.fit(traj.snapshots[:ntr], epochs=epochs)
    held = traj.snapshots[ntr:]
    heldout_prev = ws_prev.heldout_loss(held)
    log(f"gmres leg: previous-step network trained {fit_prev['steps']} Adam steps in {fit_prev['train_seconds']:.1f} s, "
        f"loss {fit_prev['loss_history'][-1]:.3e}, held-out {heldout_prev:.3e}")
    picks = np.linspace(0, len(held) - 1, min(evals, len(held))).astype(int)
    rows = warmstart.evaluate(ws, [held[i] for i in picks], [traj.snapshots[ntr + i - 1] for i in picks],
                              ws_prev=ws_prev)
    sm = warmstart.summarise(rows)
    pick = lambda k: {"ms": sm[k]["total_ms"], "iterations": sm[k]["iterations"],  # noqa: E731
                      "x0_rel_err": sm[k]["x0_rel_err"], "converged": sm[k]["all_converged"]}
    out = {"workload": f"C4 monolayer: 1024 spheres x 42 blobs = 43008 blobs, tol 1e-8, {steps}-step Brownian "
                       f"trajectory (stochastic Euler-Maruyama, dt 1e-3), trained on {ntr} steps, "
                       f"{len(picks)} held-out solves",
           "trajectory_seconds": t_traj,
           "plain_fp64": pick("zero/fp64"), "learned_fp64": pick("gnn/fp64"), "learned_mixed": pick("gnn/mixed"),
           "zero_mixed": pick("zero/mixed"), "point_field_mixed": pick("point/mixed"),
           "previous_step_fp64": pick("previous/fp64"), "previous_step_mixed": pick("previous/mixed"),
           "learned_previous_fp64": pick("gnn_prev/fp64"), "learned_previous_mixed": pick("gnn_prev/mixed"),
           "point_field_ms": float(np.mean([r["point_field_ms"] for r in rows])),
           "gnn_inference_ms": float(np.mean([r["gnn_inference_ms"] for r in rows])),
           "gnn_prev_inference_ms": float(np.mean([r["gnn_prev_inference_ms"] for r in rows])),
           "train_seconds": fit["train_seconds"], "train_seconds_previous": fit_prev["train_seconds"],
           "parameters": fit["parameters"], "parameters_previous": fit_prev["parameters"]}
    # the learned guesses against the guesses without a network, in FP64
    # GMRES alone (the network's own gain) and with the mixed solve
    out["speedup_learned_vs_plain"] = out["plain_fp64"]["ms"] / out["learned_mixed"]["ms"]
    out["speedup_learned_fp64_vs_plain"] = out["plain_fp64"]["ms"] / out["learned_fp64"]["ms"]
    out["speedup_mixed_alone_vs_plain"] = out["plain_fp64"]["ms"] / out["zero_mixed"]["ms"]
    out["speedup_learned_previous_fp64_vs_previous_fp64"] = (out["previous_step_fp64"]["ms"]
                                                             / out["learned_previous_fp64"]["ms"])
    out["speedup_learned_previous_mixed_vs_plain"] = out["plain_fp64"]["ms"] / out["learned_previous_mixed"]["ms"]
    return out


def settle(step, ms: float = 40.0) -> None:
    """Repeat `step` (untimed) until `ms` milliseconds have passed on the
    device (whole steps, four at a time): from idle, and after a change of
    kernel, the GPU needs ~30 ms of sustained load to reach its sustained
    clock (the pair kernels' first launches run up to 30 % longer)."""
    import torch
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    done = 0.0
    while done < ms:
        for _ in range(4):
            step()
        stop.record()
        stop.synchronize()
        done = start.elapsed_time(stop)


def timed_steps(advance, steps: int) -> float:
    """Mean milliseconds per call of ``advance`` over ``steps`` calls, from
    HIP events around the whole loop (device synchronised on both sides)."""
    import torch
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    start.record()
    for _ in range(steps):
        advance()
    stop.record()
    torch.cuda.synchronize()
    return start.elapsed_time(stop) / steps


def rollers_step_bench(args, log):
    """One step = one stochastic Adams-Bashforth step of the rollers
    integrator (rollers example parameters) on the 32768-roller strip:
    blob-blob + one-blob forces, the fused UF+UT product, Lanczos noise to
    the solver tolerance, the two RFD products and the position update."""
    import torch
    from rigidmultiblobswall_amd import wall_rpy, workloads
    from rigidmultiblobswall_amd.rollers_integrator import RollersIntegrator
    w = workloads.c3_rollers_strip()
    p = w["params"]
    r0 = w["r"]
    n = r0.shape[0]
    kw = dict(scheme=p["scheme"], eta=p["eta"], a=p["blob_radius"], kT=p["kT"], g=p["g"],
              repulsion_strength_wall=p["repulsion_strength_wall"], debye_length_wall=p["debye_length_wall"],
              repulsion_strength=p["repulsion_strength"], debye_length=p["debye_length"],
              omega_one_roller=p["omega_one_roller"], free_kinematics=p["free_kinematics"],
              tolerance=p["solver_tolerance"])
    integ = RollersIntegrator(r0, blob_blob_force_implementation="hip", precision=args.precision, seed=11, **kw)
    settle(lambda: integ.advance_time_step(p["dt"]))
    for _ in range(args.warmup):
        integ.advance_time_step(p["dt"])
    torch.cuda.synchronize()
    prod0, it0, inv0 = dict(integ.products), integ.stoch_iterations_count, integ.invalid_configuration_count
    advance = lambda: integ.advance_time_step(p["dt"])
    ms = timed_steps(advance, args.steps)
    per = {k: (integ.products[k] - prod0[k]) / args.steps for k in integ.products}
    lanczos_its = (integ.stoch_iterations_count - it0) / args.steps
    inv = integ.invalid_configuration_count - inv0
    # kernel times from per-launch event records, in steps of their own (the
    # records cost host time, so the step time above is taken without them)
    ksteps = max(1, min(args.steps, 10))
    prod1 = dict(integ.products)
    with wall_rpy.PairKernelTimer() as pk:
        timed_steps(advance, ksteps)
    kper = {k: (integ.products[k] - prod1[k]) / ksteps for k in integ.products}
    nw, tw, mw = pk.by_kind["wall_rpy"]
    nf, tf, mf = pk.by_kind["forces"]
    tw, tf = tw / ksteps * args.steps, tf / ksteps * args.steps
    pairs = float(n) * n
    flop = (kper["UF"] * FLOP_PER_PAIR["UF"] + kper["UF+UT"] * (FLOP_PER_PAIR["UF"] + FLOP_PER_PAIR["UT"])) \
        * args.steps * pairs
    achieved = flop / (tw * 1e-3) / 1e12
    peak = PEAK_FP64_TFLOPS if args.precision == "fp64" else PEAK_FP32_TFLOPS
    log(f"rollers step: {ms:.3f} ms/step, {per} products/step, Lanczos {lanczos_its:.1f} its, "
        f"wall-RPY kernels {tw / args.steps:.3f} ms/step, forces {tf / args.steps:.3f} ms/step")
    out = {
        "metric": "rollers time steps/s (stochastic Adams-Bashforth, 32768 rollers)",
        "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": "rollers example on 32768 rollers in its footprint (39.36 x 3224.3, z = 0.66 + "
                               "Exp(2.25), min distance 1.38): dt 0.016, kT 0.0041419464, omega (0, 62.8, 0), "
                               "free kinematics, blob-blob and wall repulsion, Lanczos tolerance 1e-3",
                   "blobs": n, "precision": args.precision + PRECISION_NOTE.get(args.precision, ""),
                   "per_step": {"products": per, "lanczos_iterations": lanczos_its,
                                "invalid_configurations": inv,
                                "wall_rpy_kernel_ms": tw / args.steps, "force_kernel_ms": tf / args.steps,
                                "pair_kernel_share": (tw + tf) / (ms * args.steps),
                                "kernel_timing_steps": ksteps}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s", "frac": achieved / peak,
                     "traffic": None, "kernel": "wall-RPY pair kernels of the step (UF, UF+UT)",
                     "kernel_mean_ms": mw, "kernel_launches": nw,
                     "flop_per_pair": {"UF": FLOP_PER_PAIR["UF"], "UF+UT": FLOP_PER_PAIR["UF"] + FLOP_PER_PAIR["UT"]}},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        from oracle import rollers_oracle as RO
        rng = np.random.default_rng(5)
        st = RO.RollersStep(r0, p["scheme"], p["eta"], p["blob_radius"], kT=p["kT"], g=p["g"],
                            eps_wall=p["repulsion_strength_wall"], b_wall=p["debye_length_wall"],
                            eps=p["repulsion_strength"], b=p["debye_length"], omega=p["omega_one_roller"],
                            free_kinematics=p["free_kinematics"], tolerance=p["solver_tolerance"], kind="fast")
        draws = iter([(rng.standard_normal(3 * n), rng.standard_normal((n, 3))) for _ in range(4)])
        t = time.perf_counter()
        attempts = st.advance(p["dt"], draws)
        dt_cpu = time.perf_counter() - t
        threads = O.num_threads("fast")
        out["cpu_baseline"] = {"value": 1.0 / dt_cpu, "unit": "steps/s", "cores": threads, "kind": "port",
                               "sample": f"1 stochastic Adams-Bashforth step ({attempts} attempt(s)) of the same "
                                         f"workload from its initial positions: FP64 C oracle products "
                                         f"(-O3 -ffast-math, OpenMP {threads} threads) and numpy Lanczos; "
                                         f"{dt_cpu:.1f} s",
                               "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {dt_cpu:.1f} s per step")
    print(json.dumps(out), flush=True)
    return 0


def lanczos_bench(args, log, ragged: bool):
    """One step = one Brownian-slip solve factor N^1/2 z by Lanczos to the
    tolerance 1e-3 with the grand (trans + rot) mobility: the C3 block of 4
    right-hand sides on 32768 rollers sharing each fused grand product
    (ragged=False), or one right-hand side per replica of the C5 batch (64
    replicas, 128-4096 rollers each) in one persistent ragged launch per
    product (ragged=True)."""
    import torch
    from rigidmultiblobswall_amd import wall_rpy, workloads
    from rigidmultiblobswall_amd.rollers import RaggedRollers, Rollers
    prec = args.precision
    if ragged:
        w = workloads.c5_ragged_rollers()
        off = w["offsets"]
        n = int(off[-1])
        Z = torch.from_numpy(np.random.default_rng(9).standard_normal((1, 6 * n))).cuda()
        R = RaggedRollers(w["r"], off, w["eta"], w["a"], precision=prec)
        pairs = float(np.sum(np.diff(off).astype(np.float64) ** 2))
        k = 1
        desc = (f"C5 ragged batch: {len(off) - 1} roller replicas of 128-4096 blobs ({n} blobs), a=0.656, eta=1e-3, "
                f"grand mobility, 1 RHS per replica, Lanczos tol 1e-3")
    else:
        w = workloads.c3_rollers()
        n = w["r"].shape[0]
        Z = torch.from_numpy(w["z"]).cuda()
        R = Rollers(w["r"], w["eta"], w["a"], precision=prec)
        pairs = float(n) * n
        k = Z.shape[0]
        desc = (f"C3 rollers: N={n}, a=0.656, eta=1e-3, grand mobility (trans + rot), block of {k} RHS sharing "
                f"each product, Lanczos tol 1e-3")
    solve = lambda: R.brownian_noise(Z, factor=1.0, tolerance=1e-3)   # noqa: E731
    settle(solve)
    for _ in range(args.warmup):
        solve()
    torch.cuda.synchronize()
    start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    with wall_rpy.PairKernelTimer() as pk:
        torch.cuda.synchronize()
        start.record()
        for _ in range(args.steps):
            _, its = solve()
        stop.record()
        torch.cuda.synchronize()
    ms = start.elapsed_time(stop) / args.steps
    nw, tw, mw = pk.by_kind["wall_rpy"]
    # algorithmic FLOP of one k-RHS grand product (314 + 72 k per pair)
    flop_product = (314 + 72 * k) * pairs
    launches_per_product = 1
    if not ragged:
        d = wall_rpy.describe(15, n, k, prec)
        launches_per_product = max(1, -(-k // max(1, d["rhs_group"])))
    nprod = nw / launches_per_product
    achieved = flop_product * nprod / (tw * 1e-3) / 1e12 if tw else 0.0
    peak = PEAK_FP64_TFLOPS if prec == "fp64" else PEAK_FP32_TFLOPS
    its_np = np.asarray(its.cpu().numpy() if torch.is_tensor(its) else its)
    log(f"lanczos: {ms:.3f} ms per solve, {nprod / args.steps:.1f} products per solve, pair kernels "
        f"{tw / args.steps:.3f} ms per solve")
    out = {
        "metric": ("Lanczos Brownian slip solves/s (C5 ragged batch)" if ragged else
                   "Lanczos Brownian slip solves/s (block of 4 RHS, 32768 rollers)"),
        "value": 1e3 / ms, "unit": "solves/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64" if prec == "fp64" else "f32", "data": "synthetic",
        "config": {"workload": desc, "blobs": n, "pairs_per_product": pairs,
                   "precision": prec + PRECISION_NOTE.get(prec, ""),
                   "per_solve": {"grand_products": nprod / args.steps, "pair_launches": nw / args.steps,
                                 "pair_kernel_ms": tw / args.steps,
                                 "lanczos_iterations_max": int(its_np.max()) if its_np.size else 0}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     "traffic": None if ragged else load_traffic("c3lz", prec, kernel="grand_k4_kernel"),
                     "kernel": "wall-RPY grand pair kernel" + (" (ragged)" if ragged else f" ({k} RHS)"),
                     "kernel_mean_ms": mw, "kernel_launches": nw, "flop_per_pair": 314 + 72 * k},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        threads = O.num_threads("fast")
        r_h = w["r"]
        nn = r_h.shape[0]
        f_h = np.random.default_rng(3).standard_normal((nn, 3))
        # bounded sample: the oracle's four grand blocks (one RHS) on a row block of the
        # largest system (C3: the whole suspension), scaled to one solve's products x RHS
        r_s = r_h if not ragged else r_h[int(off[np.argmax(np.diff(off))]): int(off[np.argmax(np.diff(off)) + 1])]
        ns = r_s.shape[0]
        f_s = f_h[:ns]
        rows = min(ns, 512)
        t = time.perf_counter()
        reps = 0
        while time.perf_counter() - t < 10.0:
            for blk in (O.UF, O.UT, O.WF, O.WT):
                O.block_apply(blk, r_s, f_s, w["eta"], w["a"], rows=(0, rows), kind="fast")
            reps += 1
        dt_cpu = time.perf_counter() - t
        rate = reps * rows * ns / dt_cpu          # grand pairs/s, one RHS
        sec = nprod / args.steps * k * pairs / rate
        out["cpu_baseline"] = {"value": 1.0 / sec, "unit": "solves/s", "cores": threads, "kind": "port",
                               "sample": f"{reps} x the four grand blocks on rows 0..{rows} of {ns} blobs (one RHS), "
                                         f"FP64 C oracle -O3 -ffast-math, OpenMP {threads} threads, {dt_cpu:.1f} s: "
                                         f"{rate:.3g} grand pairs/s, scaled to the solve's "
                                         f"{nprod / args.steps:.1f} products x {k} RHS (Lanczos algebra not timed)",
                               "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {rate:.3g} grand pairs/s")
    print(json.dumps(out), flush=True)
    return 0


def rigid_step_bench(args, log):
    """One step = one deterministic forward Euler step of the rigid-body
    integrator on the C2 suspension (256 spheres x 42 blobs) with gravity,
    wall and blob-blob repulsion: blob positions, forces and body
    force-torque sums, the block-diagonal preconditioner, the FP64 GMRES
    solve to 1e-8 (warm-started from the previous step) and the body update."""
    import torch
    from rigidmultiblobswall_amd import wall_rpy, workloads
    from rigidmultiblobswall_amd.rigid_integrator import RigidIntegrator
    w = workloads.c2_rigid_spheres()
    fp = dict(g=0.02, repulsion_strength_wall=0.02, debye_length_wall=0.1, repulsion_strength=0.02,
              debye_length=0.1)
    kw = dict(scheme="deterministic_forward_euler", eta=w["eta"], a=w["a"], blob_blob_force_implementation="hip_fp64",
              tolerance=1e-8, **fp)
    prec = "fp64" if args.precision == "fp32_blocked" else args.precision
    integ = RigidIntegrator(w["centres"], w["quats"], w["ref"], precision=prec, solver=args.rigid_solver, **kw)
    dt = 1e-3
    settle(lambda: integ.advance_time_step(dt))
    for _ in range(args.warmup):
        integ.advance_time_step(dt)
    torch.cuda.synchronize()
    it0 = integ.det_iterations_count
    advance = lambda: integ.advance_time_step(dt)
    ms = timed_steps(advance, args.steps)
    its = (integ.det_iterations_count - it0) / args.steps
    # kernel times from per-launch event records, in steps of their own (the
    # records cost host time, so the step time above is taken without them)
    ksteps = max(1, min(args.steps, 10))
    with wall_rpy.PairKernelTimer() as pk:
        timed_steps(advance, ksteps)
    nw, tw, mw = pk.by_kind["wall_rpy"]
    nw_step, tw_step = nw / ksteps, tw / ksteps
    n = w["r"].shape[0]
    pairs = float(n) * n
    peak = PEAK_FP64_TFLOPS if prec == "fp64" else PEAK_FP32_TFLOPS
    achieved = FLOP_PER_PAIR["UF"] * pairs * nw / (tw * 1e-3) / 1e12 if tw else 0.0
    log(f"rigid step ({args.rigid_solver} solver): {ms:.3f} ms/step, {its:.1f} GMRES iterations/step, wall-RPY "
        f"kernels {tw_step:.3f} ms/step ({nw_step:.1f} launches)")
    mixed = args.rigid_solver == "mixed"
    if mixed:
        # FP32-pair inner products and FP64 residuals: the FLOP rate of all
        # launches against the FP32 peak
        peak = PEAK_FP32_TFLOPS
    out = {
        "metric": "rigid-body time steps/s (deterministic forward Euler, GMRES to 1e-8"
                  + (", mixed-precision solver)" if mixed else ")"),
        "value": 1e3 / ms, "unit": "steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": ms, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64" if mixed else ("f64" if prec == "fp64" else "f32"), "data": "synthetic",
        "config": {"workload": "C2 suspension: 256 spheres x 42 blobs = 10752 blobs, a = 0.243, eta = 1, dt 1e-3, "
                               "gravity 0.02 per blob, wall and blob-blob repulsion (0.02, 0.1)",
                   "blobs": n, "precision": prec + PRECISION_NOTE.get(prec, ""),
                   "per_step": {"gmres_iterations": its, "wall_rpy_launches": nw_step,
                                "wall_rpy_kernel_ms": tw_step, "pair_kernel_share": tw_step / ms,
                                "kernel_timing_steps": ksteps}},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak, "traffic": None, "kernel": "wall-RPY UF pair kernel of the GMRES operator",
                     "kernel_mean_ms": mw, "kernel_launches": nw, "flop_per_pair": FLOP_PER_PAIR["UF"]},
    }
    if not args.no_cpu:
        from oracle import oracle as O
        from oracle import rigid_step_oracle as RS
        st = RS.RigidStep(w["centres"], w["quats"], w["ref"], "deterministic_forward_euler", w["eta"], w["a"],
                          g=fp["g"], eps_wall=fp["repulsion_strength_wall"], b_wall=fp["debye_length_wall"],
                          blob_blob=True, eps=fp["repulsion_strength"], b=fp["debye_length"], tol=1e-8)
        t = time.perf_counter()
        st.advance(dt)
        dt_cpu = time.perf_counter() - t
        threads = O.num_threads("strict")
        out["cpu_baseline"] = {"value": 1.0 / dt_cpu, "unit": "steps/s", "cores": threads, "kind": "port",
                               "sample": f"1 forward Euler step of the same suspension: FP64 C oracle products "
                                         f"(OpenMP {threads} threads), numpy preconditioner and GMRES; "
                                         f"{dt_cpu:.1f} s",
                               "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
        log(f"cpu baseline: {dt_cpu:.1f} s per step")
    print(json.dumps(out), flush=True)
    return 0


PROD_TIMING_VERTEX = {1: "blob.vertex", 12: "shell_N_12_Rg_0_7921_Rh_1.vertex",
                      42: "shell_N_42_Rg_0_8913_Rh_1.vertex", 162: "shell_N_162_Rg_0_9497_Rh_1.vertex"}


def prod_timing_workload(blobs_per_body: int):
    """The reference's four-block timing harness, multi_bodies/examples/
    Mobility_Prod_Timing/main.py:47-195 with Multiblob.inputfile: the 2048
    spheres of Const_Torque_t_15.clones (particle radius 1.0155), each a
    multiblob model of 1, 12, 42 or 162 blobs (vertex file scaled by the
    particle radius; blob radius half the smallest blob distance, or the
    particle radius for one blob), eta 0.957e-3, pseudo-periodic lengths
    (128.7923, 128.7923, 0), body centres and blob positions wrapped into
    the box, and X = N(0, 1) of 6 N_blobs values (numpy seed 0) read as rows
    [F_i T_i] per blob."""
    from scipy.spatial.distance import pdist
    from rigidmultiblobswall_amd import workloads
    from rigidmultiblobswall_amd.read_input import read_clones_file, read_vertex_file
    d = os.path.join(REPO, "tests", "golden", "mobility_prod_timing")
    radius = 1.0155
    ref = read_vertex_file(os.path.join(d, PROD_TIMING_VERTEX[blobs_per_body]))[:, :3] * radius
    a = 0.5 * float(pdist(ref).min()) if ref.shape[0] > 1 else radius
    nb, loc, quat = read_clones_file(os.path.join(d, "Const_Torque_t_15.clones"))
    L = np.array([128.7923, 128.7923, 0.0])
    loc = loc.copy()
    for i in range(2):
        loc[:, i] = np.mod(loc[:, i], L[i])
    r = workloads.bodies_blobs(loc, quat, ref)
    for i in range(2):
        r[:, i] = np.mod(r[:, i], L[i])
    X = np.random.RandomState(0).normal(0.0, 1.0, 6 * r.shape[0])    # np.random.seed(0), main.py:189-191
    FT = X.reshape(r.shape[0], 6)
    return {"r": r, "F": np.ascontiguousarray(FT[:, :3]), "T": np.ascontiguousarray(FT[:, 3:]), "eta": 0.957e-3,
            "a": a, "L": L, "bodies": nb}


def prod_timing_bench(args, log):
    """One step = the four separate block products UF F, UT T, WF F, WT T of
    the reference's timing harness (as Wall_Mobility_Mult_Parts, main.py
    :47-77) on device-resident inputs; the fused grand product (all four
    blocks in one pass) is timed beside it."""
    import torch
    from rigidmultiblobswall_amd import wall_rpy
    from rigidmultiblobswall_amd._lib import RMW_UF, RMW_UT, RMW_WF, RMW_WT
    w = prod_timing_workload(args.blobs_per_body)
    r_h, eta, a, L = w["r"], w["eta"], w["a"], w["L"]
    n = r_h.shape[0]
    dev = torch.device("cuda", 0)
    r = torch.from_numpy(r_h).to(dev)
    F = torch.from_numpy(w["F"]).to(dev)
    T = torch.from_numpy(w["T"]).to(dev)
    outs = {b: torch.empty((1, n, 3), dtype=torch.float64, device=dev) for b in ("UF", "UT", "WF", "WT")}
    prec = args.precision
    blocks = (("UF", RMW_UF, F, None, "u"), ("UT", RMW_UT, None, T, "u"), ("WF", RMW_WF, F, None, "w"),
              ("WT", RMW_WT, None, T, "w"))

    def run_block(name, mask, f, t, side):
        o = outs[name]
        wall_rpy.apply(mask, r, force=f, torque=t, eta=eta, a=a, periodic_length=L, precision=prec,
                       out_velocity=o if side == "u" else None, out_omega=o if side == "w" else None)

    def timed(fn, steps):
        torch.cuda.synchronize()
        st, sp = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        with wall_rpy.PairKernelTimer() as pk:
            st.record()
            for _ in range(steps):
                fn()
            sp.record()
            torch.cuda.synchronize()
        return st.elapsed_time(sp) / steps, pk

    settle(lambda: [run_block(*blk) for blk in blocks])
    for _ in range(args.warmup):
        for blk in blocks:
            run_block(*blk)
    per_block = {}
    for blk in blocks:
        ms, pk = timed(lambda blk=blk: run_block(*blk), args.steps)
        per_block[blk[0]] = {"ms": ms, "pair_kernel_ms": pk.mean_ms}
    step_ms = sum(v["ms"] for v in per_block.values())
    uu = torch.empty((1, n, 3), dtype=torch.float64, device=dev)
    ww = torch.empty((1, n, 3), dtype=torch.float64, device=dev)
    grand = lambda: wall_rpy.apply(RMW_UF | RMW_UT | RMW_WF | RMW_WT, r, force=F, torque=T, eta=eta, a=a,  # noqa: E731
                                   periodic_length=L, precision=prec, out_velocity=uu, out_omega=ww)
    grand()
    g_ms, g_pk = timed(grand, args.steps)
    pairs = float(n) * n
    # oracle check of each block on sampled rows (periodic images included)
    from oracle import oracle as O
    rows = oracle_rows(n, blocks=2, rows=64)
    errs = {}
    for name, ob, src in (("UF", O.UF, w["F"]), ("UT", O.UT, w["T"]), ("WF", O.WF, w["F"]), ("WT", O.WT, w["T"])):
        got = outs[name].reshape(n, 3).cpu().numpy()
        num = den = 0.0
        for t0, t1 in rows:
            ref = O.block_apply(ob, r_h, src.ravel(), eta, a, periodic_length=L, rows=(t0, t1)).reshape(n, 3)[t0:t1]
            num += float(np.sum((got[t0:t1] - ref) ** 2))
            den += float(np.sum(ref ** 2))
        errs[name] = float(np.sqrt(num / den))
    log(f"prod timing {args.blobs_per_body} blobs/body ({n} blobs): four blocks {step_ms:.3f} ms "
        f"({ {k: round(v['ms'], 3) for k, v in per_block.items()} }), fused grand {g_ms:.3f} ms, errors {errs}")
    out = {
        "metric": "wall-RPY four-block mobility product (timing harness), blob-pairs/s",
        "value": pairs / (step_ms * 1e-3), "unit": "blob-pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": step_ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64" if prec == "fp64" else "f32", "data": "reference example input",
        "config": {"workload": f"Mobility_Prod_Timing harness: 2048 spheres x {args.blobs_per_body} blobs = {n} "
                               f"blobs, blob a = {a:.6g}, eta 0.957e-3, pseudo-periodic 128.7923 x 128.7923 "
                               f"(9 images per pair), UF F, UT T, WF F, WT T as four products",
                   "blobs": n, "precision": prec + PRECISION_NOTE.get(prec, ""), "per_block": per_block,
                   "fused_grand_ms": g_ms, "fused_grand_pair_kernel_ms": g_pk.mean_ms,
                   "oracle_error_rel_l2": errs, "oracle_rows": int(sum(t1 - t0 for t0, t1 in rows))},
    }
    if not args.no_cpu:
        threads = O.num_threads("fast")
        t = time.perf_counter()
        nrow = 8 if n > 100000 else 32
        for ob, src in ((O.UF, w["F"]), (O.UT, w["T"]), (O.WF, w["F"]), (O.WT, w["T"])):
            O.block_apply(ob, r_h, src.ravel(), eta, a, periodic_length=L, rows=(0, nrow), kind="fast")
        dt_cpu = time.perf_counter() - t
        sec = dt_cpu * n / nrow
        out["cpu_baseline"] = {"value": pairs / sec, "unit": "blob-pairs/s", "cores": threads, "kind": "port",
                               "sample": f"the four blocks on rows 0..{nrow} of {n} (x {n} sources x 9 images), "
                                         f"FP64 C oracle -O3 -ffast-math, OpenMP {threads} threads, {dt_cpu:.1f} s, "
                                         f"scaled to the whole product",
                               "cpu_model": _cpu_model(), "nproc": os.cpu_count()}
    print(json.dumps(out), flush=True)
    return 0


def load_traffic(workload: str, precision: str, kernel: str = "pair_kernel_uf", launched: str = ""):
    """HBM bytes per pair-kernel launch from the committed rocprof PMC summary
    of the same kernel, workload and precision (None when there is none, or
    when the summary counted another kernel variant than `launched`)."""
    path = os.path.join(REPO, "profiles", f"{kernel}_{workload}_{precision}_traffic.json")
    if os.path.exists(path):
        with open(path) as fh:
            rec = json.load(fh)
        if launched and not launched.startswith(rec.get("kernel", "") + "<"):
            return None
        return rec.get("traffic_bytes")
    return None


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="c3")
    ap.add_argument("--precision", default="fp32_blocked")
    ap.add_argument("--domain", default="single_wall", choices=["single_wall", "no_wall", "in_plane", "free_surface"],
                    help="boundary of the UF product (workloads c1, c2, c3) or of the radii product "
                         "(c3r: single_wall, no_wall, free_surface)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-gmres", action="store_true", help="skip the learned-warm-start GMRES leg")
    ap.add_argument("--no-fp64", action="store_true", help="skip the FP64-mode leg of the UF workloads")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="only run a streaming copy of MIB MiB (counter calibration) and exit")
    ap.add_argument("--rigid-solver", default="fp64", choices=["fp64", "mixed"],
                    help="c2step: FP64 GMRES, or FP32-pair inner GMRES refined to the FP64 tolerance")
    ap.add_argument("--blobs-per-body", type=int, default=12, choices=[1, 12, 42, 162],
                    help="prodtiming: the multiblob model of the timing harness")
    ap.add_argument("--gmres-steps", type=int, default=200)
    ap.add_argument("--gmres-epochs", type=int, default=120)
    ap.add_argument("--gmres-evals", type=int, default=5)
    args = ap.parse_args(argv)
    if args.gpus != 1:
        print(f"bench.py: measures one GPU; --gpus {args.gpus} is not supported (the product runs on one "
              "device; sharding the targets over several GPUs is not built)", file=sys.stderr)
        return 2

    import torch
    from rigidmultiblobswall_amd import wall_rpy
    from rigidmultiblobswall_amd._lib import RMW_UF

    torch.cuda.set_device(0)
    if args.calibrate_copy:
        # plain streaming copy of known size (y = 1 * x, beta = 0) for scaling
        # the FETCH_SIZE / WRITE_SIZE counters of the traffic runs
        from rigidmultiblobswall_amd import _lib
        nbytes = int(args.calibrate_copy) << 20
        x = torch.ones(nbytes // 8, dtype=torch.float64, device="cuda")
        y = torch.empty_like(x)
        lib = _lib.load()
        for _ in range(5):
            _lib.check(lib.rmw_axpby(1.0, x.data_ptr(), 0.0, y.data_ptr(), x.numel(), None,
                                     torch.cuda.current_stream().cuda_stream))
        torch.cuda.synchronize()
        assert bool(torch.equal(x, y))
        print(json.dumps({"calibrate_copy_bytes": nbytes}), flush=True)
        return 0
    if args.workload == "c3step":
        return rollers_step_bench(args, lambda m: print(m, file=sys.stderr, flush=True))
    if args.workload == "c2step":
        return rigid_step_bench(args, lambda m: print(m, file=sys.stderr, flush=True))
    if args.workload == "prodtiming":
        return prod_timing_bench(args, lambda m: print(m, file=sys.stderr, flush=True))
    if args.workload in ("c3lz", "c5lz"):
        return lanczos_bench(args, lambda m: print(m, file=sys.stderr, flush=True), ragged=args.workload == "c5lz")
    dev = torch.device("cuda", 0)
    w, desc = load_workload(args.workload)
    r_h, eta, a = w["r"], w["eta"], w["a"]
    n = r_h.shape[0]
    rng = np.random.default_rng(7)
    f_h = rng.standard_normal(3 * n)
    r = torch.from_numpy(r_h).to(dev)
    f = torch.from_numpy(f_h).to(dev)
    u = torch.empty((1, n, 3), dtype=torch.float64, device=dev)

    radii = "radii" in w
    fparams = w.get("forces")
    if radii or fparams:
        u2 = torch.empty((n, 3), dtype=torch.float64, device=dev)
    if radii:
        if args.domain not in RADII_BOUNDARY:
            raise SystemExit(f"the radii product has no {args.domain} form")
        rbnd = RADII_BOUNDARY[args.domain]
        rad = torch.from_numpy(w["radii"]).to(dev)
    if fparams:
        from rigidmultiblobswall_amd import forces
        fprec = "fp64" if args.precision == "fp64" else "fp32"

    # the UF lines run the prepared product (launch shape, workspace and
    # checks settled once; the same launches as wall_rpy.apply, bit for bit),
    # so that host time per call stays far below the kernel's
    prod = None if radii or fparams else \
        wall_rpy.Product(RMW_UF, r, eta=eta, a=a, precision=args.precision, domain=args.domain)

    def step():
        if fparams:
            forces.blob_forces(r, precision=fprec, out=u2, **fparams)
        elif radii:
            wall_rpy.apply_radii_self(r, f, rad, eta, precision=args.precision, out=u2, boundary=rbnd)
        else:
            prod(force=f, out_velocity=u)

    # the FP64-mode product first (its oracle comparison comes later, so that
    # no host work separates it from the timed region)
    u64 = None
    if not radii and not fparams and args.domain == "single_wall" and args.precision != "fp64" \
            and not args.no_fp64:
        u64 = torch.empty_like(u)
        fp64_res = fp64_leg(r, f, u64, r_h, f_h, eta, a, max(5, args.steps // 2), 8, check=False)
    # then the timed product itself for 40 ms, untimed: from idle, and again
    # after the FP64 leg, the FP32 pair kernel's first launches at 32768 blobs
    # take 1.2-1.4 ms and settle at ~1.06 ms after ~35 ms of its own load
    settle(step)
    for _ in range(args.warmup):
        step()
    ms = timed_steps(step, args.steps)
    # the pair kernel's own duration: events recorded around each launch on
    # its stream, over as many products again (taken apart from the product
    # time above, which the records would otherwise include)
    with wall_rpy.PairKernelTimer() as pk:
        ms_rec = timed_steps(step, args.steps)
    pairs = float(n) * n
    value = pairs / (ms * 1e-3)
    op = ("forces" if fparams else ("radii" if args.domain == "single_wall" else f"radii_{args.domain}") if radii
          else "UF" if args.domain == "single_wall" else f"UF_{args.domain}")
    kernel_ms = pk.mean_ms
    flop = FLOP_PER_PAIR[op] * pairs
    achieved = flop / (kernel_ms * 1e-3) / 1e12
    # the force kernel computes in FP64 in both modes (FP32 only for exp)
    peak = PEAK_FP64_TFLOPS if args.precision == "fp64" or fparams else PEAK_FP32_TFLOPS
    info = wall_rpy.device_info()
    from rigidmultiblobswall_amd import autotune
    shape = autotune.lookup(RMW_UF, args.precision, 1, float(n) * n, False)
    launch = wall_rpy.describe(RMW_UF, n, 1, args.precision)
    if args.domain != "single_wall":
        sym = float(n) * n / (256.0 * 256.0) / 2 >= info.get("cu_count", 256)
        launch = {"kernel": "symmetric" if sym else "ordered", "domain": args.domain}
        kname_dom = (f"{'mutual_kernel' if sym else 'pair_kernel_f32' if args.precision != 'fp64' else 'pair_kernel_f64'}"
                     f"<UF,K=1,{args.precision},{args.domain}>")
    launch["source"] = "autotuner cache" if shape is not None else "built-in heuristic"
    if shape is not None:
        launch["max_wgs"] = shape.max_wgs
    kname = {("symmetric", True): "mutual_kernel<UF,K=1,{p}>", ("symmetric", False): "mutual_kernel<UF,K=1,fp64>",
             ("ordered", True): "pair_kernel_f32<UF,K=1,{p}>", ("ordered", False): "pair_kernel_f64<UF,K=1>"}[
        (launch["kernel"], args.precision != "fp64")].format(p=args.precision)
    if args.domain != "single_wall" and not radii and not fparams:
        kname = kname_dom
        desc = desc.replace("UF block", f"UF block, {args.domain.replace('_', ' ')}")
    if radii:
        launch = wall_rpy.describe_radii_self(n, args.precision)
        kname = f"{'radii_mutual_kernel' if launch['kernel'] == 'symmetric' else 'radii_kernel'}<{args.precision}>"
        if args.domain != "single_wall":
            launch["boundary"] = rbnd
            kname = kname[:-1] + f",{rbnd}>"
            desc = desc.replace("different-radii wall RPY", f"different-radii RPY, {rbnd.replace('_', ' ')}")
    if fparams:
        launch = forces.describe(n, fprec)
        kname = (f"{'force_mutual_kernel' if launch['kernel'] == 'symmetric' else 'force_kernel'}"
                 f"<{'fp64' if fprec == 'fp64' else 'fp32 exp'}>")
    out = {
        "metric": METRIC,
        "value": value,
        "unit": "blob-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.precision == "fp64" else "f32",
        "data": "synthetic",
        "config": {"workload": desc, "blobs": n, "pairs_per_step": pairs,
                   "precision": args.precision + PRECISION_NOTE.get(args.precision, ""),
                   "device": info, "launch_shape": launch},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "TFLOP/s",
                     "frac": achieved / peak,
                     # the committed PMC summaries count the single-wall kernels only
                     "traffic": load_traffic(args.workload, fprec if fparams else args.precision,
                                             "force_kernel" if fparams else "radii_kernel" if radii
                                             else "pair_kernel_uf", kname)
                                if args.domain == "single_wall" else None,
                     "kernel": kname,
                     "kernel_mean_ms": kernel_ms, "kernel_median_ms": pk.median_ms, "kernel_launches": pk.launches,
                     # the product time of the kernel-timed loop (its event records included)
                     "kernel_loop_ms_per_step": ms_rec,
                     "flop_per_pair": FLOP_PER_PAIR[op]},
    }
    log = lambda m: print(m, file=sys.stderr, flush=True)  # noqa: E731
    log(f"matvec: {ms:.4f} ms/step, {value:.4g} pairs/s, pair kernel {kernel_ms:.4f} ms (median {pk.median_ms:.4f}; "
        f"{ms_rec:.4f} ms/step in its timed loop)")
    if fparams:
        out["metric"] = "blob-blob soft-repulsion forces, blob-pairs/s"
        out["dtype"] = "f64"
        out["config"]["precision"] = {"fp64": "FP64 1/r and exp; FP64 sums",
                                      "fp32": "FP64 geometry and sums, FP32 exponential"}[fprec]
    if not radii and not fparams and args.domain == "single_wall":
        out["oracle_error"] = oracle_error(u, r_h, f_h, eta, a)
        log(f"oracle error (rel L2, {out['oracle_error']['rows']} rows): {out['oracle_error']['rel_l2']:.3e}")
        if u64 is not None:
            fp64_res["oracle_error"] = oracle_error(u64, r_h, f_h, eta, a)
            out["fp64"] = fp64_res
            log(f"fp64 mode: {out['fp64']['ms_per_step']:.4f} ms/step, pair kernel {out['fp64']['kernel_mean_ms']:.4f} ms, "
                f"oracle error {out['fp64']['oracle_error']['rel_l2']:.3e}")
        out["end_to_end"] = end_to_end(r_h, f_h.reshape(-1, 3), eta, a, args.precision)
        log(f"end to end through the numpy signature: {out['end_to_end']['ms']:.3f} ms")
    if not args.no_gmres:
        out["gmres"] = gmres_leg(args.gmres_steps, args.gmres_epochs, args.gmres_evals, log)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(r_h, f_h, eta, a, radii=w.get("radii"), force_params=fparams,
                                           boundary=rbnd if radii else "wall")
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
