"""The learned previous-step warm start against the previous-step guess.

The previous step's solution is the natural initial guess of a Brownian
trajectory (the reference starts every GMRES solve from zero or from
P^-1 b: multi_bodies/multi_bodies.py:1235-1265); the 'previous' network
adds a learned correction to it from the motion of the bodies since that
step.  On the C4 monolayer as configured (1024 spheres x 42 blobs), trained
on the first steps of a stochastic Euler-Maruyama trajectory, on held-out
later steps its x0 must be closer to the FP64 solution than the
previous-step guess itself and FP64 GMRES from it must need no more
iterations, both solves converging to the same true FP64 residual.
"""
import numpy as np
import pytest

from rigidmultiblobswall_amd import workloads

pytestmark = pytest.mark.gpu


def test_previous_network_beats_previous_step_guess(gpu):
    from rigidmultiblobswall_amd.trajectory import run_c4_trajectory
    from rigidmultiblobswall_amd.warmstart import WarmStart, evaluate
    w = workloads.c4_monolayer()
    nsteps, ntr = 56, 48
    print(f"case previous-step network, C4 {w['centres'].shape[0]} bodies x {nsteps} steps", flush=True)
    traj = run_c4_trajectory(w["centres"], w["quats"], w["ref"], w["a"], w["eta"], nsteps=nsteps,
                             log=lambda m: print("  " + m, flush=True) if m.startswith("step 0:") else None)
    ws = WarmStart(w["ref"], w["a"], w["eta"], device=gpu, mode="previous")
    fit = ws.fit(traj.snapshots[:ntr], epochs=120)
    print(f"  trained {fit['steps']} Adam steps, loss {fit['loss_history'][0]:.3e} -> {fit['loss_history'][-1]:.3e}",
          flush=True)
    base = WarmStart(w["ref"], w["a"], w["eta"], device=gpu)   # evaluate() needs the point network too
    picks = [ntr + 1, ntr + 4, nsteps - 1]
    rows = evaluate(base, [traj.snapshots[i] for i in picks], [traj.snapshots[i - 1] for i in picks], ws_prev=ws)
    e_net, e_prev, it_net, it_prev = [], [], [], []
    for row in rows:
        net, prev = row["gnn_prev/fp64"], row["previous/fp64"]
        print(f"  x0 error {net['x0_rel_err']:.4f} (previous step {prev['x0_rel_err']:.4f}), FP64 iterations "
              f"{net['iterations']} ({prev['iterations']}), {net['total_ms']:.1f} ms ({prev['total_ms']:.1f} ms)",
              flush=True)
        assert net["converged"] and prev["converged"] and row["gnn_prev/mixed"]["converged"]
        e_net.append(net["x0_rel_err"]); e_prev.append(prev["x0_rel_err"])
        it_net.append(net["iterations"]); it_prev.append(prev["iterations"])
    assert np.mean(e_net) < np.mean(e_prev)
    assert max(a - b for a, b in zip(e_net, e_prev)) < 0.0
    assert np.mean(it_net) <= np.mean(it_prev)
