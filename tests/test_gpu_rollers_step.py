"""The rollers time step on the GPU against the oracle's restatement of the
reference integrator (quaternion_integrator_rollers.py), on the same random
draws: every first-order scheme, hydrodynamic and uncorrelated velocities,
free and constrained kinematics, the wall retry, and the example's
parameters (inputfile_rollers.dat)."""
import numpy as np
import pytest

from oracle import rollers_oracle as R
from rigidmultiblobswall_amd.rollers_integrator import RollersIntegrator

pytestmark = pytest.mark.gpu

ETA, A, KT, G = 1.0e-3, 0.656, 0.0041419464, 0.0024892
EPS, B = 0.0165677856, 0.0656
DT = 0.016


def _config(n, seed):
    rng = np.random.default_rng(seed)
    side = 2.2 * np.sqrt(n)
    return np.column_stack([rng.uniform(0, side, n), rng.uniform(0, side, n), rng.uniform(0.8, 3.0, n)]), rng


class Draws:
    """The same draws for both sides: per attempt one z (3N) and one dx (N, 3)."""

    def __init__(self, z_list, dx_list):
        self.z, self.dx = list(z_list), list(dx_list)
        self.iz = self.idx = 0

    def gpu(self, kind, shape):
        if kind == "z":
            self.iz += 1
            return self.z[self.iz - 1].reshape(shape)
        self.idx += 1
        return self.dx[self.idx - 1].reshape(shape)

    def oracle(self):
        return iter(zip(self.z, self.dx))


def _pair(r, scheme, hydro, free, prec, draws, omega=(0.0, 62.8, 0.0), tol=1e-10):
    common = dict(eta=ETA, a=A, kT=KT, g=G, omega=omega, free_kinematics=free, hydro_interactions=hydro,
                  tolerance=tol)
    gpu = RollersIntegrator(r, scheme=scheme, eta=ETA, a=A, kT=KT, g=G, repulsion_strength_wall=EPS,
                            debye_length_wall=B, blob_blob_force_implementation="hip_fp64" if prec == "fp64"
                            else "hip", repulsion_strength=EPS, debye_length=B, omega_one_roller=omega,
                            free_kinematics=free, hydro_interactions=hydro, tolerance=tol, precision=prec,
                            noise_source=draws.gpu)
    orc = R.RollersStep(r, scheme, ETA, A, kT=KT, g=G, eps_wall=EPS, b_wall=B, eps=EPS, b=B, omega=omega,
                        free_kinematics=free, hydro_interactions=hydro, tolerance=tol)
    return gpu, orc


CASES = [(s, h, f) for s in ("deterministic_forward_euler_rollers", "stochastic_first_order_rollers",
                             "deterministic_adams_bashforth_rollers", "stochastic_adams_bashforth_rollers")
         for h in (1, 0) for f in ("True", "False")]


@pytest.mark.parametrize("scheme,hydro,free", CASES)
def test_step_matches_oracle(gpu, scheme, hydro, free):
    n, steps = 150, 3
    r0, rng = _config(n, 1)
    draws = Draws(rng.standard_normal((steps, 3 * n)), rng.standard_normal((steps, n, 3)))
    g, o = _pair(r0, scheme, hydro, free, "fp64", draws)
    od = draws.oracle()
    print(f"case rollers step {scheme} hydro={hydro} free={free}", flush=True)
    for _ in range(steps):
        g.advance_time_step(DT)
        o.advance(DT, od)
    disp = np.linalg.norm(o.r - r0)
    assert np.linalg.norm(g.locations() - o.r) <= 1e-8 * disp
    assert g.wall_overlaps == o.wall_overlaps and g.invalid_configuration_count == 0


def test_step_fp32_products(gpu):
    """Default precision (FP32 pair terms): positions agree with the FP64
    oracle to the product accuracy."""
    n, steps = 400, 2
    r0, rng = _config(n, 2)
    draws = Draws(rng.standard_normal((steps, 3 * n)), rng.standard_normal((steps, n, 3)))
    g, o = _pair(r0, "stochastic_adams_bashforth_rollers", 1, "True", "fp32_blocked", draws, tol=1e-8)
    od = draws.oracle()
    print("case rollers step fp32_blocked", flush=True)
    for _ in range(steps):
        g.advance_time_step(DT)
        o.advance(DT, od)
    assert np.linalg.norm(g.locations() - o.r) <= 1e-5 * np.linalg.norm(o.r - r0)


def test_invalid_attempt_is_redrawn(gpu):
    n = 60
    r0, rng = _config(n, 3)
    bad = np.zeros(3 * n)
    bad[2] = -1e4                    # pushes roller 0 far below the wall
    zs = [bad, rng.standard_normal(3 * n)]
    dxs = [rng.standard_normal((n, 3)), rng.standard_normal((n, 3))]
    draws = Draws(zs, dxs)
    g, o = _pair(r0, "stochastic_adams_bashforth_rollers", 1, "True", "fp64", draws)
    print("case rollers invalid attempt", flush=True)
    g.advance_time_step(DT)
    assert o.advance(DT, draws.oracle()) == 2
    assert g.invalid_configuration_count == 1
    assert np.linalg.norm(g.locations() - o.r) <= 1e-8 * np.linalg.norm(o.r - r0)
    # a deterministic step that crosses the wall cannot be redrawn
    rd = r0.copy()
    rd[0, 2] = 0.05
    d = RollersIntegrator(rd, scheme="deterministic_forward_euler_rollers", eta=ETA, a=A, g=1e3)
    with pytest.raises(RuntimeError, match="below the wall"):
        d.advance_time_step(10.0)


def test_example_parameters_run(gpu):
    """The example's own input (first 256 rollers of its clones file,
    tolerance 1e-3, pycuda names mapped to the HIP kernels) runs and keeps
    every roller above the wall."""
    import os
    from rigidmultiblobswall_amd import read_input as RI
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rollers")
    rd = RI.ReadInput(os.path.join(gdir, "inputfile_rollers.dat"))
    _, loc, _ = RI.read_clones_file(os.path.join(gdir, "rollers_256.clones"))
    g = RollersIntegrator(loc, scheme=rd.scheme, eta=rd.eta, a=rd.blob_radius, kT=rd.kT, g=rd.g,
                          repulsion_strength_wall=rd.repulsion_strength_wall,
                          debye_length_wall=rd.debye_length_wall, blob_blob_force_implementation="hip",
                          repulsion_strength=rd.repulsion_strength, debye_length=rd.debye_length,
                          omega_one_roller=rd.omega_one_roller, free_kinematics=rd.free_kinematics,
                          tolerance=rd.solver_tolerance, seed=1)
    print("case rollers example parameters", flush=True)
    for _ in range(rd.n_steps):
        g.advance_time_step(rd.dt)
    x = g.locations()
    assert np.isfinite(x).all() and (x[:, 2] > 0).all()
    # free kinematics with omega_y > 0: the rollers translate along +x on average
    assert (x[:, 0] - loc[:, 0]).mean() > 0


@pytest.mark.parametrize("save", ["one_file", "one_file_per_step"])
def test_run_driver_writes_the_reference_layout(gpu, tmp_path, save):
    """run_rollers on the example's input (first 256 rollers, 3 steps):
    bodies_info, one frame per saved step in the .config / .clones layout."""
    import os
    import re
    from rigidmultiblobswall_amd import read_input as RI
    from rigidmultiblobswall_amd import run_rollers
    gdir = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rollers")
    text = open(os.path.join(gdir, "inputfile_rollers.dat")).read()
    text = text.replace("../../Structures/blob.vertex rollers.clones",
                        f"{gdir}/blob.vertex {gdir}/rollers_256.clones")
    text = re.sub(r"^n_steps.*$", "n_steps 3", text, flags=re.M)
    text = re.sub(r"^save_clones.*$", f"save_clones {save}", text, flags=re.M)
    text = text.replace("run_rollers", str(tmp_path / "run"))
    p = tmp_path / "in.dat"
    p.write_text(text + "\nseed 4\n")
    rd = RI.ReadInput(str(p))
    assert rd.n_steps == 3 and rd.save_clones == save
    print(f"case run_rollers driver {save}", flush=True)
    info = run_rollers.run(str(p), log=lambda m: None)
    assert info["steps"] == 3 and info["invalid_configurations"] == 0
    assert (tmp_path / "run.bodies_info").read_text().splitlines()[3].split()[-1] == "256"
    if save == "one_file":
        lines = (tmp_path / "run.rollers_256.config").read_text().splitlines()
        assert len(lines) == 4 * 257 and lines[0] == "256" and lines[257] == "256"
        first = np.array([[float(v) for v in ln.split()] for ln in lines[1:257]])
        _, loc, _ = RI.read_clones_file(os.path.join(gdir, "rollers_256.clones"))
        assert np.array_equal(first[:, :3], loc)
    else:
        for step in range(4):
            n, loc, q = RI.read_clones_file(str(tmp_path / f"run.rollers_256.{step:08d}.clones"))
            assert n == 256 and (loc[:, 2] > 0).all()
