"""Run a simulation from a reference input file on the GPU.

    python -m rigidmultiblobswall_amd.run_simulation --input-file inputfile.dat

The time loop of multi_bodies/multi_bodies.py:1112-1540 for
  * the rollers schemes (names ending in '_rollers', one blob per body):
    forward Euler, first-order stochastic, deterministic and stochastic
    Adams-Bashforth (rollers_integrator.RollersIntegrator);
  * rigid multiblob bodies: deterministic_forward_euler and
    deterministic_adams_bashforth (rigid_integrator.RigidIntegrator), with
    the structures' optional .slip files.
It reads the input file and its structures (.vertex, .clones, .slip; paths
relative to the working directory as in the reference), writes
`<output_name>.bodies_info`, saves the configuration every n_save steps in
the reference's 'one_file' (`<output_name>.<ID>.config`) or
'one_file_per_step' (`<output_name>.<ID>.<step>.clones`) layout, and prints
the reference's progress line.  The reference's implementation names
select the GPU kernels: every mobility implementation runs the HIP wall-RPY
product (precision from --precision, or from a 'hip_<precision>' name;
rigid bodies default to FP64 for the GMRES operator); blob-blob force names
'python', 'C++', 'pycuda', 'numba' and 'hip*' run the HIP pair kernel,
'tree_numba' / 'tree_hip*' its cutoff form, 'None' nothing.
"""
from __future__ import annotations

import argparse
import sys
import time

import numpy as np

from .read_input import ReadInput, read_clones_file, read_slip_file, read_vertex_file, write_clones

_DENSE_FORCE = ("python", "C++", "pycuda", "numba")


def blob_blob_name(name: str) -> str:
    """The HIP force implementation serving a reference implementation name."""
    if name == "None":
        return "None"
    if name in _DENSE_FORCE:
        return "hip"
    if name == "tree_numba":
        return "tree_hip"
    if name.startswith("hip") or name.startswith("tree_hip"):
        return name
    raise ValueError(f"blob_blob_force_implementation {name!r} has no GPU counterpart for rollers")


def mobility_precision(name: str, default: str) -> str:
    for p in ("fp32_blocked", "fp32_kahan", "fp64", "fp32"):
        if name.endswith("_" + p):
            return p
    return default


def build(read: ReadInput, precision: str):
    """(integrator, body counts per structure, IDs, orientations)."""
    from .rollers_integrator import RollersIntegrator
    if read.scheme.find("rollers") == -1:
        raise ValueError(f"scheme {read.scheme!r}: this driver runs the rollers schemes only")
    locs, quats, counts = [], [], []
    for vertex_file, clones_file in (s[:2] for s in read.structures):
        v = read_vertex_file(vertex_file)
        if v.shape[0] != 1:
            raise ValueError(f"{vertex_file}: rollers have one blob per body, found {v.shape[0]}")
        n, loc, q = read_clones_file(clones_file)
        if np.any(v[0, :3] != 0.0):
            raise ValueError(f"{vertex_file}: the rollers integrator moves blob centres; put the blob at the "
                             "body origin")
        locs.append(loc)
        quats.append(q)
        counts.append(n)
    r = np.concatenate(locs)
    seed = None if read.seed is None else int(read.seed)
    integ = RollersIntegrator(
        r, scheme=read.scheme, eta=read.eta, a=read.blob_radius, kT=read.kT, g=read.g,
        repulsion_strength_wall=read.repulsion_strength_wall, debye_length_wall=read.debye_length_wall,
        blob_blob_force_implementation=blob_blob_name(read.blob_blob_force_implementation),
        repulsion_strength=read.repulsion_strength, debye_length=read.debye_length,
        periodic_length=read.periodic_length, omega_one_roller=read.omega_one_roller,
        free_kinematics=read.free_kinematics, hydro_interactions=read.hydro_interactions,
        tolerance=read.solver_tolerance, rf_delta=read.rf_delta, domain=read.domain,
        precision=mobility_precision(read.mobility_vector_prod_implementation, precision), seed=seed)
    return integ, counts, np.concatenate(quats)


RIGID_SCHEMES = ("deterministic_forward_euler", "deterministic_adams_bashforth")


class _RigidRun:
    """Adapter giving the rigid integrator the driver's interface."""

    def __init__(self, integ):
        self.integ = integ

    def __getattr__(self, name):
        return getattr(self.integ, name)

    def locations(self):
        return self.integ.centres

    def orientations(self):
        return self.integ.quats

    @property
    def wall_overlaps(self):
        return 0

    @property
    def stoch_iterations_count(self):
        return 0


def build_rigid(read: ReadInput, precision: str):
    from .rigid_integrator import RigidIntegrator
    locs, quats, refs, slips, counts = [], [], [], [], []
    have_slip = False
    for struct in read.structures:
        ref = read_vertex_file(struct[0])[:, :3]
        n, loc, q = read_clones_file(struct[1])
        slip = None
        for extra in struct[2:]:
            if extra.endswith(".slip"):
                slip = read_slip_file(extra)
                have_slip = True
        locs.append(loc)
        quats.append(q)
        refs += [ref] * n
        slips += [np.zeros_like(ref) if slip is None else slip] * n
        counts.append(n)
    integ = RigidIntegrator(
        np.concatenate(locs), np.concatenate(quats), refs, scheme=read.scheme, eta=read.eta, a=read.blob_radius,
        g=read.g, repulsion_strength_wall=read.repulsion_strength_wall, debye_length_wall=read.debye_length_wall,
        blob_blob_force_implementation=blob_blob_name(read.blob_blob_force_implementation),
        repulsion_strength=read.repulsion_strength, debye_length=read.debye_length,
        periodic_length=read.periodic_length, tolerance=read.solver_tolerance,
        precision=mobility_precision(read.mobility_vector_prod_implementation, precision),
        slip_reference=slips if have_slip else None, domain=read.domain, update_PC=read.update_PC)
    return _RigidRun(integ), counts, None


def run(input_file: str, precision: str = "fp32_blocked", log=print) -> dict:
    read = ReadInput(input_file)
    if read.scheme in RIGID_SCHEMES:
        integ, counts, quats = build_rigid(read, "fp64" if precision == "fp32_blocked" else precision)
    elif read.scheme.find("rollers") != -1:
        integ, counts, quats = build(read, precision)
    else:
        raise ValueError(f"scheme {read.scheme!r} is not provided; rollers schemes (*_rollers) and "
                         f"{', '.join(RIGID_SCHEMES)}")
    out = read.output_name
    nblobs = integ.n if hasattr(integ, "n") else sum(
        read_vertex_file(st[0]).shape[0] * counts[i] for i, st in enumerate(read.structures))
    with open(out + ".bodies_info", "w") as f:
        f.write("num_of_body_types  " + str(len(counts)) + "\n")
        f.write("body_names         " + str(read.structures_ID) + "\n")
        f.write("body_types         " + str(counts) + "\n")
        f.write("num_bodies         " + str(sum(counts)) + "\n")
        f.write("num_blobs          " + str(nblobs) + "\n")
    files = []
    if read.save_clones == "one_file":
        files = [open(out + "." + ID + ".config", "w") for ID in read.structures_ID]
    elif read.save_clones != "one_file_per_step":
        raise ValueError(f"save_clones {read.save_clones!r} is not implemented; use one_file_per_step or one_file")

    def save(step):
        loc = integ.locations()
        qs = quats if quats is not None else integ.orientations()
        off = 0
        for i, ID in enumerate(read.structures_ID):
            sl = slice(off, off + counts[i])
            if files:
                write_clones(files[i], loc[sl], qs[sl])
            else:
                with open(out + "." + ID + "." + str(step).zfill(8) + ".clones", "w") as f:
                    write_clones(f, loc[sl], qs[sl])
            off += counts[i]

    start = time.time()
    step = read.initial_step - 1
    for step in range(read.initial_step, read.n_steps):
        if step % read.n_save == 0 and step >= 0:
            log(f"Integrator =  {read.scheme} , step =  {step} , invalid configurations "
                f"{integ.invalid_configuration_count} , wallclock time =  {time.time() - start}")
            save(step)
        integ.advance_time_step(read.dt, step=step)
    if (step + 1) % read.n_save == 0 and step >= 0:
        log(f"Integrator =  {read.scheme} , step =  {step + 1} , invalid configurations "
            f"{integ.invalid_configuration_count} , wallclock time =  {time.time() - start}")
        save(step + 1)
    for f in files:
        f.close()
    return {"steps": read.n_steps - read.initial_step, "seconds": time.time() - start,
            "invalid_configurations": integ.invalid_configuration_count, "wall_overlaps": integ.wall_overlaps,
            "stoch_iterations": integ.stoch_iterations_count, "det_iterations": integ.det_iterations_count}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--input-file", required=True)
    ap.add_argument("--precision", default="fp32_blocked",
                    help="pair-term precision of the mobility products (fp32_blocked, fp32, fp32_kahan, fp64); "
                         "rigid bodies use fp64 unless another mode is named")
    args = ap.parse_args(argv)
    info = run(args.input_file, args.precision)
    print(info, flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
