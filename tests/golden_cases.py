"""Inputs of the reference's pair_active_rods example (the only committed
numeric outputs of the rigid mobility solve), read from the data files
copied into tests/golden/pair_active_rods/.

Two active extensile rods above the wall, GMRES tolerance 1e-8, eta 1e-3,
gravity 0, body force/torque from force_*_resolution.dat, tangential slip
from slip_extensile_rod (multi_bodies/examples/pair_active_rods/
user_defined_functions.py:52-101, restated below), blob radius per
resolution from inputfile_*_resolution.dat.  Expected body velocities
(v, w) per body: run_*_res.velocity.dat.reference.
"""
import os

import numpy as np

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "pair_active_rods")

CASES = {
    "low": ("Cylinder_N_14_Lg_1_9295_Rg_0_18323", "force_low_resolution.dat", "run_low_res"),
    "mid": ("Cylinder_N_86_Lg_1_9384_Rg_0_1484", "force_mid_resolution.dat", "run_mid_res"),
    "high": ("Cylinder_N_324_Lg_2_0299_Rg_0_1554", "force_high_resolution.dat", "run_high_res"),
}


def _read_input(res):
    vals = {}
    with open(os.path.join(HERE, f"inputfile_{res}_resolution.dat")) as fh:
        for line in fh:
            line = line.split("#")[0].strip()
            if not line:
                continue
            parts = line.split()
            vals[parts[0]] = parts[1:]
    return vals


def _rotation(q):
    s, p = q[0], np.asarray(q[1:4])
    d = s * s - 0.5
    return 2.0 * np.array([[p[0] ** 2 + d, p[0] * p[1] - s * p[2], p[0] * p[2] + s * p[1]],
                           [p[1] * p[0] + s * p[2], p[1] ** 2 + d, p[1] * p[2] - s * p[0]],
                           [p[2] * p[0] - s * p[1], p[2] * p[1] + s * p[0], p[2] ** 2 + d]])


def slip_extensile_rod(r_body, location, blob_radius):
    """Tangential slip of speed 20 along the rod axis, pointing to the
    nearest end, on the part of the surface within 0.8 of the ends
    (user_defined_functions.py:52-101)."""
    speed = -20.0
    nb = r_body.shape[0]
    covering, perimeter = {14: (0, 0), 86: (1, 6), 324: (6, 12)}[nb]
    if nb > 14:
        axis = r_body[nb - 2 * covering - 2] - r_body[perimeter - 2]
    else:
        axis = r_body[nb - 1] - r_body[0]
    length_rod = np.linalg.norm(axis) + 2.0 * blob_radius
    axis = axis / np.sqrt(axis @ axis)
    lower = length_rod / 2.0 - 0.8
    upper = length_rod / 2.0
    slip = np.zeros((nb, 3))
    for i in range(nb):
        if covering > 0 and i >= nb - 2 * covering:
            continue
        d = (r_body[i] - location) @ axis
        if lower < d <= upper:
            slip[i] = -speed * axis
        elif -upper <= d < -lower:
            slip[i] = speed * axis
    return slip


def load_case(res):
    name, force_file, out = CASES[res]
    inp = _read_input(res)
    a = float(inp["blob_radius"][0])
    eta = float(inp["eta"][0])
    tol = float(inp["solver_tolerance"][0])
    vtx = np.loadtxt(os.path.join(HERE, name + ".vertex"), skiprows=1)[:, :3]
    with open(os.path.join(HERE, name + ".clones")) as fh:
        lines = [ln.split() for ln in fh if ln.strip() and not ln.startswith("#")]
    nbodies = int(lines[0][0])
    clones = np.array([[float(v) for v in ln] for ln in lines[1:1 + nbodies]])
    locations, quats = clones[:, :3], clones[:, 3:7]
    r_bodies = [vtx @ _rotation(q).T + c for c, q in zip(locations, quats)]
    slip = np.vstack([slip_extensile_rod(rb, c, a) for rb, c in zip(r_bodies, locations)])
    ft = np.loadtxt(os.path.join(HERE, force_file)).reshape(nbodies, 6)
    expected = np.loadtxt(os.path.join(HERE, out + ".velocity.dat.reference")).reshape(nbodies, 6)
    return {"locations": locations, "quats": quats, "ref": vtx, "a": a, "eta": eta, "tol": tol,
            "slip": slip, "force_torque": ft, "expected": expected}
