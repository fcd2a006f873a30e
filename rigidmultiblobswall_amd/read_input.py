"""Input formats of the reference: the key/value ``.inputfile``, ``.vertex``
and ``.clones`` files, and the ``.clones`` / ``.config`` writers of its time
loop.

  ReadInput(path)        read_input/read_input.py:9-159
  read_vertex_file(p)    read_input/read_vertex_file.py:7-31
  read_clones_file(p)    read_input/read_clones_file.py:11-52
  read_slip_file(p)      read_input/read_slip_file.py:8-35
  write_clones(...)      multi_bodies/multi_bodies.py:1430-1456 (one frame)

The option names, defaults and parsing rules (comments after '#', the first
word is the key and the rest of the line its value, repeated 'structure' /
'obstacle' / 'articulated' keys numbered in order, restart files when
initial_step > 0) are the reference's, so its input files are read
unchanged.  Orientations are returned as an (N, 4) array of unit
quaternions (s, p) rather than Quaternion objects.
"""
from __future__ import annotations

import ntpath
from typing import List, Tuple

import numpy as np


def _strip(line: str) -> str:
    if "#" in line:
        line = line.split("#", 1)[0]
    return line.strip()


def _num(kind):
    return lambda text, default: kind(text or default)


def _array(kind):
    return lambda text, default: np.array([kind(v) for v in (text or default).split()],
                                          dtype=np.float64 if kind is float else None)


def _raw(text, default):
    return text if text is not None else default


def _flag(text, default):
    return str(text or default) == "True"


# The option table of the reference's input files (name, default and how the
# value is read: read_input/read_input.py:48-102): attribute, key in the
# file, converter, default.  Empty values take the default, as there.
_OPTIONS = [
    ("n_steps", "n_steps", _num(int), 0),
    ("initial_step", "initial_step", _num(int), 0),
    ("n_save", "n_save", _num(int), 1),
    ("n_relaxation", "n_relaxation", _num(int), 0),
    ("dt", "dt", _num(float), 0.0),
    ("eta", "eta", _num(float), 1.0),
    ("g", "g", _num(float), 1.0),
    ("theta", "tilt_angle", _num(float), 0.0),
    ("blob_radius", "blob_radius", _num(float), 1.0),
    ("tracer_radius", "tracer_radius", _num(float), 0.0),
    ("kT", "kT", _num(float), 1.0),
    ("scheme", "scheme", _num(str), "deterministic_forward_euler"),
    ("output_name", "output_name", _num(str), "run"),
    ("random_state", "random_state", _raw, None),
    ("seed", "seed", _raw, None),
    ("repulsion_strength_wall", "repulsion_strength_wall", _num(float), 1.0),
    ("debye_length_wall", "debye_length_wall", _num(float), 1.0),
    ("mobility_blobs_implementation", "mobility_blobs_implementation", _num(str), "python"),
    ("mobility_vector_prod_implementation", "mobility_vector_prod_implementation", _num(str), "python"),
    ("repulsion_strength", "repulsion_strength", _num(float), 1.0),
    ("debye_length", "debye_length", _num(float), 1.0),
    ("blob_blob_force_implementation", "blob_blob_force_implementation", _num(str), "None"),
    ("body_body_force_torque_implementation", "body_body_force_torque_implementation", _num(str), "None"),
    ("save_body_mobility", "save_body_mobility", _num(str), "False"),
    ("save_blobs_mobility", "save_blobs_mobility", _num(str), "False"),
    ("save_velocities", "save_velocities", _num(str), "False"),
    ("slip_file", "slip_file", _raw, None),
    ("force_file", "force_file", _raw, None),
    ("velocity_file", "velocity_file", _raw, None),
    ("solver_tolerance", "solver_tolerance", _num(float), 1e-08),
    ("nonlinear_solver_tolerance", "nonlinear_solver_tolerance", _num(float), 1e-08),
    ("rf_delta", "rf_delta", _num(float), 1e-03),
    ("save_clones", "save_clones", _num(str), "one_file_per_step"),
    ("periodic_length", "periodic_length", _array(float), "0 0 0"),
    ("omega_one_roller", "omega_one_roller", _array(float), "0 0 0"),
    ("free_kinematics", "free_kinematics", _num(str), "True"),
    ("hydro_interactions", "hydro_interactions", _num(int), 1),
    ("update_PC", "update_PC", _num(int), 1),
    ("domain", "domain", _num(str), "single_wall"),
    ("green_particles", "green_particles", _array(int), "0 0"),
    ("cells", "cells", _array(int), "1 1"),
    ("sample_HydroGrid", "sample_HydroGrid", _num(int), 1),
    ("save_HydroGrid", "save_HydroGrid", _num(int), 0),
    ("call_HydroGrid", "call_HydroGrid", _flag, "False"),
    ("repulsion_strength_firm", "repulsion_strength_firm", _num(float), 0.0),
    ("firm_delta", "firm_delta", _num(float), 1e-02),
    ("Lub_Cut", "Lub_Cut", _num(float), 4.5),
    ("zmin", "zmin", _num(float), 0.0),
    ("zmax", "zmax", _num(float), 1e7),
    ("domType", "domType", _num(str), "RPB"),
    ("diffusion_coefficient", "diffusion_coefficient", _num(float), 1.0),
]


class ReadInput:
    """Options of a simulation input file (read_input/read_input.py:9-159)."""

    def __init__(self, entries: str):
        self.entries = entries
        self.input_file = entries
        self.options = {}
        counts = {"structure": 0, "obstacle": 0, "articulated": 0}
        with open(self.input_file, "r") as f:
            for line in f:
                line = _strip(line)
                if line == "":
                    continue
                parts = line.split(None, 1)
                option = parts[0]
                value = parts[1] if len(parts) > 1 else ""
                if option in counts:
                    idx = counts[option]
                    counts[option] += 1
                    option += str(idx)
                self.options[option] = value
        for attr, key, conv, default in _OPTIONS:
            setattr(self, attr, conv(self.options.get(key), default))
        o = self.options.get

        self.num_free_bodies = counts["structure"]
        self.structures: List[List[str]] = []
        self.structures_ID: List[str] = []
        self.articulated: List[List[str]] = []
        self.articulated_ID: List[str] = []
        for i in range(counts["structure"]):
            self.structures.append(str(o("structure" + str(i))).split())
        for i in range(counts["obstacle"]):
            self.structures.append(str(o("obstacle" + str(i))).split())
        for i in range(counts["articulated"]):
            files = str(o("articulated" + str(i))).split()
            self.articulated_ID.append(ntpath.split(files[1])[1][:-7])
            self.articulated.append(files)
        for struct in self.structures:
            self.structures_ID.append(ntpath.split(struct[1])[1][:-7])   # file name without '.clones'
        if self.initial_step > 0:
            for k, struct in enumerate(self.structures):
                struct[1] = (self.output_name + "." + self.structures_ID[k] + "." + str(self.initial_step).zfill(8)
                             + ".clones")


def read_vertex_file(name_file: str) -> np.ndarray:
    """Blob coordinates (Nblobs, 3) -- or (Nblobs, 4) when the file gives a
    radius per blob -- of a .vertex file (read_vertex_file.py:7-31)."""
    coor = []
    i = 0
    with open(name_file, "r") as f:
        for line in f:
            line = _strip(line)
            if line == "":
                continue
            if i > 0:
                coor.append([float(v) for v in line.split()])
            i += 1
    return np.array(coor)


def read_clones_file(name_file: str) -> Tuple[int, np.ndarray, np.ndarray]:
    """(number_of_bodies, locations (N, 3), orientations (N, 4) unit
    quaternions) of a .clones file (read_clones_file.py:11-52): the first
    non-comment line is the count, then x y z s p1 p2 p3 per body; lines
    beyond the count are ignored."""
    locations, orientations = [], []
    number_of_bodies = 0
    i = 0
    with open(name_file, "r") as f:
        for line in f:
            line = _strip(line)
            if line == "":
                continue
            if i == 0:
                number_of_bodies = int(line.split()[0])
            else:
                d = line.split()
                locations.append([float(d[0]), float(d[1]), float(d[2])])
                q = np.array([float(d[3]), float(d[4]), float(d[5]), float(d[6])])
                orientations.append(q / np.linalg.norm(q))
            i += 1
            if i == number_of_bodies + 1:
                break
    return number_of_bodies, np.array(locations).reshape(-1, 3), np.array(orientations).reshape(-1, 4)


def read_slip_file(name_file: str) -> np.ndarray:
    """Blob slip (Nblobs, 3) in the body's reference frame
    (read_input/read_slip_file.py:8-35): a count line, then three numbers
    per blob."""
    slip = []
    i = 0
    with open(name_file, "r") as f:
        for line in f:
            line = _strip(line)
            if line == "":
                continue
            if i > 0:
                d = line.split()
                slip.append([float(d[0]), float(d[1]), float(d[2])])
            i += 1
    return np.array(slip).reshape(-1, 3)


def write_clones(f, locations: np.ndarray, orientations: np.ndarray) -> None:
    """One frame in the .clones / .config layout of the reference's time
    loop (multi_bodies.py:1430-1456): the count, then
    'x y z s p1 p2 p3' per body with Python's shortest float repr."""
    n = locations.shape[0]
    f.write(str(n) + "\n")
    for j in range(n):
        loc, q = locations[j], orientations[j]
        f.write("%s %s %s %s %s %s %s\n" % (repr(float(loc[0])), repr(float(loc[1])), repr(float(loc[2])),
                                            repr(float(q[0])), repr(float(q[1])), repr(float(q[2])),
                                            repr(float(q[3]))))
