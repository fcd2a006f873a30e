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


def _floats(text, default: str) -> np.ndarray:
    return np.array([float(v) for v in (text or default).split()], dtype=np.float64)


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
        o = self.options.get
        self.n_steps = int(o("n_steps") or 0)
        self.initial_step = int(o("initial_step") or 0)
        self.n_save = int(o("n_save") or 1)
        self.n_relaxation = int(o("n_relaxation") or 0)
        self.dt = float(o("dt") or 0.0)
        self.eta = float(o("eta") or 1.0)
        self.g = float(o("g") or 1.0)
        self.theta = float(o("tilt_angle") or 0.0)
        self.blob_radius = float(o("blob_radius") or 1.0)
        self.tracer_radius = float(o("tracer_radius") or 0.0)
        self.kT = float(o("kT") or 1.0)
        self.scheme = str(o("scheme") or "deterministic_forward_euler")
        self.output_name = str(o("output_name") or "run")
        self.random_state = o("random_state")
        self.seed = o("seed")
        self.repulsion_strength_wall = float(o("repulsion_strength_wall") or 1.0)
        self.debye_length_wall = float(o("debye_length_wall") or 1.0)
        self.mobility_blobs_implementation = str(o("mobility_blobs_implementation") or "python")
        self.mobility_vector_prod_implementation = str(o("mobility_vector_prod_implementation") or "python")
        self.repulsion_strength = float(o("repulsion_strength") or 1.0)
        self.debye_length = float(o("debye_length") or 1.0)
        self.blob_blob_force_implementation = str(o("blob_blob_force_implementation") or "None")
        self.body_body_force_torque_implementation = str(o("body_body_force_torque_implementation") or "None")
        self.save_body_mobility = str(o("save_body_mobility") or "False")
        self.save_blobs_mobility = str(o("save_blobs_mobility") or "False")
        self.save_velocities = str(o("save_velocities") or "False")
        self.slip_file = o("slip_file")
        self.force_file = o("force_file")
        self.velocity_file = o("velocity_file")
        self.solver_tolerance = float(o("solver_tolerance") or 1e-08)
        self.nonlinear_solver_tolerance = float(o("nonlinear_solver_tolerance") or 1e-08)
        self.rf_delta = float(o("rf_delta") or 1e-03)
        self.save_clones = str(o("save_clones") or "one_file_per_step")
        self.periodic_length = _floats(o("periodic_length"), "0 0 0")
        self.omega_one_roller = _floats(o("omega_one_roller"), "0 0 0")
        self.free_kinematics = str(o("free_kinematics") or "True")
        self.hydro_interactions = int(o("hydro_interactions") or 1)
        self.update_PC = int(o("update_PC") or 1)
        self.domain = str(o("domain") or "single_wall")
        self.green_particles = np.array([int(v) for v in (o("green_particles") or "0 0").split()])
        self.cells = np.array([int(v) for v in (o("cells") or "1 1").split()])
        self.sample_HydroGrid = int(o("sample_HydroGrid") or 1)
        self.save_HydroGrid = int(o("save_HydroGrid") or 0)
        self.call_HydroGrid = str(o("call_HydroGrid") or "False") == "True"
        self.repulsion_strength_firm = float(o("repulsion_strength_firm") or 0.0)
        self.firm_delta = float(o("firm_delta") or 1e-02)
        self.Lub_Cut = float(o("Lub_Cut") or 4.5)
        self.zmin = float(o("zmin") or 0)
        self.zmax = float(o("zmax") or 1e7)
        self.domType = str(o("domType") or "RPB")
        self.diffusion_coefficient = float(o("diffusion_coefficient") or 1.0)

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
