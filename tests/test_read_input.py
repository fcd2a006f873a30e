"""The reference's input formats, read from its own rollers example files
(tests/golden/rollers: inputfile_rollers.dat, blob.vertex, the first 256
bodies of rollers.clones)."""
import io
import os

import numpy as np

from rigidmultiblobswall_amd import read_input as RI

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rollers")


def test_rollers_inputfile():
    r = RI.ReadInput(os.path.join(G, "inputfile_rollers.dat"))
    assert r.scheme == "stochastic_adams_bashforth_rollers"
    assert r.mobility_vector_prod_implementation == "pycuda"
    assert r.blob_blob_force_implementation == "pycuda"
    assert r.repulsion_strength == 0.0165677856 and r.debye_length == 0.0656
    assert r.dt == 0.016 and r.n_steps == 10 and r.n_save == 1 and r.initial_step == 0
    assert r.solver_tolerance == 1e-3
    assert r.eta == 1.0e-3 and r.g == 0.0024892 and r.blob_radius == 0.656 and r.kT == 0.0041419464
    assert np.array_equal(r.omega_one_roller, [0.0, 62.8, 0.0])
    assert r.free_kinematics == "True"
    assert r.repulsion_strength_wall == 0.0165677856 and r.debye_length_wall == 0.0656
    assert r.output_name == "run_rollers" and r.save_clones == "one_file"
    assert r.structures == [["../../Structures/blob.vertex", "rollers.clones"]]
    assert r.structures_ID == ["rollers"]
    # options absent from the file take the reference's defaults
    assert r.domain == "single_wall" and r.hydro_interactions == 1 and r.rf_delta == 1e-3
    assert np.array_equal(r.periodic_length, [0.0, 0.0, 0.0]) and r.seed is None


def test_defaults_comments_and_restart(tmp_path):
    p = tmp_path / "in.dat"
    p.write_text("# only comments\n\nscheme deterministic_forward_euler_rollers # trailing\n"
                 "structure a.vertex b.clones\nstructure c.vertex dir/d.clones\ninitial_step 12\n"
                 "output_name out\nperiodic_length 3 4 0\n")
    r = RI.ReadInput(str(p))
    assert r.scheme == "deterministic_forward_euler_rollers"
    assert r.kT == 1.0 and r.eta == 1.0 and r.solver_tolerance == 1e-8 and r.save_clones == "one_file_per_step"
    assert r.structures_ID == ["b", "d"]
    assert r.structures[1][1] == "out.d.00000012.clones"
    assert np.array_equal(r.periodic_length, [3.0, 4.0, 0.0])


def test_vertex_and_clones_files():
    v = RI.read_vertex_file(os.path.join(G, "blob.vertex"))
    assert v.shape == (1, 3) and not v.any()
    n, loc, q = RI.read_clones_file(os.path.join(G, "rollers_256.clones"))
    assert n == 256 and loc.shape == (256, 3) and q.shape == (256, 4)
    assert np.array_equal(loc[0], [31.665, 2900.8, 0.81197]) and np.array_equal(loc[1], [27.954, 682.96, 6.7172])
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0) and np.array_equal(q[:, 0], np.ones(256))


def test_clones_round_trip(tmp_path):
    n, loc, q = RI.read_clones_file(os.path.join(G, "rollers_256.clones"))
    buf = io.StringIO()
    RI.write_clones(buf, loc, q)
    p = tmp_path / "x.clones"
    p.write_text(buf.getvalue())
    n2, loc2, q2 = RI.read_clones_file(str(p))
    assert n2 == n and np.array_equal(loc2, loc) and np.array_equal(q2, q)
    assert buf.getvalue().splitlines()[1] == "31.665 2900.8 0.81197 1.0 0.0 0.0 0.0"
    # a vertex file with a radius column keeps it
    p = tmp_path / "r.vertex"
    p.write_text("2\n0 0 0 0.5\n1 0 0 0.25 # blob two\n")
    assert np.array_equal(RI.read_vertex_file(str(p)), [[0, 0, 0, 0.5], [1, 0, 0, 0.25]])


def test_slip_file(tmp_path):
    p = tmp_path / "a.slip"
    p.write_text("# slip of two blobs\n2\n0.1 0.2 0.3\n-1 0 2.5 # second\n")
    assert np.array_equal(RI.read_slip_file(str(p)), [[0.1, 0.2, 0.3], [-1.0, 0.0, 2.5]])
