"""GPU parity of the rigid-body path (FP64 operator, body preconditioner,
GMRES) against the oracle and the reference's golden solve outputs."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
import golden_cases as G
from rigidmultiblobswall_amd import krylov
from rigidmultiblobswall_amd import mobility as mob
from rigidmultiblobswall_amd import workloads
from rigidmultiblobswall_amd.multi_bodies import set_mobility_blobs, set_mobility_vector_prod
from rigidmultiblobswall_amd.rigid import RigidSuspension

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(np.asarray(b)))


def _small_spheres(nb=16, seed=21):
    w = workloads.c2_rigid_spheres(seed=seed, nbodies=nb)
    return w


def _oracle_for(w):
    bodies = [(c, q, w["ref"]) for c, q in zip(w["centres"], w["quats"])]
    return O.RigidSystem(bodies, w["eta"], w["a"])


def test_dense_wall_mobility_matches_oracle(gpu):
    rng = np.random.default_rng(3)
    n, a = 60, 0.13
    r = np.column_stack([rng.uniform(-1, 1, n), rng.uniform(-1, 1, n), rng.uniform(0.05, 0.8, n)])
    print(f"case dense N={n} (with z < a)", flush=True)
    M = set_mobility_blobs("hip")(r, 0.9, a)
    R = O.single_wall_fluid_mobility(r, 0.9, a)
    assert np.abs(M - R).max() <= 1e-13 * np.abs(R).max()


def test_body_blobs_operator_and_pc(gpu):
    w = _small_spheres()
    print("case 16 spheres x 42 blobs: blobs, operator, PC", flush=True)
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    Og = _oracle_for(w)
    assert np.abs(S.r_blobs.cpu().numpy() - Og.r).max() < 1e-13
    rng = np.random.default_rng(4)
    x = rng.standard_normal(S.size)
    xd = torch.from_numpy(x).to(gpu)
    y = torch.empty_like(xd)
    S.operator(xd, y)
    assert rel(y.cpu().numpy(), Og.apply(x)) <= 1e-12
    S.build_preconditioner()
    Og.build_pc()
    S.pc_apply(xd, y)
    assert rel(y.cpu().numpy(), Og.pc_apply(x)) <= 1e-11
    # the body blocks are the exact inverse of the isolated-body operator
    Minv0 = S.minv[: (3 * 42) ** 2].reshape(126, 126).cpu().numpy()
    assert np.abs(Minv0 - Og.Minv[0]).max() <= 1e-10 * np.abs(Og.Minv[0]).max()
    assert np.abs(S.nbody[0].cpu().numpy() - Og.N[0]).max() <= 1e-10 * np.abs(Og.N[0]).max()


@pytest.mark.parametrize("wall", [True, False])
def test_point_field_matches_oracle(gpu, wall):
    """Velocity at every blob from the net forces of all other bodies (each
    a sphere of radius R at its centre) = the oracle's different-radii
    product with the blob's own body left out; one centre below R (damped
    source)."""
    w = _small_spheres(nb=16, seed=31)
    c = w["centres"].copy()
    c[3, 2] = 1.1                       # below R = 1.2: the source is damped
    R = 1.2
    S = RigidSuspension(c, w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    rng = np.random.default_rng(31)
    ft = rng.standard_normal((16, 6))
    print(f"case point field 16 spheres wall={wall}", flush=True)
    u = S.point_field(ft, radius=R, wall=wall).cpu().numpy()
    rb = S.r_blobs.cpu().numpy()
    ref = np.empty_like(rb)
    for b in range(16):
        others = np.arange(16) != b
        ref[42 * b: 42 * (b + 1)] = O.radii_apply(c[others], rb[42 * b: 42 * (b + 1)], ft[others, :3],
                                                  np.full(15, R), np.full(42, w["a"]), w["eta"],
                                                  wall=wall).reshape(42, 3)
    assert np.abs(u - ref).max() <= 1e-12 * np.abs(ref).max()


def test_point_field_guess_is_a_block_jacobi_step(gpu):
    """P^-1 (b - [u_point; 0]) against P^-1 b and the exact block-Jacobi step
    P^-1 (2 b - A P^-1 b): most of the exact step's gain without the blob
    product."""
    w = workloads.c4_monolayer(nbodies=64, seed=5)
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    rng = np.random.default_rng(5)
    ft = np.column_stack([rng.normal(0, 0.1, (64, 2)), rng.normal(-0.8, 0.1, 64), rng.normal(0, 0.1, (64, 3))])
    rhs = S.rhs(np.zeros((S.Nblobs, 3)), ft)
    print("case point-field guess, 64-sphere monolayer", flush=True)
    x = S.solve_mobility_problem(rhs=rhs, tol=1e-10).x
    pc = torch.empty_like(rhs)
    S.pc_apply(rhs, pc)
    Apc = torch.empty_like(rhs)
    S.operator(pc, Apc)
    ex = torch.empty_like(rhs)
    S.pc_apply(2 * rhs - Apc, ex)
    pt = S.point_field_guess(rhs, ft, radius=0.9)
    err = lambda v: float(torch.linalg.norm(v - x) / torch.linalg.norm(x))   # noqa: E731
    print(f"  errors: P^-1 b {err(pc):.3f}, exact block-Jacobi {err(ex):.3f}, point field {err(pt):.3f}")
    assert err(pt) < 0.8 * err(pc)
    assert (err(pc) - err(pt)) >= 0.5 * (err(pc) - err(ex))    # at least half the exact step's gain


def test_preconditioner_ragged_bodies(gpu):
    """Bodies of 1, 7, 16, 42 and 50 blobs in one suspension: the register
    Gauss-Jordan build (up to 45 blobs, the elements of the largest body per
    thread) and the global-memory Cholesky build (50 blobs) against the
    oracle's Cholesky-based M_b^-1 and N_b, and the applied preconditioner."""
    rng = np.random.default_rng(9)
    ico = workloads.icosphere_42(1.0)
    sizes = [1, 7, 16, 42, 50]
    refs = []
    for m in sizes:
        if m <= 42:
            refs.append(ico[:m].copy())
        else:
            extra = rng.standard_normal((m - 42, 3))
            extra = 0.6 * extra / np.linalg.norm(extra, axis=1, keepdims=True)   # inside the shell
            refs.append(np.vstack([ico, extra]))
    c = np.column_stack([np.arange(len(sizes)) * 3.0, np.zeros(len(sizes)), np.full(len(sizes), 1.6)])
    q = workloads.random_quaternions(rng, len(sizes))
    print("case preconditioner build, ragged bodies 1..50 blobs", flush=True)
    S = RigidSuspension(c, q, refs, 0.8, 0.2, precision="fp64")
    S.build_preconditioner()
    Og = O.RigidSystem([(c[k], q[k], refs[k]) for k in range(len(sizes))], 0.8, 0.2)
    Og.build_pc()
    off = 0
    for k, m in enumerate(sizes):
        n = 3 * m
        Mk = S.minv[off: off + n * n].reshape(n, n).cpu().numpy()
        off += n * n
        assert np.abs(Mk - Og.Minv[k]).max() <= 1e-9 * np.abs(Og.Minv[k]).max(), m
        assert np.abs(S.nbody[k].cpu().numpy() - Og.N[k]).max() <= 1e-9 * np.abs(Og.N[k]).max(), m
    x = rng.standard_normal(S.size)
    y = torch.empty(S.size, dtype=torch.float64, device=gpu)
    S.pc_apply(torch.from_numpy(x).to(gpu), y)
    assert rel(y.cpu().numpy(), Og.pc_apply(x)) <= 1e-9


@pytest.mark.parametrize("res,tol", [("low", 1e-8), ("mid", 1e-8), ("high", 5e-8)])
def test_pair_active_rods_golden_gpu(gpu, res, tol):
    c = G.load_case(res)
    print(f"case pair_active_rods {res}: fp64 GMRES tol {c['tol']}", flush=True)
    S = RigidSuspension(c["locations"], c["quats"], c["ref"], c["eta"], c["a"], precision="fp64")
    out = S.solve_mobility_problem(c["slip"], c["force_torque"], tol=c["tol"])
    assert out.info == 0
    vel = S.velocities(out.x)
    err = rel(vel, c["expected"])
    print(f"  iterations {out.iterations}, rel err vs reference file {err:.2e}")
    assert err < tol
    # FP32 pair terms: same solve, velocities within the FP32 budget
    out32 = S.solve_mobility_problem(c["slip"], c["force_torque"], tol=1e-7, precision="fp32")
    assert rel(S.velocities(out32.x), c["expected"]) < 1e-6


def test_c2_solve_residual_and_warm_start(gpu):
    w = workloads.c2_rigid_spheres()
    print("case C2 256 spheres x 42 blobs: fp64 GMRES to 1e-8", flush=True)
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    slip = np.zeros((S.Nblobs, 3))
    out = S.solve_mobility_problem(slip, w["force_torque"], tol=1e-8)
    assert out.info == 0
    rhs = S.rhs(slip, w["force_torque"])
    r = torch.empty_like(rhs)
    S.operator(out.x, r)
    true_res = float(torch.linalg.norm(r - rhs) / torch.linalg.norm(rhs))
    print(f"  iterations {out.iterations}, true residual {true_res:.2e}")
    assert true_res <= 1.5e-8
    # warm start from the solution: converged at once
    again = S.solve_mobility_problem(slip, w["force_torque"], tol=1e-8, x0=out.x)
    assert again.info == 0 and again.iterations <= 1
    # warm start from a perturbed solution: fewer iterations than cold
    rng = np.random.default_rng(5)
    pert = out.x * (1 + 1e-4 * torch.from_numpy(rng.standard_normal(S.size)).to(gpu))
    warm = S.solve_mobility_problem(slip, w["force_torque"], tol=1e-8, x0=pert)
    assert warm.info == 0 and warm.iterations < out.iterations


def test_c2_mixed_solve_reaches_fp64_tolerance(gpu):
    """FP32-pair inner GMRES with FP64 residual refinement: the FP64 residual
    of the result meets 1e-8 from a zero start and from warm starts (whose
    first residual is formed with the FP32-pair operator)."""
    w = workloads.c2_rigid_spheres()
    print("case C2 256 spheres x 42 blobs: mixed-precision solve to 1e-8", flush=True)
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    slip = np.zeros((S.Nblobs, 3))
    rhs = S.rhs(slip, w["force_torque"])

    def true_res(x):
        r = torch.empty_like(rhs)
        S.operator(x, r, precision="fp64")
        return float(torch.linalg.norm(r - rhs) / torch.linalg.norm(rhs))

    ref = S.solve_mobility_problem(rhs=rhs, tol=1e-8)
    cold = S.solve_mixed(rhs=rhs, tol=1e-8)
    print(f"  fp64 iterations {ref.iterations}, mixed iterations {cold.iterations}, matvecs {cold.matvecs}")
    assert cold.info == 0 and true_res(cold.x) <= 1.5e-8
    assert abs(cold.final_residual - true_res(cold.x)) <= 1e-9
    assert float(torch.linalg.norm(cold.x - ref.x) / torch.linalg.norm(ref.x)) < 1e-5
    rng = np.random.default_rng(8)
    pert = ref.x * (1 + 1e-3 * torch.from_numpy(rng.standard_normal(S.size)).to(gpu))
    warm = S.solve_mixed(rhs=rhs, tol=1e-8, x0=pert)
    assert warm.info == 0 and true_res(warm.x) <= 1.5e-8 and warm.iterations < cold.iterations
    exact = S.solve_mixed(rhs=rhs, tol=1e-8, x0=ref.x)
    assert exact.info == 0 and true_res(exact.x) <= 1.5e-8 and exact.iterations <= 4
    with pytest.raises(ValueError):
        S.solve_mixed(rhs=rhs, max_outer=0)


def test_gmres_unverified_inner_solve(gpu):
    """verify=False stops on the Arnoldi estimate without the product that
    recomputes b - A x; the estimate then bounds the true residual."""
    w = workloads.c2_rigid_spheres()
    S = RigidSuspension(w["centres"], w["quats"], w["ref"], w["eta"], w["a"], precision="fp64")
    rhs = S.rhs(np.zeros((S.Nblobs, 3)), w["force_torque"])
    b = rhs / torch.linalg.norm(rhs)
    S.build_preconditioner()
    full = krylov.gmres(lambda v, o: S.operator(v, o), b, Pinv=S.pc_apply, tol=1e-6, restart=60)
    fast = krylov.gmres(lambda v, o: S.operator(v, o), b, Pinv=S.pc_apply, tol=1e-6, restart=60, verify=False)
    assert fast.info == 0 and fast.matvecs == full.matvecs - 1 and fast.iterations == full.iterations
    r = torch.empty_like(b)
    S.operator(fast.x, r)
    assert float(torch.linalg.norm(r - b)) <= 2e-6
    assert abs(fast.final_residual - float(torch.linalg.norm(r - b))) < 1e-8


def test_dispatch_names(gpu):
    f = set_mobility_vector_prod("hip_fp64")
    rng = np.random.default_rng(6)
    r = np.column_stack([rng.uniform(0, 5, 50), rng.uniform(0, 5, 50), rng.uniform(0.3, 2, 50)])
    v = rng.standard_normal(150)
    assert rel(f(r, v, 1.0, 0.2), O.uf(r, v, 1.0, 0.2)) <= 1e-12
    with pytest.raises(ValueError):
        set_mobility_vector_prod("numba")
