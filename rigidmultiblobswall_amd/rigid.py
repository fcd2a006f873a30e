"""Rigid multiblob suspensions on the GPU (FP64 solver, FP32/FP64 matvec).

RigidSuspension holds the bodies of one suspension on the device and
provides the pieces of the reference's rigid mobility solve:

  operator(x, out)        [M lambda - K U; -K^T lambda]
                          (multi_bodies/multi_bodies.py:424-477)
  build_preconditioner()  per-body Cholesky of the dense wall mobility,
                          N_b = pinv(K^T M_b^-1 K) (multi_bodies.py:752-822)
  pc_apply(x, out)        block-diagonal preconditioner (multi_bodies.py:825-849)
  solve_mobility_problem  normalised-RHS right-preconditioned GMRES
                          (quaternion_integrator_multi_bodies.py:1441-1536)
  build_stochastic_preconditioner, mobility_pc, P_inv_mult, brownian_slip
                          the stochastic block-diagonal preconditioner and
                          the Brownian slip L (P M P^T)^1/2 z it serves
                          (multi_bodies.py:480-623, stochastic_forcing.py
                          :112-262)

The blob mobility product M lambda is librmw_hip's wall-RPY kernel; its
pair arithmetic is chosen by `precision` ('fp64', 'fp32', 'fp32_blocked', ...).
The operator, K products, preconditioner and Krylov algebra are FP64.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import torch

from . import _lib, krylov, wall_rpy
from . import _stream as _stream_mod
from ._lib import RMW_UF


def _stream():
    return _stream_mod.current()


class RigidSuspension:
    """Rigid bodies sharing blob radius `a` above a no-slip wall
    (domain 'single_wall') or in unbounded fluid ('no_wall': the unbounded
    RPY product and the dense RPY tensor in the preconditioner, as the
    reference's 'numba_no_wall' / 'python_no_wall' implementations).

    locations (B, 3); orientations (B, 4) unit quaternions (s, p);
    reference: one (nb, 3) array shared by all bodies, or a list of per-body
    (nb_b, 3) arrays (blob coordinates at quaternion (1, 0, 0, 0)).
    """

    def __init__(self, locations, orientations, reference, eta: float, a: float, periodic_length=None,
                 precision: str = "fp64", device=None, domain: str = "single_wall"):
        if not torch.cuda.is_available():
            raise RuntimeError("RigidSuspension needs a ROCm GPU (there is no CPU fallback)")
        if domain not in ("single_wall", "no_wall"):
            raise ValueError("rigid bodies: domain 'single_wall' or 'no_wall'")
        self.domain = domain
        self._dom = wall_rpy.DOMAINS[domain]
        self.lib = _lib.load()
        self.device = device or torch.device("cuda", torch.cuda.current_device())
        loc = np.ascontiguousarray(np.asarray(locations, dtype=np.float64).reshape(-1, 3))
        quat = np.ascontiguousarray(np.asarray(orientations, dtype=np.float64).reshape(-1, 4))
        self.nbodies = loc.shape[0]
        if quat.shape[0] != self.nbodies:
            raise ValueError("one quaternion per body is required")
        if isinstance(reference, (list, tuple)):
            refs = [np.asarray(rf, dtype=np.float64).reshape(-1, 3)[:, :3] for rf in reference]
            if len(refs) != self.nbodies:
                raise ValueError("one reference configuration per body is required")
        else:
            rf = np.asarray(reference, dtype=np.float64).reshape(-1, 3)
            refs = [rf] * self.nbodies
        sizes = np.array([rf.shape[0] for rf in refs], dtype=np.int64)
        self.body_offsets = np.ascontiguousarray(np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64))
        self.Nblobs = int(self.body_offsets[-1])
        self.size = 3 * self.Nblobs + 6 * self.nbodies
        self.eta, self.a = float(eta), float(a)
        self.L = np.zeros(3) if periodic_length is None else np.asarray(periodic_length, dtype=np.float64)
        self.precision = precision
        dev = self.device
        self.centres = torch.from_numpy(loc).to(dev)
        self.quats = torch.from_numpy(quat).to(dev)
        self.ref = torch.from_numpy(np.ascontiguousarray(np.vstack(refs))).to(dev)
        nbytes, nminv, nscr = ctypes.c_size_t(), ctypes.c_size_t(), ctypes.c_size_t()
        _lib.check(self.lib.rmw_rigid_workspace_size(self._offp(), self.nbodies, ctypes.byref(nbytes),
                                                     ctypes.byref(nminv), ctypes.byref(nscr)))
        self.ws = torch.empty(max(nbytes.value, 256), dtype=torch.uint8, device=dev)
        self.minv_doubles = nminv.value
        self.pc_scratch_doubles = nscr.value
        _lib.check(self.lib.rmw_rigid_setup(self._offp(), self.nbodies, self.ws.data_ptr(), self.ws.numel(),
                                            _stream()))
        self.r_blobs = torch.empty((self.Nblobs, 3), dtype=torch.float64, device=dev)
        self.r_rel = torch.empty((self.Nblobs, 3), dtype=torch.float64, device=dev)
        _lib.check(self.lib.rmw_body_blobs(self.centres.data_ptr(), self.quats.data_ptr(), self.ref.data_ptr(),
                                           self._offp(), self.nbodies, self.r_blobs.data_ptr(),
                                           self.r_rel.data_ptr(), self.ws.data_ptr(), self.ws.numel(), _stream()))
        self._mlam = torch.empty((self.Nblobs, 3), dtype=torch.float64, device=dev)
        self._products = {}      # precision -> wall_rpy.Product on r_blobs
        self._ops = {}           # stream handle -> KrylovOps of the solves (vectors of self.size)
        self.minv = None
        self.pc_r_rel = None     # body-frame blob offsets the preconditioner was built with
        self.chol = None         # per body L_b (M_b = L_b L_b^T), packed like minv
        self.chol_inv = None     # per body L_b^-1

    def _offp(self):
        return self.body_offsets.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))

    def set_configuration(self, locations, orientations):
        """Move the same bodies: new centres (B, 3) and unit quaternions
        (B, 4), host arrays or device tensors.  Blob positions are
        recomputed on the device into the existing buffers and the
        preconditioner is rebuilt on its next use."""
        c = torch.as_tensor(locations, dtype=torch.float64).to(self.device).reshape(self.nbodies, 3)
        q = torch.as_tensor(orientations, dtype=torch.float64).to(self.device).reshape(self.nbodies, 4)
        self.centres.copy_(c)
        self.quats.copy_(q)
        _lib.check(self.lib.rmw_body_blobs(self.centres.data_ptr(), self.quats.data_ptr(), self.ref.data_ptr(),
                                           self._offp(), self.nbodies, self.r_blobs.data_ptr(),
                                           self.r_rel.data_ptr(), self.ws.data_ptr(), self.ws.numel(), _stream()))
        self.minv = None
        self.pc_r_rel = None
        self.chol = None
        self.chol_inv = None

    # -- blob mobility and operator -------------------------------------------------
    def mobility_times(self, lam: torch.Tensor, out: torch.Tensor, precision: Optional[str] = None):
        """out = B M(z_eff) B lam on all blobs (wall-RPY UF block)."""
        prec = precision or self.precision
        prod = self._products.get(prec)
        if prod is None:
            # r_blobs is rewritten in place when the bodies move: one prepared
            # product per precision serves every configuration
            prod = wall_rpy.Product(RMW_UF, self.r_blobs, eta=self.eta, a=self.a, periodic_length=self.L,
                                    precision=prec, domain=self.domain)
            self._products[prec] = prod
        prod(force=lam.reshape(-1).contiguous(), out_velocity=out.view(-1))

    def operator(self, x: torch.Tensor, out: torch.Tensor, precision: Optional[str] = None):
        """out = [M lambda - K U; -K^T lambda]."""
        self.mobility_times(x[: 3 * self.Nblobs], self._mlam, precision)
        _lib.check(self.lib.rmw_rigid_operator_post(self.r_rel.data_ptr(), self._offp(), self.nbodies,
                                                    x.data_ptr(), self._mlam.data_ptr(), out.data_ptr(),
                                                    self.ws.data_ptr(), self.ws.numel(), _stream()))

    # -- preconditioner ---------------------------------------------------------------
    def build_preconditioner(self):
        dev = self.device
        self.minv = torch.empty(max(self.minv_doubles, 1), dtype=torch.float64, device=dev)
        self.minvk = torch.empty((self.Nblobs * 3, 6), dtype=torch.float64, device=dev)
        self.nbody = torch.empty((self.nbodies, 6, 6), dtype=torch.float64, device=dev)
        status = torch.zeros(self.nbodies, dtype=torch.int32, device=dev)
        scratch = (torch.empty(self.pc_scratch_doubles, dtype=torch.float64, device=dev)
                   if self.pc_scratch_doubles else None)
        _lib.check(self.lib.rmw_rigid_pc_build_ex(self._dom, self.r_blobs.data_ptr(), self.r_rel.data_ptr(),
                                                  self._offp(), self.nbodies, self.eta, self.a, self.minv.data_ptr(),
                                                  self.minvk.data_ptr(), self.nbody.data_ptr(), None, None,
                                                  status.data_ptr(), scratch.data_ptr() if scratch is not None else None,
                                                  self.pc_scratch_doubles, self.ws.data_ptr(), self.ws.numel(),
                                                  _stream()))
        del scratch
        self.pc_r_rel = self.r_rel
        bad = int(status.sum().item())
        if bad:
            raise RuntimeError(f"{bad} body mobility blocks are not positive definite")

    def build_stochastic_preconditioner(self):
        """The deterministic preconditioner (M_b^-1, M_b^-1 K_b, N_b) and,
        per body, the lower Cholesky factor L_b of the dense wall mobility
        M_b = L_b L_b^T with its inverse L_b^-1: the three factorisations of
        build_block_diagonal_preconditioners_det_stoch (multi_bodies.py
        :480-560; its `M_factorization_blobs` is L_b, its
        `M_factorization_blobs_inv` is the upper U_b^-1 = L_b^-T of scipy's
        M_b = U_b^T U_b, applied here as the transpose of L_b^-1)."""
        dev = self.device
        nd = max(self.minv_doubles, 1)
        self.minv = torch.empty(nd, dtype=torch.float64, device=dev)
        self.chol = torch.empty(nd, dtype=torch.float64, device=dev)
        self.chol_inv = torch.empty(nd, dtype=torch.float64, device=dev)
        self.minvk = torch.empty((self.Nblobs * 3, 6), dtype=torch.float64, device=dev)
        self.nbody = torch.empty((self.nbodies, 6, 6), dtype=torch.float64, device=dev)
        status = torch.zeros(self.nbodies, dtype=torch.int32, device=dev)
        scratch = (torch.empty(self.pc_scratch_doubles, dtype=torch.float64, device=dev)
                   if self.pc_scratch_doubles else None)
        _lib.check(self.lib.rmw_rigid_pc_build_ex(
            self._dom, self.r_blobs.data_ptr(), self.r_rel.data_ptr(), self._offp(), self.nbodies, self.eta, self.a,
            self.minv.data_ptr(), self.minvk.data_ptr(), self.nbody.data_ptr(), self.chol.data_ptr(),
            self.chol_inv.data_ptr(), status.data_ptr(), scratch.data_ptr() if scratch is not None else None,
            self.pc_scratch_doubles, self.ws.data_ptr(), self.ws.numel(), _stream()))
        del scratch
        self.pc_r_rel = self.r_rel
        bad = int(status.sum().item())
        if bad:
            raise RuntimeError(f"{bad} body mobility blocks are not positive definite")

    def preconditioner_state(self):
        """The built preconditioner (M_b^-1, M_b^-1 K_b, N_b, the K_b offsets
        it used and, if built, the Cholesky factors), for reuse on a later
        configuration."""
        if self.minv is None:
            self.build_preconditioner()
        return (self.minv, self.minvk, self.nbody, self.pc_r_rel.clone(), self.chol, self.chol_inv)

    def adopt_preconditioner(self, state):
        """Apply a preconditioner built on an earlier configuration of the
        same bodies (multi_bodies.py:766-819 and :525-555 with update_PC > 1:
        the stale M_b^-1, N_b, K_b and factors are used until the next
        rebuild)."""
        self.minv, self.minvk, self.nbody, self.pc_r_rel, self.chol, self.chol_inv = state

    def _block_apply(self, blocks: torch.Tensor, x: torch.Tensor, y: torch.Tensor, trans: bool):
        """y_b = B_b x_b (or B_b^T x_b) for every body and each of the k rows
        of x, y (k, 3N); B_b lower triangular."""
        k = 1 if x.dim() == 1 else x.shape[0]
        _lib.check(self.lib.rmw_rigid_block_apply(self._offp(), self.nbodies, blocks.data_ptr(), int(trans), 1,
                                                  x.data_ptr(), y.data_ptr(), k, 3 * self.Nblobs, self.ws.data_ptr(),
                                                  self.ws.numel(), _stream()))

    def mobility_pc(self, X: torch.Tensor, Y: torch.Tensor, precision: Optional[str] = None):
        """Y[r] = P^T M P X[r] with P = L^-T per body (the `mobility_pc` closure
        of multi_bodies.py:591-606) for X, Y (k, 3N) contiguous; its block
        diagonal is the identity."""
        if self.chol_inv is None:
            raise RuntimeError("build_stochastic_preconditioner() first")
        k = X.shape[0]
        t1 = torch.empty_like(X)
        self._block_apply(self.chol_inv, X, t1, trans=True)
        t2 = torch.empty((k, self.Nblobs, 3), dtype=torch.float64, device=self.device)
        wall_rpy.apply(RMW_UF, self.r_blobs, force=t1.view(k, self.Nblobs, 3), eta=self.eta, a=self.a,
                       periodic_length=self.L, precision=precision or self.precision, k=k, out_velocity=t2,
                       domain=self.domain)
        self._block_apply(self.chol_inv, t2.view(k, -1), Y, trans=False)

    def P_inv_mult(self, x: torch.Tensor) -> torch.Tensor:
        """L x per body (the `P_inv_mult` closure of multi_bodies.py:608-615,
        out of place)."""
        y = torch.empty_like(x)
        self._block_apply(self.chol, x.contiguous(), y, trans=False)
        return y

    def brownian_slip(self, Z: torch.Tensor, factor: float, tolerance: float, max_iter: int = 1000,
                      precision: Optional[str] = None, strict: bool = False):
        """factor * L (P^T M P)^1/2 z for the k rows of Z (k, 3N): the blob
        noise of stochastic_EM (quaternion_integrator_multi_bodies.py:288-298),
        Lanczos on the preconditioned mobility with L_mult = L.  Its
        covariance is factor^2 L (P^T M P) L^T = factor^2 M.  Returns
        (noise (k, 3N), iterations)."""
        from . import lanczos
        if self.chol is None:
            self.build_stochastic_preconditioner()
        return lanczos.lanczos_block(lambda X, Y: self.mobility_pc(X, Y, precision), Z, factor=factor,
                                     tolerance=tolerance, max_iter=max_iter, L_mult=self.P_inv_mult, strict=strict)

    def pc_apply(self, x: torch.Tensor, out: torch.Tensor):
        if self.minv is None:
            self.build_preconditioner()
        _lib.check(self.lib.rmw_rigid_pc_apply(self._offp(), self.nbodies, self.minv.data_ptr(),
                                               self.minvk.data_ptr(), self.nbody.data_ptr(),
                                               self.pc_r_rel.data_ptr(), x.data_ptr(), out.data_ptr(),
                                               self.ws.data_ptr(), self.ws.numel(), _stream()))

    # -- initial guesses -----------------------------------------------------------------
    def point_field(self, force_torque, radius: float = 0.9, wall: Optional[bool] = None) -> torch.Tensor:
        """(Nblobs, 3) velocity at every blob from the net forces of the other
        bodies, each a sphere of `radius` at its centre (different-radii
        wall RPY)."""
        if wall is None:
            wall = self.domain == "single_wall"
        ft = torch.as_tensor(force_torque, dtype=torch.float64).to(self.device).reshape(self.nbodies, -1).contiguous()
        if ft.shape[1] < 3:
            raise ValueError("force_torque must hold at least the three force components per body")
        u = torch.empty((self.Nblobs, 3), dtype=torch.float64, device=self.device)
        _lib.check(self.lib.rmw_rigid_point_field(self.r_blobs.data_ptr(), self.centres.data_ptr(), ft.data_ptr(),
                                                  ft.shape[1], self._offp(), self.nbodies, float(radius), self.a,
                                                  self.eta, int(bool(wall)), u.data_ptr(), self.ws.data_ptr(),
                                                  self.ws.numel(), _stream()))
        return u

    def point_field_guess(self, rhs: torch.Tensor, force_torque, radius: float = 0.9) -> torch.Tensor:
        """Block-Jacobi step from P^-1 b with the other bodies as point
        forces: x = P^-1 (b - [u; 0]), u = point_field(F).  Costs one
        bodies-to-blobs product and a preconditioner application instead of
        a blob-blob product."""
        b = rhs.clone()
        b[: 3 * self.Nblobs] -= self.point_field(force_torque, radius).reshape(-1)
        x = torch.empty_like(rhs)
        self.pc_apply(b, x)
        return x

    # -- solve ------------------------------------------------------------------------
    def _krylov_ops(self) -> krylov.KrylovOps:
        # one per stream: the reductions' scratch is not shared across streams
        key = _stream_mod.current(self.device)
        ops = self._ops.get(key)
        if ops is None:
            ops = krylov.KrylovOps(self.size, 2, self.device)
            self._ops[key] = ops
        return ops

    def norm(self, v: torch.Tensor) -> float:
        """||v|| of a device vector of self.size entries (the fixed-order
        reduction the solves normalise with)."""
        return self._krylov_ops().norm(v)

    def rhs(self, slip, force_torque) -> torch.Tensor:
        """[slip; -force_torque] as in solve_mobility_problem (no constraints)."""
        s = torch.as_tensor(slip, dtype=torch.float64).reshape(-1)
        ft = torch.as_tensor(force_torque, dtype=torch.float64).reshape(-1)
        if s.numel() != 3 * self.Nblobs or ft.numel() != 6 * self.nbodies:
            raise ValueError("slip must be (Nblobs, 3) and force_torque (Nbodies, 6)")
        return torch.cat([s.to(self.device), -ft.to(self.device)])

    def solve_mobility_problem(self, slip=None, force_torque=None, rhs: Optional[torch.Tensor] = None,
                               tol: float = 1e-8, restart: int = 60, maxiter: int = 1000,
                               x0: Optional[torch.Tensor] = None, precision: Optional[str] = None,
                               callback=None, rhs_norm: Optional[float] = None) -> krylov.GmresResult:
        """Solve [M -K; -K^T 0][lambda; U] = [slip; -F] to relative residual tol.

        The RHS is normalised to unit norm and the solution rescaled, as in
        quaternion_integrator_multi_bodies.py:1515-1536; x0 (unnormalised)
        is used as a true warm start.  rhs_norm: ||rhs|| when the caller has
        it from norm() (saves a read-back).  Returns a GmresResult whose x is
        [lambda; U] (U_b = (v_b, w_b))."""
        if rhs is None:
            if slip is None:
                slip = np.zeros((self.Nblobs, 3))
            if force_torque is None:
                force_torque = np.zeros((self.nbodies, 6))
            rhs = self.rhs(slip, force_torque)
        ops = self._krylov_ops()
        rnorm = ops.norm(rhs) if rhs_norm is None else float(rhs_norm)
        if rnorm == 0.0:
            return krylov.GmresResult(x=torch.zeros_like(rhs), info=0, iterations=0, matvecs=0, final_residual=0.0)
        b = torch.empty_like(rhs)
        ops.axpby(1.0 / rnorm, rhs, 0.0, b)
        xw = None
        if x0 is not None:
            xw = torch.empty_like(rhs)
            ops.axpby(1.0 / rnorm, x0.to(self.device, torch.float64).reshape(-1), 0.0, xw)
        if self.minv is None:
            self.build_preconditioner()
        prec = precision or self.precision
        res = krylov.gmres(lambda v, o: self.operator(v, o, prec), b, Pinv=self.pc_apply, x0=xw, tol=tol,
                           restart=restart, maxiter=maxiter, callback=callback)
        out = torch.empty_like(rhs)
        ops.axpby(rnorm, res.x, 0.0, out)
        res.x = out
        return res

    def solve_mixed(self, slip=None, force_torque=None, rhs: Optional[torch.Tensor] = None, tol: float = 1e-8,
                    restart: int = 60, maxiter: int = 1000, x0: Optional[torch.Tensor] = None,
                    inner_precision: str = "fp32_blocked", inner_floor: float = 1e-8, richardson: float = 2.5,
                    max_outer: int = 8, rhs_norm: Optional[float] = None):
        """Mixed-precision solve to the FP64 tolerance: FP64 residuals
        r = b - A_64 x (outer), GMRES on A_32 d = r with FP32-pair products
        and FP64 Krylov algebra (inner), x += d.

        The inner tolerance asks for the reduction still needed, floored at
        `inner_floor`: the FP32-pair operator's own relative error on a
        solution (|(A_32 - A_64) x| / |b| is 1.2e-8 on the 256-sphere and
        1.5e-8 on the 1024-sphere suspension), below which inner iterations
        stop lowering the FP64 residual.  An FP64 residual within
        `richardson` x tol of the target is corrected by one block-diagonal
        step x += P^-1 r (no product; 0 disables it) before another inner
        solve is tried.  The residual of a warm start x0 is formed with the
        inner operator, whose error is below what the first inner solve
        reaches; it never decides convergence -- every later residual, and
        the one that ends the solve, is FP64.  A start whose inner-operator
        residual is within 10x of `inner_floor` is judged by its FP64
        residual instead.  rhs_norm: ||rhs|| when the caller has it from
        norm().  Returns a GmresResult whose iterations and
        matvecs sum over the inner solves (at most `maxiter` iterations in
        all); final_residual is the FP64 one, and info is 0 only when it meets
        tol (else a positive count, as scipy's gmres)."""
        if rhs is None:
            rhs = self.rhs(np.zeros((self.Nblobs, 3)) if slip is None else slip,
                           np.zeros((self.nbodies, 6)) if force_torque is None else force_torque)
        if self.minv is None:
            self.build_preconditioner()
        ops = self._krylov_ops()
        bnorm = ops.norm(rhs) if rhs_norm is None else float(rhs_norm)
        x = torch.zeros_like(rhs) if x0 is None else x0.to(self.device, torch.float64).reshape(-1).clone()
        r = torch.empty_like(rhs)
        out = krylov.GmresResult(x=x, info=0, iterations=0, matvecs=0)
        if bnorm == 0.0:
            out.final_residual = 0.0
            return out
        if max_outer < 1:
            raise ValueError("max_outer must be at least 1 (the FP64 residual that ends the solve)")

        def fp64_residual():
            self.operator(x, r, precision="fp64")
            out.matvecs += 1
            ops.axpby(1.0, rhs, -1.0, r)
            rel = ops.norm(r) / bnorm
            out.final_residual = rel
            out.residuals.append(rel)
            return rel

        # residual of the start: b for a zero start, the inner operator's for a warm start
        if x0 is None:
            ops.axpby(1.0, rhs, 0.0, r)
            rel = 1.0
        else:
            self.operator(x, r, precision=inner_precision)
            out.matvecs += 1
            ops.axpby(1.0, rhs, -1.0, r)
            rel = ops.norm(r) / bnorm
            if rel <= 10.0 * inner_floor:
                # near the inner operator's own error: only the FP64 residual tells
                rel = fp64_residual()
                if rel <= tol:
                    return out
        last_step = "start"
        for outer in range(max_outer + 1):
            if outer > 0:
                rel = fp64_residual()
                if rel <= tol or outer == max_outer or out.iterations >= maxiter:
                    # info: 0 when converged, else a positive iteration count
                    out.info = 0 if rel <= tol else max(out.iterations, 1)
                    break
            if rel == 0.0:
                continue             # the start solves the inner operator exactly: judge it in FP64
            if outer > 0 and rel <= richardson * tol and last_step != "richardson":
                d = torch.empty_like(r)
                self.pc_apply(r, d)
                ops.axpby(1.0, d, 1.0, x)
                last_step = "richardson"
                continue
            inner_tol = max(0.5 * tol / rel, inner_floor)
            rn = rel * bnorm
            rs = torch.empty_like(r)
            ops.axpby(1.0 / rn, r, 0.0, rs)
            # the inner solves share the caller's iteration budget
            res = krylov.gmres(lambda v, o: self.operator(v, o, inner_precision), rs, Pinv=self.pc_apply, tol=inner_tol,
                               restart=restart, maxiter=max(1, maxiter - out.iterations), verify=False)
            out.iterations += res.iterations
            out.matvecs += res.matvecs
            ops.axpby(rn, res.x, 1.0, x)
            last_step = "gmres"
        out.x = x
        return out

    def velocities(self, sol: torch.Tensor) -> np.ndarray:
        """(Nbodies, 6) body velocities [v, w] from a solution vector."""
        return sol[3 * self.Nblobs:].reshape(self.nbodies, 6).cpu().numpy()

    def constraint_forces(self, sol: torch.Tensor) -> np.ndarray:
        return sol[: 3 * self.Nblobs].reshape(self.Nblobs, 3).cpu().numpy()
