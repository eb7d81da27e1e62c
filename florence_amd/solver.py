"""FEMSolver / LinearSolver: the incremental Newton-Raphson driver on the GPU.

Semantics restated from Florence/Solver/FEMSolver.py:
* constructor flags used by the static DF path (:34-90; the solution tolerance
  defaults to 10 x newton_raphson_tolerance, :174);
* ``Solve`` (:518-714): Dirichlet data, Eulerx = copy(points), first ``Assemble`` at
  the undeformed state, then ``StaticSolver``;
* ``StaticSolver`` (:841-929): per increment, the predictor residual
  K[in, out] u_D * LoadFactor, NormForces at increment 0, ``NewtonRaphson``,
  TotalDisp[:, :, inc] = Eulerx - X, stop on failure;
* ``NewtonRaphson`` (:932-1044): Eulerx += incremental Dirichlet values; loop { solve
  K_b dU = -R_b; Eulerx += dU; re-assemble; R_in = T_in - F_in; NormForces re-based at
  Iter 0; exits on |R| < tol, ||dU|| <= solution tolerance, max iterations (failure),
  NaN or norm > 1e6 (failure), stagnation, user functions }.

Everything between the iterations stays in HBM: assembly (libflorence_amd
``fl_df_assemble``), the masked reduction and the linear solve (``fl_pcg``, block-Jacobi
CG, or ``multigrid.MultigridSolver``'s p-multigrid flexible CG) -- the reference's
sparse direct solve has no GPU counterpart here.  Host work per iteration is the scalar
bookkeeping (a few norms).
"""
import time

import numpy as np

from . import _lib
from .assembly import assemble_device, device_mesh_data
from .sparse import DeviceCSR

__all__ = ["FEMSolver", "LinearSolver", "NewtonResult"]


class LinearSolver(object):
    """Device linear solver of the Newton steps: block-Jacobi PCG on the masked
    tangent (replaces LinearSolver.Solve, LinearSolver.py:208-300)."""

    def __init__(self, linear_solver="pcg", rtol=1e-12, maxit=20000, check_every=10):
        self.linear_solver = linear_solver
        self.rtol = rtol
        self.maxit = maxit
        self.check_every = check_every
        self.iterations = []
        self._ws = None

    def Solve(self, dm, V, free_mask, b, x, precision=None):
        import torch
        lib = _lib.load()
        if self._ws is None or self._ws.numel() < lib.fl_pcg_workspace_size(dm.nnode):
            self._ws = torch.empty(int(lib.fl_pcg_workspace_size(dm.nnode)), dtype=torch.uint8, device=dm.device)
        it = _lib.ctypes.c_int32(0)
        rel = _lib.ctypes.c_double(0.0)
        fn = lib.fl_pcg_v32 if V.dtype == torch.float32 else lib.fl_pcg
        _lib.check(fn(_lib.ptr(dm.indptr), _lib.ptr(dm.node_nbr_ptr), _lib.ptr(dm.node_nbrs), _lib.ptr(V),
                      dm.nnode, _lib.ptr(free_mask), _lib.ptr(b), _lib.ptr(x), float(self.rtol),
                      int(self.maxit), int(self.check_every), _lib.ptr(self._ws), self._ws.numel(),
                      _lib.ctypes.byref(it), _lib.ctypes.byref(rel), _lib.stream_ptr()),
                   "fl_pcg")
        self.iterations.append(int(it.value))
        if not (rel.value <= self.rtol) and it.value >= self.maxit:
            raise _lib.FlorenceAMDError("PCG did not reach rtol={} in {} iterations (rel={})".format(
                self.rtol, it.value, rel.value))
        return x


def default_linear_solver(dm):
    """The Newton-step solver FEMSolver takes when none is given: p-multigrid flexible CG
    for Lagrange hexes of degree >= 2, block-Jacobi PCG otherwise (rtol 1e-12)."""
    from .multigrid import MultigridSolver, coarse_degrees
    et = getattr(dm, "element_type", None)
    p1 = int(round(dm.npe ** (1.0 / 3.0)))
    if et == "hex" and p1 ** 3 == dm.npe and coarse_degrees("hex", p1 - 1):
        return MultigridSolver(rtol=1e-12)
    return LinearSolver(rtol=1e-12)


def spmv(dm, V, x, y, free_mask=None):
    """y = K x (masked operator when free_mask is given); V float64 or float32."""
    import torch
    lib = _lib.load()
    fn = lib.fl_csr_spmv_v32 if V.dtype == torch.float32 else lib.fl_csr_spmv
    _lib.check(fn(_lib.ptr(dm.indptr), _lib.ptr(dm.node_nbr_ptr), _lib.ptr(dm.node_nbrs), _lib.ptr(V),
                  dm.nnode, _lib.ptr(free_mask), _lib.ptr(x), _lib.ptr(y), _lib.stream_ptr()),
               "fl_csr_spmv")
    return y


class NewtonResult(object):
    """What Solve returns: TotalDisp (nnode, 3, increments), NRConvergence, timings."""

    def __init__(self, TotalDisp, NRConvergence, iterations, cg_iterations, wall_time, assembly_time, failed):
        self.TotalDisp = TotalDisp
        self.NRConvergence = NRConvergence
        self.newton_iterations = iterations
        self.cg_iterations = cg_iterations
        self.wall_time = wall_time
        self.assembly_time = assembly_time
        self.newton_raphson_failed_to_converge = failed

    def GetSolutionVectors(self):
        return self.TotalDisp


class FEMSolver(object):

    def __init__(self, has_low_level_dispatcher=False, optimise=False, analysis_type="static",
                 analysis_nature="nonlinear", requires_geometry_update=True, number_of_load_increments=1,
                 load_factor=None, newton_raphson_tolerance=1.0e-6, newton_raphson_solution_tolerance=None,
                 maximum_iteration_for_newton_raphson=50, break_at_stagnation=True, user_defined_break_func=None,
                 user_defined_stop_func=None, report_log_level=0, linear_solver=None, warm_start=None,
                 tangent_precision="fp64", csr_precision="fp64", **kwargs):
        if has_low_level_dispatcher != optimise:
            has_low_level_dispatcher = True
        self.has_low_level_dispatcher = has_low_level_dispatcher
        if analysis_type != "static" or analysis_nature != "nonlinear":
            raise NotImplementedError("the GPU path covers static nonlinear analysis")
        self.analysis_type = analysis_type
        self.analysis_nature = analysis_nature
        self.requires_geometry_update = requires_geometry_update
        self.number_of_load_increments = number_of_load_increments
        self.load_factor = load_factor
        self.newton_raphson_tolerance = newton_raphson_tolerance
        self.newton_raphson_solution_tolerance = newton_raphson_solution_tolerance
        if self.newton_raphson_solution_tolerance is None:
            self.newton_raphson_solution_tolerance = 10. * self.newton_raphson_tolerance
        self.maximum_iteration_for_newton_raphson = maximum_iteration_for_newton_raphson
        self.break_at_stagnation = break_at_stagnation
        self.user_defined_break_func = user_defined_break_func
        self.user_defined_stop_func = user_defined_stop_func
        self.report_log_level = report_log_level
        # None: chosen in Solve -- the p-multigrid flexible CG for hexes of p >= 2 (nested
        # coarse spaces; 12-16 iterations per Newton step on the 24^3 twisted hex p=4 cube), else
        # block-Jacobi PCG; both to rtol 1e-12, where the reference solves directly
        self.linear_solver = linear_solver
        self.warm_start = warm_start
        # "fp64": the reference's arithmetic; "fp32": inexact-Jacobian Newton with the
        # tangent contracted on FP32 MFMA (kinematics, residual and update stay FP64)
        self.tangent_precision = tangent_precision
        # "fp32" (with an fp32/bf16 tangent): CSR values stored as float32 for the solve
        self.csr_precision = csr_precision
        self.assembly_time = 0.0
        self.newton_raphson_failed_to_converge = False
        for k, v in kwargs.items():
            setattr(self, k, v)

    # ------------------------------------------------------------------ driver
    def Solve(self, formulation=None, mesh=None, material=None, boundary_condition=None, function_spaces=None,
              solver=None, Eulerx=None, Eulerp=None):
        import torch
        if not material.has_low_level_dispatcher:
            raise RuntimeError("material has no low-level dispatcher")
        fs = (function_spaces or formulation.function_spaces)[0]
        t0 = time.perf_counter()
        boundary_condition.GetDirichletBoundaryConditions(formulation, mesh, material, solver, self)
        dm = device_mesh_data(self, mesh, fs)
        if self.linear_solver is None:
            self.linear_solver = default_linear_solver(dm)
        dev = dm.device
        ndof = 3 * dm.nnode
        free = torch.as_tensor(boundary_condition.free_mask(ndof), device=dev)
        free_f = free.to(torch.float64)
        cols_out = torch.as_tensor(boundary_condition.columns_out, device=dev)
        applied = torch.as_tensor(boundary_condition.applied_dirichlet, dtype=torch.float64, device=dev)
        X = torch.as_tensor(np.ascontiguousarray(mesh.points, dtype=np.float64), device=dev)
        Eulerx = X.clone().reshape(-1)
        NodalForces = torch.zeros(ndof, dtype=torch.float64, device=dev)
        V = torch.empty(dm.nnz, dtype=torch.float32 if self.csr_precision == "fp32" else torch.float64, device=dev)
        T = torch.empty(ndof, dtype=torch.float64, device=dev)
        tol = self.newton_raphson_tolerance
        nincr = self.number_of_load_increments
        TotalDisp = torch.zeros((dm.nnode, 3, nincr), dtype=torch.float64, device=dev)
        self.NRConvergence = {"Increment_" + str(i): [] for i in range(nincr)}
        iters = []
        ta = 0.0

        def assemble(x):
            nonlocal ta
            t = time.perf_counter()
            assemble_device(dm, material, x, V, T, int(self.requires_geometry_update),
                            precision=self.tangent_precision, csr=self.csr_precision)
            ta += time.perf_counter() - t

        assemble(Eulerx)   # first assembly at the undeformed state (FEMSolver.py:641-642)
        Residual = torch.zeros(ndof, dtype=torch.float64, device=dev)
        dUD = torch.zeros(ndof, dtype=torch.float64, device=dev)
        KuD = torch.empty(ndof, dtype=torch.float64, device=dev)
        dU = torch.zeros(ndof, dtype=torch.float64, device=dev)
        rhs = torch.empty(ndof, dtype=torch.float64, device=dev)
        history = []   # converged Eulerx per increment (for warm starts)
        self.newton_raphson_failed_to_converge = False
        for Increment in range(nincr):
            LoadFactor = 1. / nincr if self.load_factor is None else self.load_factor[Increment]
            # predictor residual: Residual_in = K[in, out] (LoadFactor u_D)  (FEMSolver.py:857-862)
            dUD.zero_()
            dUD[cols_out] = LoadFactor * applied
            spmv(dm, V, dUD, KuD)
            Residual = KuD * free_f
            if Increment == 0:
                self.NormForces = float(torch.linalg.norm(Residual))
                if np.isclose(self.NormForces, 0.0):
                    self.NormForces = 1e-14
            self.norm_residual = float(torch.linalg.norm(Residual)) / self.NormForces
            Eulerx = Eulerx + dUD    # UpdateFixDoFs (FEMSolver.py:943-946)
            if self.warm_start is not None:
                Eulerx = self.warm_start(self, Increment, Eulerx, history, dm, material, free_f, V, T, assemble,
                                         Residual)
            it = self._newton(Increment, dm, material, free, free_f, Eulerx, V, T, Residual, NodalForces, dU, rhs,
                              assemble, tol)
            iters.append(it)
            Eulerx = self._Eulerx
            TotalDisp[:, :, Increment] = (Eulerx - X.reshape(-1)).reshape(-1, 3)
            history.append(Eulerx.clone())
            if self.newton_raphson_failed_to_converge:
                TotalDisp = TotalDisp[:, :, :Increment]
                break
            self._save_increment(TotalDisp, Increment)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        self.assembly_time = ta
        return NewtonResult(TotalDisp, self.NRConvergence, iters, list(self.linear_solver.iterations), wall, ta,
                            self.newton_raphson_failed_to_converge)

    def _save_increment(self, TotalDisp, Increment):
        """save_incremental_solution (FEMSolver.py:1613-1627): every
        incremental_solution_save_frequency-th increment's displacement to
        <incremental_solution_filename>_<Increment>.mat as {'solution': (nnode, 3)}."""
        if not getattr(self, "save_incremental_solution", False):
            return
        if Increment % int(getattr(self, "incremental_solution_save_frequency", 50)) != 0:
            return
        filename = getattr(self, "incremental_solution_filename", None)
        if filename is None:
            raise ValueError("No file name provided to save incremental solution")
        if ".mat" in filename:
            filename = filename.split(".")[0]
        from scipy.io import savemat
        # the extension written out: scipy only appends it when the bare name cannot be opened
        savemat(filename + "_" + str(Increment) + ".mat", {"solution": TotalDisp[:, :, Increment].cpu().numpy()},
                do_compression=True)

    def _newton(self, Increment, dm, material, free, free_f, Eulerx, V, T, Residual, NodalForces, dU, rhs, assemble,
                tol):
        import torch
        Iter = 0
        hist = []
        self._Eulerx = Eulerx
        while self.norm_residual > tol or Iter == 0:
            torch.neg(Residual, out=rhs)
            self.linear_solver.Solve(dm, V, free, rhs, dU, precision=self.tangent_precision)
            Eulerx = Eulerx + dU
            self._Eulerx = Eulerx
            assemble(Eulerx)
            Residual = (T - NodalForces) * free_f
            self.abs_norm_residual = float(torch.linalg.norm(Residual))
            if Iter == 0:
                self.NormForces = self.abs_norm_residual
                if np.isclose(self.NormForces, 0.):
                    self.NormForces = 1e-14
            self.norm_residual = abs(self.abs_norm_residual / self.NormForces)
            self.NRConvergence["Increment_" + str(Increment)].append(self.norm_residual)
            if self.report_log_level:
                import sys
                print("Increment {} Iteration {}: residual {:.3e} (abs {:.3e}), CG iterations {}".format(
                    Increment, Iter, self.norm_residual, self.abs_norm_residual, self.linear_solver.iterations[-1]),
                    file=sys.stderr, flush=True)
            if abs(self.abs_norm_residual) < tol:
                break
            if Iter > 0 and float(torch.linalg.norm(dU)) <= self.newton_raphson_solution_tolerance:
                break
            Iter += 1
            if Iter == self.maximum_iteration_for_newton_raphson:
                self.newton_raphson_failed_to_converge = True
                break
            if np.isnan(self.norm_residual) or self.norm_residual > 1e06:
                self.newton_raphson_failed_to_converge = True
                break
            if self.break_at_stagnation:
                hist.append(self.norm_residual)
                if Iter >= 6 and np.mean(hist) < 1. and self.abs_norm_residual < 0.001:
                    break
            if self.user_defined_break_func is not None and \
                    self.user_defined_break_func(Increment, Iter, self.norm_residual, self.abs_norm_residual, tol):
                break
            if self.user_defined_stop_func is not None and \
                    self.user_defined_stop_func(Increment, Iter, self.norm_residual, self.abs_norm_residual, tol):
                self.newton_raphson_failed_to_converge = True
                break
        return Iter + 1
