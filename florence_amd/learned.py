"""Learned warm start of the incremental Newton iteration.

No reference counterpart: Florence starts every load increment from the previous
converged state plus the Dirichlet increment and takes its first step with the
previous tangent (FEMSolver.py:857-862, 932-946).  This module adds

* ``TwistStretchFamily`` -- the trajectories: hex p=4 unit cubes of n^3 cells (nodes
  jittered by +-0.03 h per member seed), Mooney-Rivlin(0.5, 0.5, 20), 6^3
  Gauss-Lobatto points, x=0 clamped, the x=1 face twisted by theta ~ U[0, 60] degrees
  about the x axis through (y, z) = (0.5, 0.5) and moved axially by s - 1,
  s ~ U[0.8, 1.5], applied over ``nincr`` proportional increments (the Dirichlet
  values times k / nincr, as the reference's LoadFactor does);
* ``LockstepNewton`` -- many members solved together: their meshes are concatenated
  into one node-disjoint mesh (one assembly call, one block-diagonal CSR, one PCG
  solve per Newton iteration) and every member is tested against its own FP64
  residual: ||R_free,i|| <= tol * R0_i, R0_i = the member's predictor residual at the
  first increment (the reference's NormForces before its per-increment rebase);
* ``WarmStartMLP`` -- the predictor: per node, inputs (X, u_{k-1}, u_{k-1} - u_{k-2},
  load factor) standardised, output the correction to the linear extrapolation
  2 u_{k-1} - u_{k-2}; trained on the device (libflorence_amd fl_mlp_* kernels, BF16
  MFMA products, FP32 master weights, Adam) on the trajectories the solver generates.

Start modes of LockstepNewton.solve:
  "previous"  the reference's: u_{k-1} + Dirichlet increment, first step with the
              previous tangent on the predictor residual K[:, out] du_D;
  "extrap"    2 u_{k-1} - u_{k-2} (u_0 = u_{-1} = 0), full Newton from there;
  "extrap2"   3 u_{k-1} - 3 u_{k-2} + u_{k-3} from the third increment on (the quadratic
              through the last three converged states; "extrap" before);
  "extrap3"   4 u_{k-1} - 6 u_{k-2} + 4 u_{k-3} - u_{k-4} from the fourth increment on
              (the cubic through the last four; the lower orders before);
  "poly"      per member, whichever of the linear, quadratic and cubic extrapolations
              (those the increment has states for) has the smallest FP64 residual;
  "learned"   the MLP's prediction u_L, the "poly" start u_E, or the point
              u_E + a* (u_L - u_E) that minimises the linearised residual along the
              predicted correction -- whichever has the smallest FP64 residual, per
              member (u_E is the fallback).
Each Newton iteration of the warm-started modes assembles K and T at the current
state, solves, and evaluates the residual-only FP64 kernel at the new state; K is
assembled again only if another step is needed.

Linear solves (``linear="bj"``: block-Jacobi PCG; ``"mg"``: the p-multigrid flexible CG
of multigrid.py): ``forcing="fixed"`` solves every step to ``lin_rtol`` (the stand-in for
the reference's direct solve); ``forcing="adaptive"`` is an inexact Newton method
whose PCG tolerance per step follows the model |R_next| ~ max(C |R|^2, eta |R|), C
measured on the first (tightly solved) step and refined by every later one (see
``_eta``), capped at ``eta_max``.  Convergence is always judged on the FP64 residual.
"""
import time

import numpy as np

from . import _lib
from .assembly import DeviceMeshData, assemble_device, residual_device
from .batch import _Concat
from .function_space import FunctionSpace
from .materials import MooneyRivlin
from .quadrature import QuadratureRule
from .solver import LinearSolver, spmv
from .workloads import cube_mesh, twist_field

__all__ = ["TwistStretchFamily", "LockstepNewton", "WarmStartMLP", "build_dataset", "train"]

NIN, IP, H, OP = 10, 32, 128, 32


class TwistStretchFamily(object):

    def __init__(self, n=2, p=4, nincr=30, theta_range=(0.0, 60.0), stretch_range=(0.8, 1.5), jitter=0.03):
        self.n, self.p, self.nincr = n, p, nincr
        self.theta_range, self.stretch_range, self.jitter = theta_range, stretch_range, jitter
        self.material = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
        self.rule = QuadratureRule(qtype="lobatto", norder=p + 2, mesh_type="hex")

    def member(self, seed):
        """(mesh, theta in degrees, stretch) of member ``seed``."""
        rng = np.random.Generator(np.random.PCG64(seed))
        theta = float(rng.uniform(*self.theta_range))
        stretch = float(rng.uniform(*self.stretch_range))
        return cube_mesh("hex", self.p, self.n, jitter=self.jitter, seed=seed), theta, stretch

    def function_space(self, mesh):
        return FunctionSpace(mesh, self.rule, p=self.p)

    @staticmethod
    def dirichlet(mesh, theta, stretch, tol=1e-10):
        """(free mask uint8 [3 nnode], total Dirichlet displacement [3 nnode], 0 on free dofs)."""
        X = np.asarray(mesh.points)
        u = np.zeros_like(X)
        fixed = np.zeros(X.shape, dtype=bool)
        left = np.abs(X[:, 0]) < tol
        right = np.abs(X[:, 0] - 1.0) < tol
        fixed[left] = True
        fixed[right] = True
        u[right] = twist_field(X[right], np.deg2rad(theta))
        u[right, 0] = stretch - 1.0
        return (~fixed).reshape(-1).astype(np.uint8), u.reshape(-1)


class LockstepNewton(object):
    """Incremental Newton of many family members as one device problem."""

    def __init__(self, family, seeds, precision="fp64", tol=1e-8, lin_rtol=1e-10, maxit=25, forcing="fixed",
                 eta_max=1e-2, csr="fp64", device="cuda", linear="bj"):
        import torch
        if forcing not in ("fixed", "adaptive"):
            raise ValueError(forcing)
        self.family, self.seeds = family, list(seeds)
        self.precision, self.tol, self.maxit = precision, tol, maxit
        self.forcing, self.lin_rtol, self.eta_max = forcing, lin_rtol, eta_max
        self.csr = csr
        members = [family.member(s) for s in self.seeds]
        meshes = [m for m, _, _ in members]
        npm = meshes[0].points.shape[0]
        if any(m.points.shape[0] != npm for m in meshes):
            raise ValueError("members must share one topology")
        self.nmember, self.nnode_member = len(meshes), npm
        cm = _Concat(meshes)
        self.dm = DeviceMeshData(cm, family.function_space(meshes[0]), device=device, build_indices=False)
        dev = self.dm.device
        bc = [family.dirichlet(m, th, s) for m, th, s in members]
        self.params = [(th, s) for _, th, s in members]
        self.free = torch.as_tensor(np.concatenate([f for f, _ in bc]), device=dev)
        self.free_f = self.free.to(torch.float64)
        self.u_total = torch.as_tensor(np.concatenate([u for _, u in bc]), device=dev)
        self.X = self.dm.points.reshape(-1)
        self.V = torch.empty(self.dm.nnz, dtype=torch.float32 if csr == "fp32" else torch.float64, device=dev)
        self.T = torch.empty(3 * self.dm.nnode, dtype=torch.float64, device=dev)
        # "bj": block-Jacobi PCG; "mg": p-multigrid flexible CG (its coarse operators come
        # from the point records of the assembly that produced V)
        self.setup_s = 0.0
        if linear == "mg":
            from .multigrid import MultigridSolver
            self.lin = MultigridSolver(rtol=lin_rtol, maxit=2000)
            # the coarse levels are per-mesh set-up, as the sparsity pattern: built here,
            # outside the timed solve (setup_s)
            t = time.perf_counter()
            self.lin.hierarchy(self.dm)
            torch.cuda.synchronize()
            self.setup_s = time.perf_counter() - t
        else:
            self.lin = LinearSolver(rtol=lin_rtol, maxit=50000, check_every=10)
        self.linear = linear

    def _norms(self, R):
        import torch
        return torch.linalg.norm(R.view(self.nmember, -1), dim=1)

    def _per_dof(self, v):
        return v.repeat_interleave(3 * self.nnode_member)

    def _assemble(self, u):
        assemble_device(self.dm, self.family.material, self.X + u, self.V, self.T, precision=self.precision,
                        csr=self.csr)

    def _residual(self, u, out):
        residual_device(self.dm, self.family.material, self.X + u, out)
        out.mul_(self.free_f)
        return out

    def _solve(self, rhs, du, done=None):
        """du = K^-1 rhs on the free dofs of the members not yet converged (``done``: a
        bool per member); a converged member is held, so its rows drop out of every
        SpMV, as if each member were solved alone."""
        import torch
        du.zero_()
        mask = self.free
        if done is not None and done.any():
            live = torch.as_tensor(~done, device=self.free.device)
            mask = self.free * self._per_dof(live).to(self.free.dtype)
        self.lin.Solve(self.dm, self.V, mask, rhs, du, precision=self.precision)

    def _eta(self, r, C):
        """Linear tolerance of the next step at (worst member's) relative residual r,
        with the step's result modelled as max(C r^2, eta r): a step that can be the
        last solves to 0.5 tol / r; a step that can leave a residual from which one
        more step finishes (r_t = sqrt(0.5 tol / C)) solves just to r_t / r; any other
        step only to C r, its Newton error."""
        if self.forcing == "fixed" or C is None:
            return self.lin_rtol
        half = 0.5 * self.tol
        if C * r * r <= half:
            eta = half / r
        else:
            r_t = np.sqrt(half / C)
            eta = r_t / r if C * r * r <= 0.5 * r_t else C * r
        return float(min(self.eta_max, max(self.lin_rtol, eta)))

    def solve(self, start="previous", predictor=None, record=False, log=None):
        """Run every increment; returns a dict of iteration counts, timings, start
        choices and (record=True) the converged displacements u_0..u_nincr on the host.
        ``log``: optional callable given one line per increment."""
        import torch
        if start not in ("previous", "extrap", "extrap2", "extrap3", "poly", "learned"):
            raise ValueError(start)
        if start == "learned" and predictor is None:
            raise ValueError("start='learned' needs a predictor")
        dev = self.dm.device
        n3 = 3 * self.dm.nnode
        nincr = self.family.nincr
        fixed_f = 1.0 - self.free_f
        z = lambda: torch.zeros(n3, dtype=torch.float64, device=dev)
        u_prev, u_prev2, u_prev3, u_prev4, u, du, rhs, R, RL, uL = (z() for _ in range(10))
        # order of the polynomial extrapolation through the last converged states (the
        # load factors are equally spaced, u_0 = 0 at load 0 included)
        order = {"extrap2": 2, "extrap3": 3, "poly": 3, "learned": 3}.get(start, 1)
        # "poly" and "learned": the FP64 residual picks the order per member
        pick_order = start in ("poly", "learned")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        t_lin = t_asm = t_res = t_pred = 0.0

        def timed(fn, *a):
            t = time.perf_counter()
            fn(*a)
            torch.cuda.synchronize()
            return time.perf_counter() - t

        t_asm += timed(self._assemble, u)
        # R0_i: predictor residual norm of the first increment
        spmv(self.dm, self.V, self.u_total * fixed_f / nincr, rhs)
        R0 = self._norms(rhs * self.free_f)
        R0 = torch.where(R0 > 0, R0, torch.full_like(R0, 1e-14))
        thresh = self.tol * R0
        iters = np.zeros((nincr, self.nmember), dtype=np.int32)      # solves until each member converged
        lock_iters = np.zeros(nincr, dtype=np.int32)                  # solves of the lockstep batch
        chose_learned = np.zeros(nincr, dtype=np.int32)
        start_ratio = []                                              # |R(start)| / R0 per increment (mean)
        history = []                                                  # worst |R| / R0 before/after each step
        traj = [np.zeros(n3)] if record else None
        failed = False
        C = None      # quadratic-convergence constant; the first step is solved tightly to measure it
        for k in range(1, nincr + 1):
            lam = k / float(nincr)
            ubc = self.u_total * (lam * fixed_f)
            first_linear = start == "previous" or k == 1
            if first_linear:
                # reference predictor: Eulerx += du_D; first rhs = -K[:, out] du_D (free rows)
                torch.add(u_prev * self.free_f, ubc, out=u)
                spmv(self.dm, self.V, u - u_prev, rhs)
                rhs.mul_(-self.free_f)
                conv = torch.zeros(self.nmember, dtype=torch.bool, device=dev)
                r_cur = float((self._norms(rhs) / R0).max())
            else:
                extrap = [lambda: (2.0 * u_prev - u_prev2) * self.free_f + ubc,
                          lambda: (3.0 * u_prev - 3.0 * u_prev2 + u_prev3) * self.free_f + ubc,
                          lambda: (4.0 * u_prev - 6.0 * u_prev2 + 4.0 * u_prev3 - u_prev4) * self.free_f + ubc]
                q = min(order, k - 1)
                uE = extrap[q - 1]()
                t_res += timed(self._residual, uE, R)
                nE = self._norms(R)
                if pick_order:
                    # the lower orders, kept per member where their residual is smaller
                    for qq in range(q - 1, 0, -1):
                        uq, Rq = extrap[qq - 1](), torch.empty_like(R)
                        t_res += timed(self._residual, uq, Rq)
                        nq = self._norms(Rq)
                        better = nq < nE
                        uE = torch.where(self._per_dof(better), uq, uE)
                        R = torch.where(self._per_dof(better), Rq, R)
                        nE = torch.where(better, nq, nE)
                if start == "learned":
                    qb = min(predictor.base_order, k - 1)
                    base = extrap[qb - 1]() if qb > 1 else None
                    t_pred += timed(predictor.predict, self.X, u_prev, u_prev2, lam, self.free, ubc, uL, base)
                    t_res += timed(self._residual, uL, RL)
                    nL = self._norms(RL)
                    # step length along the predicted correction d = uL - uE: with R
                    # linear in a small step, |R(uE + a d)| is least at
                    # a* = -<R_E, R_L - R_E> / |R_L - R_E|^2 (per member, kept in [0, 1.5])
                    dR = (RL - R).view(self.nmember, -1)
                    den = (dR * dR).sum(1)
                    a = torch.where(den > 0, -(R.view(self.nmember, -1) * dR).sum(1) / den, torch.zeros_like(den))
                    a = a.clamp(0.0, 1.5)
                    ua = uE + self._per_dof(a) * (uL - uE)
                    Ra = torch.empty_like(R)
                    t_res += timed(self._residual, ua, Ra)
                    na = self._norms(Ra)
                    # the FP64 residual decides, member by member: extrapolation,
                    # prediction or the line-search point
                    best = torch.stack([nE, nL, na]).argmin(0)
                    chose_learned[k - 1] = int((best > 0).sum())
                    pick = lambda x0, x1, x2: torch.where(self._per_dof(best == 1), x1,
                                                          torch.where(self._per_dof(best == 2), x2, x0))
                    u.copy_(pick(uE, uL, ua))
                    R.copy_(pick(R, RL, Ra))
                    nS = torch.stack([nE, nL, na]).gather(0, best[None])[0]
                else:
                    u.copy_(uE)
                    nS = nE
                start_ratio.append(float((nS / R0).mean()))
                conv = nS <= thresh
                r_cur = float((nS / R0).max())
                if bool(conv.all()):
                    lock_iters[k - 1] = 0
                    history.append([r_cur])
                    u_prev4, u_prev3, u_prev2, u_prev = u_prev3, u_prev2, u_prev, u.clone()
                    if record:
                        traj.append(u.cpu().numpy())
                    continue
                t_asm += timed(self._assemble, u)
                torch.neg(R, out=rhs)
            it = 0
            done = conv.cpu().numpy().copy()
            hk = [r_cur]
            while True:
                eta = self._eta(r_cur, C)
                self.lin.rtol = eta
                t_lin += timed(self._solve, rhs, du, done)
                u.add_(du)
                it += 1
                if first_linear:
                    t_asm += timed(self._assemble, u)
                    torch.mul(self.T, self.free_f, out=R)
                else:
                    t_res += timed(self._residual, u, R)
                nr = self._norms(R)
                live = torch.as_tensor(~done, device=dev)
                r_new = float((nr / R0)[live].max())
                if r_cur > 0:
                    bound = max(1e-12, r_new / (r_cur * r_cur))   # C <= bound; = C when the solve was tight
                    C = bound if (C is None or eta * r_cur <= 0.3 * r_new) else min(C, bound)
                r_cur = float((nr / R0).max())
                hk.append(r_cur)
                c = (nr <= thresh).cpu().numpy()
                newly = c & ~done
                iters[k - 1, newly] = it
                done |= c
                if done.all():
                    break
                if it >= self.maxit or not bool(torch.isfinite(nr).all()):
                    failed = True
                    iters[k - 1, ~done] = -1
                    break
                if not first_linear:
                    t_asm += timed(self._assemble, u)
                torch.neg(R, out=rhs)
            lock_iters[k - 1] = it
            history.append(hk)
            if log is not None:
                log("increment {}/{}: {} Newton iterations, worst |R|/R0 {}, {:.1f}s".format(
                    k, nincr, it, ["{:.1e}".format(v) for v in hk], time.perf_counter() - t0))
            # in the reference's mode V now holds K at the converged state, which the
            # next increment's predictor uses; the warm-started modes assemble at
            # their start instead
            u_prev4, u_prev3, u_prev2, u_prev = u_prev3, u_prev2, u_prev, u.clone()
            if record:
                traj.append(u.cpu().numpy())
            if failed:
                break
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
        out = {"start": start, "precision": self.precision, "csr": self.csr, "wall_s": wall, "linear_s": t_lin, "assembly_s": t_asm,
               "residual_s": t_res, "predict_s": t_pred, "lockstep_iterations": lock_iters.tolist(),
               "member_iterations": iters, "chose_learned": chose_learned.tolist(), "start_residual": start_ratio,
               "pcg_iterations": list(self.lin.iterations), "residual_history": history,
               "forcing": self.forcing, "failed": failed}
        self.lin.iterations = []
        self.lin.rtol = self.lin_rtol
        if record:
            out["trajectory"] = np.stack(traj)     # (nincr + 1, 3 nnode)
        return out


class WarmStartMLP(object):
    """Device-resident predictor: parameters, Adam moments and normalisation."""

    def __init__(self, device="cuda", seed=0):
        import torch
        lib = _lib.load()
        self.device = torch.device(device)
        self.nparam = int(lib.fl_mlp_param_count())
        self.base_order = 1     # the extrapolation the network's output corrects
        self.P = torch.as_tensor(self.initial_params(seed), device=self.device)
        self.m = torch.zeros_like(self.P)
        self.v = torch.zeros_like(self.P)
        self.step = 0
        self.norm_in = torch.as_tensor(np.concatenate([np.zeros(NIN), np.ones(NIN)]).astype(np.float32),
                                       device=self.device)
        self.norm_out = torch.as_tensor(np.array([1, 1, 1, 0, 0, 0], np.float32), device=self.device)
        self._buf = {}

    @staticmethod
    def initial_params(seed):
        """Weights U(-1/sqrt(fan_in), 1/sqrt(fan_in)), biases 0, in the library's order."""
        rng = np.random.Generator(np.random.PCG64(seed))
        parts = []
        for fan_in, fan_out in ((IP, H), (H, H), (H, OP)):
            b = 1.0 / np.sqrt(fan_in)
            parts.append(rng.uniform(-b, b, size=fan_in * fan_out))
            parts.append(np.zeros(fan_out))
        return np.concatenate(parts).astype(np.float32)

    def _buffers(self, n):
        import torch
        if self._buf.get("n", 0) < n:
            f32 = dict(dtype=torch.float32, device=self.device)
            self._buf = {"n": n, "H1": torch.empty((n, H), **f32), "H2": torch.empty((n, H), **f32),
                         "Y": torch.empty((n, OP), **f32), "F": torch.empty((n, IP), **f32)}
        return self._buf

    def forward(self, F, n):
        lib = _lib.load()
        b = self._buffers(n)
        _lib.check(lib.fl_mlp_forward(_lib.ptr(self.P), _lib.ptr(F), n, _lib.ptr(b["H1"]), _lib.ptr(b["H2"]),
                                      _lib.ptr(b["Y"]), _lib.stream_ptr()), "fl_mlp_forward")
        return b

    def features(self, X, u1, u2, lam, F):
        lib = _lib.load()
        nn = X.numel() // 3
        _lib.check(lib.fl_warmstart_features(_lib.ptr(X), _lib.ptr(u1), _lib.ptr(u2), float(lam), nn,
                                             _lib.ptr(self.norm_in), _lib.ptr(F), _lib.stream_ptr()),
                   "fl_warmstart_features")
        return F

    def predict(self, X, u1, u2, lam, free_u8, ubc, out, base=None):
        """out = prediction of the next converged displacement (Dirichlet dofs = ubc):
        the network's correction added to its base, the linear extrapolation
        2 u1 - u2, or ``base`` (the extrapolation of order base_order the network was
        trained against; fl_warmstart_apply forms 2 a - b + correction, so a = b = base)."""
        lib = _lib.load()
        nn = X.numel() // 3
        b = self._buffers(nn)
        self.features(X, u1, u2, lam, b["F"])
        self.forward(b["F"], nn)
        a1, a2 = (u1, u2) if base is None else (base, base)
        _lib.check(lib.fl_warmstart_apply(_lib.ptr(a1), _lib.ptr(a2), _lib.ptr(b["Y"]), nn, _lib.ptr(self.norm_out),
                                          _lib.ptr(free_u8), _lib.ptr(ubc), _lib.ptr(out), _lib.stream_ptr()),
                   "fl_warmstart_apply")
        return out

    def train_step(self, F, T, n, lr, nout=3, ws=None, G=None, loss=None):
        """One Adam step on the minibatch (F, T) of n rows (n % 32 == 0); returns the
        device loss tensor."""
        import torch
        lib = _lib.load()
        b = self.forward(F, n)
        if G is None:
            G = torch.empty(self.nparam, dtype=torch.float32, device=self.device)
        if loss is None:
            loss = torch.empty(1, dtype=torch.float32, device=self.device)
        if ws is None:
            ws = torch.empty(int(lib.fl_mlp_workspace_size(n)), dtype=torch.uint8, device=self.device)
        _lib.check(lib.fl_mlp_backward(_lib.ptr(self.P), _lib.ptr(F), _lib.ptr(b["H1"]), _lib.ptr(b["H2"]),
                                       _lib.ptr(b["Y"]), _lib.ptr(T), n, nout, _lib.ptr(G), _lib.ptr(loss),
                                       _lib.ptr(ws), ws.numel(), _lib.stream_ptr()), "fl_mlp_backward")
        self.step += 1
        _lib.check(lib.fl_mlp_adam(_lib.ptr(self.P), _lib.ptr(G), _lib.ptr(self.m), _lib.ptr(self.v), float(lr),
                                   0.9, 0.999, 1e-8, self.step, _lib.stream_ptr()), "fl_mlp_adam")
        return loss

    def state(self):
        return {"P": self.P.cpu().numpy(), "norm_in": self.norm_in.cpu().numpy(),
                "norm_out": self.norm_out.cpu().numpy(), "base_order": np.int32(self.base_order)}

    def save(self, path, **meta):
        np.savez(path, **self.state(), **{k: np.asarray(v) for k, v in meta.items()})

    @classmethod
    def load(cls, path, device="cuda"):
        import torch
        d = np.load(path, allow_pickle=False)
        net = cls(device=device)
        net.P.copy_(torch.as_tensor(d["P"]))
        net.norm_in.copy_(torch.as_tensor(d["norm_in"]))
        net.norm_out.copy_(torch.as_tensor(d["norm_out"]))
        net.base_order = int(d["base_order"]) if "base_order" in d.files else 1
        return net


def _free_nodes(free_u8):
    return free_u8.view(-1, 3).bool().all(dim=1)


# polynomial extrapolation to the next of equally spaced states u_{k-1}, u_{k-2}, ...
EXTRAP_COEF = {1: (2.0, -1.0), 2: (3.0, -3.0, 1.0), 3: (4.0, -6.0, 4.0, -1.0)}


def extrapolate(history, order):
    """sum_j EXTRAP_COEF[order][j] history[j] (history: u_{k-1}, u_{k-2}, ...)."""
    c = EXTRAP_COEF[order]
    out = c[0] * history[0]
    for j in range(1, len(c)):
        out = out + c[j] * history[j]
    return out


def _base(U, k, order):
    """The extrapolated start of increment k from the recorded states U[0..k-1]."""
    q = min(order, k - 1)
    return extrapolate([U[k - 1 - j] for j in range(q + 1)], q)


def build_dataset(net, family, runs):
    """Training rows from recorded trajectories.

    ``runs`` = [(X [3 nnode], free mask uint8 [3 nnode], trajectory (nincr + 1, 3 nnode)),
    ...] (X and the mask as device tensors, e.g. a LockstepNewton's ``X``/``free``).  For every
    increment k >= 2 and every node whose three dofs are free: inputs (X, u_{k-1},
    u_{k-1} - u_{k-2}, k / nincr), target u_k minus the polynomial extrapolation of
    order min(net.base_order, k - 1) (order 1: 2 u_{k-1} - u_{k-2}).  Sets the
    network's input/output standardisation from these rows and returns device
    (F [N, 32], T [N, 32]) with N rounded down to a multiple of 32."""
    import torch
    dev = net.device
    nincr = family.nincr
    raw_s = torch.zeros(NIN, dtype=torch.float64, device=dev)
    raw_q = torch.zeros(NIN, dtype=torch.float64, device=dev)
    tgt_s = torch.zeros(3, dtype=torch.float64, device=dev)
    tgt_q = torch.zeros(3, dtype=torch.float64, device=dev)
    count = 0
    pieces = []
    for Xf, free, traj in runs:
        U = torch.as_tensor(traj, device=dev)
        X = Xf.view(-1, 3)
        sel = _free_nodes(free)
        for k in range(2, nincr + 1):
            u1, u2, uk = U[k - 1].view(-1, 3), U[k - 2].view(-1, 3), U[k].view(-1, 3)
            raw = torch.cat([X, u1, u1 - u2, torch.full_like(X[:, :1], k / float(nincr))], dim=1)[sel]
            tgt = (uk - _base(U, k, net.base_order).view(-1, 3))[sel]
            raw_s += raw.sum(0)
            raw_q += (raw * raw).sum(0)
            tgt_s += tgt.sum(0)
            tgt_q += (tgt * tgt).sum(0)
            count += raw.shape[0]
        pieces.append((Xf, U, sel))
    mean = raw_s / count
    std = torch.sqrt(torch.clamp(raw_q / count - mean * mean, min=0.0))
    inv = torch.where(std > 1e-12, 1.0 / std, torch.ones_like(std))
    tmean = tgt_s / count
    tstd = torch.sqrt(torch.clamp(tgt_q / count - tmean * tmean, min=1e-30))
    net.norm_in.copy_(torch.cat([mean, inv]).to(torch.float32))
    net.norm_out.copy_(torch.cat([tstd, tmean]).to(torch.float32))
    n = (count // 32) * 32
    F = torch.empty((count, IP), dtype=torch.float32, device=dev)
    T = torch.zeros((count, OP), dtype=torch.float32, device=dev)
    o = 0
    for Xf, U, sel in pieces:
        nn = Xf.numel() // 3
        Fk = torch.empty((nn, IP), dtype=torch.float32, device=dev)
        for k in range(2, nincr + 1):
            net.features(Xf, U[k - 1], U[k - 2], k / float(nincr), Fk)
            uk = U[k].view(-1, 3)
            tgt = ((uk - _base(U, k, net.base_order).view(-1, 3))[sel] - tmean) / tstd
            m = tgt.shape[0]
            F[o:o + m] = Fk[sel]
            T[o:o + m, :3] = tgt.to(torch.float32)
            o += m
    return F[:n], T[:n]


def train(net, F, T, epochs=8, batch=8192, lr=2e-3, lr_final=2e-5, seed=0, log=None):
    """Adam over shuffled minibatches (cosine-decayed learning rate); returns the mean
    loss of each epoch (standardised units)."""
    import torch
    lib = _lib.load()
    N = F.shape[0]
    batch = min(batch, (N // 32) * 32)
    nb = N // batch
    total = epochs * nb
    G = torch.empty(net.nparam, dtype=torch.float32, device=net.device)
    ws = torch.empty(int(lib.fl_mlp_workspace_size(batch)), dtype=torch.uint8, device=net.device)
    losses = torch.zeros(nb, dtype=torch.float32, device=net.device)
    gen = torch.Generator(device=net.device)
    gen.manual_seed(seed)
    hist = []
    step = 0
    for ep in range(epochs):
        perm = torch.randperm(N, device=net.device, generator=gen)
        for b in range(nb):
            idx = perm[b * batch:(b + 1) * batch]
            Fb = F.index_select(0, idx)
            Tb = T.index_select(0, idx)
            c = 0.5 * (1 + np.cos(np.pi * step / max(total - 1, 1)))
            net.train_step(Fb, Tb, batch, lr_final + (lr - lr_final) * c, ws=ws, G=G, loss=losses[b:b + 1])
            step += 1
        hist.append(float(losses.mean()))
        if log is not None:
            log("epoch {} loss {:.4e}".format(ep, hist[-1]))
    return hist
