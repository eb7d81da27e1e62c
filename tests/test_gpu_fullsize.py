"""The benchmark workload at its full size: the 24^3 hex p=4 cube (13,824 elements,
2,738,019 dofs, 1.73e9 CSR entries) in the bench's twisted state, assembled on the
device, checked through properties that do not depend on the mesh size (the CPU oracle
cannot assemble this size in test time; it is compared directly at the smaller sizes of
test_gpu_assembly.py, and carries the same properties, tests/test_oracle.py
test_rigid_translation_and_equilibrium):

* rigid translations are in the kernel of K: every element block is
  sum_q grad N_a . D . grad N_b, and sum_b grad N_b = 0 (partition of unity), so
  K e_k = 0 for e_k = 1 on dof component k of every node -- a contribution scattered to
  a wrong dof component, dropped or doubled breaks it;
* the internal forces sum to zero per component (sum_a grad N_a = 0 again);
* K is symmetric (the Mooney-Rivlin tangent is an energy Hessian): entry (3a+p, 3b+q)
  equals entry (3b+q, 3a+p) -- a contribution placed in a wrong column breaks it;
* the direct scatter and the element-block path agree (their sums run in different
  orders: 1e-13 of max |K|), T bitwise (both sum it in element order).

The Newton solve on the same mesh is checked by its own FP64 residual test and by the
agreement of two solver paths that share nothing but the assembly's FP64 residual.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_hex_p4_24cube_invariants(cuda):
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.solver import spmv
    from florence_amd.workloads import cube_hex_p4, twisted_state
    mesh, fs, mat = cube_hex_p4(24)
    x = torch.as_tensor(twisted_state(mesh, 15.0), device="cuda")
    dm = DeviceMeshData(mesh, fs, build_indices=False)
    assert dm.nnz == 1728900297
    V, T, st = assemble_device(dm, mat, x)
    V, T = V.clone(), T.clone()
    torch.cuda.synchronize()
    assert dm.scatter is not None and int(st[0].item()) == 0
    vmax = float(V.abs().max())

    # rigid translations and the force balance
    e = torch.zeros(3 * dm.nnode, dtype=torch.float64, device="cuda")
    y = torch.empty_like(e)
    for k in range(3):
        e.zero_()
        e[k::3] = 1.0
        spmv(dm, V, e, y)
        ky = float(y.abs().max())
        tk = float(T[k::3].sum().abs())
        tabs = float(T[k::3].abs().sum())
        print("component {}: max |K e_k| {:.2e} of max |K| {:.2e}; |sum T_k| {:.2e} of sum |T_k| {:.2e}".format(
            k, ky, vmax, tk, tabs), flush=True)
        assert ky <= 1e-11 * vmax
        assert tk <= 1e-12 * tabs

    # symmetry, node pair by node pair (chunks of the node-level pattern)
    nn = dm.nnode
    ptr = dm.node_nbr_ptr.long()
    nbrs = dm.node_nbrs.long()
    indptr = dm.indptr.long()
    deg = ptr[1:] - ptr[:-1]
    row = torch.repeat_interleave(torch.arange(nn, device="cuda"), deg)
    keys = row * nn + nbrs                    # ascending: rows in order, neighbours sorted
    worst = 0.0
    chunk = 1 << 24
    for s in range(0, nbrs.numel(), chunk):
        k = torch.arange(s, min(s + chunk, nbrs.numel()), device="cuda")
        a, b = row[k], nbrs[k]
        kt = torch.searchsorted(keys, b * nn + a)
        assert torch.equal(keys[kt], b * nn + a)    # the pattern is structurally symmetric
        r, rt = k - ptr[a], kt - ptr[b]
        for p in range(3):
            for q in range(3):
                v = V[indptr[3 * a + p] + 3 * r + q]
                vt = V[indptr[3 * b + q] + 3 * rt + p]
                worst = max(worst, float((v - vt).abs().max()))
    print("max |K_ij - K_ji| {:.2e} of max |K| {:.2e}".format(worst, vmax), flush=True)
    assert worst <= 1e-12 * vmax
    del row, keys

    # the element-block path
    V2, T2, st2 = assemble_device(dm, mat, x, direct=False)
    torch.cuda.synchronize()
    assert int(st2[0].item()) == 0
    dv = float((V2 - V).abs().max())
    print("direct vs block path: max |dK| {:.2e} of max |K| {:.2e}".format(dv, vmax), flush=True)
    assert dv <= 1e-13 * vmax
    assert torch.equal(T2, T)
    del V, V2, T, T2, dm
    torch.cuda.empty_cache()


def test_twisted_cube_24_newton(cuda):
    """The 24^3 twisted-cube Newton solve (10 increments to |R_free| <= 1e-8 R0, the
    p-multigrid flexible CG): the fast path (FP32 tangent, FP32-stored CSR, polynomial
    start, adaptive linear tolerances) and the plain FP64 path (the reference's
    predictor start, adaptive tolerances) both converge every increment and end at the
    same displacements (1e-8 relative: both stop at the residual tolerance, not at the
    same iterate)."""
    import torch
    from florence_amd.learned import LockstepNewton, TwistStretchFamily
    fam = TwistStretchFamily(n=24, nincr=10, theta_range=(30.0, 30.0), stretch_range=(1.0, 1.0))
    finals = {}
    for name, prec, start, csr in (("fast", "fp32", "poly", "fp32"), ("plain", "fp64", "previous", "fp64")):
        ln = LockstepNewton(fam, [0], precision=prec, forcing="adaptive", csr=csr, linear="mg")
        r = ln.solve(start, record=True)
        print(name, "Newton", r["lockstep_iterations"], "CG", sum(r["pcg_iterations"]),
              "wall {:.2f}s".format(r["wall_s"]), flush=True)
        assert not r["failed"]
        assert all(0 < it <= 6 for it in r["lockstep_iterations"])
        assert max(h[-1] for h in r["residual_history"]) <= 1e-8
        finals[name] = r["trajectory"][-1]
        del ln
        torch.cuda.empty_cache()
    ref = finals["plain"]
    assert np.abs(finals["fast"] - ref).max() <= 1e-8 * np.abs(ref).max()
