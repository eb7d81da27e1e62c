"""Host tables of the p-multigrid hierarchy (florence_amd/multigrid.py), without a GPU:
coarse node selection, coarse bases and the transfer tables, pinned by what the
interpolation must do."""
import numpy as np
import pytest
import scipy.sparse as sp

from cases import cube_mesh, function_space
from florence_amd.multigrid import _coarse_locals, coarse_basis, hierarchy_tables, local_node_coordinates


@pytest.mark.parametrize("et,p", [("hex", 2), ("hex", 3), ("hex", 4), ("tet", 2), ("tet", 3)])
def test_coarse_bases_are_lagrange_and_reproduce(et, p):
    npe = (p + 1) ** 3 if et == "hex" else (p + 1) * (p + 2) * (p + 3) // 6
    xi = local_node_coordinates(et, npe)
    m = cube_mesh(et, p, 1)
    fs = function_space(m, p)
    tabs = hierarchy_tables(et, m.elements, m.nnode, fs.quadrature.points)
    expect = {("hex", 4): [27, 8]}.get((et, p), [8] if et == "hex" else [4])
    assert [t["elements"].shape[1] for t in tabs] == expect
    rng = np.random.default_rng(0)
    pts = rng.uniform(-0.6, 0.3, (7, 3))
    fine_xi = xi
    for t in tabs:
        pc = t["degree"]
        # the coarse nodes are the fine nodes on the coarse lattice: Kronecker there
        cxi = fine_xi[_coarse_locals(et, pc, fine_xi)]
        N, _ = coarse_basis(et, pc, cxi, cxi)
        assert np.abs(N - np.eye(cxi.shape[0])).max() < 1e-12
        # partition of unity and gradients (central differences)
        N, G = coarse_basis(et, pc, cxi, pts)
        assert np.abs(N.sum(axis=1) - 1).max() < 1e-12
        h = 1e-6
        for d in range(3):
            e = np.zeros(3)
            e[d] = h
            fd = (coarse_basis(et, pc, cxi, pts + e)[0] - coarse_basis(et, pc, cxi, pts - e)[0]) / (2 * h)
            assert np.abs(fd - G[d]).max() < 1e-7
        # JmCt: the gradients at the fine quadrature points
        _, Gq = coarse_basis(et, pc, cxi, fs.quadrature.points)
        assert np.array_equal(t["JmCt"], np.transpose(Gq, (1, 2, 0)))
        fine_xi = cxi


@pytest.mark.parametrize("et,p,n", [("hex", 4, 2), ("hex", 2, 3), ("tet", 2, 2), ("tet", 3, 1)])
def test_prolongation_interpolates_linear_fields(et, p, n):
    """On a straight-sided mesh the nodal interpolation of every coarse level reproduces
    a linear field exactly (coordinates are trilinear / linear in the reference
    coordinates), R is P^T, and every fine node's weights sum to one."""
    m = cube_mesh(et, p, n)
    fs = function_space(m, p)
    X = np.asarray(m.points)
    tabs = hierarchy_tables(et, m.elements, m.nnode, fs.quadrature.points)
    A = np.array([[0.3, -1.2, 2.0], [0.7, 0.1, -0.4], [1.5, 0.2, 0.9]])
    fine_pos = X
    for t in tabs:
        ptr, idx, w = t["P"]
        nfine = ptr.size - 1
        P = sp.csr_matrix((w, idx, ptr), shape=(nfine, t["fine_ids"].size))
        uc = X[t["fine_ids"]] @ A.T
        assert np.abs(P @ uc - fine_pos @ A.T).max() < 1e-12
        assert np.abs(np.asarray(P.sum(axis=1)).ravel() - 1.0).max() < 1e-12
        rptr, ridx, rw = t["R"]
        R = sp.csr_matrix((rw, ridx, rptr), shape=(t["fine_ids"].size, nfine))
        assert abs(R - P.T).max() == 0.0
        # every row of R lists fine nodes in ascending order (fixed summation order)
        assert all(np.all(np.diff(ridx[rptr[c]:rptr[c + 1]]) > 0) for c in range(rptr.size - 1))
        fine_pos = X[t["fine_ids"]]
