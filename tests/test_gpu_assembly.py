"""GPU parity of the DF assembly against the FP64 oracle (oracle/df_numpy.py).

Bar: CSR indptr/indices bit-exact; K data and residual T within 1e-12 relative
(infinity norm over the whole vector); two runs bitwise identical.
"""
import numpy as np
import pytest

from cases import tet_p2_cube, hex_p4_cube, cube_mesh, deformed, function_space
from florence_amd.materials import MooneyRivlin, NeoHookean
from oracle import df_numpy as O

pytestmark = pytest.mark.gpu
RTOL = 1e-12


def _gpu_assemble(mesh, fs, mat, x):
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    dm = DeviceMeshData(mesh, fs)
    V, T, status = assemble_device(dm, mat, x)
    torch.cuda.synchronize()
    return dm, V.cpu().numpy(), T.cpu().numpy(), status.cpu().numpy()


def _check(mesh, fs, mat, x, label):
    dm, V, T, status = _gpu_assemble(mesh, fs, mat, x)
    indptr_ref, indices_ref = O.sparsity_pattern(mesh.elements, mesh.nnode)
    assert np.array_equal(dm.indptr.cpu().numpy(), indptr_ref), label
    assert np.array_equal(dm.indices.cpu().numpy(), indices_ref), label
    Vr, Tr = O.assemble_csr(mesh.points, mesh.elements, x, fs.Jm, fs.AllGauss, mat, indptr_ref, indices_ref)
    ek = np.abs(V - Vr).max() / np.abs(Vr).max()
    et = np.abs(T - Tr).max() / np.abs(Tr).max()
    print("{}: nnz={} |dK|/|K|={:.3e} |dT|/|T|={:.3e} status={}".format(label, Vr.size, ek, et, status), flush=True)
    assert ek <= RTOL, label
    assert et <= RTOL, label
    assert status[0] == 0, label
    return dm, V, T


@pytest.mark.parametrize("et,p,n,mat", [
    ("tet", 1, 2, "NH"), ("tet", 2, 2, "MR"), ("tet", 3, 2, "NH"), ("tet", 4, 1, "MR"),
    ("hex", 1, 3, "MR"), ("hex", 2, 2, "NH"), ("hex", 3, 2, "MR"),
])
def test_small_meshes(cuda, et, p, n, mat):
    m = cube_mesh(et, p, n, jitter=0.05, seed=p)
    fs = function_space(m, p)
    material = NeoHookean(3, mu=1.0, lamb=10.0) if mat == "NH" else MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    _check(m, fs, material, deformed(m.points, 0.05, seed=p), "{}{} n={} {}".format(et, p, n, mat))


def test_tet_p2_cube_full(cuda):
    """3072 p=2 tets, 11-point rule, Neo-Hookean, deformed state."""
    m, fs, mat = tet_p2_cube()
    x = deformed(m.points, 0.04, seed=11)
    dm, V, T = _check(m, fs, mat, x, "tet_p2_cube")
    # determinism: second call bitwise identical
    _, V2, T2, _ = _gpu_assemble(m, fs, mat, x)
    assert np.array_equal(V, V2) and np.array_equal(T, T2)


def test_hex_p4_cube_n2(cuda):
    """Hex p=4, 6^3 Gauss-Lobatto points, Mooney-Rivlin on a 2^3 cube (8 elements)."""
    m, fs, mat = hex_p4_cube(n=2)
    _check(m, fs, mat, deformed(m.points, 0.05, seed=5), "hex p4 cube n=2")


def test_residual_only_matches(cuda):
    import torch
    from florence_amd.assembly import DeviceMeshData, residual_device, assemble_device
    m, fs, mat = hex_p4_cube(n=2)
    x = deformed(m.points, 0.05, seed=6)
    dm = DeviceMeshData(m, fs)
    T, st = residual_device(dm, mat, x)
    _, Tfull, _ = assemble_device(dm, mat, x)
    torch.cuda.synchronize()
    Tr = O.residual_only(m.points, m.elements, x, fs.Jm, fs.AllGauss, mat)
    # the reference-shaped entry (no tensor basis: dense point and traction kernels)
    from florence_amd import _lib
    lib = _lib.load()
    Td = torch.empty_like(T)
    xd = torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    ws = dm.workspace(residual_only=True)
    prm, _keep = _lib.double_array(mat.device_parameters())
    _lib.check(lib.fl_df_residual(mat.material_id, prm, _lib.ptr(dm.points), _lib.ptr(dm.elements), _lib.ptr(xd),
                                  _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem, dm.npe, dm.nnode,
                                  _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem), _lib.ptr(Td), _lib.ptr(ws),
                                  ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr()), "fl_df_residual")
    torch.cuda.synchronize()
    assert dm.tensor_basis is not None
    T, Tfull, Td = T.cpu().numpy(), Tfull.cpu().numpy(), Td.cpu().numpy()
    for label, t in (("factorised", T), ("assembly", Tfull), ("dense", Td)):
        err = np.abs(t - Tr).max() / np.abs(Tr).max()
        print("residual-only {}: |dT|/|T| = {:.3e}".format(label, err))
        assert err <= RTOL, label
    assert np.abs(T - Tfull).max() / np.abs(Tr).max() <= RTOL


@pytest.mark.parametrize("et,p,n", [("tet", 3, 2), ("hex", 4, 2)])
def test_reference_entry_point_equals_rank_table_path(cuda, et, p, n):
    """fl_df_assemble (the reference-shaped entry: no rank table, neighbour lists
    searched per call, dense contraction) and the rank-table path give bitwise the same
    V and T (the rank-table side built without the tensor basis, so both use the same
    stiffness kernel)."""
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    mesh = cube_mesh(et, p, n, jitter=0.05, seed=p)
    fs = function_space(mesh, p)
    mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = deformed(mesh.points, 0.05, seed=p)
    dm = DeviceMeshData(mesh, fs, tensor_basis=False)
    assert dm.ranks is not None
    V1, T1, _ = assemble_device(dm, mat, x)
    V1, T1 = V1.clone(), T1.clone()
    lib = _lib.load()
    V2 = torch.empty_like(V1)
    T2 = torch.empty_like(T1)
    xd = torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    ws = dm.workspace()
    prm, _keep = _lib.double_array(mat.device_parameters())
    _lib.check(lib.fl_df_assemble(mat.material_id, prm, _lib.ptr(dm.points), _lib.ptr(dm.elements), _lib.ptr(xd),
                                  _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem, dm.npe, dm.nnode, 1,
                                  _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem), _lib.ptr(dm.node_nbr_ptr),
                                  _lib.ptr(dm.node_nbrs), _lib.ptr(dm.node_order), dm.bucket_offsets,
                                  dm.bucket_degree, _lib.ptr(dm.indptr), _lib.ptr(V2), _lib.ptr(T2), _lib.ptr(ws),
                                  ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr()), "fl_df_assemble")
    torch.cuda.synchronize()
    assert torch.equal(V1, V2) and torch.equal(T1, T2)
    assert int(dm.status[0].item()) == 0


@pytest.mark.parametrize("variant", [0, 1, 2])
def test_hex_p4_tile_variants(cuda, variant):
    """Every tile-shape variant of the dense (MFMA) contraction the autotuner may pick
    for 128-node elements meets the oracle bar (p=4 hex cube, 2^3 elements); the
    sum-factorised path is switched off for it."""
    from florence_amd import _lib
    lib = _lib.load()
    key = _lib.FL_TUNE_STIFF_VARIANT_BASE + 4
    _lib.check(lib.fl_set_tuning(key, variant + 1), "fl_set_tuning")
    _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 1), "fl_set_tuning")
    try:
        m, fs, mat = hex_p4_cube(n=2)
        _check(m, fs, mat, deformed(m.points, 0.05, seed=9), "hex p4 variant {}".format(variant))
    finally:
        lib.fl_set_tuning(key, 0)
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)


_CLASS_CASES = [("tet", 3, 1, v) for v in range(4)] + [("tet", 4, 2, v) for v in range(3)] + \
    [("hex", 3, 3, v) for v in range(5)]


@pytest.mark.parametrize("et,p,npc,variant", _CLASS_CASES,
                         ids=["{}{}-class{}-v{}".format(*c) for c in _CLASS_CASES])
def test_dense_tile_variants(cuda, et, p, npc, variant):
    """Every tile-shape variant of the dense contraction for the 32-, 48- and 64-node
    classes (tet p=3, tet p=4, hex p=3 with the factorised path off) meets the oracle bar,
    in FP64, and in FP32 within the FP32 bar of the FP64 oracle."""
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    lib = _lib.load()
    assert lib.fl_stiffness_variant_count(npc) == _lib.STIFF_VARIANTS[npc] > variant
    key = _lib.FL_TUNE_STIFF_VARIANT_BASE + npc
    _lib.check(lib.fl_set_tuning(key, variant + 1), "fl_set_tuning")
    _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 1), "fl_set_tuning")
    try:
        m = cube_mesh(et, p, 2 if et == "hex" else 1, jitter=0.05, seed=11)
        fs = function_space(m, p)
        mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
        x = deformed(m.points, 0.05, seed=12)
        _check(m, fs, mat, x, "{} p{} dense variant {}".format(et, p, variant))
        dm = DeviceMeshData(m, fs)
        V32, _, st = assemble_device(dm, mat, x, precision="fp32")
        torch.cuda.synchronize()
        V64, _, _ = assemble_device(dm, mat, x)
        e32 = float((V32 - V64).abs().max() / V64.abs().max())
        assert e32 <= 1e-5 and int(st[0]) == 0, e32
    finally:
        lib.fl_set_tuning(key, 0)
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)


@pytest.mark.parametrize("p", [1, 2, 3, 4])
def test_sum_factorised_equals_dense(cuda, p):
    """Tensor-product hexes: the sum-factorised point, block and traction kernels (default
    when the tensor basis is present) and the dense ones both meet the oracle bar; their
    blocks agree to 1e-13."""
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    if p == 4:
        m, fs, mat = hex_p4_cube(n=2)
    else:
        m = cube_mesh("hex", p, 2, jitter=0.05, seed=p)
        fs = function_space(m, p)
        mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = deformed(m.points, 0.05, seed=20 + p)
    dm, V, T = _check(m, fs, mat, x, "hex p{} sum-factorised".format(p))
    assert dm.tensor_basis is not None
    lib = _lib.load()
    _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 1), "fl_set_tuning")
    try:
        V2, T2, _ = assemble_device(dm, mat, x)
        torch.cuda.synchronize()
    finally:
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
    V2, T2 = V2.cpu().numpy(), T2.cpu().numpy()
    assert np.abs(V - V2).max() <= 1e-13 * np.abs(V2).max()
    # T sums element forces that partly cancel, and the two paths also take the point
    # Jacobians' node sums in different orders: the oracle bar
    assert np.abs(T - T2).max() <= RTOL * np.abs(T2).max()
    # the all-VALU form of the factorised contraction (FL_TUNE_STIFF_SF = 3) and the
    # form with one type per MFMA reduction (5)
    for form in (3, 5):
        _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, form), "fl_set_tuning")
        try:
            V3, T3, _ = assemble_device(dm, mat, x)
            torch.cuda.synchronize()
        finally:
            lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
        print("form {}: |dV|/|V| = {:.3e}".format(form, np.abs(V - V3.cpu().numpy()).max() / np.abs(V2).max()))
        assert np.abs(V - V3.cpu().numpy()).max() <= 1e-13 * np.abs(V2).max(), form
        assert np.array_equal(T, T3.cpu().numpy()), form


@pytest.mark.parametrize("et,p,n", [("hex", 4, 2), ("hex", 2, 3), ("tet", 3, 2)])
def test_gather_forms_agree(cuda, et, p, n):
    """Every gather form (FL_TUNE_GATHER bits: one element at a time, element-interior
    nodes through the general kernel) and the factorised stiffness with and without the
    interior-row fusion (forms 0 and 5) give the same V and T: bitwise for the gather
    forms over the same element blocks, to 1e-13 across stiffness forms, and all within
    the oracle bar."""
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    if (et, p) == ("hex", 4):
        m, fs, mat = hex_p4_cube(n=n)
    else:
        m = cube_mesh(et, p, n, jitter=0.05, seed=p)
        fs = function_space(m, p)
        mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = deformed(m.points, 0.05, seed=31 + p)
    dm = DeviceMeshData(m, fs)
    lib = _lib.load()
    out = {}
    try:
        for sf in (0, 5):
            for gf in (0, 1, 2, 3):
                _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, sf), "fl_set_tuning")
                _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_GATHER, gf), "fl_set_tuning")
                V, T, st = assemble_device(dm, mat, x, direct=False)
                torch.cuda.synchronize()
                assert int(st[0].item()) == 0, (sf, gf)
                out[(sf, gf)] = (V.cpu().numpy(), T.cpu().numpy())
    finally:
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
        lib.fl_set_tuning(_lib.FL_TUNE_GATHER, 0)
    for sf in (0, 5):
        V0, T0 = out[(sf, 0)]
        for gf in (1, 2, 3):
            V1, T1 = out[(sf, gf)]
            assert np.array_equal(V0, V1) and np.array_equal(T0, T1), (sf, gf)
    Va, Ta = out[(0, 0)]
    Vb, Tb = out[(5, 0)]
    assert np.abs(Va - Vb).max() <= 1e-13 * np.abs(Vb).max()
    assert np.abs(Ta - Tb).max() <= RTOL * np.abs(Tb).max()
    indptr_ref, indices_ref = O.sparsity_pattern(m.elements, m.nnode)
    Vr, Tr = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, mat, indptr_ref, indices_ref)
    assert np.abs(Va - Vr).max() <= RTOL * np.abs(Vr).max()
    assert np.abs(Ta - Tr).max() <= RTOL * np.abs(Tr).max()


def test_direct_scatter_table(cuda):
    """fl_pattern_scatter_table against a restatement from the oracle's CSR pattern, slot
    by slot in the stiffness kernel's store order: the column block of every (row node,
    column node) pair of an element, negative (offset by -32768) exactly when an element
    of a lower colour holds both nodes, 0 in unused slots."""
    from florence_amd.assembly import DeviceMeshData
    for p, n in ((2, 3), (4, 2)):
        m = cube_mesh("hex", p, n, jitter=0.05, seed=2)
        fs = function_space(m, p)
        dm = DeviceMeshData(m, fs)
        assert dm.scatter is not None and dm.ncolour == 8
        p1, npe = p + 1, dm.npe
        pp, mt, nt = p1 * p1, (p1 * p1 + 15) // 16, (npe + 15) // 16
        tab = dm.scatter.cpu().numpy().reshape(dm.nelem, p1, nt, 64, mt, 4)
        ints = np.frombuffer(dm.tensor_basis.cpu().numpy().tobytes(), dtype=np.int32)
        q3 = int(ints[1]) ** 3
        loc = ints[4 + q3:4 + q3 + npe]
        el = np.asarray(m.elements)
        colour = dm.colour_of_elem.cpu().numpy()
        indptr_ref, indices_ref = O.sparsity_pattern(el, m.nnode)
        node_sets = [set(map(int, r)) for r in el]
        nadd = nused = 0
        for e in range(dm.nelem):
            for az in range(p1):
                for st in range(nt):
                    for lane in range(64):
                        for tm in range(mt):
                            for r in range(4):
                                row, col = 16 * tm + lane // 16 + 4 * r, 16 * st + lane % 16
                                got = tab[e, az, st, lane, tm, r]
                                if row >= pp or col >= npe:
                                    assert got == 0
                                    continue
                                ab, bb = divmod(row, p1)
                                ae, be, bz = col // pp, (col // p1) % p1, col % p1
                                a = int(el[e, loc[(ab * p1 + ae) * p1 + az]])
                                b = int(el[e, loc[(bb * p1 + be) * p1 + bz]])
                                nbrs = indices_ref[indptr_ref[3 * a]:indptr_ref[3 * a + 1]:3] // 3
                                rank = int(np.searchsorted(nbrs, b))
                                assert nbrs[rank] == b
                                add = any(colour[f] < colour[e] and a in node_sets[f] and b in node_sets[f]
                                          for f in range(dm.nelem))
                                nadd += add
                                nused += 1
                                assert got == (rank - 32768 if add else rank), (p, e, az, st, lane, tm, r)
        # every (row node, column node) pair of every element once; each CSR entry is
        # stored once and added to by each further element holding it
        assert nused == dm.nelem * npe * npe
        assert nadd == dm.nelem * npe * npe - dm.nnz // 9


@pytest.mark.parametrize("p,n", [(4, 2), (1, 4), (3, 3)])
def test_direct_scatter_equals_gather(cuda, p, n):
    """The direct scatter (colour-ordered stores and adds from the stiffness kernel)
    meets the oracle bar, repeats bitwise, and equals the gather over stored element
    blocks to 1e-13 in V (sums of three or more contributions taken in another order)
    and bitwise in T."""
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    if p == 4:
        m, fs, mat = hex_p4_cube(n=n)
    else:
        m = cube_mesh("hex", p, n, jitter=0.05, seed=p)
        fs = function_space(m, p)
        mat = NeoHookean(3, mu=1.0, lamb=10.0)
    x = deformed(m.points, 0.05, seed=41 + p)
    dm, V, T = _check(m, fs, mat, x, "direct scatter hex p{} n={}".format(p, n))
    assert dm.scatter is not None
    V2, T2, _ = assemble_device(dm, mat, x)
    torch.cuda.synchronize()
    assert np.array_equal(V, V2.cpu().numpy()) and np.array_equal(T, T2.cpu().numpy())
    Vg, Tg, st = assemble_device(dm, mat, x, direct=False)
    torch.cuda.synchronize()
    assert int(st[0].item()) == 0
    Vg, Tg = Vg.cpu().numpy(), Tg.cpu().numpy()
    print("direct vs gather: |dV|/|V| = {:.3e}".format(np.abs(V - Vg).max() / np.abs(Vg).max()))
    assert np.abs(V - Vg).max() <= 1e-13 * np.abs(Vg).max()
    assert np.array_equal(T, Tg)


@pytest.mark.parametrize("shape", ["cube, shuffled elements", "cylinder"])
def test_direct_scatter_irregular(cuda, shape):
    """The direct scatter on colourings that are not the 2 x 2 x 2 parity classes: a
    p=2 hex cube with its elements in a random order (greedy colours then follow no
    lattice) and a p=2 hex cylinder (curved elements).  Oracle bar, and the block path
    to 1e-13."""
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.mesh import Mesh
    if shape == "cylinder":
        m = Mesh()
        m.Cylinder(radius=0.5, length=2.0, nlong=4, nrad=1, element_type="hex")
        m.GetHighOrderMesh(p=2)
    else:
        m = cube_mesh("hex", 2, 3, jitter=0.05, seed=11)
        perm = np.random.default_rng(12).permutation(m.nelem)
        m.elements = np.ascontiguousarray(np.asarray(m.elements)[perm])
    fs = function_space(m, 2)
    mat = NeoHookean(3, mu=1.0, lamb=10.0)
    x = deformed(m.points, 0.04, seed=13)
    dm, V, T = _check(m, fs, mat, x, "direct scatter, {}".format(shape))
    assert dm.scatter is not None and dm.ncolour >= 8
    print("{}: {} colours".format(shape, dm.ncolour))
    Vg, Tg, _ = assemble_device(dm, mat, x, direct=False)
    torch.cuda.synchronize()
    Vg = Vg.cpu().numpy()
    assert np.abs(V - Vg).max() <= 1e-13 * np.abs(Vg).max()
    assert np.array_equal(T, Tg.cpu().numpy())


@pytest.mark.parametrize("et,p,direct", [("hex", 2, True), ("hex", 2, False), ("tet", 2, False)])
def test_reference_configuration(cuda, et, p, direct):
    """requires_geometry_update = 0 (integration on the reference configuration, the
    reference's `_KinematicMeasures_.h:94-99` branch): K meets the oracle bar on the direct
    scatter, the block path and the dense tet path, and T is zero."""
    import torch
    from florence_amd.assembly import DeviceMeshData, assemble_device
    m = cube_mesh(et, p, 2, jitter=0.05, seed=17)
    fs = function_space(m, p)
    mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = deformed(m.points, 0.05, seed=18)
    dm = DeviceMeshData(m, fs)
    V, T, st = assemble_device(dm, mat, x, requires_geometry_update=0, direct=direct)
    torch.cuda.synchronize()
    assert int(st[0].item()) == 0
    indptr_ref, indices_ref = O.sparsity_pattern(m.elements, m.nnode)
    Vr, Tr = O.assemble_csr(m.points, m.elements, x, fs.Jm, fs.AllGauss, mat, indptr_ref, indices_ref,
                            requires_geometry_update=0)
    V = V.cpu().numpy()
    err = np.abs(V - Vr).max() / np.abs(Vr).max()
    print("{}{} update=0 direct={}: |dK|/|K| = {:.3e}".format(et, p, direct, err))
    assert err <= RTOL
    assert np.all(Tr == 0.0) and np.all(T.cpu().numpy() == 0.0)


def test_direct_scatter_contract(cuda):
    """fl_df_assemble_cs through the C-ABI with the smaller workspace of
    fl_df_workspace_size_cs equals the Python path bitwise; a colour plan that does not
    cover the elements is refused before any launch; a colouring with two elements of
    one colour sharing a node is reported by fl_pattern_scatter_table."""
    import ctypes
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    m = cube_mesh("hex", 2, 3, jitter=0.05, seed=7)
    fs = function_space(m, 2)
    mat = MooneyRivlin(3, mu1=0.5, mu2=0.5, lamb=20.0)
    x = deformed(m.points, 0.05, seed=8)
    dm = DeviceMeshData(m, fs)
    V1, T1, _ = assemble_device(dm, mat, x)
    V1, T1 = V1.clone(), T1.clone()
    lib = _lib.load()
    ws = torch.empty(int(lib.fl_df_workspace_size_cs(dm.ngauss, dm.nelem, dm.npe)), dtype=torch.uint8,
                     device="cuda")
    assert ws.numel() < int(lib.fl_df_workspace_size(dm.ngauss, dm.nelem, dm.npe))
    V2, T2 = torch.empty_like(V1), torch.empty_like(T1)
    xd = torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    prm, _keep = _lib.double_array(mat.device_parameters())

    def call(offsets, ncolour):
        return lib.fl_df_assemble_cs(mat.material_id, prm, None, None, _lib.ptr(dm.points), _lib.ptr(dm.elements),
                                     _lib.ptr(xd), _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss), dm.ngauss, dm.nelem,
                                     dm.npe, dm.nnode, 1, _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem),
                                     _lib.ptr(dm.node_nbr_ptr), _lib.ptr(dm.node_nbrs), _lib.ptr(dm.node_order),
                                     dm.bucket_offsets, dm.bucket_degree, _lib.ptr(dm.indptr), _lib.ptr(V2),
                                     _lib.ptr(T2), _lib.FL_PREC_FP64, _lib.ptr(dm.ranks), _lib.ptr(dm.tensor_basis),
                                     ncolour, offsets, _lib.ptr(dm.colour_elems), _lib.ptr(dm.scatter),
                                     _lib.ptr(ws), ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr())

    _lib.check(call(dm.colour_offsets, dm.ncolour), "fl_df_assemble_cs")
    torch.cuda.synchronize()
    assert torch.equal(V1, V2) and torch.equal(T1, T2)
    assert int(dm.status[0].item()) == 0
    short = (ctypes.c_int32 * 65)(*list(dm.colour_offsets))
    short[dm.ncolour] = dm.nelem - 1
    assert call(short, dm.ncolour) == _lib.FL_E_ARG
    assert b"colour offsets" in lib.fl_last_error()
    assert call(dm.colour_offsets, 0) == _lib.FL_E_ARG
    # one colour for every element: neighbours share it, so the table build reports it
    bad = torch.zeros(dm.nelem, dtype=torch.int32, device="cuda")
    table = torch.empty_like(dm.scatter)
    dm.status.zero_()
    _lib.check(lib.fl_pattern_scatter_table(_lib.ptr(dm.elements), dm.nelem, dm.npe, _lib.ptr(dm.node_elem_ptr),
                                            _lib.ptr(dm.node_elem), _lib.ptr(dm.node_nbr_ptr),
                                            _lib.ptr(dm.node_nbrs), _lib.ptr(bad), _lib.ptr(dm.tensor_basis),
                                            _lib.ptr(table), _lib.ptr(dm.status), _lib.stream_ptr()),
               "fl_pattern_scatter_table")
    torch.cuda.synchronize()
    assert int(dm.status[0].item()) & _lib.FL_S_PATTERN
    dm.status.zero_()


def test_orphan_node(cuda):
    """A point no element uses (its CSR rows are empty, its T entries zero) leaves the
    assembly of the other nodes unchanged."""
    import copy
    m = cube_mesh("hex", 2, 2, jitter=0.05, seed=3)
    m2 = copy.deepcopy(m)
    m2.points = np.vstack([np.asarray(m.points), [[2.0, 2.0, 2.0]]])
    m2.nnode = m2.points.shape[0]
    fs = function_space(m2, 2)
    mat = NeoHookean(3, mu=1.0, lamb=10.0)
    x = deformed(m2.points, 0.05, seed=4)
    dm, V, T = _check(m2, fs, mat, x, "hex p2 with an orphan point")
    assert np.all(T[-3:] == 0.0)
