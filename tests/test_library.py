"""The C-ABI library loads on the CPU host and exports every entry point that
include/florence_amd.h declares, with the argument lists the binding uses
(no device calls: this runs without a GPU)."""
import ctypes
import os
import re

import numpy as np
import pytest

from cases import cube_mesh
from florence_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    h = open(os.path.join(ROOT, "include", "florence_amd.h")).read()
    h = re.sub(r"/\*.*?\*/", "", h, flags=re.S)
    return re.findall(r"^[\w][\w\s\*]*?\b(fl_\w+)\(([^;]*)\);", h, re.M)


def test_every_declared_symbol_is_exported():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    decl = _declared()
    assert len(decl) >= 10
    for name, _ in decl:
        assert hasattr(lib, name), name


def test_binding_matches_header_arity():
    for name, args in _declared():
        n = 0 if args.strip() in ("", "void") else len(args.split(","))
        assert name in _lib.SIGNATURES, name
        assert len(_lib.SIGNATURES[name][1]) == n, (name, n, len(_lib.SIGNATURES[name][1]))


def test_host_only_queries():
    lib = _lib.load()
    assert lib.fl_version() >= 100
    assert lib.fl_pattern_workspace_size(1000) > 0
    # workspace grows with elements; unsupported element sizes report 0
    assert lib.fl_df_workspace_size(216, 8, 125) > lib.fl_df_workspace_size(216, 1, 125) > 0
    assert lib.fl_df_workspace_size(8, 1, 200) == 0
    assert lib.fl_df_residual_workspace_size(11, 10, 10) > 0
    assert lib.fl_set_tuning(99, 1) != 0
    assert b"unknown key" in lib.fl_last_error()
    # the direct scatter: no element blocks in its workspace; its table has one int16 per
    # slot of the kernel's store order (p=4: 5 slices x 8 column tiles x 64 lanes x 8 rows)
    assert 0 < lib.fl_df_workspace_size_cs(216, 8, 125) < lib.fl_df_workspace_size(216, 8, 125)
    assert lib.fl_pattern_scatter_table_bytes(8, 125) == 8 * 5 * 8 * 64 * 8 * 2
    assert lib.fl_pattern_scatter_table_bytes(8, 27) == 8 * 3 * 2 * 64 * 4 * 2
    assert lib.fl_pattern_scatter_table_bytes(8, 10) == 0     # not a hex of p1^3 nodes
    # tile-shape variants of the dense stiffness per padded node-count class, as the
    # autotuner enumerates them
    assert [lib.fl_stiffness_variant_count(c) for c in range(6)] == \
        [_lib.STIFF_VARIANTS[c] for c in range(5)] + [0]
    # per-element material rows: 4 doubles, 8 for the anisotropic law (its fibre direction)
    assert [lib.fl_material_table_stride(m) for m in range(7)] == [0, 4, 4, 4, 4, 8, 0]
    # the rows the host puts in such a table (the batch groups) have the library's stride
    from florence_amd.materials import MATERIAL_IDS, LinearElastic, MooneyRivlin, NeoHookean, \
        NearlyIncompressibleMooneyRivlin
    for mat in (NeoHookean(3, mu=1.5, lamb=7.0), MooneyRivlin(3, mu1=0.5, mu2=0.25, lamb=9.0),
                LinearElastic(3, mu=1.0, lamb=2.0), NearlyIncompressibleMooneyRivlin(3, mu1=0.4, mu2=0.3, lamb=5.0)):
        row = mat.device_table_row()
        assert row.shape == (lib.fl_material_table_stride(MATERIAL_IDS[type(mat).__name__]),)
        assert np.array_equal(row, mat.device_parameters()[:4])
    # the residual pass without allocation refuses (on the host, before any device call)
    # meshes whose workspace exceeds its fixed array: 200k p=2 tets on the dense path
    prm, _keep = _lib.double_array(NeoHookean(3, mu=1.0, lamb=10.0).device_parameters())
    assert lib.fl_df_residual_workspace_size(11, 200000, 10) > lib_noalloc_max()
    rc = lib.fl_df_residual_noalloc(1, prm, None, None, None, None, None, None, None, 11, 200000, 10, 40000,
                                    None, None, None, None, None, None)
    assert rc == -3 and b"fixed maximum" in lib.fl_last_error()


def test_matrix_free_rejections():
    """fl_mf_apply refuses, on the host before any device call, records that are not the
    factorised path's C tensors and elements that are not tensor-product hexes."""
    from florence_amd.multigrid import FlMgMatrixFree
    lib = _lib.load()
    fake = 256   # never dereferenced: every case is refused before the first launch
    def op(records, npe, ngauss):
        return FlMgMatrixFree(fake, 1 << 30, records, fake, fake, 8, npe, ngauss, 125, fake, fake, fake, None)
    y = ctypes.c_void_p(fake)
    rc = lib.fl_mf_apply(ctypes.byref(op(_lib.FL_RECORDS_D, 125, 216)), y, y, y, None)
    assert rc == _lib.FL_E_ARG and b"C point records" in lib.fl_last_error()
    rc = lib.fl_mf_apply(ctypes.byref(op(_lib.FL_RECORDS_C, 10, 11)), y, y, y, None)
    assert rc == _lib.FL_E_ARG and b"tensor-product hex" in lib.fl_last_error()
    rc = lib.fl_mf_apply(None, y, y, y, None)
    assert rc == _lib.FL_E_ARG
    rc = lib.fl_mf_apply(ctypes.byref(op(_lib.FL_RECORDS_C, 125, 216)), None, y, y, None)
    assert rc == _lib.FL_E_ARG


def lib_noalloc_max():
    """FL_NOALLOC_SCRATCH_BYTES of include/florence_amd.h."""
    with open(os.path.join(ROOT, "include", "florence_amd.h")) as f:
        m = re.search(r"#define FL_NOALLOC_SCRATCH_BYTES \((\d+)u << (\d+)\)", f.read())
    return int(m.group(1)) << int(m.group(2))


def _colour(el, nnode, max_colours=64):
    lib = _lib.load()
    el = np.ascontiguousarray(el, dtype=np.int32)
    nelem, npe = el.shape
    colour = np.full(nelem, -1, dtype=np.int32)
    elems = np.full(nelem, -1, dtype=np.int32)
    offsets = (ctypes.c_int32 * (max_colours + 1))()
    nc = ctypes.c_int32(0)
    rc = lib.fl_pattern_colour(el.ctypes.data_as(ctypes.c_void_p), nelem, npe, nnode, max_colours,
                               colour.ctypes.data_as(ctypes.c_void_p), elems.ctypes.data_as(ctypes.c_void_p),
                               offsets, ctypes.byref(nc))
    return rc, int(nc.value), colour, elems, np.array(offsets[:nc.value + 1])


def _first_fit(el, nnode):
    """Greedy first-fit colouring in element order, by brute force over earlier
    elements (no two elements sharing a node get one colour)."""
    nodes = [set(map(int, r)) for r in el]
    colour = []
    for e, s in enumerate(nodes):
        taken = {colour[f] for f in range(e) if nodes[f] & s}
        c = 0
        while c in taken:
            c += 1
        colour.append(c)
    return np.array(colour, dtype=np.int32)


@pytest.mark.parametrize("et,p,n", [("hex", 1, 4), ("hex", 4, 3), ("tet", 2, 2)])
def test_colouring_is_greedy_first_fit(et, p, n):
    """fl_pattern_colour (host only): the first-fit colours of a brute-force greedy pass,
    no node shared inside a colour, elements grouped by colour in ascending order; a
    structured hex grid takes 8 colours (its 2 x 2 x 2 parity classes)."""
    m = cube_mesh(et, p, n, jitter=0.0, seed=0)
    el = np.asarray(m.elements)
    rc, nc, colour, elems, offsets = _colour(el, m.nnode)
    assert rc == 0
    assert np.array_equal(colour, _first_fit(el, m.nnode))
    assert nc == colour.max() + 1
    if et == "hex":
        assert nc == 8
    for c in range(nc):
        grp = elems[offsets[c]:offsets[c + 1]]
        assert np.all(colour[grp] == c) and np.all(np.diff(grp) > 0)
        nodes = el[grp].ravel()
        assert np.unique(nodes).size == nodes.size, c
    assert offsets[0] == 0 and offsets[-1] == el.shape[0]
    assert np.array_equal(np.sort(elems), np.arange(el.shape[0]))


def test_colouring_limits():
    m = cube_mesh("hex", 1, 3, jitter=0.0, seed=0)
    el = np.asarray(m.elements)
    rc, _, _, _, _ = _colour(el, m.nnode, max_colours=4)
    assert rc == _lib.FL_E_LIMIT
    assert b"more than 4 colours" in _lib.load().fl_last_error()
    bad = el.copy()
    bad[0, 0] = m.nnode
    rc, _, _, _, _ = _colour(bad, m.nnode)
    assert rc == _lib.FL_E_ARG
