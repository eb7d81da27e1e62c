"""Batched form: several meshes of mixed (type, p) assembled through one captured HIP
graph; each mesh's CSR block must equal assembling that mesh alone (same kernel
variant: bitwise), and every graph replay must equal the eager batch.  Each mesh of
the first test is also checked against the FP64 oracle: pattern bit-exact, K and T
within 1e-12 relative, or within twice the oracle's own FP64 spread where that is
larger (T of a small-strain state at p=3: the reference's parent gradients sum absolute
coordinates, and the same oracle on a translated copy of the mesh moves T by up to
1.8e-12 relative)."""
import numpy as np
import pytest

from cases import oracle_with_spread, within_oracle

pytestmark = pytest.mark.gpu


def test_batch_graph_matches_per_mesh(cuda):
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.batch import BatchedAssembly, random_batch, _np_class

    batch = random_batch(14, seed=3, nelem_range=(20, 400))
    ba = BatchedAssembly(batch)
    disp = lambda X: 0.02 * np.sin(np.pi * X[:, [1, 2, 0]]) + 0.01 * X
    ba.set_state(disp)
    ba.assemble()
    torch.cuda.synchronize()
    eager = [(g["V"].clone(), g["T"].clone()) for g in ba.groups]
    ba.capture()
    for rep in range(3):
        for g in ba.groups:
            g["V"].zero_()
            g["T"].zero_()
        ba.replay()
        torch.cuda.synchronize()
        for g, (V0, T0) in zip(ba.groups, eager):
            assert torch.equal(g["V"], V0) and torch.equal(g["T"], T0), "replay {} differs".format(rep)
    for i, (m, key, mat) in enumerate(batch):
        g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
        dm = DeviceMeshData(m, g["fs"])
        _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + _np_class(dm.npe), g["variant"] + 1)
        _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 2 if g.get("sum_factorised", True) else 1)
        X = np.asarray(m.points)
        V, T, st = assemble_device(dm, mat, X + disp(X))
        torch.cuda.synchronize()
        ip, ix, data, Tb = ba.mesh_result(i)
        assert np.array_equal(ip, dm.indptr.cpu().numpy())
        assert np.array_equal(ix, dm.indices.cpu().numpy())
        assert np.array_equal(data, V.cpu().numpy()), (i, key)
        assert np.array_equal(Tb, T.cpu().numpy()), (i, key)
    assert all(s[0] == 0 for s in ba.statuses())
    _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
    # and every mesh of the batch against the FP64 oracle
    from oracle import df_numpy as O
    items, got = [], []
    for i, (m, key, mat) in enumerate(batch):
        g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
        ip, ix, data, Tb = ba.mesh_result(i)
        ipr, ixr = O.sparsity_pattern(m.elements, m.nnode)
        assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), (i, key)
        X = np.asarray(m.points)
        items.append((m, g["fs"], mat, X + disp(X), ipr, ixr))
        got.append((data, Tb))
    for i, (Vr, Tr, dV, dT), (data, Tb) in zip(range(len(batch)), oracle_with_spread(items), got):
        ok, ek, et, bk, bt = within_oracle(data, Tb, Vr, Tr, dV, dT)
        assert ok, (i, batch[i][1], ek, et, bk, bt)


def test_tuned_batch_matches_per_mesh(cuda, tmp_path):
    """The autotuner's choice per group (dense tile variant, factorised element blocks +
    gather, or the direct scatter) is recorded in the cache, and each mesh of the tuned
    batch equals that mesh assembled alone on the chosen path, bitwise."""
    import torch
    from florence_amd import _lib, autotune
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.batch import BatchedAssembly, random_batch, _np_class

    batch = random_batch(10, seed=5, nelem_range=(20, 300))
    ba = BatchedAssembly(batch)
    disp = lambda X: 0.02 * np.sin(np.pi * X[:, [1, 2, 0]]) + 0.01 * X
    ba.set_state(disp)
    cache = autotune.tune_batch(ba, repeats=1, cache={})
    for g in ba.groups:
        entry = cache[ba.tuning_key(g["key"], g["fs"])]
        assert entry["direct"] == g["direct"] and entry["sum_factorised"] == g["sum_factorised"]
        assert len(entry["assembly_ms"]) == len(entry["candidates"])
        if g["dm"].scatter is not None:
            assert "sum-factorised, direct scatter" in entry["candidates"]
    ba.assemble()
    torch.cuda.synchronize()
    try:
        for i, (m, key, mat) in enumerate(batch):
            g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
            dm = DeviceMeshData(m, g["fs"])
            _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + _np_class(dm.npe), g["variant"] + 1)
            _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 2 if g["sum_factorised"] else 1)
            X = np.asarray(m.points)
            V, T, st = assemble_device(dm, mat, X + disp(X), direct=g["direct"])
            torch.cuda.synchronize()
            _, _, data, Tb = ba.mesh_result(i)
            assert np.array_equal(data, V.cpu().numpy()), (i, key, g["direct"])
            assert np.array_equal(Tb, T.cpu().numpy()), (i, key)
    finally:
        _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
        for npc in range(5):
            _lib.load().fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + npc, 0)
    assert all(s[0] == 0 for s in ba.statuses())
