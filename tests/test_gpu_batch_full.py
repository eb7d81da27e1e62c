"""Batching form at its stated size: 256 seeded meshes (unit cube or cylinder; tet
p1-4 and hex p1-3; element counts log-uniform in [200, 20000], 1,142,248 elements in
all; Neo-Hookean mu ~ U[0.5, 2], lamb ~ U[5, 50]) assembled by one captured HIP graph.

* every graph replay equals the eager batch, bitwise;
* every mesh's slice of the batch equals that mesh assembled alone on the same data
  path (bitwise), for all 256 meshes;
* every mesh of at most 2000 elements (119 of them) equals the FP64 oracle (the C
  restatement of _GlobalAssemblyDF_): CSR pattern bit-exact against the oracle's
  ComputeSparsityPattern restatement, K and T within 1e-12 relative (inf-norm), or
  within twice the oracle's own FP64 spread (the same oracle on the mesh translated by
  a quarter unit) where that is larger.
"""
import numpy as np
import pytest

from cases import oracle_with_spread, within_oracle

pytestmark = pytest.mark.gpu

KEYS = [("tet", 1), ("tet", 2), ("tet", 3), ("tet", 4), ("hex", 1), ("hex", 2), ("hex", 3)]
ORACLE_MAX_ELEMENTS = 2000


def _disp(X):
    return 0.02 * np.sin(np.pi * X[:, [1, 2, 0]]) + 0.01 * X


@pytest.fixture(scope="module")
def full_batch(cuda):
    import torch
    from florence_amd.batch import BatchedAssembly, random_batch

    batch = random_batch(256, seed=0)
    assert sum(m.nelem for m, _, _ in batch) == 1142248
    ba = BatchedAssembly(batch)
    ba.set_state(_disp)
    ba.assemble()
    torch.cuda.synchronize()
    eager = [(g["V"].clone(), g["T"].clone()) for g in ba.groups]
    ba.capture()
    for rep in range(2):
        for g in ba.groups:
            g["V"].zero_()
            g["T"].zero_()
        ba.replay()
        torch.cuda.synchronize()
        for g, (V0, T0) in zip(ba.groups, eager):
            assert torch.equal(g["V"], V0) and torch.equal(g["T"], T0), "replay {} differs".format(rep)
    del eager
    assert all(s[0] == 0 for s in ba.statuses())
    return batch, ba


@pytest.mark.parametrize("key", KEYS, ids=["{}{}".format(*k) for k in KEYS])
def test_full_batch_equals_each_mesh_alone(full_batch, key):
    import torch
    from florence_amd import _lib
    from florence_amd.assembly import DeviceMeshData, assemble_device
    from florence_amd.batch import _np_class

    batch, ba = full_batch
    lib = _lib.load()
    ids = [i for i, (_, k, _) in enumerate(batch) if k == key]
    assert ids
    try:
        for i in ids:
            m, _, mat = batch[i]
            g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
            dm = DeviceMeshData(m, g["fs"])
            lib.fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + _np_class(dm.npe), g["variant"] + 1)
            lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 2 if g["sum_factorised"] else 1)
            X = np.asarray(m.points)
            V, T, st = assemble_device(dm, mat, X + _disp(X), direct=g["direct"])
            torch.cuda.synchronize()
            ip, ix, data, Tb = ba.mesh_result(i)
            assert np.array_equal(ip, dm.indptr.cpu().numpy()) and np.array_equal(ix, dm.indices.cpu().numpy())
            assert np.array_equal(data, V.cpu().numpy()), (i, key, m.nelem)
            assert np.array_equal(Tb, T.cpu().numpy()), (i, key, m.nelem)
            assert int(st[0]) == 0
            del dm, V, T
    finally:
        lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 0)
        for npc in range(5):
            lib.fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + npc, 0)


@pytest.mark.parametrize("key", KEYS, ids=["{}{}".format(*k) for k in KEYS])
def test_full_batch_matches_oracle(full_batch, key):
    from oracle import df_numpy as O

    batch, ba = full_batch
    ids = [i for i, (m, k, _) in enumerate(batch) if k == key and m.nelem <= ORACLE_MAX_ELEMENTS]
    assert ids, key
    items, results = [], []
    for i in ids:
        m, _, mat = batch[i]
        g = next(gg for gg in ba.groups if i in gg["mesh_ids"])
        ip, ix, data, Tb = ba.mesh_result(i)
        ipr, ixr = O.sparsity_pattern(m.elements, m.nnode)
        assert np.array_equal(ip, ipr) and np.array_equal(ix, ixr), (i, key)
        X = np.asarray(m.points)
        items.append((m, g["fs"], mat, X + _disp(X), ipr, ixr))
        results.append((data, Tb))
    worst = [0.0, 0.0]
    for i, (Vr, Tr, dV, dT), (data, Tb) in zip(ids, oracle_with_spread(items), results):
        ok, ek, et, bk, bt = within_oracle(data, Tb, Vr, Tr, dV, dT)
        assert ok, (i, key, batch[i][0].nelem, ek, et, bk, bt)
        worst = [max(worst[0], ek), max(worst[1], et)]
    print("{}{}: {} meshes against the oracle, worst K {:.2e}, T {:.2e}".format(key[0], key[1], len(ids), *worst))
