"""Batched form: many independent meshes of mixed element type and order assembled as
one device program.

Meshes are grouped by (element type, p, quadrature points); each group is
concatenated into one node-disjoint mesh whose elements carry a per-element material
row (``fl_df_assemble_mt``), so every group is a single call of the ordinary assembly
kernels.  The calls of all groups are captured once into a HIP graph
(``torch.cuda.CUDAGraph`` on the library's stream) and the whole batch is then one
graph launch.  The stiffness-kernel variant of each group comes from the autotuner
(florence_amd/autotune.py).  Per-mesh CSR blocks are slices of the group's CSR: rows
[3 n0, 3 n1) of the concatenated pattern, columns shifted by 3 n0, values identical
to assembling the mesh alone with the same kernel variant (same per-element
arithmetic, same ascending-element summation order).

No reference counterpart: Florence assembles one mesh per call.
"""
import numpy as np

from . import _lib
from .assembly import DeviceMeshData
from .function_space import FunctionSpace
from .materials import MATERIAL_IDS, NeoHookean
from .mesh import Mesh
from .quadrature import QuadratureRule

__all__ = ["random_batch", "BatchedAssembly", "default_rule"]

TYPES = [("tet", 1), ("tet", 2), ("tet", 3), ("tet", 4), ("hex", 1), ("hex", 2), ("hex", 3)]


def default_rule(element_type, p):
    """The reference's default volume rule (VariationalPrinciple.py:53-72)."""
    C = p - 1
    if element_type == "tet":
        return QuadratureRule(norder=2 * C if C > 0 else 1, mesh_type="tet")
    return QuadratureRule(norder=C + 2, mesh_type="hex")


def random_batch(nmesh=256, seed=0, nelem_range=(200, 20000), mu_range=(0.5, 2.0), lamb_range=(5.0, 50.0)):
    """``nmesh`` seeded meshes: unit cube or cylinder (radius 0.5, length 2), element
    (type, p) uniform over tet p1-4 and hex p1-3, element count log-uniform in
    ``nelem_range`` (rounded to the structured generator), Neo-Hookean with mu, lamb
    uniform in the given ranges.  Returns a list of (mesh, (type, p), material)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nmesh):
        et, p = TYPES[int(rng.integers(len(TYPES)))]
        target = float(np.exp(rng.uniform(np.log(nelem_range[0]), np.log(nelem_range[1]))))
        shape = "cube" if rng.uniform() < 0.5 else "cylinder"
        mu, lamb = rng.uniform(*mu_range), rng.uniform(*lamb_range)
        tf = 6.0 if et == "tet" else 1.0
        m = Mesh()
        if shape == "cube":
            n = max(1, int(round((target / tf) ** (1.0 / 3.0))))
            m.Cube(n=n, element_type=et)
        else:
            nrad = max(1, int(round((target / (16.0 * tf)) ** (1.0 / 3.0))))
            m.Cylinder(radius=0.5, length=2.0, nlong=4 * nrad, nrad=nrad, element_type=et)
        m.GetHighOrderMesh(p=p)
        out.append((m, (et, p), NeoHookean(3, mu=float(mu), lamb=float(lamb))))
    return out


class _Concat(object):
    """A node-disjoint union of meshes of one element type/order (duck-types Mesh)."""

    def __init__(self, meshes):
        self.element_type = meshes[0].element_type
        offs = np.cumsum([0] + [m.points.shape[0] for m in meshes])
        eoffs = np.cumsum([0] + [m.elements.shape[0] for m in meshes])
        self.points = np.concatenate([np.asarray(m.points) for m in meshes])
        self.elements = np.concatenate([np.asarray(m.elements, dtype=np.int64) + offs[i]
                                        for i, m in enumerate(meshes)])
        self.node_offsets = offs
        self.elem_offsets = eoffs
        self.nnode = self.points.shape[0]
        self.nelem = self.elements.shape[0]




class BatchedAssembly(object):
    """Assemble a list of (mesh, (type, p), material) as one device program.

    Meshes of one (type, p) are concatenated into groups capped so that a group's CSR
    stays inside int32 indices (sum of element-block entries <= max_block_entries) and
    its element-block workspace stays small.  The groups are spread over ``branches``
    streams (longest first onto the least-loaded branch), each with its own workspace,
    so independent groups run concurrently: in the captured graph the branches fork from
    the capture stream and join it again, and the batch stays one graph launch.
    """

    def __init__(self, batch, precision="fp64", tuning=None, device="cuda", max_block_entries=1.5e9,
                 direct=True, branches=4):
        import torch
        self.precision = precision
        self.direct = direct
        self.device = torch.device(device)
        self.batch = batch
        keys = {}
        for i, (m, (et, p), mat) in enumerate(batch):
            keys.setdefault((et, p), []).append(i)
        self.groups = []
        for (et, p), allidx in sorted(keys.items()):
            rule = default_rule(et, p)
            probe = Mesh()
            m0 = batch[allidx[0]][0]
            probe.element_type, probe.elements, probe.points = et, m0.elements, m0.points
            fs = FunctionSpace(probe, rule, p=p)
            for idx in self._split(allidx, max_block_entries):
                self.groups.append(self._make_group((et, p), idx, fs, tuning))
        lib = _lib.load()
        nb = max(1, min(int(branches), len(self.groups)))
        load = [0.0] * nb
        for g in sorted(self.groups, key=lambda g: -float(g["dm"].nelem) * g["dm"].npe ** 2):
            b = int(np.argmin(load))
            g["branch"] = b
            load[b] += float(g["dm"].nelem) * g["dm"].npe ** 2
        self.workspaces = []
        for b in range(nb):
            need = max(int(lib.fl_df_workspace_size(g["dm"].ngauss, g["dm"].nelem, g["dm"].npe))
                       for g in self.groups if g["branch"] == b)
            self.workspaces.append(torch.empty(need, dtype=torch.uint8, device=self.device))
        self.workspace = self.workspaces[0]
        self.streams = [torch.cuda.Stream(device=self.device) for _ in range(nb)]   # default priority
        self.graph = None

    def _split(self, idx, cap):
        out, cur, acc = [], [], 0.0
        for i in idx:
            m = self.batch[i][0]
            w = float(m.nelem) * (3 * m.elements.shape[1]) ** 2
            if cur and acc + w > cap:
                out.append(cur)
                cur, acc = [], 0.0
            cur.append(i)
            acc += w
        if cur:
            out.append(cur)
        return out

    def _make_group(self, key, idx, fs, tuning):
        import torch
        batch = self.batch
        cm = _Concat([batch[i][0] for i in idx])
        dm = DeviceMeshData(cm, fs, device=self.device, build_colours=self.direct)
        elem_mat = np.repeat(np.arange(len(idx), dtype=np.int32), np.diff(cm.elem_offsets))
        mat_ids = {MATERIAL_IDS[type(batch[i][2]).__name__] for i in idx}
        if len(mat_ids) != 1:
            raise ValueError("one material law per (type, p) group")
        stride = int(_lib.load().fl_material_table_stride(next(iter(mat_ids))))
        table = np.array([batch[i][2].device_table_row() for i in idx], dtype=np.float64)
        if table.shape[1] != stride:
            raise ValueError("material table rows of {} values, the library reads {}".format(table.shape[1], stride))
        tkey = self.tuning_key(key, fs)
        return {"key": key, "mesh_ids": idx, "concat": cm, "fs": fs, "dm": dm,
                "elem_mat": torch.as_tensor(elem_mat, device=self.device),
                "table": torch.as_tensor(table, device=self.device), "material": mat_ids.pop(),
                "params": batch[idx[0]][2].device_parameters(),
                "x": torch.as_tensor(np.ascontiguousarray(cm.points), device=self.device).reshape(-1).clone(),
                "V": torch.zeros(dm.nnz, dtype=torch.float64, device=self.device),
                "T": torch.zeros(3 * dm.nnode, dtype=torch.float64, device=self.device),
                "direct": self.direct and bool(tuning[tkey].get("direct", True)) if tuning and tkey in tuning
                else self.direct,
                "variant": int(tuning[tkey]["variant"]) if tuning and tkey in tuning else 0,
                "sum_factorised": bool(tuning[tkey].get("sum_factorised", True)) if tuning and tkey in tuning
                else True}

    def tuning_key(self, key, fs):
        et, p = key
        return "{}-p{}-G{}-{}".format(et, p, fs.AllGauss.shape[0], self.precision)

    @property
    def nelem(self):
        return sum(g["dm"].nelem for g in self.groups)

    @property
    def nnz(self):
        return sum(g["dm"].nnz for g in self.groups)

    def set_state(self, displacement_fn):
        """Set every group's current coordinates to X + displacement_fn(X)."""
        import torch
        for g in self.groups:
            X = g["concat"].points
            g["x"].copy_(torch.as_tensor(np.ascontiguousarray(X + displacement_fn(X)), device=self.device).reshape(-1))

    def _launch_group(self, g):
        lib = _lib.load()
        dm = g["dm"]
        npc = _np_class(dm.npe)
        _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_VARIANT_BASE + npc, int(g["variant"]) + 1),
                   "fl_set_tuning")
        _lib.check(lib.fl_set_tuning(_lib.FL_TUNE_STIFF_SF, 2 if g.get("sum_factorised", True) else 1),
                   "fl_set_tuning")
        prm, _keep = _lib.double_array(g["params"])
        ws = self.workspaces[g.get("branch", 0)]
        args = (g["material"], prm, _lib.ptr(g["elem_mat"]), _lib.ptr(g["table"]), _lib.ptr(dm.points),
                _lib.ptr(dm.elements), _lib.ptr(g["x"]), _lib.ptr(dm.Jm), _lib.ptr(dm.AllGauss), dm.ngauss,
                dm.nelem, dm.npe, dm.nnode, 1, _lib.ptr(dm.node_elem_ptr), _lib.ptr(dm.node_elem),
                _lib.ptr(dm.node_nbr_ptr), _lib.ptr(dm.node_nbrs), _lib.ptr(dm.node_order), dm.bucket_offsets,
                dm.bucket_degree, _lib.ptr(dm.indptr), _lib.ptr(g["V"]), _lib.ptr(g["T"]),
                _lib.PRECISIONS[self.precision], _lib.ptr(dm.ranks), _lib.ptr(dm.tensor_basis))
        tail = (_lib.ptr(ws), ws.numel(), _lib.ptr(dm.status), _lib.stream_ptr())
        if g.get("direct", self.direct) and dm.scatter is not None:
            # the direct scatter where it applies (FP64 factorised hexes), else the same
            # call as fl_df_assemble_mt
            _lib.check(lib.fl_df_assemble_cs(*(args + (dm.ncolour, dm.colour_offsets, _lib.ptr(dm.colour_elems),
                                                       _lib.ptr(dm.scatter)) + tail)), "fl_df_assemble_cs")
        else:
            _lib.check(lib.fl_df_assemble_mt(*(args + tail)), "fl_df_assemble_mt")

    def assemble(self):
        """All groups (one library call per group), each on its branch's stream; the
        branches start after the current stream's earlier work and the current stream
        waits for all of them."""
        import torch
        cur = torch.cuda.current_stream()
        for st in self.streams:
            st.wait_stream(cur)
        for g in self.groups:
            with torch.cuda.stream(self.streams[g["branch"]]):
                self._launch_group(g)
        for st in self.streams:
            cur.wait_stream(st)

    def capture(self):
        """Record the batch into one HIP graph (call assemble() once before, so every
        kernel attribute is set outside the capture)."""
        import torch
        self.graph = torch.cuda.CUDAGraph()
        s = torch.cuda.Stream()
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            with torch.cuda.graph(self.graph, stream=s):
                self.assemble()
        torch.cuda.current_stream().wait_stream(s)
        return self.graph

    def replay(self):
        self.graph.replay()

    def mesh_result(self, i):
        """(indptr, indices, data, T) of mesh i in its own numbering, as host arrays."""
        for g in self.groups:
            if i in g["mesh_ids"]:
                k = g["mesh_ids"].index(i)
                cm, dm = g["concat"], g["dm"]
                n0, n1 = cm.node_offsets[k], cm.node_offsets[k + 1]
                ip = dm.indptr[3 * n0:3 * n1 + 1].cpu().numpy().astype(np.int64)
                lo, hi = int(ip[0]), int(ip[-1])
                indices = dm.indices[lo:hi].cpu().numpy() - 3 * n0
                data = g["V"][lo:hi].cpu().numpy()
                T = g["T"][3 * n0:3 * n1].cpu().numpy()
                return (ip - lo).astype(np.int32), indices.astype(np.int32), data, T
        raise KeyError(i)

    def statuses(self):
        return [g["dm"].status.cpu().numpy().tolist() for g in self.groups]


def _np_class(npe):
    for c, n in enumerate((16, 32, 48, 64, 128)):
        if npe <= n:
            return c
    raise ValueError(npe)
