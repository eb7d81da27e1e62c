"""Mesh: the connectivity/coordinate contract the assembler consumes.

Keeps the reference's mesh contract (Florence/MeshGeneration/Mesh.py):
``points`` float64 C-contiguous (nnode, 3), ``elements`` uint64 C-contiguous
(nelem, npe) after ``ChangeType`` (Mesh.py:7684-7697), ``element_type`` in
{"tet", "hex"}.

Generators restated for the workloads of this path:

* ``Parallelepiped`` / ``Cube`` -- structured linear hex mesh with the reference's
  node numbering (x fastest, then y, then z) and element numbering and vertex order
  (Mesh.py:5843-5920);
* ``ConvertHexesToTets`` -- the 6-tet split of each linear hex sharing the 0-7
  diagonal, all first tets first (Mesh.py:8757-8802, NodeArrangement.py:696-726);
* ``GetHighOrderMesh`` -- p-refinement placing GLL nodes in hexes and Fekete nodes in
  tets through the linear map of each element (HigherOrderMeshingHex.py:19-133,
  HigherOrderMeshingTet.py:19-133).  ``numbering="reference"`` runs the device
  generator (libflorence_amd fl_mesh_refine), which numbers the nodes as the
  reference does (rank of the rounded coordinates in numpy's raw-byte order).  The
  default ``numbering="first"`` (host) merges nodes on the same rounded coordinates
  but keeps the vertices' linear numbers and numbers new nodes in order of first
  appearance (element by element, local node order): the same mesh up to a node
  permutation, with element-coherent node ids (the reference's order scatters an
  element's nodes over the whole id range, which costs the CSR rows their locality).
* ``Cylinder`` -- a structured hex/tet discretisation of a cylinder along x used by
  the batched workloads.
"""
import numpy as np

from .bases import GaussLobattoPointsHex, FeketePointsTet, LinearHexBases, LinearTetBases

__all__ = ["Mesh"]

# 6-tet split of a hexahedron around its 0-7 diagonal (NodeArrangement.py:696-726, C=0)
_HEX_TO_TET = np.array([[0, 1, 2, 7], [0, 2, 3, 7], [0, 1, 5, 7], [0, 5, 4, 7], [1, 6, 2, 7], [1, 5, 6, 7]],
                       dtype=np.int64)


class Mesh(object):

    def __init__(self):
        self.points = None
        self.elements = None
        self.element_type = None
        self.nelem = 0
        self.nnode = 0
        self.degree = None

    # ------------------------------------------------------------------ queries
    def InferPolynomialDegree(self):
        npe = self.elements.shape[1]
        if self.element_type == "hex":
            p = int(round(npe ** (1.0 / 3.0))) - 1
            if (p + 1) ** 3 != npe:
                raise ValueError("not a Lagrange hexahedron: {} nodes".format(npe))
        elif self.element_type == "tet":
            p = None
            for q in range(1, 20):
                if (q + 1) * (q + 2) * (q + 3) // 6 == npe:
                    p = q
                    break
            if p is None:
                raise ValueError("not a Lagrange tetrahedron: {} nodes".format(npe))
        else:
            raise NotImplementedError(self.element_type)
        self.degree = p
        return p

    def InferNumberOfNodesPerElement(self, p=None, element_type=None):
        return self.elements.shape[1]

    def InferSpatialDimension(self):
        return 3

    def ChangeType(self):
        """elements -> uint64, points -> contiguous float64 (Mesh.py:7684-7697)."""
        self.elements = np.ascontiguousarray(self.elements, dtype=np.uint64)
        self.points = np.ascontiguousarray(self.points, dtype=np.float64)
        self.nelem = self.elements.shape[0]
        self.nnode = self.points.shape[0]

    # --------------------------------------------------------------- generators
    def Parallelepiped(self, lower_left_rear_point=(0, 0, 0), upper_right_front_point=(2, 4, 10),
                       nx=2, ny=4, nz=10, element_type="hex"):
        if element_type not in ("hex", "tet"):
            raise ValueError("Can only generate parallelepiped mesh using tetrahedrals or hexahedrals")
        lo, hi = np.asarray(lower_left_rear_point, float), np.asarray(upper_right_front_point, float)
        if np.any(lo > hi):
            raise ValueError("Incorrect coordinate for lower left rear and upper right front vertices")
        nx, ny, nz = int(nx), int(ny), int(nz)
        if nx <= 0 or ny <= 0 or nz <= 0:
            raise ValueError("Number of discretisation cannot be zero or negative")
        x = np.linspace(lo[0], hi[0], nx + 1)
        y = np.linspace(lo[1], hi[1], ny + 1)
        z = np.linspace(lo[2], hi[2], nz + 1)
        Z, Y, X = np.meshgrid(z, y, x, indexing="ij")
        self.points = np.stack([X.ravel(), Y.ravel(), Z.ravel()], axis=1)
        ez, ey, ex = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        n0 = (ez * (ny + 1) + ey) * (nx + 1) + ex
        n0 = n0.ravel()
        layer = (nx + 1) * (ny + 1)
        base = np.stack([n0, n0 + 1, n0 + nx + 2, n0 + nx + 1], axis=1)
        self.elements = np.concatenate([base, base + layer], axis=1).astype(np.int64)
        self.element_type = "hex"
        self.degree = 1
        self.nelem = self.elements.shape[0]
        self.nnode = self.points.shape[0]
        if element_type == "tet":
            self.ConvertHexesToTets()

    def Cube(self, lower_left_rear_point=(0., 0., 0.), side_length=1, nx=5, ny=5, nz=5, n=None, element_type="hex"):
        if n is not None:
            nx = ny = nz = n
        lo = np.asarray(lower_left_rear_point, float)
        self.Parallelepiped(lo, lo + side_length, nx, ny, nz, element_type)

    def Cylinder(self, center=(0., 0., 0.), radius=1., length=2., nlong=4, nrad=2, element_type="hex"):
        """Cylinder along +x: a square core blended to a circular rim (O-grid free
        mapping of a (2*nrad)^2 x nlong block), linear elements."""
        m = 2 * nrad
        self.Parallelepiped((-1, -1, 0), (1, 1, 1), m, m, nlong, "hex")
        s, t, u = self.points[:, 0], self.points[:, 1], self.points[:, 2]
        # square -> disk (Shirley-Chiu style elliptical map), radius scaled
        ys = s * np.sqrt(np.maximum(0.0, 1.0 - 0.5 * t * t))
        zs = t * np.sqrt(np.maximum(0.0, 1.0 - 0.5 * s * s))
        c = np.asarray(center, float)
        self.points = np.stack([c[0] + length * u, c[1] + radius * ys, c[2] + radius * zs], axis=1)
        if element_type == "tet":
            self.ConvertHexesToTets()

    def ConvertHexesToTets(self):
        if self.element_type == "tet":
            return
        if self.element_type != "hex" or self.elements.shape[1] != 8:
            raise ValueError("only linear hexahedral meshes are split into tetrahedra")
        el = np.asarray(self.elements, dtype=np.int64)
        self.elements = np.concatenate([el[:, _HEX_TO_TET[k]] for k in range(6)], axis=0)
        self.element_type = "tet"
        self.degree = 1
        self.nelem = self.elements.shape[0]

    def GetHighOrderMesh(self, p=1, silent=True, Decimals=10, numbering="first", **kwargs):
        if not isinstance(p, int) or p < 1:
            raise ValueError("Value of p={} is not acceptable. Provide p>=1.".format(p))
        if numbering not in ("first", "reference"):
            raise ValueError("numbering is 'first' or 'reference'")
        if self.degree is None:
            self.InferPolynomialDegree()
        if self.degree == p:
            return
        if self.degree != 1:
            raise NotImplementedError("p-refinement starts from a linear mesh")
        C = p - 1
        el = np.asarray(self.elements, dtype=np.int64)
        X = np.asarray(self.points, dtype=np.float64)
        Neval, nv = self._linear_bases_at_nodes(C)
        npe = Neval.shape[1]
        if numbering == "reference":
            from .meshgen import refine_on_device
            self.points, self.elements = refine_on_device(X, el, Neval, Decimals)
            self.degree = p
            self.ChangeType()
            return
        nelem = el.shape[0]
        # physical coordinates of the non-vertex nodes of every element
        vert = X[el]                                               # (nelem, nv, 3)
        newp = np.einsum("va,evd->ead", Neval[:, nv:], vert)       # (nelem, npe-nv, 3)
        flat = newp.reshape(-1, 3)
        key = np.round(flat, decimals=Decimals)
        key[np.abs(key) < 1e-14] = 0.0
        _, first, inverse = np.unique(key, axis=0, return_index=True, return_inverse=True)
        inverse = inverse.ravel()
        # number unique nodes by first appearance
        order = np.argsort(first, kind="stable")
        rank = np.empty_like(order)
        rank[order] = np.arange(order.shape[0])
        nnode_lin = X.shape[0]
        new_ids = nnode_lin + rank[inverse]
        elements = np.empty((nelem, npe), dtype=np.int64)
        elements[:, :nv] = el
        elements[:, nv:] = new_ids.reshape(nelem, npe - nv)
        points = np.empty((nnode_lin + order.shape[0], 3))
        points[:nnode_lin] = X
        points[nnode_lin:] = flat[first[order]]
        self.points, self.elements = points, elements
        self.degree = p
        self.ChangeType()

    def _linear_bases_at_nodes(self, C):
        """Neval (nv, npe): the linear basis at the degree C+1 nodes, values below 1e-12
        set to 0 (HigherOrderMeshingHex.py:30-46, HigherOrderMeshingTet.py:32-49)."""
        if self.element_type == "hex":
            Neval, nv = LinearHexBases(GaussLobattoPointsHex(C)), 8
        elif self.element_type == "tet":
            Neval, nv = LinearTetBases(FeketePointsTet(C)), 4
        else:
            raise NotImplementedError(self.element_type)
        Neval = np.array(Neval, dtype=np.float64)
        Neval[np.abs(Neval) < 1e-12] = 0.0
        return Neval, nv

    # ------------------------------------------------------------------ helpers
    def JitterInteriorNodes(self, amplitude, seed=0, lower=None, upper=None, tol=1e-12):
        """Add U(-amplitude, amplitude)^3 noise (numpy PCG64(seed)) to every node strictly
        inside the box [lower, upper] (default: the mesh bounding box).  Noise is drawn
        for all nodes in node order so the result depends only on (mesh, seed)."""
        lo = self.points.min(axis=0) if lower is None else np.asarray(lower, float)
        hi = self.points.max(axis=0) if upper is None else np.asarray(upper, float)
        rng = np.random.Generator(np.random.PCG64(seed))
        noise = rng.uniform(-amplitude, amplitude, size=self.points.shape)
        inside = np.all((self.points > lo + tol) & (self.points < hi - tol), axis=1)
        self.points = np.ascontiguousarray(self.points + noise * inside[:, None])
        return inside
