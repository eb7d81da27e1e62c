"""High-order mesh generation with the reference's node numbering, restated in numpy --
TEST INFRASTRUCTURE (see oracle/__init__.py).

Follows HighOrderMeshHex (Florence/MeshGeneration/HigherOrderMeshing/
HigherOrderMeshingHex.py:19-120) and HighOrderMeshTet_SEMISTABLE
(HigherOrderMeshingTet.py:19-139), called by Mesh.GetHighOrderMesh (Mesh.py:1339-1409)
with their defaults (Decimals=10, check_duplicates=True):

1. the linear basis at the p-nodes, values below 1e-12 set to 0 (Neval, :30-46);
2. every element's non-vertex nodes at Neval^T x_vertices (GetInteriorCoordinates.py:3-13),
   appended element by element after the linear nodes (:72-90);
3. the new nodes rounded (makezero at 1e-14, then np.round to Decimals) and merged by
   unique2d(order=False) (Tensor/Tensors.py:15-74): numpy's unique of the rows viewed as
   raw bytes, i.e. sorted by the bytes of the rounded doubles, first occurrence kept;
   the linear nodes stay first, the merged new nodes follow in that byte order (:96-115);
4. check_duplicates: remove_duplicates_2D(points, decimals=Decimals-2)
   (Tensors.py:428-441) on ALL nodes -- makezero in place, round, unique2d again -- which
   renumbers every node, vertices included, by the byte order of its coordinates rounded
   to 8 decimals (:122-133).

The ordering facts it relies on (numpy's byte-view unique sorts by memcmp of the
little-endian doubles; np.round(x, d) = rint(x 10^d) / 10^d) are checked by
tests/test_oracle_mesh.py, with the properties that pin the result.
"""
import numpy as np


def makezero(A, tol=1e-14):
    """Tensor/Numeric.pyx:107-120: entries with |a| < tol set to 0, in place."""
    A[np.abs(A) < tol] = 0.0
    return A


def unique2d_rows(arr):
    """unique2d(arr, order=False, consider_sort=False, return_index=True,
    return_inverse=True) of Tensor/Tensors.py:15-74."""
    b = np.ascontiguousarray(arr).view(np.dtype((np.void, arr.dtype.itemsize * arr.shape[1])))
    _, idx, inv = np.unique(b, return_index=True, return_inverse=True)
    return arr[idx], idx, inv.ravel()


def high_order_mesh(points, elements, element_type, C, Decimals=10):
    """(points, elements) of the degree C+1 mesh built from a linear one."""
    from florence_amd.bases import FeketePointsTet, GaussLobattoPointsHex, LinearHexBases, LinearTetBases
    points = np.asarray(points, dtype=np.float64)
    elements = np.asarray(elements, dtype=np.int64)
    if element_type == "hex":
        eps, nv = GaussLobattoPointsHex(C), 8
        Neval = LinearHexBases(eps)
    elif element_type == "tet":
        eps, nv = FeketePointsTet(C), 4
        Neval = LinearTetBases(eps)
    else:
        raise NotImplementedError(element_type)
    Neval = makezero(np.array(Neval, dtype=np.float64), tol=1e-12)
    npe = eps.shape[0]
    nelem, nlin = elements.shape[0], points.shape[0]
    iesize = npe - nv
    repoints = np.zeros((nlin + iesize * nelem, 3))
    repoints[:nlin] = points
    for e in range(nelem):
        repoints[nlin + e * iesize:nlin + (e + 1) * iesize] = np.dot(Neval[:, nv:].T, points[elements[e]])
    reelements = np.empty((nelem, npe), dtype=np.int64)
    reelements[:, :nv] = elements
    reelements[:, nv:] = nlin + np.arange(nelem * iesize).reshape(nelem, iesize)
    # step 3: merge the new nodes
    rounded = makezero(repoints[nlin:].copy())
    rounded = np.round(rounded, decimals=Decimals)
    _, idx, inv = unique2d_rows(rounded)
    repoints = repoints[np.concatenate((np.arange(nlin), idx + nlin))]
    # HigherOrderMeshingHex.py:111-115: the new ids were consecutive (element-major), so
    # np.unique of them is the identity and every new node takes nlin + its merged row
    reelements[:, nv:] = (nlin + inv).reshape(nelem, iesize)
    # step 4: remove duplicates over all nodes at Decimals - 2
    deci = int(Decimals) - 2 if Decimals >= 6 else Decimals
    makezero(repoints)
    _, idx, inv = unique2d_rows(np.round(repoints, decimals=deci))
    repoints = repoints[idx]
    reelements = inv[reelements]
    return repoints, reelements
