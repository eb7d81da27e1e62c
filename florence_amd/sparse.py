"""Device CSR matrix returned by ``Assemble``.

Holds the reference's CSR arrays (int32 ``indptr``/``indices``, float64 ``data``) as
torch tensors in HBM.  ``to_scipy()`` copies them into the ``scipy.sparse.csr_matrix``
the reference returns (Florence/FiniteElements/Assembly/_LowLevelAssembly_.py:74-78).
"""
import numpy as np

__all__ = ["DeviceCSR"]


class DeviceCSR(object):

    def __init__(self, indptr, indices, data, shape):
        self.indptr = indptr
        self.indices = indices
        self.data = data
        self.shape = tuple(shape)
        self.dtype = np.float64

    @property
    def nnz(self):
        return int(self.data.numel())

    def to_scipy(self):
        import scipy.sparse as sp
        if self.indices is None:
            raise ValueError("pattern was built without column indices")
        return sp.csr_matrix((self.data.cpu().numpy(), self.indices.cpu().numpy(), self.indptr.cpu().numpy()),
                             shape=self.shape)

    def __repr__(self):
        return "DeviceCSR(shape={}, nnz={}, device={})".format(self.shape, self.nnz, self.data.device)
