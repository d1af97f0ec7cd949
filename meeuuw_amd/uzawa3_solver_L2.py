"""Drop-in uzawa3_solver_L2 (uzawa3_solver_L2.py:21-99) and cd_u_eta
(cd_u_eta.py:13-99) on the GPU.

Same arguments (SciPy sparse matrices and NumPy right-hand sides) and return
tuples as the reference.  The SuperLU factorisations of K, M_p and M_eta are
replaced by Jacobi-preconditioned CG to a relative residual ``inner_rtol``
(default 1e-13), so the outer iterates follow the reference's.
"""
import ctypes

import numpy as np
import scipy.sparse as sparse

from . import _dev as dev
from . import _lib


def _csr(A):
    A = sparse.csr_matrix(A)
    A.sort_indices()
    return (dev.to_dev(A.indptr.astype(np.int32), np.int32),
            dev.to_dev(A.indices.astype(np.int32), np.int32),
            dev.to_dev(A.data, np.float64))


def _ptrs(triplet):
    return [dev.ptr(t) for t in triplet]


def uzawa3_solver_L2(K_mat, G_mat, GT_mat, MP_mat, H_mat, f_rhs, h_rhs, Nfem_P,
                     inner_rtol=1e-13):
    """-> (solV, solP, niter, array_xiV, array_xiP, array_alpha)"""
    nV, nP = K_mat.shape[0], int(Nfem_P)
    mats = [_csr(M) for M in (K_mat, G_mat, GT_mat, MP_mat, H_mat)]
    f, h = dev.to_dev(f_rhs, np.float64), dev.to_dev(h_rhs, np.float64)
    solV, solP = dev.empty(nV), dev.empty(nP)
    niter = ctypes.c_int32(0)
    xiV, xiP, al = np.zeros(100), np.zeros(100), np.zeros(100)
    args = [p for m in mats for p in _ptrs(m)]
    _lib.call("mw_uzawa3_solver_L2", nV, nP, *args, dev.ptr(f), dev.ptr(h), dev.ptr(solV),
              dev.ptr(solP), float(inner_rtol), ctypes.addressof(niter), xiV.ctypes.data,
              xiP.ctypes.data, al.ctypes.data, dev.stream())
    return dev.host(solV), dev.host(solP), int(niter.value), xiV, xiP, al


def cd_u_eta(K, G, GT, M, M_eta, H, f, h, Nfem_P, inner_rtol=1e-13):
    """-> (solV, solP, niter, array_xiV, array_xiP, array_alpha, array_beta)"""
    nV, nP = K.shape[0], int(Nfem_P)
    mats = [_csr(A) for A in (K, G, GT, M, M_eta, H)]
    fd, hd = dev.to_dev(f, np.float64), dev.to_dev(h, np.float64)
    solV, solP = dev.empty(nV), dev.empty(nP)
    niter = ctypes.c_int32(0)
    xiV, xiP, al, be = np.zeros(100), np.zeros(100), np.zeros(100), np.zeros(100)
    args = [p for m in mats for p in _ptrs(m)]
    _lib.call("mw_cd_u_eta", nV, nP, *args, dev.ptr(fd), dev.ptr(hd), dev.ptr(solV),
              dev.ptr(solP), float(inner_rtol), ctypes.addressof(niter), xiV.ctypes.data,
              xiP.ctypes.data, al.ctypes.data, be.ctypes.data, dev.stream())
    return dev.host(solV), dev.host(solP), int(niter.value), xiV, xiP, al, be
