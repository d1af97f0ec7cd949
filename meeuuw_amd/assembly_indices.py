"""COO index arrays of the assembled matrices (meeuuw.py:874-1015), on the GPU.

The reference fills these with pure-Python triple loops once before the time
loop; here each entry is one GPU thread.  Orders and values are identical.
"""
import numpy as np

from . import _dev as dev
from . import _lib


def local_to_globalV(icon_V):
    """meeuuw.py:874-881 -> int32 (18, nel)."""
    numpy_out = not dev.is_dev(icon_V)
    ic = dev.to_dev(icon_V, np.int32)
    nel = ic.shape[1]
    l2g = dev.empty(18 * nel, np.int32)
    _lib.call("mw_local_to_global_V", nel, dev.ptr(ic), dev.ptr(l2g), dev.stream())
    return dev.out(l2g, numpy_out, (18, nel))


def fill_II_JJ_Stokes(local_to_globalV, icon_P, Nfem_V):
    """meeuuw.py:885-912 -> (II_Stokes, JJ_Stokes), nel*468 entries each."""
    numpy_out = not dev.is_dev(local_to_globalV, icon_P)
    l2g = dev.to_dev(local_to_globalV, np.int32)
    ip = dev.to_dev(icon_P, np.int32)
    nel = l2g.shape[1]
    II, JJ = dev.empty(468 * nel, np.int32), dev.empty(468 * nel, np.int32)
    _lib.call("mw_coo_stokes", nel, dev.ptr(l2g), dev.ptr(ip), int(Nfem_V), dev.ptr(II),
              dev.ptr(JJ), dev.stream())
    return dev.out(II, numpy_out), dev.out(JJ, numpy_out)


def fill_II_JJ_K(local_to_globalV):
    """meeuuw.py:916-935 -> (II_K, JJ_K), nel*324 entries each."""
    numpy_out = not dev.is_dev(local_to_globalV)
    l2g = dev.to_dev(local_to_globalV, np.int32)
    nel = l2g.shape[1]
    II, JJ = dev.empty(324 * nel, np.int32), dev.empty(324 * nel, np.int32)
    _lib.call("mw_coo_K", nel, dev.ptr(l2g), dev.ptr(II), dev.ptr(JJ), dev.stream())
    return dev.out(II, numpy_out), dev.out(JJ, numpy_out)


def fill_II_JJ_G(local_to_globalV, icon_P):
    """meeuuw.py:939-970 -> (II_G, JJ_G, II_GT, JJ_GT), nel*72 entries each."""
    numpy_out = not dev.is_dev(local_to_globalV, icon_P)
    l2g = dev.to_dev(local_to_globalV, np.int32)
    ip = dev.to_dev(icon_P, np.int32)
    nel = l2g.shape[1]
    a = [dev.empty(72 * nel, np.int32) for _ in range(4)]
    _lib.call("mw_coo_G", nel, dev.ptr(l2g), dev.ptr(ip), *[dev.ptr(t) for t in a], dev.stream())
    return tuple(dev.out(t, numpy_out) for t in a)


def fill_II_JJ_scalar(icon):
    """II_MP/JJ_MP (icon_P, meeuuw.py:974-993) or II_T/JJ_T (icon_T, meeuuw.py:997-1015)."""
    numpy_out = not dev.is_dev(icon)
    ic = dev.to_dev(icon, np.int32)
    m, nel = ic.shape
    II, JJ = dev.empty(m * m * nel, np.int32), dev.empty(m * m * nel, np.int32)
    _lib.call("mw_coo_scalar", nel, m, dev.ptr(ic), dev.ptr(II), dev.ptr(JJ), dev.stream())
    return dev.out(II, numpy_out), dev.out(JJ, numpy_out)
