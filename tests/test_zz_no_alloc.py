"""mw_sort_particles_by_element_noalloc (scratch in the library's fixed device
array) against mw_sort_particles_by_element (caller's workspace), both through
the C ABI: the same element ranges and the same stable order of each element's
particles, on two calls in a row, the second with a smaller input (the array
keeps the first call's contents).  Sizes beyond the array are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(nel, n, seed):
    rng = np.random.default_rng(seed)
    iel = rng.integers(0, nel, n).astype(np.int32)
    iel[rng.random(n) < 0.01] = nel + 3
    active = rng.random(n) > 0.05
    return active, iel


def _both(nel, active, iel):
    from meeuuw_amd import _dev as dev
    from meeuuw_amd import _lib
    n = len(iel)
    a, e = dev.to_dev(active, np.uint8), dev.to_dev(iel, np.int32)
    o1, s1 = dev.empty(n, np.int32), dev.empty(nel + 1, np.int32)
    o2, s2 = dev.empty(n, np.int32), dev.empty(nel + 1, np.int32)
    work = dev.empty((_lib.lib.mw_sort_workspace_bytes(n, nel) + 7) // 8)
    _lib.call("mw_sort_particles_by_element", n, nel, dev.ptr(a), dev.ptr(e), dev.ptr(o1),
              dev.ptr(s1), dev.ptr(work), dev.stream())
    _lib.call("mw_sort_particles_by_element_noalloc", n, nel, dev.ptr(a), dev.ptr(e), dev.ptr(o2),
              dev.ptr(s2), dev.stream())
    return dev.host(o1), dev.host(s1), dev.host(o2), dev.host(s2)


def _check(nel, n, seed):
    active, iel = _inputs(nel, n, seed)
    o1, s1, o2, s2 = _both(nel, active, iel)
    np.testing.assert_array_equal(s2, s1)
    na = s1[nel]
    np.testing.assert_array_equal(o2[:na], o1[:na])
    assert sorted(o2[na:]) == sorted(o1[na:])
    key = np.where(active & (iel < nel), iel, nel)
    np.testing.assert_array_equal(o2[:na], np.argsort(key, kind="stable")[:na])


def test_noalloc_sort_matches_workspace_sort():
    print("no-alloc sort: 65536 elements x 25, then 1000 x 9", flush=True)
    _check(65536, 65536 * 25, 11)
    _check(1000, 9000, 12)


def test_noalloc_sort_refuses_oversize():
    from meeuuw_amd import _lib
    from meeuuw_amd import _dev as dev
    x = dev.empty(4, np.int32)
    rc = _lib.lib.mw_sort_particles_by_element_noalloc((1 << 22) + 1, 10, dev.ptr(x), dev.ptr(x),
                                                       dev.ptr(x), dev.ptr(x), dev.stream())
    assert rc == 6  # MW_ERANGE
    assert "at most" in _lib.lib.mw_last_error().decode()
