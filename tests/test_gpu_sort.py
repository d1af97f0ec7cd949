"""Stable sort of particles by element (mw_sort_particles_by_element) against
NumPy's stable argsort of the same keys: the element index of an active particle
whose index lies in [0, nel), nel for every other particle (placed last, in
unspecified order).  Edge cases: empty elements, one element holding thousands
of particles, inactive particles, element indices outside [0, nel)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def expected(nel, active, iel):
    key = np.where(active & (iel >= 0) & (iel < nel), iel, nel)
    order = np.argsort(key, kind="stable").astype(np.int32)
    seg = np.searchsorted(key[order], np.arange(nel + 1), side="left").astype(np.int32)
    return order, seg


def gpu_sort(nel, active, iel):
    from meeuuw_amd import _dev as dev
    from meeuuw_amd import _lib
    n = len(iel)
    a, e = dev.to_dev(active, np.uint8), dev.to_dev(iel, np.int32)
    order, seg = dev.empty(n, np.int32), dev.empty(nel + 1, np.int32)
    work = dev.empty((_lib.lib.mw_sort_workspace_bytes(n, nel) + 7) // 8)
    _lib.call("mw_sort_particles_by_element", n, nel, dev.ptr(a), dev.ptr(e), dev.ptr(order),
              dev.ptr(seg), dev.ptr(work), dev.stream())
    return dev.host(order), dev.host(seg)


def check(nel, active, iel):
    o, s = gpu_sort(nel, active, iel)
    eo, es = expected(nel, active, iel)
    np.testing.assert_array_equal(s, es)
    na = es[nel]
    np.testing.assert_array_equal(o[:na], eo[:na])  # the elements' particles: stable
    assert sorted(o[na:]) == sorted(eo[na:])         # the rest: the same particles


@pytest.mark.parametrize("nel,ppe,seed", [(7, 5, 0), (1024, 16, 1), (65536, 25, 2)])
def test_sort_random(nel, ppe, seed):
    rng = np.random.default_rng(seed)
    n = nel * ppe
    iel = rng.integers(0, nel, n).astype(np.int32)
    active = rng.random(n) > 0.05
    print(f"sort nel {nel} n {n}", flush=True)
    check(nel, active, iel)


def test_sort_edge_cases():
    rng = np.random.default_rng(3)
    nel, n = 300, 20000
    iel = rng.integers(0, nel, n).astype(np.int32)
    iel[::7] = 5                         # one crowded element (~2900 particles)
    iel[rng.random(n) < 0.01] = -3       # indices outside the mesh
    iel[rng.random(n) < 0.01] = nel + 10
    iel[(iel >= 100) & (iel < 120)] = 0  # 20 empty elements
    active = rng.random(n) > 0.1
    check(nel, active, iel)
    check(nel, np.zeros(n, bool), iel)   # nothing active: every element empty
    check(1, np.ones(3, bool), np.zeros(3, np.int32))
