"""The oracle's free-slip rotation on curved boundaries (build_matrix_stokes.py:163-203)
pinned by what the rotation means, not by its formula: the rotated, reduced
system acting on a velocity field given in (normal, tangential) components at
the boundary nodes equals the unrotated system acting on the same field in
Cartesian components, rotated afterwards; a rigid rotation (tangential on both
circles, strain-rate free) is in its null space; and its divergence is
unchanged.  A wrong sign in the rotation, a wrong boundary node list or a
missing elimination fails these checks."""
import numpy as np
import pytest
from scipy import sparse

from conftest import curved_stokes_inputs


def _systems(O, geometry, nelx, nelz):
    args, cm = curved_stokes_inputs(O, geometry, nelx, nelz, seed=3)
    plain = list(args)
    plain[33] = np.zeros_like(args[33])  # no Dirichlet rows
    plain[37] = plain[38] = False  # no free slip
    rot = O.build_matrix_stokes(*args, True)
    ref = O.build_matrix_stokes(*plain, True)
    m = cm["mesh"]
    coo = O.coo_indices(m)
    K = lambda o: sparse.coo_matrix((o[3], coo["K"]), shape=(m.Nfem_V, m.Nfem_V)).tocsr()
    G = lambda o: sparse.coo_matrix((o[4], coo["G"]), shape=(m.Nfem_V, m.Nfem_P)).tocsr()
    return args, cm, (K(rot), G(rot), rot[1]), (K(ref), G(ref), ref[1])


def _rotation(cm, fix):
    """Global rotation to (normal, tangential) at the free boundary nodes and
    the mask of the dofs the reduced system eliminates."""
    m = cm["mesh"]
    th = cm["theta_V"]
    boundary = (m.z_V < 1e-8) | (m.z_V > 1 - 1e-8)
    free = ~fix[0::2] & ~fix[1::2]
    rotn = boundary & free
    c, s = np.where(rotn, np.cos(th), 1.0), np.where(rotn, np.sin(th), 0.0)
    elim = fix.copy()
    elim[0::2] |= rotn

    def to_rot(x):  # Cartesian -> (n, t) at the rotated nodes
        u, w = x[0::2], x[1::2]
        y = np.empty_like(x)
        y[0::2], y[1::2] = c * u + s * w, -s * u + c * w
        return y

    def to_cart(y):
        n, t = y[0::2], y[1::2]
        x = np.empty_like(y)
        x[0::2], x[1::2] = c * n - s * t, s * n + c * t
        return x

    return to_rot, to_cart, elim, rotn


@pytest.mark.parametrize("geometry,nelx,nelz", [("half", 10, 4), ("quarter", 7, 3),
                                                ("eighth", 5, 2)])
def test_rotated_system_equals_rotated_cartesian_system(oracle, geometry, nelx, nelz):
    args, cm, (Kr, Gr, fr), (K0, G0, f0) = _systems(oracle, geometry, nelx, nelz)
    fix = np.asarray(args[33], dtype=bool)
    to_rot, to_cart, elim, rotn = _rotation(cm, fix)
    assert rotn.sum() > 0
    rng = np.random.default_rng(7)
    y = rng.standard_normal(fix.size)
    y[elim] = 0.0  # no normal velocity on the circles, fixed dofs zero
    x = to_cart(y)
    lhs = Kr @ y
    rhs = to_rot(K0 @ x)
    rhs[elim] = 0.0
    assert np.linalg.norm(lhs - rhs) <= 1e-13 * np.linalg.norm(rhs)
    # G^T (divergence) does not see the rotation
    d0, dr = G0.T @ x, Gr.T @ y
    assert np.linalg.norm(dr - d0) <= 1e-13 * np.linalg.norm(d0)
    # the pressure gradient rotates with the rows and leaves the eliminated rows
    q = rng.standard_normal(G0.shape[1])
    ge = to_rot(G0 @ q)
    ge[elim] = 0.0
    assert np.linalg.norm(Gr @ q - ge) <= 1e-13 * np.linalg.norm(ge)
    # the body force rotates with the rows (zero Dirichlet values)
    fe = to_rot(f0)
    fe[elim] = 0.0
    assert np.linalg.norm(fr - fe) <= 1e-14 * np.linalg.norm(fe)
    # the eliminated rows keep only their diagonal
    Kd = Kr.tocoo()
    off = (Kd.row != Kd.col) & (np.abs(Kd.data) > 0)
    assert not np.any(elim[Kd.row[off]]) and not np.any(elim[Kd.col[off]])


def test_rigid_rotation_is_in_the_null_space(oracle):
    args, cm, (Kr, _, _), (K0, _, _) = _systems(oracle, "half", 12, 4)
    fix = np.asarray(args[33], dtype=bool)
    to_rot, _, elim, _ = _rotation(cm, fix)
    X, Z = cm["x_V"], cm["z_V"]
    x = np.empty(2 * X.size)
    x[0::2], x[1::2] = -Z, X  # omega = 1 about the origin
    y = to_rot(x)
    assert np.abs(y[elim & ~fix]).max() < 1e-14  # tangential on both circles
    # the sides' fixed dofs are not rotation-compatible: rows within an element
    # of the sides are left out
    m = cm["mesh"]
    inner = (m.x_V > 0.2) & (m.x_V < 0.8)
    rows = np.repeat(inner, 2) & ~fix & ~elim
    scale = np.abs(K0).max() * np.abs(x).max()
    r = (Kr @ np.where(fix, 0.0, y))[rows]
    assert r.size > 0 and np.abs(r).max() < 1e-11 * scale
    assert np.abs(K0 @ x).max() < 1e-11 * scale  # the unrotated system agrees
