"""Drop-in build_velocity_mesh (build_velocity_mesh.py:17-227) on the GPU.

The Q2 velocity space (m_V = 9) of the box, the curved sectors (eighth,
quarter, half) and the annulus; the Q1+ space (m_V = 5) is not on this path.
Returns the reference's 19-tuple as NumPy arrays (device=True: torch tensors
on the GPU for the arrays).
"""
import numpy as np

from . import _dev as dev
from . import _lib

GEOMETRY = {"box": 0, "eighth": 1, "quarter": 2, "half": 3, "annulus": 4}


def build_velocity_mesh(geometry, m_V, nn_V, nelx, nelz, Lx, Lz, Rinner, Router, opening_angle,
                        debug_ascii, device=False):
    if m_V == 5:
        raise NotImplementedError("build_velocity_mesh: the Q1+ space (m_V=5) is not on the GPU path")
    if m_V != 9:
        raise ValueError("build_velocity_mesh: unknown m_V value")
    if geometry not in GEOMETRY:  # the node-count match comes first in the reference
        raise ValueError("build_velocity_mesh: unknown geometry for m_V=9")
    nelx, nelz = int(nelx), int(nelz)
    hx = Lx / nelx
    hz = Lz / nelz
    nnx = 2 * nelx if geometry == "annulus" else 2 * nelx + 1
    nnz = 2 * nelz + 1
    nn = nnx * nnz
    if int(nn_V) != nn:
        raise ValueError(f"build_velocity_mesh: nn_V={nn_V} but the mesh has {nn} nodes")
    dev.require_gpu()
    x, z, rad, theta = dev.empty(nn), dev.empty(nn), dev.empty(nn), dev.empty(nn)
    flags = dev.empty(7 * nn, np.uint8)
    _lib.call("mw_build_velocity_mesh", GEOMETRY[geometry], nelx, nelz, float(Lx), float(Lz),
              float(Rinner), float(Router), float(opening_angle), dev.ptr(x), dev.ptr(z),
              dev.ptr(rad), dev.ptr(theta), dev.ptr(flags), dev.stream())
    numpy_out = not device
    f = flags.reshape(7, nn).to(dtype=dev.torch.bool)
    arrays = [dev.out(t, numpy_out) for t in (x, z, rad, theta)]
    fl = [dev.out(f[k], numpy_out) for k in range(7)]  # top, bot, left, right, mH, mV, hull
    if geometry == "annulus":
        corners = (0, 0, 0, 0)
    else:
        corners = (0, nnx - 1, (nnz - 1) * nnx, nnx * nnz - 1)  # BL, BR, TL, TR
    if debug_ascii:
        xs, zs = (a if numpy_out else dev.host(a) for a in arrays[:2])
        np.savetxt("DEBUG/mesh_V.ascii", np.array([xs, zs]).T, header="# x,z")
    return (hx, hz, nnx, nnz, *arrays, *fl, *corners)
