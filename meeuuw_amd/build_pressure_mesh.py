"""Drop-in build_pressure_mesh (build_pressure_mesh.py:10-88) on the GPU."""
import numpy as np

from . import _dev as dev
from . import _lib
from .build_velocity_mesh import GEOMETRY


def build_pressure_mesh(geometry, nn_P, nelx, nelz, hx, hz, Rinner, Router, opening_angle, debug,
                        device=False):
    if geometry not in GEOMETRY:
        raise ValueError("build_pressure_mesh: unknown geometry")
    nelx, nelz = int(nelx), int(nelz)
    nnx = nelx if geometry == "annulus" else nelx + 1
    nn = nnx * (nelz + 1)
    if int(nn_P) != nn:
        raise ValueError(f"build_pressure_mesh: nn_P={nn_P} but the mesh has {nn} nodes")
    dev.require_gpu()
    x, z, rad, theta = dev.empty(nn), dev.empty(nn), dev.empty(nn), dev.empty(nn)
    flags = dev.empty(5 * nn, np.uint8)
    _lib.call("mw_build_pressure_mesh", GEOMETRY[geometry], nelx, nelz, float(hx), float(hz),
              float(Rinner), float(Router), float(opening_angle), dev.ptr(x), dev.ptr(z),
              dev.ptr(rad), dev.ptr(theta), dev.ptr(flags), dev.stream())
    numpy_out = not device
    f = flags.reshape(5, nn).to(dtype=dev.torch.bool)
    arrays = [dev.out(t, numpy_out) for t in (x, z, rad, theta)]
    if debug:
        xs, zs = (a if numpy_out else dev.host(a) for a in arrays[:2])
        np.savetxt("DEBUG/mesh_P.ascii", np.array([xs, zs]).T, header="# x,z")
    return (*arrays, *(dev.out(f[k], numpy_out) for k in range(5)))  # left, right, top, bot, hull
