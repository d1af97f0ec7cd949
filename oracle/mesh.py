"""CPU oracle of the mesh and connectivity builders (NumPy).

TEST INFRASTRUCTURE ONLY (checker).  Restates, for the Q2 x Q1 spaces:
  build_velocity_mesh (m_V = 9)     build_velocity_mesh.py:17-46, 110-200
  build_velocity_connectivity (Q2)  build_velocity_connectivity.py:14-136
  build_pressure_mesh               build_pressure_mesh.py:10-88
  build_pressure_connectivity       build_pressure_connectivity.py:10-48
on the whole node / element lattice at once, with the reference's element
expressions (i * hx / 2, the icon formulas, the sector and annulus maps).
The reference's per-node loops are restated separately, for small meshes, as
``*_loops`` (a second, independent reading used by the tests).
"""
import numpy as np

EPS = 1e-9  # constants.py:7
GEOMETRIES = ("box", "eighth", "quarter", "half", "annulus")


def _curved(geometry, x, z, Rinner, Router, opening_angle):
    """build_velocity_mesh.py:185-200 / build_pressure_mesh.py:62-79."""
    rad = np.zeros_like(x)
    theta = np.zeros_like(x)
    if geometry in ("eighth", "quarter", "half"):
        rad = Rinner + z * (Router - Rinner)
        theta = np.pi / 2 - x * opening_angle
        x, z = rad * np.cos(theta), rad * np.sin(theta)
    elif geometry == "annulus":
        t = x * 2 * np.pi
        r = Rinner + z * (Router - Rinner)
        x, z = np.cos(t) * r, np.sin(t) * r
        rad = r
        theta = np.arctan2(z, x)
        theta = np.where(theta < 0.0, theta + 2.0 * np.pi, theta)
    return x, z, rad, theta


def velocity_mesh(geometry, nelx, nelz, Lx, Lz, Rinner=0.0, Router=1.0, opening_angle=0.0):
    """(hx, hz, nnx, nnz, x, z, rad, theta, top, bot, left, right, middleH,
    middleV, hull, cornerBL, cornerBR, cornerTL, cornerTR)"""
    hx, hz = Lx / nelx, Lz / nelz
    ann = geometry == "annulus"
    nnx = 2 * nelx if ann else 2 * nelx + 1
    nnz = 2 * nelz + 1
    j, i = np.divmod(np.arange(nnx * nnz), nnx)
    x = i * hx / 2
    z = j * hz / 2
    top, bot = j == 2 * nelz, j == 0
    f = np.zeros(nnx * nnz, dtype=bool)
    left = f.copy() if ann else i == 0
    right = f.copy() if ann else i == 2 * nelx
    mv = f.copy() if ann else np.abs(x / Lx - 0.5) < EPS
    mh = f.copy() if ann else np.abs(z / Lz - 0.5) < EPS
    hull = top | bot | left | right
    x, z, rad, theta = _curved(geometry, x, z, Rinner, Router, opening_angle)
    corners = (0, 0, 0, 0) if ann else (0, nnx - 1, (nnz - 1) * nnx, nnx * nnz - 1)
    return (hx, hz, nnx, nnz, x, z, rad, theta, top, bot, left, right, mh, mv, hull, *corners)


def velocity_connectivity(geometry, nelx, nelz, middleH, middleV):
    """(icon_V (9, nel) int32, top, bot, left, right, middleH, middleV)"""
    nel = nelx * nelz
    j, i = np.divmod(np.arange(nel), nelx)
    e = np.arange(nel)
    icon = np.zeros((9, nel), dtype=np.int32)
    top, bot = j == nelz - 1, j == 0
    f = np.zeros(nel, dtype=bool)
    if geometry == "annulus":
        nelt = nelx
        icon[0] = 2 * e + 2 + 2 * j * nelt
        icon[1] = 2 * e + 2 * j * nelt
        icon[2] = icon[1] + 4 * nelt
        icon[3] = icon[1] + 4 * nelt + 2
        icon[4] = icon[0] - 1
        icon[5] = icon[1] + 2 * nelt
        icon[6] = icon[2] + 1
        icon[7] = icon[5] + 2
        icon[8] = icon[5] + 1
        last = i == nelt - 1
        for k in (0, 7, 3):
            icon[k, last] -= 2 * nelt
        return icon, top, bot, f.copy(), f.copy(), f.copy(), f.copy()
    nnx = 2 * nelx + 1
    icon[0] = i * 2 + 1 + j * 2 * nnx - 1
    icon[1] = i * 2 + 3 + j * 2 * nnx - 1
    icon[2] = i * 2 + 3 + j * 2 * nnx + nnx * 2 - 1
    icon[3] = i * 2 + 1 + j * 2 * nnx + nnx * 2 - 1
    icon[4] = i * 2 + 2 + j * 2 * nnx - 1
    icon[5] = i * 2 + 3 + j * 2 * nnx + nnx - 1
    icon[6] = i * 2 + 2 + j * 2 * nnx + nnx * 2 - 1
    icon[7] = i * 2 + 1 + j * 2 * nnx + nnx - 1
    icon[8] = i * 2 + 2 + j * 2 * nnx + nnx - 1
    mh = np.asarray(middleH, dtype=bool)
    mv = np.asarray(middleV, dtype=bool)
    return (icon, top, bot, i == 0, i == nelx - 1, mh[icon[0]] | mh[icon[2]],
            mv[icon[0]] | mv[icon[1]])


def pressure_mesh(geometry, nelx, nelz, hx, hz, Rinner=0.0, Router=1.0, opening_angle=0.0):
    """(x, z, rad, theta, left, right, top, bot, hull)"""
    ann = geometry == "annulus"
    nnx = nelx if ann else nelx + 1
    j, i = np.divmod(np.arange(nnx * (nelz + 1)), nnx)
    x = i * hx
    z = j * hz
    top, bot = j == nelz, j == 0
    f = np.zeros(i.size, dtype=bool)
    left = f.copy() if ann else i == 0
    right = f.copy() if ann else i == nelx
    hull = f.copy() if ann else top | bot | left | right
    x, z, rad, theta = _curved(geometry, x, z, Rinner, Router, opening_angle)
    return x, z, rad, theta, left, right, top, bot, hull


def pressure_connectivity(geometry, nelx, nelz):
    nel = nelx * nelz
    j, i = np.divmod(np.arange(nel), nelx)
    icon = np.zeros((4, nel), dtype=np.int32)
    if geometry == "annulus":
        nelt = nelx
        a1 = np.arange(nel)
        a2 = a1 + 1
        a3 = i + (j + 1) * nelt + 1
        a4 = i + (j + 1) * nelt
        last = i == nelt - 1
        a2 = np.where(last, a2 - nelt, a2)
        a3 = np.where(last, a3 - nelt, a3)
        icon[0], icon[1], icon[2], icon[3] = a2, a1, a4, a3
        return icon
    icon[0] = i + j * (nelx + 1)
    icon[1] = i + 1 + j * (nelx + 1)
    icon[2] = i + 1 + (j + 1) * (nelx + 1)
    icon[3] = i + (j + 1) * (nelx + 1)
    return icon


# --- the reference's loops, restated for small meshes --------------------------------------

def velocity_connectivity_loops(geometry, nelx, nelz):
    """build_velocity_connectivity.py:76-128, element by element."""
    nel = nelx * nelz
    icon = np.zeros((9, nel), dtype=np.int32)
    counter = 0
    if geometry == "annulus":
        nelt, nelr = nelx, nelz
        for j in range(nelr):
            for i in range(nelt):
                icon[0, counter] = 2 * counter + 2 + 2 * j * nelt
                icon[1, counter] = 2 * counter + 2 * j * nelt
                icon[2, counter] = icon[1, counter] + 4 * nelt
                icon[3, counter] = icon[1, counter] + 4 * nelt + 2
                icon[4, counter] = icon[0, counter] - 1
                icon[5, counter] = icon[1, counter] + 2 * nelt
                icon[6, counter] = icon[2, counter] + 1
                icon[7, counter] = icon[5, counter] + 2
                icon[8, counter] = icon[5, counter] + 1
                if i == nelt - 1:
                    icon[0, counter] -= 2 * nelt
                    icon[7, counter] -= 2 * nelt
                    icon[3, counter] -= 2 * nelt
                counter += 1
        return icon
    nnx = 2 * nelx + 1
    for j in range(nelz):
        for i in range(nelx):
            base = i * 2 + j * 2 * nnx
            for k, (di, dj) in enumerate(((0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1),
                                          (1, 2), (0, 1), (1, 1))):
                icon[k, counter] = base + di + dj * nnx
            counter += 1
    return icon
