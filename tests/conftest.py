import json
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs a ROCm GPU (MI355X) and the built HIP library")


@pytest.fixture(scope="session")
def ref_results():
    with open(os.path.join(GOLDEN, "reference_results.json")) as f:
        return json.load(f)


@pytest.fixture(scope="session")
def oracle():
    from oracle import oracle as O
    O.build()
    return O


def donea_huerta_bodyforce(x, y):
    """gravity_model of experiment19.py (Donea & Huerta manufactured solution)."""
    gx = ((12.0 - 24.0 * y) * x ** 4 + (-24.0 + 48.0 * y) * x * x * x
          + (-48.0 * y + 72.0 * y * y - 48.0 * y * y * y + 12.0) * x * x
          + (-2.0 + 24.0 * y - 72.0 * y * y + 48.0 * y * y * y) * x
          + 1.0 - 4.0 * y + 12.0 * y * y - 8.0 * y * y * y)
    gz = ((8.0 - 48.0 * y + 48.0 * y * y) * x * x * x + (-12.0 + 72.0 * y - 72.0 * y * y) * x * x
          + (4.0 - 24.0 * y + 48.0 * y * y - 48.0 * y * y * y + 24.0 * y ** 4) * x
          - 12.0 * y * y + 24.0 * y * y * y - 12.0 * y ** 4)
    return gx, gz


def donea_huerta_exact(x, y):
    """ms_velocity_x/z and ms_pressure for case 19 (postprocessors.py:255-303)."""
    u = x * x * (1.0 - x) ** 2 * (2.0 * y - 6.0 * y * y + 4 * y * y * y)
    w = -y * y * (1.0 - y) ** 2 * (2.0 * x - 6.0 * x * x + 4 * x * x * x)
    p = x * (1.0 - x) - 1.0 / 6.0
    return u, w, p


def stokes_inputs(O, mesh, rhoq, etaq, gxq, gzq, bc):
    """Positional argument tuple of build_matrix_stokes for a box mesh."""
    m = mesh
    fix, val = bc
    eta_e = etaq.mean(axis=1)
    return (m.bignb_Stokes, m.bignb_K, m.bignb_P, m.bignb_G, m.nel, 9, 9, 4, 2, m.Nfem_V, m.Nfem_P,
            18, m.icon_V, m.icon_P, rhoq, etaq, m.JxWq, m.local_to_globalV, gxq, gzq, m.N_V, m.N_P,
            m.dNdr_V, m.dNdt_V, m.dNdr_P, m.dNdt_P, m.jcbi00q, m.jcbi01q, m.jcbi10q, m.jcbi11q,
            eta_e, 1.0, m.L_ref, fix, val, m.bot_element, m.top_element, False, False, "box", None,
            False, m.xq)


def random_swarm(O, mesh, ppe, seed):
    r, t = O.random_reduced_coords(mesh.nel, ppe, seed)
    x, z, iel = O.place_particles(mesh, r, t, ppe)
    return x, z, r, t, iel
