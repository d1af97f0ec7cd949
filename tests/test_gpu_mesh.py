"""Mesh and connectivity drop-ins (build_velocity_mesh, build_velocity_connectivity,
build_pressure_mesh, build_pressure_connectivity) against the oracle: integer
outputs and box coordinates bit-exact; the curved coordinates, which go
through the device's cos, sin and atan2, within 4 ulp of the radius."""
import numpy as np
import pytest

from oracle import mesh as M

pytestmark = pytest.mark.gpu

ANGLE = {"box": 0.0, "eighth": np.pi / 4, "quarter": np.pi / 2, "half": np.pi, "annulus": 0.0}


def close(a, b, scale):
    assert a.shape == b.shape
    np.testing.assert_allclose(a, b, rtol=0, atol=4 * np.finfo(np.float64).eps * scale)


@pytest.mark.parametrize("geometry", M.GEOMETRIES)
@pytest.mark.parametrize("nelx,nelz", [(3, 2), (64, 16), (1024, 128), (256, 256)])
def test_mesh_dropins_match_the_oracle(geometry, nelx, nelz):
    from meeuuw_amd.build_velocity_mesh import build_velocity_mesh
    from meeuuw_amd.build_velocity_connectivity import build_velocity_connectivity
    from meeuuw_amd.build_pressure_mesh import build_pressure_mesh
    from meeuuw_amd.build_pressure_connectivity import build_pressure_connectivity
    Lx, Lz = (4.0, 1.0) if geometry == "box" else (1.0, 1.0)
    Rin, Rout, ang = 0.55, 1.0, ANGLE[geometry]
    nnx = 2 * nelx if geometry == "annulus" else 2 * nelx + 1
    nn_V = nnx * (2 * nelz + 1)
    got = build_velocity_mesh(geometry, 9, nn_V, nelx, nelz, Lx, Lz, Rin, Rout, ang, False)
    ref = M.velocity_mesh(geometry, nelx, nelz, Lx, Lz, Rin, Rout, ang)
    assert got[:4] == ref[:4] and got[15:] == ref[15:]
    for k in range(8, 15):
        np.testing.assert_array_equal(got[k], ref[k])
        assert got[k].dtype == np.bool_
    if geometry == "box":
        for k in range(4, 8):
            np.testing.assert_array_equal(got[k], ref[k])
    else:
        for k in (4, 5, 6):
            close(got[k], ref[k], Rout)
        dth = np.abs(got[7] - ref[7])
        dth = np.minimum(dth, 2 * np.pi - dth)  # an angle of 0 may come back as 2 pi
        assert dth.max() <= 8 * np.finfo(np.float64).eps * 2 * np.pi
    gc = build_velocity_connectivity(geometry, 9, nelx, nelz, got[2], got[3], got[12], got[13])
    rc = M.velocity_connectivity(geometry, nelx, nelz, ref[12], ref[13])
    assert gc[0].dtype == np.int32 and gc[0].shape == (9, nelx * nelz)
    for a, b in zip(gc, rc):
        np.testing.assert_array_equal(a, b)
    nn_P = (nelx if geometry == "annulus" else nelx + 1) * (nelz + 1)
    gp = build_pressure_mesh(geometry, nn_P, nelx, nelz, got[0], got[1], Rin, Rout, ang, False)
    rp = M.pressure_mesh(geometry, nelx, nelz, ref[0], ref[1], Rin, Rout, ang)
    for k in range(4, 9):
        np.testing.assert_array_equal(gp[k], rp[k])
    if geometry == "box":
        for k in range(4):
            np.testing.assert_array_equal(gp[k], rp[k])
    else:
        for k in range(3):
            close(gp[k], rp[k], Rout)
    ip = build_pressure_connectivity(geometry, nelx, nelz, nelx * nelz, 4)
    np.testing.assert_array_equal(ip, M.pressure_connectivity(geometry, nelx, nelz))


def test_box_mesh_feeds_the_other_dropins_on_the_device():
    """device=True keeps the arrays on the GPU; the connectivity then takes
    the device flags and returns device tensors."""
    import torch
    from meeuuw_amd.build_velocity_mesh import build_velocity_mesh
    from meeuuw_amd.build_velocity_connectivity import build_velocity_connectivity
    v = build_velocity_mesh("box", 9, 65 * 33, 32, 16, 2.0, 1.0, 0, 0, 0, False, device=True)
    assert isinstance(v[4], torch.Tensor) and v[4].is_cuda
    c = build_velocity_connectivity("box", 9, 32, 16, v[2], v[3], v[12], v[13])
    assert isinstance(c[0], torch.Tensor) and tuple(c[0].shape) == (9, 512)
    ref = M.velocity_connectivity("box", 32, 16, v[12].cpu().numpy(), v[13].cpu().numpy())
    np.testing.assert_array_equal(c[0].cpu().numpy(), ref[0])
