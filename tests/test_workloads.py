"""Synthetic workload inputs (meeuuw_amd.workloads): the seeded Gaussian random
field of the two-material workload and its thresholding."""
import numpy as np


def test_random_field_statistics():
    """Unit variance, zero mean, and the squared-exponential correlation
    exp(-d^2 / (2 l^2)) at lags l and 2 l (a filter with the spectral density
    instead of its square root gives exp(-d^2 / (4 l^2)): 0.78 at lag l)."""
    from meeuuw_amd.workloads import gaussian_random_field
    g, hx, hz = gaussian_random_field(8.0, 4.0, corr_len=0.1, seed=3)
    assert abs(g.mean()) < 1e-12 and abs(g.std() - 1.0) < 1e-12
    step = int(round(0.1 / hx))
    assert abs(step * hx - 0.1) < 1e-12
    for lag, expect in ((step, np.exp(-0.5)), (2 * step, np.exp(-2.0))):
        c = np.mean(g[:, :-lag] * g[:, lag:]) / np.mean(g * g)
        assert abs(c - expect) < 0.05, (lag, c, expect)


def test_blob_materials_seeded_and_balanced():
    from meeuuw_amd.workloads import blob_materials, box_particles
    r, t, x, z = box_particles(64, 16, 4.0, 1.0, 9, seed=5)
    assert x.min() > 0 and x.max() < 4.0 and z.min() > 0 and z.max() < 1.0
    m1 = blob_materials(x, z, 4.0, 1.0, seed=7)
    m2 = blob_materials(x, z, 4.0, 1.0, seed=7)
    m3 = blob_materials(x, z, 4.0, 1.0, seed=8)
    assert m1.dtype == np.uint8 and np.array_equal(m1, m2) and not np.array_equal(m1, m3)
    assert 0.3 < m1.mean() < 0.7
    # blobs are coherent over the correlation length: neighbours in an element agree
    same = np.mean(m1.reshape(-1, 9) == m1.reshape(-1, 9)[:, :1])
    assert same > 0.85


def test_box_particles_match_the_placement_formula(oracle):
    """The host positions of the seeded swarm are the oracle's placement
    (x = sum N_V(r, t) x_V over the element) to rounding."""
    from meeuuw_amd.workloads import box_particles
    m = oracle.Mesh(12, 4, Lx=3.0, Lz=1.0)
    r, t, x, z = box_particles(12, 4, 3.0, 1.0, 9, seed=11)
    r2, t2 = oracle.random_reduced_coords(m.nel, 9, 11)
    assert np.array_equal(r, r2) and np.array_equal(t, t2)
    xo, zo, _ = oracle.place_particles(m, r, t, 9)
    assert np.abs(x - xo).max() < 1e-14 and np.abs(z - zo).max() < 1e-14
