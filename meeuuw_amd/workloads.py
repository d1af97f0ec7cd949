"""Synthetic inputs of the benchmark workloads (host side, NumPy, seeded).

The two-material workload puts stiff, slightly denser blobs in a weak matrix:
the blobs are where a Gaussian random field with squared-exponential
covariance exp(-d^2 / (2 l^2)), correlation length l = 0.1, is positive.  The
field is drawn on a regular grid by spectral filtering of white noise (FFT of
N(0,1) samples times the square root of the covariance's spectral density,
exp(-|k|^2 l^2 / 4), on a grid twice the domain so the periodic wrap does not
correlate opposite walls), scaled to unit variance, and read at the particles
by bilinear interpolation.  The reference has no such generator: its material
layouts are per-experiment plugin functions (experiment21.py:94-107 style), and
this one plays that role for the workload.
"""
import numpy as np


def gaussian_random_field(Lx, Lz, corr_len=0.1, seed=20261203, pts_per_len=8):
    """(grid, hx, hz): a unit-variance Gaussian random field on the
    (nx + 1) x (nz + 1) nodes of [0, Lx] x [0, Lz], spacing <= corr_len / pts_per_len."""
    nx = int(np.ceil(Lx / corr_len * pts_per_len))
    nz = int(np.ceil(Lz / corr_len * pts_per_len))
    hx, hz = Lx / nx, Lz / nz
    Nx, Nz = 2 * (nx + 1), 2 * (nz + 1)  # padded periodic grid
    rng = np.random.Generator(np.random.PCG64(seed))
    w = rng.standard_normal((Nz, Nx))
    kx = 2 * np.pi * np.fft.fftfreq(Nx, d=hx)
    kz = 2 * np.pi * np.fft.fftfreq(Nz, d=hz)
    K2 = kz[:, None] ** 2 + kx[None, :] ** 2
    f = np.real(np.fft.ifft2(np.fft.fft2(w) * np.exp(-K2 * corr_len ** 2 / 4.0)))
    f = f[: nz + 1, : nx + 1]
    f = (f - f.mean()) / f.std()
    return f, hx, hz


def sample_field(grid, hx, hz, x, z):
    """Bilinear interpolation of a node-grid field (rows z, columns x) at (x, z)."""
    nz, nx = grid.shape[0] - 1, grid.shape[1] - 1
    fx = np.clip(np.asarray(x) / hx, 0.0, nx)
    fz = np.clip(np.asarray(z) / hz, 0.0, nz)
    i = np.minimum(fx.astype(np.int64), nx - 1)
    j = np.minimum(fz.astype(np.int64), nz - 1)
    a, b = fx - i, fz - j
    return ((1 - a) * (1 - b) * grid[j, i] + a * (1 - b) * grid[j, i + 1]
            + (1 - a) * b * grid[j + 1, i] + a * b * grid[j + 1, i + 1])


def blob_materials(x, z, Lx, Lz, corr_len=0.1, seed=20261203):
    """Material of each particle: 1 where the thresholded random field is
    positive (the blobs), 0 elsewhere; uint8."""
    g, hx, hz = gaussian_random_field(Lx, Lz, corr_len, seed)
    return (sample_field(g, hx, hz, x, z) > 0.0).astype(np.uint8)


def box_particles(nelx, nelz, Lx, Lz, ppe, seed):
    """Reduced coordinates (r, t) ~ U(-1, 1), element-major, r before t, from
    PCG64(seed) (the seeded placement the timestep uses), and the particles'
    positions x = x_e + (r + 1) hx / 2, z = z_e + (t + 1) hz / 2."""
    nel = nelx * nelz
    rng = np.random.Generator(np.random.PCG64(seed))
    rt = rng.uniform(-1.0, 1.0, size=(nel, ppe, 2))
    r, t = rt[:, :, 0].ravel(), rt[:, :, 1].ravel()
    e = np.repeat(np.arange(nel), ppe)
    hx, hz = Lx / nelx, Lz / nelz
    x = (e % nelx) * hx + (r + 1) * hx / 2
    z = (e // nelx) * hz + (t + 1) * hz / 2
    return r, t, x, z


def two_material_config(nelx=512, nelz=128, aspect=4, ppe=36, eta_ratio=1e6, drho=0.01,
                        **kw):
    """Configuration of the two-material workload: nelx x nelz elements on
    [0, aspect] x [0, 1], blobs (material 1) of viscosity eta_ratio and density
    1 + drho in a matrix of viscosity 1 and density 1, gravity 1, no thermal
    buoyancy (alphaT 0; the temperature is still advected and diffused), RK4."""
    from .timestep import make_config
    args = dict(nelx=nelx, nelz=nelz, Lx=float(aspect), Lz=1.0, ppe=ppe, rk=4, alphaT=0.0,
                material_mode=1, mat_eta=(1.0, float(eta_ratio)), mat_rho=(1.0, 1.0 + drho),
                gravity=1.0)
    args.update(kw)
    return make_config(**args)
