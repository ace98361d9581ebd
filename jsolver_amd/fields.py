"""Seeded synthetic inputs for the benchmark configurations (host numpy, FP64).

Gaussian random field GRF(l, sigma): i.i.d. N(0,1) white noise on the full
padded array, filtered in Fourier space by exp(-2 pi^2 l^2 |k|^2) with
k = fftfreq(n, d=h) (a periodic Gaussian kernel of standard deviation l),
then standardised to mean 0 and scaled to standard deviation sigma.
Seeds: numpy.random.default_rng([20261203, config_id, field_id]).
"""
import numpy as np

from .grid_1d import grid_1D

SEED = 20261203


def rng_for(config_id, field_id):
    return np.random.default_rng([SEED, int(config_id), int(field_id)])


def grf(shape, h, corr_len, sigma, rng):
    noise = rng.standard_normal(shape)
    f = np.fft.rfftn(noise)
    k2 = np.zeros(f.shape)
    for ax, n in enumerate(shape):
        if ax == len(shape) - 1:
            k = np.fft.rfftfreq(n, d=h[ax])
        else:
            k = np.fft.fftfreq(n, d=h[ax])
        sh = [1] * len(shape)
        sh[ax] = k.size
        k2 = k2 + (k.reshape(sh) ** 2)
    f *= np.exp(-2.0 * np.pi ** 2 * corr_len ** 2 * k2)
    g = np.fft.irfftn(f, s=shape, axes=tuple(range(len(shape))))
    g -= g.mean()
    sd = g.std()
    return g * (sigma / sd if sd > 0 else 0.0)


def anisotropic_3d(m, config_id=3, corr_len=0.05, sigma=1.0, dzz_scale=10.0, lam_max=5.0,
                   noise=0.05):
    """3D anisotropic diffusion-absorption problem on an m^3 grid: log-normal
    D_xx, D_yy, D_zz (D_zz scaled by dzz_scale), lambda ~ U[0, lam_max],
    rhs = sin(pi x) sin(2 pi y) sin(pi z) + N(0, noise^2)."""
    g = grid_1D(m)
    shape = (m + 3,) * 3
    h = (g.del_,) * 3
    dxx = np.exp(grf(shape, h, corr_len, sigma, rng_for(config_id, 0)))
    dyy = np.exp(grf(shape, h, corr_len, sigma, rng_for(config_id, 1)))
    dzz = dzz_scale * np.exp(grf(shape, h, corr_len, sigma, rng_for(config_id, 2)))
    lam = rng_for(config_id, 3).uniform(0.0, lam_max, shape)
    c = g.centers
    x = c[None, None, :]
    y = c[None, :, None]
    z = c[:, None, None]
    rhs = np.sin(np.pi * x) * np.sin(2 * np.pi * y) * np.sin(np.pi * z)
    rhs = rhs + noise * rng_for(config_id, 4).standard_normal(shape)
    return (g, g, g), rhs, lam, dxx, dyy, dzz


def anisotropic_2d(m, config_id=2, corr_len=0.05, sigma=1.0, ratio=100.0, lam_max=10.0,
                   noise=0.1, seed_offset=0):
    """2D anisotropic problem: D_xx = exp(g1), D_yy = ratio * exp(g2),
    lambda ~ U[0, lam_max], rhs = sin(2 pi x) sin(3 pi y) + N(0, noise^2)."""
    g = grid_1D(m)
    shape = (m + 3,) * 2
    h = (g.del_,) * 2
    o = 10 * seed_offset
    dxx = np.exp(grf(shape, h, corr_len, sigma, rng_for(config_id, o + 0)))
    dyy = ratio * np.exp(grf(shape, h, corr_len, sigma, rng_for(config_id, o + 1)))
    lam = rng_for(config_id, o + 2).uniform(0.0, lam_max, shape)
    c = g.centers
    x = c[None, :]
    y = c[:, None]
    rhs = np.sin(2 * np.pi * x) * np.sin(3 * np.pi * y)
    rhs = rhs + noise * rng_for(config_id, o + 3).standard_normal(shape)
    return (g, g), rhs, lam, dxx, dyy
