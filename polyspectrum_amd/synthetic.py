"""Seeded synthetic input fields (host side, numpy/scipy).

The reference reads raw N^3 grids from disk (grid.c:121-228); benchmarks and
tests here use fields generated from a seed instead:

* Gaussian random field g: white noise, multiplied in Fourier space by
  |m|^(-alpha/2) (m in units of the fundamental mode, m = 0 set to 0),
  transformed back and standardised to zero mean and unit variance;
* log-normal density contrast delta = exp(sigma*g - sigma^2/2) - 1 (float32);
* binary ionisation map: a GRF with alpha = 2 smoothed by exp(-k^2 R^2 / 2)
  (k in radians per cell, R in cells), thresholded at its median (float64).
"""
import numpy as np

try:
    import scipy.fft as _fft

    def _rfftn(a):
        return _fft.rfftn(a, workers=-1)

    def _irfftn(a, s):
        return _fft.irfftn(a, s=s, workers=-1)
except ImportError:  # pragma: no cover
    def _rfftn(a):
        return np.fft.rfftn(a)

    def _irfftn(a, s):
        return np.fft.irfftn(a, s=s)


def _kgrid(n):
    m = np.fft.fftfreq(n, d=1.0 / n)
    mz = np.arange(n // 2 + 1, dtype=np.float64)
    return m[:, None, None], m[None, :, None], mz[None, None, :]


def gaussian_field(n, seed, alpha, smooth_cells=0.0):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal((n, n, n))
    wk = _rfftn(w)
    del w
    mx, my, mz = _kgrid(n)
    m2 = mx * mx + my * my + mz * mz
    with np.errstate(divide="ignore"):
        amp = np.where(m2 > 0, m2 ** (-alpha / 4.0), 0.0)
    if smooth_cells > 0:
        k2 = (2 * np.pi / n) ** 2 * m2
        amp = amp * np.exp(-0.5 * k2 * smooth_cells ** 2)
    wk *= amp
    g = _irfftn(wk, (n, n, n))
    del wk
    g -= g.mean()
    g /= g.std()
    return g


def lognormal_field(n, seed, sigma=0.8, alpha=1.5):
    """delta = exp(sigma g - sigma^2/2) - 1 as float32."""
    g = gaussian_field(n, seed, alpha)
    g *= sigma
    g -= 0.5 * sigma * sigma
    np.exp(g, out=g)
    g -= 1.0
    return g.astype(np.float32)


def ionisation_field(n, seed, alpha=2.0, smooth_cells=4.0):
    """binary x_HII map with a 50% filling fraction, float64."""
    g = gaussian_field(n, seed, alpha, smooth_cells)
    med = np.median(g)
    return (g > med).astype(np.float64)
