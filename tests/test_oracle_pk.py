"""The CorrMeasure restatement (Cosmos/Correlation.c:13-116), pinned by properties of the
estimator: linear bin weights sum to one per mode, a single Fourier mode lands in the two bins
around |k| with the closed-form amplitude, and sigma_8 and xi(r) agree with the same quantities
measured directly in real space on a Gaussian random field. CPU only."""
import numpy as np
import pytest

from oracle import pyoracle as orc


def tsc_window(n):
    k = np.fft.fftfreq(n) * n
    x = np.pi * k / n
    s = np.where(k == 0, 1.0, np.sin(x) / np.where(k == 0, 1, x)) ** 3
    return s


def spectrum_of(den):
    return np.fft.rfftn(den)  # [k][j][i], i <= n/2, unnormalised like FFTW/rocFFT


def test_bin_weights_sum_to_one_per_mode():
    n = 16
    g = np.random.default_rng(1)
    spec = spectrum_of(g.standard_normal((n, n, n)))
    pk, npk, cor, s8, sl = orc.power_spectrum(n, spec, boxsize=64.0)
    modes = n * n * (n // 2)  # i < n/2, all j, k
    # modes with |k| = 0 put weight 0 on bin -1 (dropped), all others put 1 in bins <= n
    assert np.isclose(npk.sum(), modes, rtol=0, atol=1e-9)


@pytest.mark.parametrize("mode", [(3, 0, 0), (2, 1, 2), (1, 5, 3)])
def test_single_mode_lands_in_the_two_neighbouring_bins(mode):
    n = 16
    i, j, k = mode
    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    A = 0.7
    den = A * np.cos(2 * np.pi * (i * x + j * y + k * z) / n)
    pk, npk, cor, s8, sl = orc.power_spectrum(n, spectrum_of(den), boxsize=100.0)
    W = tsc_window(n)
    # |delta_k|^2 / n^6 = A^2 / 4 at the mode (the partner at -k has i < 0 or i = 0 with j, k mirrored)
    d = (A * A / 4) / (W[i] * W[j] * W[k]) ** 2
    mirrored = i == 0  # then (0, -j, -k) is stored too and counts as a second mode
    wave = np.sqrt(i * i + j * j + k * k)
    b1 = int(wave)
    t = wave - b1
    got = pk * npk  # un-normalised bin sums
    want = np.zeros(n + 1)
    want[b1] += (1 - t) * d * (2 if mirrored else 1)
    if t > 0:
        want[b1 + 1] += t * d * (2 if mirrored else 1)
    assert np.allclose(got[1:], want[1:], rtol=1e-9, atol=1e-12 * d)


def gaussian_field(n, seed, index=-1.5):
    g = np.random.default_rng(seed)
    kz = np.fft.fftfreq(n)[:, None, None]
    ky = np.fft.fftfreq(n)[None, :, None]
    kx = np.fft.rfftfreq(n)[None, None, :]
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    amp = np.where(kk > 0, np.power(np.where(kk > 0, kk, 1.0), index / 2), 0.0)
    w = np.fft.rfftn(g.standard_normal((n, n, n))) * amp
    return np.fft.irfftn(w, s=(n, n, n), axes=(0, 1, 2))


def test_sigma8_matches_the_real_space_top_hat_variance():
    """sigma^2 = variance of the field smoothed with a sphere of radius 8 Mpc/h, computed in real
    space on the grid (the window applied to the TSC-deconvolved field)."""
    n, boxsize = 64, 128.0
    den = gaussian_field(n, 3)
    spec = spectrum_of(den)
    pk, npk, cor, s8, sl = orc.power_spectrum(n, spec, boxsize)
    W = tsc_window(n)
    deconv = spec / (W[:, None, None] * W[None, :, None] * W[None, None, : n // 2 + 1])
    rth = 8.0 / boxsize * n
    c = np.fft.fftfreq(n) * n
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    ball = (xx * xx + yy * yy + zz * zz <= rth * rth).astype(float)
    ball /= ball.sum()
    smooth = np.fft.irfftn(deconv * np.fft.rfftn(ball), s=(n, n, n), axes=(0, 1, 2))
    assert s8 == pytest.approx(smooth.std(), rel=0.05)


def test_correlation_function_matches_the_real_space_pair_average():
    """xi(r) from the binned spectrum vs <delta(x) delta(x + r)> averaged over a shell of radius
    r on the grid, for a field whose power is concentrated below k = 4 (the estimator's sine sum
    over noisy high-k shells would otherwise dominate)."""
    n = 64
    g = np.random.default_rng(5)
    kz = np.fft.fftfreq(n)[:, None, None] * n
    ky = np.fft.fftfreq(n)[None, :, None] * n
    kx = np.fft.rfftfreq(n)[None, None, :] * n
    kk = np.sqrt(kx * kx + ky * ky + kz * kz)
    w = np.fft.rfftn(g.standard_normal((n, n, n))) * np.exp(-((kk / 4.0) ** 2) / 2)
    den = np.fft.irfftn(w, s=(n, n, n), axes=(0, 1, 2))
    spec = spectrum_of(den)
    pk, npk, cor, s8, sl = orc.power_spectrum(n, spec, 100.0)
    W = tsc_window(n)
    deconv = spec / (W[:, None, None] * W[None, :, None] * W[None, None, : n // 2 + 1])
    xi_grid = np.fft.irfftn(np.abs(deconv) ** 2, s=(n, n, n), axes=(0, 1, 2)) / n ** 3  # <delta(x) delta(x+r)>
    c = np.fft.fftfreq(n) * n
    zz, yy, xx = np.meshgrid(c, c, c, indexing="ij")
    r = np.sqrt(xx * xx + yy * yy + zz * zz)
    for ri in (1, 2, 3, 4, 6):
        shell = np.abs(r - ri) < 0.5
        assert cor[ri] == pytest.approx(xi_grid[shell].mean(), rel=0.1)
