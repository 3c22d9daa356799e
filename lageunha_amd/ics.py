"""Synthetic initial conditions for the TreePM workloads (numpy, FP64 positions in cells).

Random numbers come from numpy's Philox4x64 counter-based generator, seeded per workload.

perturbed_lattice   lattice sites displaced by a Zel'dovich field psi(k) = i k delta(k) / k^2 of
                    a Gaussian field with P(k) ~ k^index; the 3-D rms |psi| is set to `rms`.
plummer_halos       Plummer spheres (scale radius a, truncated at 10 a) with particle counts drawn
                    from dN/dM ~ M^-alpha on [mmin, mmax], rescaled to the requested total.
clustered_box       the clustered mix: a fraction of seeded-shuffled lattice sites perturbed as
                    above, the rest in Plummer halos with uniform centres.
"""
from __future__ import annotations

import numpy as np
from scipy import fft as sfft


def rng(seed: int) -> np.random.Generator:
    return np.random.Generator(np.random.Philox(seed))


def zeldovich_displacement(n: int, rms: float, seed: int, index: float = -1.0) -> np.ndarray:
    """Displacement field (3, n, n, n) [z][y][x] with rms |psi| = rms."""
    g = rng(seed)
    white = g.standard_normal((n, n, n))
    wk = sfft.rfftn(white, workers=-1)
    kz = np.fft.fftfreq(n) * 2 * np.pi
    ky = np.fft.fftfreq(n) * 2 * np.pi
    kx = np.fft.rfftfreq(n) * 2 * np.pi
    KZ, KY, KX = np.meshgrid(kz, ky, kx, indexing="ij")
    k2 = KX ** 2 + KY ** 2 + KZ ** 2
    k2[0, 0, 0] = 1.0
    amp = np.power(k2, index / 4.0)          # sqrt(P(k)) with P ~ k^index
    amp[0, 0, 0] = 0.0
    delta = wk * amp
    psi = np.empty((3, n, n, n))
    for a, K in enumerate((KX, KY, KZ)):
        psi[a] = sfft.irfftn(1j * K * delta / k2, s=(n, n, n), axes=(0, 1, 2), workers=-1)
    cur = np.sqrt(np.mean(np.sum(psi ** 2, axis=0)))
    return psi * (rms / cur)


def perturbed_lattice(n: int, rms: float = 0.3, seed: int = 1, index: float = -1.0,
                      box: float | None = None) -> np.ndarray:
    """n^3 particles on lattice sites (i + 1/2 not used: sites at integers, as EUNHA's XofP)."""
    box = float(n) if box is None else box
    psi = zeldovich_displacement(n, rms, seed, index)
    k, j, i = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    s = box / n
    pos = np.stack([(i + psi[0]) * s, (j + psi[1]) * s, (k + psi[2]) * s], axis=-1).reshape(-1, 3)
    return np.mod(pos, box)


def halo_counts(nhalo: int, total: int, alpha: float, mmin: float, mmax: float, g: np.random.Generator) -> np.ndarray:
    u = g.random(nhalo)
    e = 1.0 - alpha
    m = (mmin ** e + u * (mmax ** e - mmin ** e)) ** (1.0 / e)
    m *= total / m.sum()
    c = np.floor(m).astype(np.int64)
    short = total - int(c.sum())
    if short > 0:
        frac = m - c
        order = np.argsort(-frac, kind="stable")
        c[order[:short]] += 1
    return c


def plummer_halos(box: float, total: int, nhalo: int = 2000, a: float = 0.5, alpha: float = 1.9,
                  mmin: float = 32, mmax: float = 1e5, seed: int = 2) -> np.ndarray:
    g = rng(seed)
    counts = halo_counts(nhalo, total, alpha, mmin, mmax, g)
    centres = g.random((nhalo, 3)) * box
    umax = 1000.0 / 101.0 ** 1.5            # enclosed mass fraction at r = 10 a
    u = g.random(total) * umax
    u = np.clip(u, 1e-300, None)
    r = a / np.sqrt(u ** (-2.0 / 3.0) - 1.0)
    cost = 2 * g.random(total) - 1
    phi = 2 * np.pi * g.random(total)
    sint = np.sqrt(1 - cost ** 2)
    d = np.stack([r * sint * np.cos(phi), r * sint * np.sin(phi), r * cost], axis=1)
    which = np.repeat(np.arange(nhalo), counts)
    return np.mod(centres[which] + d, box)


def clustered_box(n: int, halo_fraction: float = 0.3, nhalo: int = 2000, rms: float = 0.3,
                  seed: int = 1) -> np.ndarray:
    """n^3 particles in a box of n cells: (1 - f) perturbed lattice sites + f in Plummer halos."""
    N = n ** 3
    nh = int(round(halo_fraction * N))
    lat = perturbed_lattice(n, rms=rms, seed=seed)
    keep = rng(seed + 1000).permutation(N)[: N - nh]
    keep.sort()
    parts = [lat[keep]]
    if nh:
        parts.append(plummer_halos(float(n), nh, nhalo=nhalo, seed=seed + 2000))
    return np.concatenate(parts, axis=0)


def uniform_random(N: int, box: float, seed: int = 3) -> np.ndarray:
    return rng(seed).random((N, 3)) * box
