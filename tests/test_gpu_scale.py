"""Parity at the single-box configurations' full sizes, through the C ABI (run with -m gpu).

  256^3: 16,777,216 particles, 70 % on perturbed lattice sites, 30 % in 2000 Plummer halos.
  512^3: 134,217,728 = 2^27 particles (the most a box holds: the walk's lists pack a particle
         index << 5 in 32 bits), 40 % in 16,000 halos, deeper trees.
Both: 256^3 / 512^3 mesh, erfc split r_s = 1.25, cutoff 5.625, TSC^2 deconvolution, theta 0.5.

At these sizes the oracle cannot give every particle's exact force (a halo core holds 10^5..10^6
particles within the cutoff of each other), so each property is checked on a seeded sample:
  * theta -> 0: a sampled particle walked alone gets exactly its pair sum within the cutoff (the
    erfc law in FP64, oracle pp_direct_analytic); a missing or doubled source, a wrong image shift
    or a particle out of order shows at O(1e-3) or more;
  * PM: the PM kick of sampled particles against the oracle's float PM chain over the whole box
    (same fixed-point deposit, FP64 FFT rounded to float, same Green factor and FDA4 stencil);
  * precision: the full FP32 force (PM + tree, production groups) against the plain FP64 kernels,
    on a 1 % subsample, within the 1e-3 gate with a wide margin.
"""
import time

import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu

CASES = {256: (0.3, 2000), 512: (0.4, 16000)}


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def p99(a, b):
    d = np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)
    return float(np.percentile(d, 99))


def say(msg):
    print(msg, flush=True)


@pytest.fixture(scope="module", params=[256, 512], ids=["n256", "n512"])
def big(request):
    n = request.param
    hf, nh = CASES[n]
    t = time.perf_counter()
    say(f"[scale] n={n}: generating {n ** 3} particles ({100 * hf:.0f} % in {nh} halos)")
    pos = ics.clustered_box(n, halo_fraction=hf, nhalo=nh, seed=1)
    assert len(pos) == n ** 3
    say(f"[scale] n={n}: initial conditions {time.perf_counter() - t:.1f} s")
    yield n, pos


def test_theta_zero_sampled_targets_are_the_exact_pair_sum(ctx, big):
    n, pos = big
    ids = np.sort(np.random.default_rng(11).choice(len(pos), 256, replace=False))
    p = lagpm.make_params(n, theta=1e-6)
    t = time.perf_counter()
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        a = b.tree_accel_sample(ids)
        st = b.tree_stats()
    say(f"[scale] n={n}: sampled theta->0 walk {time.perf_counter() - t:.1f} s, nodes {st.n_nodes}")
    t = time.perf_counter()
    ref = orc.pp_direct_analytic(1, n, p.epsilon, p.efold / 2, pos, p.rsphere, targets=ids)
    err = rel_rms(a, ref)
    say(f"[scale] n={n}: oracle pair sums {time.perf_counter() - t:.1f} s; rms rel {err:.2e} p99 {p99(a, ref):.2e}")
    assert np.all(np.isfinite(a))
    assert err < 1e-4


def test_pm_kick_of_sampled_particles_matches_the_oracle_chain(ctx, big):
    n, pos = big
    ids = np.sort(np.random.default_rng(12).choice(len(pos), 4096, replace=False))
    p = lagpm.make_params(n, theta=0.5)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
        a = b.velocities()[ids]
    t = time.perf_counter()
    ref = orc.pm_accel(n, pos, mass=p.mass, efold=p.efold, deconvolve=p.deconvolve, prec=0, targets=ids)
    err = rel_rms(a, ref)
    say(f"[scale] n={n}: oracle PM chain {time.perf_counter() - t:.1f} s; PM kick rms rel {err:.2e} p99 {p99(a, ref):.2e}")
    assert err < 1e-4


def test_fp32_force_vs_plain_fp64_on_a_one_percent_subsample(ctx, big):
    n, pos = big
    sel = np.random.default_rng(13).random(len(pos)) < 0.01
    a = {}
    for prec in (lagpm.PREC_FP32, lagpm.PREC_FP64):
        t = time.perf_counter()
        r = lagpm.treepm_accel(ctx, lagpm.make_params(n, theta=0.5, precision=prec), pos)
        a[prec] = r.total[sel]
        say(f"[scale] n={n}: {'FP32' if prec == lagpm.PREC_FP32 else 'FP64'} force {time.perf_counter() - t:.1f} s")
    err, e99 = rel_rms(a[lagpm.PREC_FP32], a[lagpm.PREC_FP64]), p99(a[lagpm.PREC_FP32], a[lagpm.PREC_FP64])
    say(f"[scale] n={n}: FP32 vs FP64 on {sel.sum()} particles: rms {err:.2e} p99 {e99:.2e}")
    assert err < 1e-4 and e99 < 1e-3
