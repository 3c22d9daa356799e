"""Total TreePM force accuracy of the GPU path, stated against physics and against the reference
algorithm (run with -m gpu).

Yardsticks (both in the oracle, FP64):
  * Ewald: the periodic force of Plummer-softened point masses (eps = 0.1 cell), the law that the
    PM + tree split reproduces (long range: Newtonian; short range: Plummer), summed by Ewald's
    method to ~1e-9.
  * Reference CPU TreePM: the oracle's restatement of the reference path on the same particles and
    parameters -- the float PM chain (TSC deposit, FFT, Green, FDA4 gradient and TSC
    interpolation, Cosmos/pmmain.c:390-527) plus the per-cell octree walk of the tree correction
    (Cosmos/treecorrection.c:225-475, OST/ost.ver2.c).

Errors are per-particle relative errors |a - a_ref| / |a_ref|, stated as the rms over particles
(sqrt(sum |a - a_ref|^2 / sum |a_ref|^2)) and the 99th percentile."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def errors(a, ref):
    d = np.linalg.norm(a - ref, axis=1)
    r = np.linalg.norm(ref, axis=1)
    return float(np.sqrt(np.sum(d ** 2) / np.sum(r ** 2))), float(np.percentile(d / np.maximum(r, 1e-300), 99))


def gpu_total(ctx, p, pos):
    return lagpm.treepm_accel(ctx, p, pos).total


def oracle_total(n, p, pos, theta):
    """The reference CPU TreePM (restated) with the box's law, split, cutoff and opening angle."""
    pm = orc.pm_accel(n, pos, mass=p.mass, efold=p.efold, deconvolve=p.deconvolve, prec=0)
    diff, slope = orc.force_spline(p.law, n, p.epsilon, p.efold / 2)
    tr = orc.treecorrection_ref(n, pos, diff, slope, p.rsphere, theta, cellw=max(4.0, p.rsphere), mass=p.mass)
    return pm + tr


@pytest.fixture(scope="module")
def clustered32():
    n = 32
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=16, seed=3)
    ewald = orc.ewald_converged(pos, n, eps=0.1)
    return n, pos, ewald


@pytest.mark.parametrize("theta", [0.5, 1e-6])
def test_total_force_vs_ewald_at_least_as_accurate_as_the_reference_algorithm(ctx, clustered32, theta):
    """Accurate mode (erfc split r_s = 1.25, TSC^2 deconvolution, cutoff 5.625): the GPU TreePM total
    force is no less accurate against Ewald than the reference algorithm with the same parameters.
    Measured (32^3 clustered, 30 % in 16 halos): theta = 0.5: rms ~1.6e-3; theta -> 0: ~4.4e-4, the
    split's own error (mesh anisotropy, truncation), independent of the tree."""
    n, pos, ewald = clustered32
    p = lagpm.make_params(n, theta=theta)
    a = gpu_total(ctx, p, pos)
    ref = oracle_total(n, p, pos, theta)
    e_gpu, e_ref = errors(a, ewald), errors(ref, ewald)
    print(f"theta={theta}: GPU vs Ewald rms {e_gpu[0]:.3e} p99 {e_gpu[1]:.3e}; reference CPU TreePM vs Ewald "
          f"rms {e_ref[0]:.3e} p99 {e_ref[1]:.3e}", flush=True)
    # group-conservative opening: every target sees a stricter test than the reference's
    assert e_gpu[0] <= 1.02 * e_ref[0]
    assert e_gpu[1] <= 1.05 * e_ref[1]
    assert e_gpu[0] < (3e-3 if theta > 0.1 else 1e-3)


def test_total_force_vs_ewald_reference_parameters(ctx, clustered32):
    """Reference mode (the reference's fitted short-range law, efold 0.9, RSPHERE 4, no
    deconvolution): same comparison. The fitted law is the less accurate split (rms ~3e-3)."""
    n, pos, ewald = clustered32
    p = lagpm.reference_params(n, theta=0.5)
    a = gpu_total(ctx, p, pos)
    ref = oracle_total(n, p, pos, 0.5)
    e_gpu, e_ref = errors(a, ewald), errors(ref, ewald)
    print(f"reference mode: GPU vs Ewald rms {e_gpu[0]:.3e} p99 {e_gpu[1]:.3e}; reference CPU TreePM "
          f"rms {e_ref[0]:.3e} p99 {e_ref[1]:.3e}", flush=True)
    assert e_gpu[0] <= 1.02 * e_ref[0]
    assert e_gpu[1] <= 1.05 * e_ref[1]
    assert e_gpu[0] < 6e-3


@pytest.mark.parametrize("mode", ["accurate", "reference"])
def test_total_force_vs_the_reference_cpu_treepm(ctx, clustered32, mode):
    """GPU vs the reference CPU TreePM on the same particles and parameters. At theta -> 0 both are
    the same sums (PM chain bit-exact up to the FFT; tree = every pair within the cutoff), so the
    difference is FP32 rounding and table interpolation; at theta = 0.5 the trees differ (one global
    octree with group opening vs a per-cell octree with per-target opening), and the two differ by
    about the reference walk's own approximation error, no more than its error against Ewald."""
    n, pos, ewald = clustered32
    for theta in (1e-6, 0.5):
        p = lagpm.make_params(n, theta=theta) if mode == "accurate" else lagpm.reference_params(n, theta=theta)
        a = gpu_total(ctx, p, pos)
        ref = oracle_total(n, p, pos, theta)
        rms, p99 = errors(a, ref)
        print(f"{mode} theta={theta}: GPU vs reference CPU TreePM rms {rms:.3e} p99 {p99:.3e}", flush=True)
        if theta < 0.1:
            assert rms < 5e-5 and p99 < 3e-4
        else:
            assert rms < 1.2 * errors(ref, ewald)[0]
