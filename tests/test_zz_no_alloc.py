"""lagpm_power_spectrum_noalloc (bins in the device's static scratch, no allocation) against
lagpm_power_spectrum (bins in a buffer it allocates), through the C ABI on the GPU.

Both run the same deposit, r2c and binning kernels; the bins are summed with double atomics whose
order varies, so the two agree to rounding (rtol 1e-12), not bit for bit."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm

pytestmark = pytest.mark.gpu


def same(a, b):
    pa, npa, ca, s8a, sla = a
    pb, npb, cb, s8b, slb = b
    assert np.allclose(npa, npb, rtol=1e-12, atol=1e-12 * np.abs(npb).max())
    assert np.allclose(pa, pb, rtol=1e-12, atol=1e-12 * np.abs(pb).max())
    assert np.allclose(ca, cb, rtol=1e-10, atol=1e-12 * np.abs(cb).max())
    assert s8a == pytest.approx(s8b, rel=1e-12)
    assert sla == pytest.approx(slb, rel=1e-10, abs=1e-12)


def test_noalloc_power_spectrum_equals_the_allocating_one_on_two_calls(ctx):
    """Two boxes; the allocating entry point first on each (it makes the FFT plan and execution
    info), then the allocation-free one on the larger box and right after on the smaller box, whose
    bins must not carry anything of the first call."""
    boxes, ref = [], []
    try:
        for n, seed in ((48, 31), (32, 32)):
            pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=20, seed=seed)
            b = lagpm.Box(ctx, lagpm.make_params(n), len(pos))
            b.upload_pos(pos)
            boxes.append(b)
            ref.append(b.power_spectrum(boxsize=100.0))
        got = [b.power_spectrum_noalloc(boxsize=100.0) for b in boxes]
        for n, g, r in zip((48, 32), got, ref):
            print(f"n={n}: sigma8 noalloc {g[3]:.15g} allocating {r[3]:.15g}", flush=True)
            assert len(g[0]) == n + 1
            same(g, r)
        # and again after the smaller box, on the larger one: the same values
        same(boxes[0].power_spectrum_noalloc(boxsize=100.0), ref[0])
    finally:
        for b in boxes:
            b.close()


def test_noalloc_power_spectrum_refuses_a_box_without_fft_setup(ctx):
    """Creating the FFT execution info allocates, so the allocation-free entry point refuses a box
    that has none yet, with a clear error and no change to the box."""
    n = 40
    pos = ics.perturbed_lattice(n, rms=0.3, seed=5)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        with pytest.raises(lagpm.LagpmError, match="execution info"):
            b.power_spectrum_noalloc(boxsize=100.0)
        r = b.power_spectrum(boxsize=100.0)
        same(b.power_spectrum_noalloc(boxsize=100.0), r)
