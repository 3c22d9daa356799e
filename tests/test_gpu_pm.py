"""GPU parity of the PM half through the C ABI against the oracle (run with -m gpu)."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def edge_positions(n):
    g = np.random.default_rng(21)
    pos = g.random((3000, 3)) * n
    pos[:8] = [[0, 0, 0], [n - 1e-12, n - 1e-12, n - 1e-12], [0.5, 1.5, 2.5], [n - 0.5, 3.0, 4.0],
               [1.25, 2.75, n - 0.25], [7.0, 7.0, 7.0], [n / 2, n / 2, n / 2], [0.4999999999, 0.5000000001, 1e-9]]
    return pos


def cases(n):
    return {
        "uniform": ics.uniform_random(20000, n, seed=5),
        "lattice": ics.perturbed_lattice(n, rms=0.3, seed=1),
        "clustered": ics.clustered_box(n, halo_fraction=0.3, nhalo=40, seed=3),
        "edges": edge_positions(n),
    }


@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
@pytest.mark.parametrize("case", ["uniform", "lattice", "clustered", "edges"])
def test_tsc_deposit_bit_exact(ctx, case, prec):
    n = 32
    pos = cases(n)[case]
    print(f"deposit case={case} prec={prec} N={len(pos)}", flush=True)
    p = lagpm.make_params(n, precision=prec)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        acc, bits = b.tsc_fixed()
        den = b.tsc().ravel()
    if prec == lagpm.PREC_FP32:
        oacc, obits = orc.tsc_fixed(n, pos, mass=1.0)
        assert bits == obits
        assert np.array_equal(acc, oacc)
        assert np.array_equal(den, orc.fixed_to_density(n, oacc, obits))
    else:
        # FP64 weights: compare with the FP32-weight oracle within float rounding
        oacc, obits = orc.tsc_fixed(n, pos, mass=1.0)
        assert np.allclose(acc * 2.0 ** -bits, oacc * 2.0 ** -obits, atol=2e-5)
    assert abs(acc.sum() * 2.0 ** -bits - len(pos)) < 1e-3


@pytest.mark.parametrize("n,npart", [(16, 50000), (64, 500)])
def test_tsc_deposit_repeated_in_one_box(ctx, n, npart):
    """Two deposits of different particles into one box: the second equals the oracle bit for bit,
    so the accumulator is fully cleared between calls, with more particles than cells (16^3 mesh,
    50000 particles) and far fewer (64^3 mesh, 500 particles)."""
    g = np.random.default_rng(n)
    print(f"repeated deposit n={n} N={npart}", flush=True)
    with lagpm.Box(ctx, lagpm.make_params(n), npart) as b:
        for rep in range(2):
            pos = g.random((npart, 3)) * n
            if rep == 0:
                pos[: npart // 2] = n / 3 + g.standard_normal((npart // 2, 3))  # a clump the second set lacks
                pos %= n
            b.upload_pos(pos)
            acc, bits = b.tsc_fixed()
            oacc, obits = orc.tsc_fixed(n, pos, mass=1.0)
            assert bits == obits
            assert np.array_equal(acc, oacc), f"deposit {rep}"


@pytest.mark.parametrize("n", [48, 96, 100])
def test_tsc_deposit_chunking_by_morton_block(ctx, n):
    """The chunk list is made per Morton block (side n / 2^l <= 13 cells: 12 at n = 48 and 96,
    12.5 at n = 100): blocks holding many chunks (a dense clump), empty blocks and a non-integer
    block side all give the oracle's sum bit for bit."""
    g = np.random.default_rng(n + 5)
    bg = g.random((3 * n * n, 3)) * n                                   # a sparse background: many empty blocks
    clump = (n / 3 + g.standard_normal((40000, 3)) * 0.6) % n           # thousands of particles in a few blocks
    edge = np.stack([np.full(2000, n - 1e-9), g.random(2000) * n, g.random(2000) * n], axis=1)  # at the face
    pos = np.concatenate([bg, clump, edge])
    print(f"deposit chunking n={n} N={len(pos)}", flush=True)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        acc, bits = b.tsc_fixed()
    oacc, obits = orc.tsc_fixed(n, pos, mass=1.0)
    assert bits == obits
    assert np.array_equal(acc, oacc)


def test_tsc_deposit_dense_halo_falls_back_and_stays_exact(ctx):
    n = 32
    g = np.random.default_rng(9)
    # one chunk spans the whole box (forces the global-atomic path) + a dense core
    pos = np.concatenate([g.random((700, 3)) * n, 16 + g.standard_normal((30000, 3)) * 0.05])
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        acc, bits = b.tsc_fixed()
    oacc, _ = orc.tsc_fixed(n, pos)
    assert np.array_equal(acc, oacc)


@pytest.mark.parametrize("n", [32, 300, 514])  # rows of 17, 151 and 258 (two segments) cells
def test_green_multiply_bit_exact(ctx, n):
    g = np.random.default_rng(22)
    spec = (g.standard_normal((n, n, n // 2 + 1)) + 1j * g.standard_normal((n, n, n // 2 + 1))).astype(np.complex64)
    for efold, dec in ((0.9, 0), (2.5, 1)):
        out = lagpm.green_apply(ctx, n, efold, dec, spec)
        ref = orc.psolver_f32(n, efold, dec, spec)
        assert np.array_equal(out, ref), (efold, dec)


@pytest.mark.parametrize("n", [32, 36, 300])  # a row in one workgroup, a partial one, two segments
def test_fda4_gradient_and_kick_bit_exact(ctx, n):
    g = np.random.default_rng(23)
    phi = np.fft.irfftn(np.fft.rfftn(g.standard_normal((n, n, n))) * np.exp(-np.arange(n // 2 + 1) / 4.0)).astype(np.float32)
    pos = edge_positions(n)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        b.set_potential(phi)
        b.fda4(lagpm.evolfact(0.75), lagpm.KICK)
        grad = b.get_gradient()
        dv = b.velocities()
    assert np.array_equal(grad, orc.fda4_gradient(n, phi))
    ref = orc.fda4_kick(n, phi, pos, fact2=0.75).astype(np.float64)
    assert np.array_equal(dv, ref)


def test_pm_factor_selection(ctx):
    """KICK uses fact2, PULL fact2_pull, PUSH fact2_push / fact1_push (Cosmos/fda4.c:49-63)."""
    n = 32
    pos = ics.uniform_random(5000, n, seed=8)
    out = {}
    for st, ef in ((lagpm.KICK, lagpm.evolfact(2.0)),
                   (lagpm.PULL, lagpm.evolfact(9.0, fact2_pull=0.5)),
                   (lagpm.PUSH, lagpm.evolfact(9.0, fact2_push=1.5, fact1_push=3.0))):
        with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
            b.upload_pos(pos)
            b.pmmain(ef, st)
            out[st] = b.velocities()
    assert np.allclose(out[lagpm.PULL] * 4.0, out[lagpm.KICK], rtol=1e-6, atol=1e-12)
    assert np.allclose(out[lagpm.PUSH] * 4.0, out[lagpm.KICK], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("mode", ["reference", "accurate"])
def test_pmmain_matches_oracle_chain(ctx, mode):
    n = 64
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=60, seed=4)
    if mode == "reference":
        p = lagpm.reference_params(n)
        efold, dec = p.efold, 0
    else:
        p = lagpm.make_params(n, efold=2.5, deconvolve=1)
        efold, dec = 2.5, 1
    print(f"pmmain mode={mode} N={len(pos)}", flush=True)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
        a = b.velocities()
    ref = orc.pm_accel(n, pos, mass=1.0, efold=efold, deconvolve=dec, prec=0)
    err = np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2))
    print(f"  rms rel vs oracle FP32-storage chain: {err:.3e}")
    assert err < 2e-5
    ref64 = orc.pm_accel(n, pos, mass=1.0, efold=efold, deconvolve=dec, prec=1)
    err64 = np.sqrt(np.sum((a - ref64) ** 2) / np.sum(ref64 ** 2))
    print(f"  rms rel vs oracle FP64 chain: {err64:.3e}")
    assert err64 < 1e-4


def test_pmmain_fp64_matches_oracle_fp64(ctx):
    n = 32
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=30, seed=6)
    p = lagpm.make_params(n, efold=2.5, deconvolve=1, precision=lagpm.PREC_FP64)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
        a = b.velocities()
    ref = orc.pm_accel(n, pos, mass=1.0, efold=2.5, deconvolve=1, prec=1)
    err = np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2))
    assert err < 1e-9, err


def test_pm_momentum_conservation_on_device(ctx):
    n = 64
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=50, seed=10)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
        a = b.velocities()
    assert np.max(np.abs(a.sum(axis=0))) < 1e-4 * np.abs(a).sum() / len(a) * np.sqrt(len(a))


def test_eunha_layout_roundtrip(ctx):
    """32-byte GOTPM+XYZDBL records (eunha.h:372-400): u64 lattice index in the low 56 bits with a
    flag byte on top, f32 displacement, f32 velocity. Position = fmod(d + i + n, n) (eunha.h:1228-1232)."""
    n = 16
    g = np.random.default_rng(8)
    npart = n ** 3
    rec = np.zeros(npart, dtype=[("u4if", "<u8"), ("d", "<f4", 3), ("v", "<f4", 3)])
    idx = g.permutation(npart).astype(np.uint64)
    flags = g.integers(0, 256, npart).astype(np.uint64)
    rec["u4if"] = idx | (flags << np.uint64(56))
    rec["d"] = (g.standard_normal((npart, 3)) * 2.0).astype(np.float32)
    rec["d"][:4] = [[-0.5, 0.0, 15.9], [-16.0, 16.0, -1e-7], [0.25, -0.75, 3.5], [-3.0, -15.99, 0.0]]
    rec["v"] = g.standard_normal((npart, 3)).astype(np.float32)
    ii = (idx % n).astype(np.float64)
    jj = ((idx // n) % n).astype(np.float64)
    kk = (idx // (n * n)).astype(np.float64)
    d = rec["d"].astype(np.float64)
    want = np.stack([np.fmod(d[:, 0] + ii + n, n), np.fmod(d[:, 1] + jj + n, n), np.fmod(d[:, 2] + kk + n, n)], 1)
    p = lagpm.make_params(n)
    with lagpm.Box(ctx, p, npart) as b:
        b.upload_eunha(rec)
        pos, vel = b.download()
        assert np.array_equal(pos, want)
        assert np.array_equal(vel, rec["v"].astype(np.float64))
        b.hubble_drag(0.5)
        out = rec.copy()
        b.download_eunha_vel(out)
    assert np.array_equal(out["u4if"], rec["u4if"]) and np.array_equal(out["d"], rec["d"])
    assert np.array_equal(out["v"], (rec["v"].astype(np.float64) * 0.5).astype(np.float32))


def test_power_spectrum_matches_the_oracle(ctx):
    """CorrMeasure on the device (deposit, r2c, binning; Cosmos/Correlation.c:13-116) against the
    oracle's FP64 FFT of the same fixed-point density."""
    n = 32
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=30, seed=17)
    p = lagpm.make_params(n)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        pk, npk, cor, s8, sl = b.power_spectrum(boxsize=64.0)
    acc, bits = orc.tsc_fixed(n, pos)
    den = orc.fixed_to_density(n, acc, bits).astype(np.float64)
    opk, onpk, ocor, os8, osl = orc.power_spectrum(n, orc.r2c(n, den), 64.0)
    print(f"power spectrum: sigma8 {s8:.6g} (oracle {os8:.6g}), slope {sl:.4g} ({osl:.4g})", flush=True)
    assert np.allclose(npk, onpk, rtol=1e-12, atol=1e-9)
    m = onpk > 0
    assert np.allclose(pk[m], opk[m], rtol=2e-5, atol=1e-6 * np.abs(opk).max())
    assert s8 == pytest.approx(os8, rel=1e-5)
    assert np.allclose(cor, ocor, rtol=1e-4, atol=1e-5 * np.abs(ocor).max())
