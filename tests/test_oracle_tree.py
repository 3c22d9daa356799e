"""Pins of the oracle's short-range half: force tables, direct pair sum, the reference tree
walk, the Ewald sum and the kick factors. CPU only."""
import numpy as np
import pytest
from scipy import integrate
from scipy.special import erfc

from lageunha_amd import ics
from oracle import pyoracle as orc


def rel_rms(a, b):
    return np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2))


# ---------------- force tables ----------------
@pytest.mark.parametrize("law", [0, 1])
def test_table_nodes_equal_law_over_r(law):
    n, eps, rs = 64, 0.1, 1.25
    diff, slope = orc.force_spline(law, n, eps, rs)
    ran = 6.0 / 4096
    for i in (1, 7, 100, 1000, 3000, 4000):
        r = i * ran
        fr = orc.pmcf(n, eps, r) if law == 0 else orc.erfc_fr(n, eps, rs, r)
        assert np.isclose(diff[i, 0], -fr / r, rtol=1e-14)
    assert np.all(diff[4095] == 0) and np.all(slope[4095] == 0)
    assert np.all(diff[0] == diff[1]) and np.all(slope[0] == slope[1])


def test_treeforce_variant_table():
    """ForceLaw/treeforce.c:134-186: the fitted law with the softening fixed at EPSILON = 0.1 (the
    Epsilon argument is not used by its pmfit), derivatives by central differences of step 0.2 bin,
    entry 0 zero with the slope up to entry 1."""
    n = 64
    ran = 6.0 / 4096
    d2, s2 = orc.force_spline(2, n, 0.1, 0.45)
    d2b, s2b = orc.force_spline(2, n, 0.37, 0.45)
    assert np.array_equal(d2, d2b) and np.array_equal(s2, s2b)  # eps is not an input of this law
    assert np.all(d2[0] == 0) and np.allclose(s2[0], d2[1] / ran, rtol=1e-15)
    assert np.all(d2[4095] == 0) and np.all(s2[4095] == 0)
    for i in (1, 50, 700, 2600):
        r = i * ran
        assert np.isclose(d2[i, 0], -orc.pmcf(n, 0.1, r) / r, rtol=1e-14)
    # the derivative step: a central difference of step h misses f' by h^2 f'''(r) / 6, so the
    # step-0.2 table's f' error is four times the step-0.1 table's (f' recovered from diff1 and f)
    d0, _ = orc.force_spline(0, n, 0.1, 0.45)

    def fprime_from_table(d, i):
        r = i * ran
        fr = orc.pmcf(n, 0.1, r)
        return (fr - 2 * r ** 3 * d[i, 1]) / r

    for i in (200, 400, 800):
        r, h = i * ran, 1e-3
        f = lambda x: orc.pmcf(n, 0.1, x)  # noqa: E731
        exact = (8 * (f(r + h) - f(r - h)) - (f(r + 2 * h) - f(r - 2 * h))) / (12 * h)  # O(h^4)
        ratio = (fprime_from_table(d2, i) - exact) / (fprime_from_table(d0, i) - exact)
        assert abs(ratio - 4.0) < 0.05, (i, ratio)
    # the step parameter with the reference law's entry rules gives the reference table
    d0b, _ = orc.force_spline(0, n, 0.1, 0.45, dr_bins=0.1)
    assert np.array_equal(d0b, d0)


def test_erfc_law_is_newton_times_split_complement():
    n, rs = 64, 1.25
    for r in (0.5, 1.0, 2.0, 4.0):
        u = r / (2 * rs)
        split = erfc(u) + r / (rs * np.sqrt(np.pi)) * np.exp(-u * u)
        assert np.isclose(orc.erfc_fr(n, 0.0, rs, r) * n * n * r * r, split, rtol=1e-12)


def test_table_derivative_terms_match_finite_differences():
    """diff1 = -(f' r - f)/(2 r^3), diff2 = -(f'' r^2 - 3 f' r + 3 f)/(2 r^5) with f = law(r)."""
    n, eps, rs = 64, 0.1, 1.25
    diff, _ = orc.force_spline(1, n, eps, rs)
    ran = 6.0 / 4096
    for i in (300, 1200, 2500):
        r = i * ran
        h = 1e-4
        f = lambda x: orc.erfc_fr(n, eps, rs, x)  # noqa: E731
        fp = (f(r + h) - f(r - h)) / (2 * h)
        fpp = (f(r + h) - 2 * f(r) + f(r - h)) / h ** 2
        d1 = -(fp * r - f(r)) / (2 * r ** 3)
        d2 = -(fpp * r * r - 3 * fp * r + 3 * f(r)) / (2 * r ** 5)
        assert np.isclose(diff[i, 1], d1, rtol=1e-5)
        assert np.isclose(diff[i, 2], d2, rtol=1e-3)


def test_point_grav_interpolates_the_law():
    n, eps, rs = 64, 0.1, 1.25
    diff, slope = orc.force_spline(1, n, eps, rs)
    for r in np.linspace(0.2, 5.5, 37):
        got = orc.point_grav(diff, slope, float(np.float32(r)))
        r32 = float(np.float32(r))
        want = -orc.erfc_fr(n, eps, rs, r32) / r32
        assert abs(got - want) < 2e-4 * abs(want)


def test_pmcf_vanishes_beyond_the_reference_cutoff():
    n = 64
    newton = 1.0 / (4.0 ** 2) / n ** 2
    assert abs(orc.pmcf(n, 0.1, 4.0)) < 0.05 * newton
    assert abs(orc.pmcf(n, 0.1, 0.05)) > 0.5 * 0.05 / (0.05 ** 2 + 0.01) ** 1.5 / n ** 2


# ---------------- direct sum ----------------
def test_direct_table_sum_matches_analytic_law_and_momentum():
    n, eps, rs, rc = 32, 0.1, 1.25, 5.625
    pos = ics.uniform_random(3000, n, seed=7)
    diff, slope = orc.force_spline(1, n, eps, rs)
    a_tab = orc.pp_direct(n, pos, diff, slope, rc)
    a_ana = orc.pp_direct_analytic(1, n, eps, rs, pos, rc)
    assert rel_rms(a_tab, a_ana) < 2e-4
    assert np.max(np.abs(a_ana.sum(axis=0))) < 1e-12 * np.abs(a_ana).sum()


# ---------------- reference tree walk ----------------
def clustered(n, N, seed):
    g = np.random.default_rng(seed)
    half = N // 2
    a = g.random((half, 3)) * n
    c = g.random((8, 3)) * n
    b = (c[g.integers(0, 8, N - half)] + g.standard_normal((N - half, 3)) * 0.7) % n
    return np.concatenate([a, b])


@pytest.mark.parametrize("law,rc,cellw", [(0, 4.0, 4.0), (1, 5.625, 5.625)])
def test_reference_walk_theta_to_zero_is_the_direct_sum(law, rc, cellw):
    n = 24 if law == 0 else 34
    pos = clustered(n, 12000, 11)
    diff, slope = orc.force_spline(law, n, 0.1, 1.25)
    direct = orc.pp_direct(n, pos, diff, slope, rc)
    tree = orc.treecorrection_ref(n, pos, diff, slope, rc, 1e-6, cellw=cellw)
    assert rel_rms(tree, direct) < 1e-5


def test_reference_walk_theta_half_error_is_small_and_d1_is_large():
    n, rc = 24, 4.0
    pos = clustered(n, 12000, 12)
    diff, slope = orc.force_spline(0, n, 0.1, 1.25)
    direct = orc.pp_direct(n, pos, diff, slope, rc)
    fixed = orc.treecorrection_ref(n, pos, diff, slope, rc, 0.5)
    e_fixed = rel_rms(fixed, direct)
    quirky = orc.treecorrection_ref(n, pos, diff, slope, rc, 0.5, quirks=1)
    e_quirk = rel_rms(quirky, direct)
    assert e_fixed < 2e-2
    assert e_quirk > 3 * e_fixed  # D1 drops y/z node forces and triples x


def test_node_quadrupole_error_is_third_order():
    diff, slope = orc.force_spline(1, 64, 0.0, 1.25)
    g = np.random.default_rng(13)
    src = g.standard_normal((40, 3)) * 0.08
    m = np.ones(40, dtype=np.float32)
    errs = []
    for d in (0.8, 1.6):
        t = np.array([d, 0.3 * d, -0.2 * d])
        node = orc.node_accel(diff, slope, 5.9, src, m, t)
        exact = np.zeros(3)
        for s in src:
            dv = t - s
            r = np.linalg.norm(dv)
            exact += dv * (-orc.erfc_fr(64, 0.0, 1.25, r) / r)
        errs.append(np.linalg.norm(node - exact) / np.linalg.norm(exact))
    assert errs[1] < errs[0] / 4  # O((s/d)^3): 8x for 2x distance, allow table noise
    assert errs[0] < 2e-3


# ---------------- Ewald ----------------
def test_ewald_alpha_independent_and_antisymmetric():
    L, G = 10.0, 1.0
    pos = np.random.default_rng(14).random((12, 3)) * L
    m = np.random.default_rng(15).random(12) + 0.5
    a1 = orc.ewald_accel(pos, m, L, G, alpha=0.35, nreal=3, kmax=12)
    a2 = orc.ewald_accel(pos, m, L, G, alpha=0.5, nreal=3, kmax=14)
    assert rel_rms(a1, a2) < 1e-9
    assert np.max(np.abs((m[:, None] * a1).sum(axis=0))) < 1e-10


def test_ewald_lattice_is_force_free_and_close_pair_is_newtonian():
    L = 8.0
    g = np.arange(2) * 4.0
    pos = np.array([[a, b, c] for a in g for b in g for c in g]) + 0.3
    a = orc.ewald_accel(pos, np.ones(8), L, 1.0, alpha=0.6, nreal=3, kmax=12)
    assert np.max(np.abs(a)) < 1e-11
    pair = np.array([[1.0, 1.0, 1.0], [1.01, 1.0, 1.0]])
    a = orc.ewald_accel(pair, np.ones(2), L, 1.0, alpha=0.6, nreal=3, kmax=12)
    assert np.isclose(a[0, 0], 1.0 / 0.01 ** 2, rtol=1e-4)


# ---------------- kick factors ----------------
def test_evolfactor_einstein_de_sitter_closed_form():
    n, amax = 256, 48.0
    for a1, a2 in ((1.0, 1.25), (3.0, 3.125), (20.0, 20.25)):
        f2 = orc.fact2(1.0, 0.0, amax, n, a1, a2)
        closed = 3 * n / (4 * np.pi) * (np.sqrt(a2) - np.sqrt(a1)) * np.sqrt(a2) / a2 ** 2
        assert np.isclose(f2, closed, rtol=1e-12)
        assert np.isclose(orc.fact1(1.0, 0.0, amax, a1, a2), (a1 / a2) ** 1.5, rtol=1e-13)


def test_evolfactor_lcdm_quadrature():
    om, ol, amax, n = 0.26, 0.74, 48.0, 128

    def ap(a):
        zp1 = amax / a
        h2 = zp1 ** 3 * om + zp1 ** 2 * (1 - om - ol) + ol
        return np.sqrt(8 * np.pi / 3 / (amax * om) / zp1 ** 2 * h2)

    a1, a2 = 2.0, 2.25
    val, _ = integrate.quad(lambda a: 1 / (a * ap(a)), a1, a2, epsabs=0, epsrel=1e-13)
    assert np.isclose(orc.fact2(om, ol, amax, n, a1, a2), n / (a2 ** 2 * ap(a2)) * val, rtol=1e-11)
    ef = orc.evolfactor(om, ol, amax, n, 2.0, 0.25)
    assert np.isclose(ef.fact2, orc.fact2(om, ol, amax, n, 1.875, 2.125), rtol=1e-14)
    assert np.isclose(ef.pfact, 0.25 * n)
