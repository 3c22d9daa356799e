"""ctypes binding of the CPU oracle (oracle/liblagpm_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg, as the checker or the timed CPU baseline — never by the product path.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblagpm_oracle.so")

_D = ctypes.POINTER(ctypes.c_double)
_F = ctypes.POINTER(ctypes.c_float)
_I64 = ctypes.POINTER(ctypes.c_int64)
_lib = None


class OrcEvolFact(ctypes.Structure):
    _fields_ = [(k, ctypes.c_double) for k in (
        "afact", "bfact", "pfact", "fact1", "fact2", "fact1_push", "fact2_push", "fact1_pull", "fact2_pull")]


def build():
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        build()
    L = ctypes.CDLL(LIB_PATH)
    sz, i, d, f = ctypes.c_size_t, ctypes.c_int, ctypes.c_double, ctypes.c_float
    sig = {
        "orc_num_threads": (i, []), "orc_set_threads": (None, [i]),
        "orc_tsc_fixed": (None, [i, sz, _D, _D, _D, f, i, _I64]),
        "orc_tsc_float": (None, [i, sz, _D, _D, _D, f, _F]),
        "orc_fixed_scale_bits": (i, [d]),
        "orc_fixed_to_density": (None, [i, _I64, i, _F]),
        "orc_fft3d": (None, [i, _D, i]),
        "orc_r2c": (None, [i, _D, _D]), "orc_c2r": (None, [i, _D, _D]),
        "orc_psolver_f32": (None, [i, d, i, _F]), "orc_psolver_f64": (None, [i, d, i, _D]),
        "orc_fda4_gradient": (None, [i, _F, _F]),
        "orc_fda4_kick": (None, [i, _F, sz, _D, _D, _D, f, _F, _F, _F]),
        "orc_fda4_kick_f64": (None, [i, _D, sz, _D, _D, _D, d, _D]),
        "orc_pm_accel": (None, [i, sz, _D, _D, _D, d, d, i, i, _D]),
        "orc_pm_accel_timed": (None, [i, sz, _D, _D, _D, d, d, i, i, _D, _D]),
        "orc_pm_accel_at": (None, [i, sz, _D, _D, _D, d, d, i, i, sz, _I64, _D]),
        "orc_force_spline": (None, [i, i, d, d, _D, _D]),
        "orc_force_spline_dr": (None, [i, i, d, d, d, _D, _D]),
        "orc_pmcf": (d, [i, d, d]), "orc_erfc_fr": (d, [i, d, d, d]),
        "orc_point_grav": (f, [_D, _D, f]),
        "orc_pp_direct": (None, [i, sz, _D, _D, _D, f, _D, _D, d, sz, _I64, _D]),
        "orc_pp_direct_analytic": (None, [i, i, d, d, sz, _D, _D, _D, d, d, sz, _I64, _D]),
        "orc_treecorrection_ref": (i, [i, sz, _D, _D, _D, f, _D, _D, d, d, d, f, i, _F]),
        "orc_treecorrection_ref_sel": (i, [i, sz, _D, _D, _D, f, _D, _D, d, d, d, f, i,
                                           ctypes.POINTER(ctypes.c_uint8), _F]),
        "orc_node_accel": (None, [_D, _D, d, sz, _D, _D, _D, _F, _D, i, _F]),
        "orc_ewald_accel": (None, [sz, _D, _D, _D, _D, d, d, d, i, i, _D]),
        "orc_ewald_accel_soft": (None, [sz, _D, _D, _D, _D, d, d, d, i, i, d, _D]),
        "orc_hofez": (d, [d, d, d, d]), "orc_ap": (d, [d, d, d, d]),
        "orc_fact1": (d, [d, d, d, d, d]), "orc_fact2": (d, [d, d, d, i, d, d]),
        "orc_evolfactor": (None, [d, d, d, i, d, d, ctypes.POINTER(OrcEvolFact)]),
        "orc_power_spectrum": (None, [i, _D, d, _D, _D, _D, _D, _D]),
        "orc_prefof": (sz, [i, sz, _D, _D, _D, f, i, ctypes.POINTER(ctypes.c_uint8)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _dp(a):
    return a.ctypes.data_as(_D)


def set_threads(n: int):
    lib().orc_set_threads(int(n))


def num_threads() -> int:
    return lib().orc_num_threads()


def split(pos):
    pos = np.asarray(pos, dtype=np.float64)
    return _d(pos[:, 0]), _d(pos[:, 1]), _d(pos[:, 2])


def scale_bits(total_mass: float) -> int:
    return lib().orc_fixed_scale_bits(total_mass)


def tsc_fixed(n, pos, mass=1.0, bits=None):
    x, y, z = split(pos)
    if bits is None:
        bits = scale_bits(len(x) * mass)
    acc = np.empty(n ** 3, dtype=np.int64)
    lib().orc_tsc_fixed(n, len(x), _dp(x), _dp(y), _dp(z), mass, bits, acc.ctypes.data_as(_I64))
    return acc, bits


def fixed_to_density(n, acc, bits):
    den = np.empty(n ** 3, dtype=np.float32)
    acc = np.ascontiguousarray(acc, dtype=np.int64)
    lib().orc_fixed_to_density(n, acc.ctypes.data_as(_I64), bits, den.ctypes.data_as(_F))
    return den


def tsc_float(n, pos, mass=1.0):
    x, y, z = split(pos)
    den = np.empty(n ** 3, dtype=np.float32)
    lib().orc_tsc_float(n, len(x), _dp(x), _dp(y), _dp(z), mass, den.ctypes.data_as(_F))
    return den


def r2c(n, real):
    real = _d(real).ravel()
    out = np.empty(2 * n * n * (n // 2 + 1))
    lib().orc_r2c(n, _dp(real), _dp(out))
    return out.view(np.complex128).reshape(n, n, n // 2 + 1)


def c2r(n, spec):
    spec = np.ascontiguousarray(spec, dtype=np.complex128)
    out = np.empty(n ** 3)
    lib().orc_c2r(n, spec.view(np.float64).ctypes.data_as(_D), _dp(out))
    return out


def psolver_f32(n, efold, deconvolve, spec):
    s = np.ascontiguousarray(spec, dtype=np.complex64).copy()
    lib().orc_psolver_f32(n, efold, deconvolve, s.view(np.float32).ctypes.data_as(_F))
    return s


def psolver_f64(n, efold, deconvolve, spec):
    s = np.ascontiguousarray(spec, dtype=np.complex128).copy()
    lib().orc_psolver_f64(n, efold, deconvolve, s.view(np.float64).ctypes.data_as(_D))
    return s


def fda4_gradient(n, phi):
    phi = np.ascontiguousarray(phi, dtype=np.float32).ravel()
    g = np.empty(4 * n ** 3, dtype=np.float32)
    lib().orc_fda4_gradient(n, phi.ctypes.data_as(_F), g.ctypes.data_as(_F))
    return g.reshape(-1, 4)


def fda4_kick(n, phi, pos, fact2=1.0):
    phi = np.ascontiguousarray(phi, dtype=np.float32).ravel()
    x, y, z = split(pos)
    dv = np.empty((3, len(x)), dtype=np.float32)
    lib().orc_fda4_kick(n, phi.ctypes.data_as(_F), len(x), _dp(x), _dp(y), _dp(z), fact2,
                        dv[0].ctypes.data_as(_F), dv[1].ctypes.data_as(_F), dv[2].ctypes.data_as(_F))
    return dv.T.copy()


def fda4_kick_f64(n, phi, pos, fact2=1.0):
    phi = _d(phi).ravel()
    x, y, z = split(pos)
    dv = np.empty(3 * len(x))
    lib().orc_fda4_kick_f64(n, _dp(phi), len(x), _dp(x), _dp(y), _dp(z), fact2, _dp(dv))
    return dv.reshape(-1, 3)


def pm_accel(n, pos, mass=1.0, efold=0.9, deconvolve=0, prec=0, targets=None):
    """PM acceleration of every particle, or (targets: indices) of those particles only, from the
    mesh of all of them."""
    x, y, z = split(pos)
    if targets is not None:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        acc = np.empty(3 * len(t))
        lib().orc_pm_accel_at(n, len(x), _dp(x), _dp(y), _dp(z), mass, efold, deconvolve, prec, len(t),
                              t.ctypes.data_as(_I64), _dp(acc))
        return acc.reshape(-1, 3)
    acc = np.empty(3 * len(x))
    lib().orc_pm_accel(n, len(x), _dp(x), _dp(y), _dp(z), mass, efold, deconvolve, prec, _dp(acc))
    return acc.reshape(-1, 3)


def pm_accel_timed(n, pos, mass=1.0, efold=0.9, deconvolve=0, prec=0):
    """pm_accel with the wall time of each phase: {TSC, F-FFT, psolv, B-FFT, 4FDA} in seconds."""
    x, y, z = split(pos)
    acc = np.empty(3 * len(x))
    sec = np.zeros(5)
    lib().orc_pm_accel_timed(n, len(x), _dp(x), _dp(y), _dp(z), mass, efold, deconvolve, prec, _dp(acc), _dp(sec))
    return acc.reshape(-1, 3), dict(zip(("TSC", "F-FFT", "psolv", "B-FFT", "4FDA"), (float(v) for v in sec)))


def force_spline(law, n, eps, rs, dr_bins=None):
    """Force tables (4096, 3) diff and slope; law 0 reference, 1 erfc, 2 ForceLaw/treeforce.c.
    dr_bins overrides the derivative step (in bins)."""
    diff = np.empty((4096, 3))
    slope = np.empty((4096, 3))
    if dr_bins is None:
        lib().orc_force_spline(law, n, eps, rs, _dp(diff), _dp(slope))
    else:
        lib().orc_force_spline_dr(law, n, eps, rs, dr_bins, _dp(diff), _dp(slope))
    return diff, slope


def pmcf(n, eps, r):
    return lib().orc_pmcf(n, eps, r)


def erfc_fr(n, eps, rs, r):
    return lib().orc_erfc_fr(n, eps, rs, r)


def point_grav(diff, slope, dist):
    return lib().orc_point_grav(_dp(_d(diff)), _dp(_d(slope)), dist)


def pp_direct(n, pos, diff, slope, rcut, mass=1.0, targets=None):
    x, y, z = split(pos)
    diff, slope = _d(diff), _d(slope)
    if targets is None:
        nt, tp = len(x), None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        nt, tp = len(t), t.ctypes.data_as(_I64)
    acc = np.empty(3 * nt)
    lib().orc_pp_direct(n, len(x), _dp(x), _dp(y), _dp(z), mass, _dp(diff), _dp(slope), rcut, nt, tp, _dp(acc))
    return acc.reshape(-1, 3)


def pp_direct_analytic(law, n, eps, rs, pos, rcut, mass=1.0, targets=None):
    x, y, z = split(pos)
    if targets is None:
        nt, tp = len(x), None
    else:
        t = np.ascontiguousarray(targets, dtype=np.int64)
        nt, tp = len(t), t.ctypes.data_as(_I64)
    acc = np.empty(3 * nt)
    lib().orc_pp_direct_analytic(law, n, eps, rs, len(x), _dp(x), _dp(y), _dp(z), mass, rcut, nt, tp, _dp(acc))
    return acc.reshape(-1, 3)


def treecorrection_ref(n, pos, diff, slope, rsphere, theta, cellw=None, mass=1.0, fact2=1.0, quirks=0,
                       targets=None):
    """Reference tree correction kicks (N, 3); with `targets` (indices), the kicks of those
    particles only, shape (len(targets), 3), computed without the trees of other cells."""
    x, y, z = split(pos)
    diff, slope = _d(diff), _d(slope)
    cellw = max(4.0, rsphere) if cellw is None else cellw
    dv = np.empty(3 * len(x), dtype=np.float32)
    if targets is None:
        rc = lib().orc_treecorrection_ref(n, len(x), _dp(x), _dp(y), _dp(z), mass, _dp(diff), _dp(slope),
                                          rsphere, theta, cellw, fact2, quirks, dv.ctypes.data_as(_F))
    else:
        sel = np.zeros(len(x), dtype=np.uint8)
        sel[np.asarray(targets)] = 1
        rc = lib().orc_treecorrection_ref_sel(n, len(x), _dp(x), _dp(y), _dp(z), mass, _dp(diff), _dp(slope),
                                              rsphere, theta, cellw, fact2, quirks,
                                              sel.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                              dv.ctypes.data_as(_F))
    if rc != 0:
        raise RuntimeError("orc_treecorrection_ref failed")
    out = dv.reshape(-1, 3).astype(np.float64)
    return out if targets is None else out[np.asarray(targets)]


def node_accel(diff, slope, rcut, src, masses, target, quirks=0):
    x, y, z = split(src)
    m = np.ascontiguousarray(masses, dtype=np.float32)
    t = _d(target)
    out = np.empty(3, dtype=np.float32)
    lib().orc_node_accel(_dp(_d(diff)), _dp(_d(slope)), rcut, len(x), _dp(x), _dp(y), _dp(z),
                         m.ctypes.data_as(_F), _dp(t), quirks, out.ctypes.data_as(_F))
    return out.astype(np.float64)


def ewald_accel(pos, masses, L, G, alpha, nreal=2, kmax=10, eps=0.0):
    """Periodic force by Ewald summation; eps > 0 softens the minimum-image pair force (Plummer)."""
    x, y, z = split(pos)
    m = _d(masses)
    acc = np.empty(3 * len(x))
    lib().orc_ewald_accel_soft(len(x), _dp(x), _dp(y), _dp(z), _dp(m), L, G, alpha, nreal, kmax, eps, _dp(acc))
    return acc.reshape(-1, 3)


def ewald_converged(pos, n, eps=0.1, tol=1e-9):
    """Plummer-softened periodic force in code units (box side n cells, G = 1/n^2, unit masses),
    with the Ewald parameters chosen so both sums are converged to about tol: real space on the
    minimum image only (alpha n/2 > 6), reciprocal modes to exp(-k^2/4 alpha^2) < tol."""
    L = float(n)
    alpha = 12.5 / L
    kmax = int(np.ceil(np.sqrt(-4 * alpha ** 2 * np.log(tol)) * L / (2 * np.pi)))
    return ewald_accel(pos, np.ones(len(pos)), L, 1.0 / (L * L), alpha, nreal=0, kmax=kmax, eps=eps)


def evolfactor(omep, omeplam, amax, n, anow, astep):
    ef = OrcEvolFact()
    lib().orc_evolfactor(omep, omeplam, amax, n, anow, astep, ctypes.byref(ef))
    return ef


def fact2(omep, omeplam, amax, n, a1, a2):
    return lib().orc_fact2(omep, omeplam, amax, n, a1, a2)


def fact1(omep, omeplam, amax, a1, a2):
    return lib().orc_fact1(omep, omeplam, amax, a1, a2)


def ap(omep, omeplam, amax, a):
    return lib().orc_ap(omep, omeplam, amax, a)


def power_spectrum(n, spec, boxsize):
    """CorrMeasure restatement (Cosmos/Correlation.c:13-116) on an r2c spectrum of shape
    (n, n, n//2+1) [k][j][i]: returns pk, npk, cor (n+1 each), sigma8, slope."""
    sp = np.ascontiguousarray(np.asarray(spec, dtype=np.complex128)).view(np.float64).ravel()
    pk, npk, cor = np.empty(n + 1), np.empty(n + 1), np.empty(n + 1)
    s8, sl = ctypes.c_double(), ctypes.c_double()
    lib().orc_power_spectrum(n, _dp(sp), boxsize, _dp(pk), _dp(npk), _dp(cor), ctypes.byref(s8), ctypes.byref(sl))
    return pk, npk, cor, s8.value, sl.value


def prefof(n, pos, fof_link=0.2, minnum=20):
    """PreFoF halo-candidate flags (uint8 per particle) and their count."""
    x, y, z = split(pos)
    flag = np.zeros(len(x), dtype=np.uint8)
    cnt = lib().orc_prefof(n, len(x), _dp(x), _dp(y), _dp(z), fof_link, minnum,
                           flag.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return flag, int(cnt)
