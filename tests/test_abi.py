"""The C-ABI library loads, exports every symbol include/lagpm.h declares, and its host-only
entry points (force tables, kick factors, size estimate) agree with the oracle. CPU only."""
import ctypes
import os
import re

import numpy as np

from lageunha_amd import lagpm
from oracle import pyoracle as orc

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "lagpm.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lagpm_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(lagpm.LIB_PATH)
    syms = declared_symbols()
    assert len(syms) >= 25
    missing = [s for s in syms if not hasattr(lib, s)]
    assert not missing, missing
    # and the Python mirror binds exactly those
    assert sorted(lagpm.SIGNATURES) == syms


def test_abi_version():
    hdr = open(os.path.join(ROOT, "include", "lagpm.h")).read()
    want = int(re.search(r"#define LAGPM_ABI_VERSION (\d+)", hdr).group(1))
    assert lagpm.load().lagpm_abi_version() == want


def test_force_spline_matches_oracle_tables():
    tf = lagpm.reference_params(256)
    tf.law, tf.epsilon = lagpm.LAW_TREEFORCE, 0.25  # the variant fixes its softening at 0.1
    for law, p in ((0, lagpm.reference_params(256)), (1, lagpm.make_params(256, efold=2.5)), (2, tf)):
        diff, slope = lagpm.i_force_spline(p)
        od, os_ = orc.force_spline(law, 256, p.epsilon, p.efold / 2)
        assert np.allclose(diff, od, rtol=1e-13, atol=0)
        assert np.allclose(slope, os_, rtol=1e-9, atol=1e-9 * np.abs(os_).max())


def test_evolfactor_matches_oracle_quadrature():
    p = lagpm.make_params(256, omep=0.26, omeplam=0.74, amax=48.0)
    for anow in (1.0, 5.0, 30.0):
        ef = lagpm.get_evol_factor(p, anow, 0.25)
        o = orc.evolfactor(0.26, 0.74, 48.0, 256, anow, 0.25)
        for k in ("fact1", "fact2", "fact1_push", "fact2_push", "fact1_pull", "fact2_pull", "pfact"):
            assert np.isclose(getattr(ef, k), getattr(o, k), rtol=2e-7), k


def test_invalid_parameters_are_rejected_without_a_device():
    bad = lagpm.make_params(256, rsphere=7.0)
    d = np.empty((4096, 3))
    s = np.empty((4096, 3))
    rc = lagpm.load().lagpm_force_spline(ctypes.byref(bad), d.ctypes.data_as(lagpm._D), s.ctypes.data_as(lagpm._D))
    assert rc == -1


def test_box_byte_estimate_scales():
    p = lagpm.make_params(256)
    b1 = lagpm.estimate_box_bytes(p, 256 ** 3)
    b2 = lagpm.estimate_box_bytes(lagpm.make_params(128), 128 ** 3)
    assert 6 < b1 / b2 < 10
    assert 2e9 < b1 < 20e9


def test_c_translation_unit_compiles_links_and_agrees(tmp_path):
    """tests/abi_check.c, a C program as a maintainer's EUNHA shim would be, compiles against
    include/lagpm.h in strict C11, links against liblagpm.so with the address of every declared
    entry point, and its host-only calls print what the Python binding computes."""
    import shutil
    import subprocess
    src = os.path.join(ROOT, "tests", "abi_check.c")
    text = open(src).read()
    table = text[text.index("entry_points[]"):text.index("};", text.index("entry_points[]"))]
    assert sorted(re.findall(r"\b(lagpm_[a-z0-9_]+)\b", table)) == declared_symbols()
    cc = shutil.which("gcc") or shutil.which("cc")
    assert cc, "no C compiler"
    exe = str(tmp_path / "abi_check")
    libdir = os.path.dirname(lagpm.LIB_PATH)
    subprocess.run([cc, "-std=c11", "-Wall", "-Wextra", "-Werror", "-pedantic", "-I", os.path.join(ROOT, "include"), src,
                    "-L", libdir, "-llagpm", f"-Wl,-rpath,{libdir}", "-o", exe], check=True)
    out = dict(ln.split(" ", 1) for ln in subprocess.run([exe], check=True, capture_output=True, text=True).stdout.splitlines())
    p = lagpm.Params(64, lagpm.LAW_PMFIT, lagpm.PREC_FP32, 0, 1.0, 0.9, 0.1, 4.0, 0.5, 48.0, 0.26, 0.74)
    diff, slope = lagpm.i_force_spline(p)
    assert [float(v) for v in out["diff"].split()] == list(diff[100])
    assert float(out["slope"]) == slope[2000, 0]
    ef = lagpm.get_evol_factor(p, 1.0, 0.25)
    # %.9g of a float names that float exactly
    assert [np.float32(float(v)) for v in out["evolfact"].split()] == [np.float32(x) for x in (ef.fact1, ef.fact2, ef.pfact)]
    assert int(out["bytes"]) == lagpm.estimate_box_bytes(p, 64 ** 3)
    assert out["abi"].split()[0] == str(lagpm.load().lagpm_abi_version())
    # the ctypes mirrors have the C structs' sizes and member offsets
    for name, cls, last in (("params", lagpm.Params, "omeplam"), ("evolfact", lagpm.EvolFact, "fact2_pull"),
                            ("tree_stats", lagpm.TreeStats, "n_pp_within"), ("timings", lagpm.Timings, "ms_walk"),
                            ("snapshot_info", lagpm.SnapshotInfo, "file_bytes")):
        size, off = (int(v) for v in out["layout_" + name].split())
        assert (ctypes.sizeof(cls), getattr(cls, last).offset) == (size, off), name
