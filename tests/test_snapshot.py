"""GOTPM particle files (jread/jwrite, Cosmos/CosmosRW/kjhrw.c:51-220; header format
Params/header.c:28-105, Params/params.h:49-139,249,500-518) and the 32-byte XYZDBL records
(eunha.h:389-399, PINDX :1212, XofP :1228-1232) through the C ABI's host-only functions.

Fixture: tests/golden/params.dat is the reference's own Params/params.dat, a header written in
this format by the reference's parameter writer (garbage bytes included: its unused string fields
were never initialised). The particle files below are built from it in the format kjhrw.c reads.
"""
import os

import numpy as np
import pytest

from lageunha_amd import lagpm

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "params.dat")
CLOSING = b"#End of Ascii Header\n"


def _header_with(updates):
    """The fixture's header lines, with 'define <key> = <value>' lines replaced for `updates`."""
    out = []
    for ln in open(GOLDEN, "rb").read().split(b"\n")[:-1]:
        ln += b"\n"
        if ln.startswith(b"define ") and b"=" in ln:
            key = ln[7:ln.index(b"=")].strip().decode("latin-1")
            if key in updates:
                ln = ln[:ln.index(b"=") + 1] + f" {updates[key]}\n".encode()
        out.append(ln)
    return b"".join(out)


def _records(n, nx, seed):
    g = np.random.default_rng(seed)
    r = np.zeros(n, dtype=lagpm.DM_RECORD)
    idx = g.permutation(nx ** 3)[:n].astype(np.uint64)
    flags = g.integers(0, 256, n).astype(np.uint64) << np.uint64(56)  # the flag byte: not index bits
    r["u4if"] = idx | flags
    for k in ("x", "y", "z"):
        r[k] = g.normal(0, 0.7, n).astype(np.float32)
    for k in ("vx", "vy", "vz"):
        r[k] = g.normal(0, 2.0, n).astype(np.float32)
    return r


def _write_file(path, n, nx, seed, ptype=0, tail=b"", extra=None):
    rec = _records(n, nx, seed)
    upd = {"Np": n, "DNp": n, "Nx": nx, "Ny": nx, "Nz": nx, "Particle Type": ptype}
    upd.update(extra or {})
    body = rec.tobytes()
    if ptype == 1:  # treedmparticletype: a leading 8-byte list pointer per record
        t = np.zeros(n, dtype=[("next", "<u8"), ("rec", lagpm.DM_RECORD)])
        t["next"] = np.arange(n, dtype=np.uint64) * np.uint64(40) + np.uint64(0x7F0000000000)
        t["rec"] = rec
        body = t.tobytes()
    with open(path, "wb") as f:
        f.write(_header_with(upd) + body + tail)
    return rec


def test_header_of_the_reference_fixture():
    info = lagpm.read_snapshot_header(GOLDEN)
    raw = open(GOLDEN, "rb").read()
    assert info.header_bytes == raw.index(CLOSING) + len(CLOSING) == info.file_bytes
    assert (info.nx, info.ny, info.nz) == (256, 256, 256)
    assert (info.amax, info.anow, info.astep) == (94.0, 1.0, 1e-4)
    assert info.theta == pytest.approx(0.3) and info.epsilon == pytest.approx(0.1)
    assert info.stepcount == 1 and info.ptype == 0 and info.record_bytes == 32
    assert (info.np_sum, info.np_dm, info.np_sph, info.np_star, info.np_agn) == (0, 0, 0, 0, 0)
    info2, rec = lagpm.read_snapshot(GOLDEN)
    assert rec.shape == (0,)


@pytest.mark.parametrize("ptype", [0, 1])
def test_read_pm_and_tree_records(tmp_path, ptype):
    p = str(tmp_path / "snap.00000")
    tail = bytes(range(256)) * 3
    rec = _write_file(p, 1000, 16, seed=3, ptype=ptype, tail=tail)
    info, got = lagpm.read_snapshot(p)
    assert info.np_dm == 1000 and info.ptype == ptype and info.record_bytes == (40 if ptype else 32)
    assert info.file_bytes == info.header_bytes + 1000 * info.record_bytes + len(tail)
    assert got.tobytes() == rec.tobytes()


def test_write_keeps_header_lines_and_drops_the_trailing_block(tmp_path):
    """The written file is the template's header lines (time lines rewritten) and the records,
    and ends there. jread reads a SimParameters struct back over the parsed header after the
    particles (Cosmos/CosmosRW/kjhrw.c:118): a carried template block would make a restart resume
    at the template's time and particle type, so none is written and the header stands."""
    src, dst = str(tmp_path / "in.00000"), str(tmp_path / "out.00000")
    tail = np.random.default_rng(5).integers(0, 256, 5000, dtype=np.uint8).tobytes()
    rec = _write_file(src, 512, 8, seed=4, tail=tail)
    new = rec.copy()
    new["x"] += np.float32(0.25)
    anow = 1.0 + 1.0 / 3.0
    lagpm.write_snapshot(src, dst, new, anow, 7)
    info, got = lagpm.read_snapshot(dst)
    assert info.anow == anow and info.stepcount == 7 and info.np_dm == 512
    assert got.tobytes() == new.tobytes()
    a, b = open(src, "rb").read(), open(dst, "rb").read()
    assert len(b) == info.header_bytes + 512 * 32 == info.file_bytes
    assert a.endswith(tail) and not b.endswith(tail[-64:])
    ha, hb = a[:a.index(CLOSING)].split(b"\n"), b[:b.index(CLOSING)].split(b"\n")
    assert len(ha) == len(hb)
    changed = [x for x, y in zip(ha, hb) if x != y]
    assert [c.split(b"=")[0].strip() for c in changed] == [b"define Anow", b"define Stepcount", b"define Stepnum"]


def test_file_errors(tmp_path):
    with pytest.raises(lagpm.LagpmError) as e:
        lagpm.read_snapshot_header(str(tmp_path / "missing"))
    assert e.value.code == -7
    p = str(tmp_path / "short")
    _write_file(p, 100, 8, seed=1)
    raw = open(p, "rb").read()
    open(p, "wb").write(raw[:-10])  # ends inside the particle array
    with pytest.raises(lagpm.LagpmError) as e:
        lagpm.read_snapshot_header(p)
    assert e.value.code == -7
    open(p, "wb").write(raw[:200])  # no closing line
    with pytest.raises(lagpm.LagpmError):
        lagpm.read_snapshot_header(p)
    q = str(tmp_path / "bad")
    _write_file(q, 10, 8, seed=1, extra={"Nx": "12x"})
    with pytest.raises(lagpm.LagpmError) as e:
        lagpm.read_snapshot_header(q)
    assert e.value.code == -7
    s = str(tmp_path / "stars")
    rec = _write_file(s, 10, 8, seed=1, extra={"SNp": 3})
    with pytest.raises(lagpm.LagpmError) as e:
        lagpm.write_snapshot(s, str(tmp_path / "o"), rec, 1.0, 1)
    assert e.value.code == -1
    good = str(tmp_path / "good")
    rec = _write_file(good, 10, 8, seed=1)
    with pytest.raises(lagpm.LagpmError) as e:
        lagpm.write_snapshot(good, str(tmp_path / "o"), rec[:9], 1.0, 1)
    assert e.value.code == -1
    with pytest.raises(lagpm.LagpmError):
        lagpm.write_snapshot(good, good, rec, 1.0, 1)


def test_decode_lattice_sites_and_flag_byte():
    nx = 12
    idx = np.arange(nx ** 3, dtype=np.uint64)
    r = np.zeros(idx.size, dtype=lagpm.DM_RECORD)
    r["u4if"] = idx | (np.uint64(0xA5) << np.uint64(56))
    pos, vel = lagpm.eunha_decode(r, nx)
    i, j, k = np.meshgrid(np.arange(nx), np.arange(nx), np.arange(nx), indexing="ij")
    # idx = i + nx (j + nx k): x fastest
    want = np.stack([k.ravel(), j.ravel(), i.ravel()], axis=1).astype(np.float64)
    assert np.array_equal(pos, want) and not vel.any()
    # displacements across the box edge wrap periodically: site 0 moved by -0.5 is at n - 0.5
    r2 = np.zeros(2, dtype=lagpm.DM_RECORD)
    r2["u4if"] = [0, nx - 1]
    r2["x"] = [-0.5, 1.25]
    p2, _ = lagpm.eunha_decode(r2, nx)
    assert p2[0, 0] == nx - 0.5 and p2[1, 0] == 0.25


def test_encode_decode_round_trip():
    nx = 32
    rec = _records(4096, nx, seed=9)
    pos, vel = lagpm.eunha_decode(rec, nx)
    out = rec.copy()
    lagpm.eunha_encode(out, nx, pos, vel)
    assert np.array_equal(out["u4if"], rec["u4if"])
    p2, v2 = lagpm.eunha_decode(out, nx)
    assert np.array_equal(v2, vel)
    d = np.abs(p2 - pos)
    d = np.minimum(d, nx - d)  # periodic distance
    assert d.max() <= 4 * np.finfo(np.float32).eps * nx
    # positions anywhere in the box: displacement is the image nearest the site, |d| <= n/2
    g = np.random.default_rng(2)
    anywhere = g.uniform(0, nx, (4096, 3))
    lagpm.eunha_encode(out, nx, anywhere, vel)
    assert np.abs(out["x"]).max() <= nx / 2
    p3, _ = lagpm.eunha_decode(out, nx)
    d = np.abs(p3 - anywhere)
    d = np.minimum(d, nx - d)
    assert d.max() <= 4 * np.finfo(np.float32).eps * nx
