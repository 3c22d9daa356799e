"""A GOTPM particle file through a whole step on the device: jread-format file -> records ->
box (lagpm_box_upload_eunha) -> lagpm_kdk_step -> records (lagpm_box_download_eunha) -> jwrite-
format file -> a second box. The reference's own loop does the same around its step
(Cosmos/CosmosRW/kjhrw.c:51,155; Cosmos/globalT.c:221-306)."""
import numpy as np
import pytest

from lageunha_amd import lagpm
from test_snapshot import _write_file

pytestmark = pytest.mark.gpu


def test_file_step_file_round_trip(ctx, tmp_path):
    n = 24  # the smallest mesh for the default cutoff (n >= 4 rsphere)
    src, dst = str(tmp_path / "rv.00000"), str(tmp_path / "rv.00001")
    _write_file(src, n ** 3, n, seed=11, tail=b"SimParameters block" * 10)
    info, rec = lagpm.read_snapshot(src)
    p = lagpm.make_params(n)
    with lagpm.Box(ctx, p, info.np_dm) as b, lagpm.Box(ctx, p, info.np_dm) as b2:
        b.upload_eunha(rec)
        pos0, vel0 = b.download()
        want0, wv0 = lagpm.eunha_decode(rec, n)
        assert np.array_equal(pos0, want0) and np.array_equal(vel0, wv0)
        b.kdk_step(1.0, 0.25)
        pos1, vel1 = b.download()
        assert not np.array_equal(pos1, pos0)
        out = rec.copy()
        b.download_eunha(out)
        assert np.array_equal(out["u4if"], rec["u4if"])
        assert np.array_equal(out["vx"], vel1[:, 0].astype(np.float32))
        lagpm.write_snapshot(src, dst, out, 1.25, 1)
        info2, rec2 = lagpm.read_snapshot(dst)
        assert info2.anow == 1.25 and info2.stepcount == 1 and rec2.tobytes() == out.tobytes()
        b2.upload_eunha(rec2)
        pos2, vel2 = b2.download()
    # positions come back within the rounding of an f32 displacement
    d = np.abs(pos2 - pos1)
    d = np.minimum(d, n - d)
    assert d.max() <= 4 * np.finfo(np.float32).eps * n
    assert np.array_equal(vel2, vel1.astype(np.float32).astype(np.float64))
