"""PreFoF halo-candidate flags on the device (lagpm_prefof) against the oracle's restatement of
Cosmos/PreFoF.c:207-386 / OST/fof.c:190-293, bit for bit, on clustered boxes (halo cells with
thousands of particles), the hand-built configuration of the oracle's own test, and argument
errors."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n,hf,nh,seed,link", [(16, 0.4, 6, 1, 0.2), (24, 0.3, 10, 2, 0.35), (32, 0.5, 4, 3, 0.2),
                                               (48, 0.3, 30, 4, 0.17), (64, 0.3, 60, 5, 0.2)])
def test_prefof_matches_oracle(ctx, n, hf, nh, seed, link):
    pos = ics.clustered_box(n, halo_fraction=hf, nhalo=nh, seed=seed)
    print(f"prefof n={n} np={len(pos)} link={link}", flush=True)
    want, wcnt = orc.prefof(n, pos, link, 20)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        got, cnt = b.prefof(link, 20)
        got5, cnt5 = b.prefof(link, 5)
    assert cnt == wcnt and np.array_equal(got, want)
    want5, w5 = orc.prefof(n, pos, link, 5)
    assert cnt5 == w5 and np.array_equal(got5, want5)


def test_prefof_known_configuration(ctx):
    n, link = 16, 0.2

    def chain(start, count, step):
        p = np.tile(np.asarray(start, dtype=np.float64), (count, 1))
        p[:, 0] += step * np.arange(count)
        return p

    pos = np.concatenate([chain([5.0, 6.0, 6.0], 20, 0.1), chain([9.0, 10.0, 6.0], 19, 0.1),
                          chain([8.05, 2.0, 2.0], 2, 0.1), chain([13.0, 13.0, 13.0], 2, 0.25),
                          chain([3.7, 10.0, 10.0], 6, 0.1)])
    want = np.concatenate([np.ones(20), np.zeros(19), np.ones(2), np.zeros(2), np.ones(6)]).astype(np.uint8)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        got, cnt = b.prefof(link, 20)
        with pytest.raises(lagpm.LagpmError):
            b.prefof(0.0, 20)
        with pytest.raises(lagpm.LagpmError):
            b.prefof(0.2, 0)
    assert cnt == 28 and np.array_equal(got, want)


@pytest.mark.parametrize("link", [0.5, 1.3, 2.5, 4.5])
def test_prefof_long_links_and_cell_faces(ctx, link):
    """Links from 0.5 to beyond the cell width (one sub-cell per cell), with particles placed
    exactly on cell faces (x = 4k) and corners: the flags equal the oracle's."""
    n = 24
    g = np.random.default_rng(int(link * 10))
    pos = g.random((3000, 3)) * n
    faces = np.stack(np.meshgrid(np.arange(0, n, 4.0), np.arange(0, n, 4.0), [2.0, 6.5]), -1).reshape(-1, 3)
    pos = np.concatenate([pos, faces, faces + [0.0, 0.0, 0.3]])
    want, wcnt = orc.prefof(n, pos, link, 20)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        got, cnt = b.prefof(link, 20)
    assert cnt == wcnt and np.array_equal(got, want)
