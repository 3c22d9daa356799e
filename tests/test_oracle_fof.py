"""The oracle's PreFoF (Cosmos/PreFoF.c:207-386, OST/fof.c:190-293), pinned by
(1) scipy's kd-tree pair search and graph components on the same cells (an independent FoF),
(2) hand-built configurations whose flags follow from the rules: a 20-member interior group is
kept and a 19-member one is not, any group within the link of a cell face is kept, groups never
link across cells, and a pair just beyond the link stays apart.
Runs on the CPU."""
import numpy as np
import pytest
from scipy.sparse import coo_matrix
from scipy.sparse.csgraph import connected_components
from scipy.spatial import cKDTree

from lageunha_amd import ics
from oracle import pyoracle as orc


def scipy_prefof(n, pos, link, minnum):
    pos = np.asarray(pos, dtype=np.float64)
    cw = 4.0
    mx = int(np.ceil(n / cw))
    c3 = (pos // cw).astype(np.int64)
    cell = c3[:, 0] + mx * (c3[:, 1] + mx * c3[:, 2])
    flag = np.zeros(len(pos), dtype=np.uint8)
    order = np.argsort(cell, kind="stable")
    bounds = np.flatnonzero(np.diff(cell[order])) + 1
    for idx in np.split(order, bounds):
        if idx.size == 0:
            continue
        r = pos[idx] - cw * c3[idx[0]]
        pairs = cKDTree(r).query_pairs(link, output_type="ndarray")
        m = idx.size
        g = coo_matrix((np.ones(len(pairs)), (pairs[:, 0], pairs[:, 1])), shape=(m, m)) if len(pairs) else \
            coo_matrix((m, m))
        ncomp, lab = connected_components(g, directed=False)
        for c in range(ncomp):
            sel = lab == c
            lo, hi = r[sel].min(axis=0), r[sel].max(axis=0)
            inside = bool(np.all(lo > link) and np.all(hi < cw - link))
            flag[idx[sel]] = 1 if (not inside or sel.sum() >= minnum) else 0
    return flag


@pytest.mark.parametrize("n,seed,link", [(16, 1, 0.2), (20, 2, 0.35), (24, 3, 0.2)])
def test_oracle_prefof_matches_scipy_components(n, seed, link):
    pos = ics.clustered_box(n, halo_fraction=0.4, nhalo=6, seed=seed)
    flag, cnt = orc.prefof(n, pos, link, 20)
    want = scipy_prefof(n, pos, np.float64(np.float32(link)), 20)
    assert np.array_equal(flag, want)
    assert cnt == int(want.sum()) and 0 < cnt < len(pos)


def _chain(start, count, step, axis=0):
    p = np.tile(np.asarray(start, dtype=np.float64), (count, 1))
    p[:, axis] += step * np.arange(count)
    return p


def test_oracle_prefof_known_answers():
    n, link = 16, 0.2
    groups = [
        _chain([5.0, 6.0, 6.0], 20, 0.1),   # cell (1,1,1): interior, 20 members -> kept
        _chain([9.0, 10.0, 6.0], 19, 0.1),  # cell (2,2,1): interior, 19 members -> dropped
        _chain([8.05, 2.0, 2.0], 2, 0.1),   # cell (2,0,0): within the link of its x face -> kept
        _chain([13.0, 13.0, 13.0], 2, 0.25),  # cell (3,3,3): a pair beyond the link: two singletons, dropped
        _chain([3.7, 10.0, 10.0], 6, 0.1),  # crosses x = 4: two groups, each at a face -> all kept
    ]
    pos = np.concatenate(groups)
    flag, cnt = orc.prefof(n, pos, link, 20)
    want = np.concatenate([np.ones(20), np.zeros(19), np.ones(2), np.zeros(2), np.ones(6)]).astype(np.uint8)
    assert np.array_equal(flag, want) and cnt == 28
    # the same flags whatever the particle order
    perm = np.random.default_rng(0).permutation(len(pos))
    f2, _ = orc.prefof(n, pos[perm], link, 20)
    assert np.array_equal(f2, want[perm])
    # minnum 19 keeps the 19-member group too
    f3, c3 = orc.prefof(n, pos, link, 19)
    assert c3 == 47 and f3[20:39].all()
