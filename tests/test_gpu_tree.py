"""GPU parity of the tree correction through the C ABI (run with -m gpu).

theta -> 0 must reproduce the oracle's exact short-range pair sum (the erfc law analytically in
FP64; the reference's fitted law through its own 4096-entry table); at theta = 0.5 the FP32
wavefront walk must agree with the plain FP64 walk of the same tree to well below 1e-3
and stay within the reference walk's own approximation error of the exact sum."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def rel_rms(a, b):
    return float(np.sqrt(np.sum((a - b) ** 2) / np.sum(b ** 2)))


def tree_accel(ctx, params, pos, fact=1.0):
    with lagpm.Box(ctx, params, len(pos)) as b:
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(fact))
        v = b.velocities()
        st = b.tree_stats()
    return v, st


def exact_erfc(n, p, pos):
    """The erfc short-range law evaluated analytically in FP64 for every pair within the cutoff."""
    return orc.pp_direct_analytic(1, n, p.epsilon, p.efold / 2, pos, p.rsphere)


def workload(n, kind, seed=3):
    if kind == "uniform":
        return ics.uniform_random(n ** 3, n, seed=seed)
    if kind == "lattice":
        return ics.perturbed_lattice(n, rms=0.3, seed=seed)
    return ics.clustered_box(n, halo_fraction=0.3, nhalo=max(8, n ** 3 // 2000), seed=seed)


@pytest.mark.parametrize("kind", ["uniform", "clustered"])
@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_tree_theta_zero_is_the_direct_sum(ctx, kind, prec):
    n = 32
    pos = workload(n, kind)
    p = lagpm.make_params(n, theta=1e-6, precision=prec)
    print(f"tree theta0 kind={kind} prec={prec} N={len(pos)}", flush=True)
    a, st = tree_accel(ctx, p, pos)
    ref = exact_erfc(n, p, pos)
    err = rel_rms(a, ref)
    print(f"  rms rel {err:.3e}  n_pp {st.n_pp} n_pc {st.n_pc} nodes {st.n_nodes} groups {st.n_groups}")
    assert st.n_pc == 0
    # FP32: ~3e3 table pairs per target summed in FP32 with strong cancellation -> ~1e-5;
    # a single missing pair would show at ~1e-3
    assert err < 3e-5


def test_tree_reference_law_theta_zero(ctx):
    n = 32
    pos = workload(n, "clustered", seed=5)
    p = lagpm.reference_params(n, theta=1e-6)
    a, _ = tree_accel(ctx, p, pos)
    diff, slope = orc.force_spline(0, n, p.epsilon, 0.45)
    ref = orc.pp_direct(n, pos, diff, slope, p.rsphere)
    assert rel_rms(a, ref) < 3e-5


def test_tree_treeforce_variant_law_theta_zero(ctx):
    """The ForceLaw/treeforce.c table (softening 0.1, derivative step 0.2 bin, entry 0 zero) through
    the walk: theta -> 0 gives the pair sum with the oracle's table of that law."""
    n = 32
    pos = workload(n, "clustered", seed=6)
    p = lagpm.reference_params(n, theta=1e-6)
    p.law = lagpm.LAW_TREEFORCE
    a, _ = tree_accel(ctx, p, pos)
    diff, slope = orc.force_spline(2, n, 0.1, 0.45)
    ref = orc.pp_direct(n, pos, diff, slope, p.rsphere)
    assert rel_rms(a, ref) < 3e-5
    # a different law from the reference one: the tables differ at the 1e-6 level near r = 0
    d0, s0 = orc.force_spline(0, n, 0.1, 0.45)
    assert not np.array_equal(diff, d0)


@pytest.mark.parametrize("kind", ["lattice", "clustered"])
def test_tree_theta_half_fp32_vs_fp64_and_exact(ctx, kind):
    n = 48
    pos = workload(n, kind, seed=7)
    p32 = lagpm.make_params(n, theta=0.5)
    p64 = lagpm.make_params(n, theta=0.5, precision=lagpm.PREC_FP64)
    a32, st32 = tree_accel(ctx, p32, pos)
    a64, st64 = tree_accel(ctx, p64, pos)
    diff, slope = orc.force_spline(1, n, p32.epsilon, p32.efold / 2)
    exact = exact_erfc(n, p32, pos)
    e_prec = rel_rms(a32, a64)
    e_tree = rel_rms(a64, exact)
    ref_walk = orc.treecorrection_ref(n, pos, diff, slope, p32.rsphere, 0.5)
    e_ref = rel_rms(ref_walk, exact)
    print(f"theta=0.5 {kind}: fp32 vs fp64 {e_prec:.2e}; gpu tree vs exact {e_tree:.2e}; "
          f"reference walk vs exact {e_ref:.2e}; pp {st32.n_pp} pc {st32.n_pc}", flush=True)
    assert e_prec < 1e-4
    # group-conservative opening: each target sees a stricter opening test than the reference's
    # per-target test, so the GPU tree is at least as accurate as the reference walk
    assert e_tree <= 1.02 * e_ref
    assert rel_rms(a32, exact) <= 1.02 * e_ref
    assert st32.n_pc > 0


def test_tree_reference_mode_theta_03_close_to_reference_walk(ctx):
    n = 32
    pos = workload(n, "clustered", seed=9)
    p = lagpm.reference_params(n, theta=0.3)
    a, _ = tree_accel(ctx, p, pos)
    diff, slope = orc.force_spline(0, n, p.epsilon, 0.45)
    exact = orc.pp_direct(n, pos, diff, slope, p.rsphere)
    ref_walk = orc.treecorrection_ref(n, pos, diff, slope, p.rsphere, 0.3)
    e_gpu, e_ref = rel_rms(a, exact), rel_rms(ref_walk, exact)
    print(f"reference mode theta=0.3: gpu tree vs exact {e_gpu:.2e}; reference walk vs exact {e_ref:.2e}", flush=True)
    assert e_gpu <= 1.02 * e_ref
    assert rel_rms(a, ref_walk) < 5e-3


def test_tree_is_deterministic_and_scales_with_fact(ctx):
    n = 32
    pos = workload(n, "clustered", seed=11)
    p = lagpm.make_params(n, theta=0.5)
    a1, _ = tree_accel(ctx, p, pos, fact=1.0)
    a2, _ = tree_accel(ctx, p, pos, fact=1.0)
    assert np.array_equal(a1, a2)
    a3, _ = tree_accel(ctx, p, pos, fact=0.5)
    assert np.allclose(a3, 0.5 * a1, rtol=1e-6, atol=1e-12)


def test_tree_small_and_empty_inputs(ctx):
    n = 24
    p = lagpm.make_params(n, theta=0.5)
    for npart in (0, 1, 2, 17, 100):
        pos = ics.uniform_random(npart, n, seed=npart + 1)
        a, st = tree_accel(ctx, p, pos)
        assert a.shape == (npart, 3)
        if npart:
            ref = exact_erfc(n, p, pos)
            # per particle: the tabulated law is within ~3e-5 of the analytic one pair by pair, so
            # a force vector may differ by that much of its pair forces (not of its own norm
            # when they cancel): bound by 1e-4 of the norm plus 1e-5 of the largest force
            d = np.linalg.norm(a - ref, axis=1)
            r = np.linalg.norm(ref, axis=1)
            assert np.all(d <= 1e-4 * r + 1e-5 * max(r.max(), 1e-300)), (npart, float((d / np.maximum(r, 1e-300)).max()))


def test_tree_rejects_box_smaller_than_four_cutoffs(ctx):
    p = lagpm.make_params(16, rsphere=5.625)
    with lagpm.Box(ctx, p, 10) as b:
        b.upload_pos(ics.uniform_random(10, 16))
        with pytest.raises(lagpm.LagpmError):
            b.treecorrection(lagpm.evolfact(1.0))


def test_coincident_particles_beyond_group_size(ctx):
    """200 particles at one point (identical Morton keys at depth 16) plus a background."""
    n = 32
    g = np.random.default_rng(4)
    pos = np.concatenate([np.tile([[10.3, 11.7, 12.1]], (200, 1)), g.random((3000, 3)) * n])
    p = lagpm.make_params(n, theta=1e-6)
    a, st = tree_accel(ctx, p, pos)
    assert rel_rms(a, exact_erfc(n, p, pos)) < 3e-5


def test_autotune_keeps_velocities_and_a_valid_tuning(ctx):
    n = 32
    pos = workload(n, "clustered", seed=13)
    vel = np.random.default_rng(13).standard_normal(pos.shape) * 0.01
    p = lagpm.make_params(n, theta=0.5)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos, vel)
        nl, g, mode = b.autotune()
        assert (nl, g, mode) == b.tuning() and 4 <= nl <= 256 and 8 <= g <= 128
        assert mode in (lagpm.GROUPS_TREE, lagpm.GROUPS_RUNS, lagpm.GROUPS_PACKED)
        assert np.array_equal(b.velocities(), vel)
        # and the tuned tree still gives the exact sum at theta -> 0
    p0 = lagpm.make_params(n, theta=1e-6)
    with lagpm.Box(ctx, p0, len(pos)) as b:
        b.set_tuning(nl, g, mode)
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(1.0))
        assert rel_rms(b.velocities(), exact_erfc(n, p0, pos)) < 3e-5


def test_tree_large_box_sampled_targets_are_the_direct_sum(ctx):
    """8.5M particles: the radix sort's offsets come from column scans of more than 2048 tiles per
    digit (several chunks per column), and the tree built on that order must still give the exact
    pair sum at theta -> 0 for a sample of targets."""
    n, npart = 128, 8_500_000
    pos = ics.uniform_random(npart, n, seed=13)
    p = lagpm.make_params(n, theta=1e-6)
    print(f"large box N={npart} n={n}", flush=True)
    a, st = tree_accel(ctx, p, pos)
    tg = np.random.default_rng(13).choice(npart, 256, replace=False)
    ref = orc.pp_direct_analytic(1, n, p.epsilon, p.efold / 2, pos, p.rsphere, targets=tg)
    err = rel_rms(a[tg], ref)
    print(f"  rms rel {err:.3e}  n_pp {st.n_pp} nodes {st.n_nodes}", flush=True)
    assert st.n_pc == 0
    # ~6.7e3 FP32 table pairs per target at this density -> ~4e-5; one pair missing from one of the
    # 256 targets would show at ~8e-4, and particles out of order break the tree at O(1)
    assert err < 1e-4


@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_node_moments_match_brute_force_and_leaves_partition_the_particles(ctx, prec):
    """Make_Tree3 / divide_node3 (OST/ost.ver2.c:36-144) pinned node by node: every node's mass,
    centre of mass, quadrupole about it (Qxx..Qyz, trQ) and rmax against FP64 sums over its own
    particles; the leaves partition the particles; a node's particles lie in its octree cell and are
    exactly its children's particles. The FP32 tree sums with Kahan compensation, so its moments
    stay within a few float ulps of the box scale."""
    n = 32
    pos = np.mod(workload(n, "clustered", seed=21), n)
    p = lagpm.make_params(n, theta=0.5, precision=prec)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(0.0))
        t = b.tree_nodes()
    info, com, quad, order = t["info"], t["com"], t["quad"], t["order"].astype(np.int64)
    assert np.array_equal(np.sort(order), np.arange(len(pos)))
    sp = pos[order]  # positions in the box's order
    first, count, nchild, level = info[:, 0], info[:, 1], info[:, 2], info[:, 3]
    print(f"nodes {len(info)} (leaves {(nchild == 0).sum()}), levels {level.min()}..{level.max()}", flush=True)
    # leaves partition [0, N)
    leaf = nchild == 0
    cover = np.zeros(len(pos), dtype=np.int64)
    np.add.at(cover, np.concatenate([np.arange(f, f + c) for f, c in zip(first[leaf], count[leaf])]), 1)
    assert np.all(cover == 1)
    # cell roots partition [0, N) as well
    roots = level == level.min()
    assert count[roots].sum() == len(pos)
    tol_x = 2e-5 if prec == lagpm.PREC_FP32 else 1e-10
    width = n / 2.0 ** level
    for v in range(len(info)):
        f, c = first[v], count[v]
        x = sp[f:f + c]
        # the node is an octree cell: one cell index per axis at its level
        cell = np.floor(x / width[v])
        assert np.all(cell == cell[0]), v
        cm = x.mean(axis=0)
        d = x - cm
        q = np.array([(d[:, 0] ** 2).sum(), (d[:, 1] ** 2).sum(), (d[:, 2] ** 2).sum(), (d[:, 0] * d[:, 1]).sum(),
                      (d[:, 0] * d[:, 2]).sum(), (d[:, 1] * d[:, 2]).sum()])
        rmax = np.sqrt((d ** 2).sum(axis=1).max())
        assert com[v, 3] == c * p.mass
        assert np.max(np.abs(com[v, :3] - cm)) <= tol_x, (v, com[v], cm)
        qscale = c * (width[v] ** 2)
        assert np.max(np.abs(quad[v, :6] - q)) <= (1e-5 if prec == lagpm.PREC_FP32 else 1e-12) * qscale, (v, quad[v], q)
        assert abs(quad[v, 6] - q[:3].sum()) <= (1e-5 if prec == lagpm.PREC_FP32 else 1e-12) * qscale
        if nchild[v] == 0 or c <= 128:
            assert abs(quad[v, 7] - rmax) <= tol_x, (v, quad[v, 7], rmax)
        else:
            assert quad[v, 7] >= rmax - tol_x  # an upper bound of the largest distance
    # a parent's particles are exactly its children's: every node below the cell level lies inside
    # one node of the level above, and per parent the children's counts add up to its count
    kids = np.zeros(len(info), dtype=np.int64)
    kid_particles = np.zeros(len(info), dtype=np.int64)
    for lv in range(level.min() + 1, level.max() + 1):
        up = np.nonzero(level == lv - 1)[0]
        up = up[np.argsort(first[up])]
        for u in np.nonzero(level == lv)[0]:
            k = up[np.searchsorted(first[up], first[u], side="right") - 1]
            assert first[k] <= first[u] and first[u] + count[u] <= first[k] + count[k], u
            kids[k] += 1
            kid_particles[k] += count[u]
    assert np.array_equal(kids, nchild)
    inner = nchild > 0
    assert np.array_equal(kid_particles[inner], count[inner])


@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_sampled_walk_theta_zero_is_the_direct_sum_and_the_full_walk(ctx, prec):
    """lagpm_tree_accel_sample at theta -> 0: each sampled particle alone gets exactly its pair sum
    within the cutoff, the same as the full walk gives it; velocities stay untouched."""
    n = 32
    pos = workload(n, "clustered", seed=7)
    p = lagpm.make_params(n, theta=1e-6, precision=prec)
    ids = np.random.default_rng(4).choice(len(pos), 300, replace=False)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        a = b.tree_accel_sample(ids)
        v = b.velocities()
        b.treecorrection(lagpm.evolfact(1.0))
        full = b.velocities()[ids]
    assert np.all(v == 0.0)
    ref = exact_erfc(n, p, pos)[ids]
    e_ref, e_full = rel_rms(a, ref), rel_rms(a, full)
    print(f"sampled theta0 prec={prec}: vs exact {e_ref:.2e}, vs full walk {e_full:.2e}", flush=True)
    # both precisions interpolate the tabulated law (~2e-6 from the analytic one); the FP64 walk of a
    # lone particle adds the same nonzero terms in the same order as its group's walk
    assert e_ref < 3e-5
    assert e_full < (3e-5 if prec == lagpm.PREC_FP32 else 1e-12)


def test_sampled_walk_theta_half_matches_the_per_particle_reference_walk(ctx):
    """At theta = 0.5 a sampled particle is opened with the reference's per-particle test; its
    error against the exact sum is the same size as the reference walk's (a different octree
    topology), and FP32 and FP64 sampled walks agree to float rounding."""
    n = 48
    pos = workload(n, "clustered", seed=8)
    ids = np.sort(np.random.default_rng(5).choice(len(pos), 2000, replace=False))
    a = {}
    for prec in (lagpm.PREC_FP32, lagpm.PREC_FP64):
        p = lagpm.make_params(n, theta=0.5, precision=prec)
        with lagpm.Box(ctx, p, len(pos)) as b:
            b.upload_pos(pos)
            a[prec] = b.tree_accel_sample(ids)
    exact = exact_erfc(n, p, pos)[ids]
    diff, slope = orc.force_spline(1, n, p.epsilon, p.efold / 2)
    ref_walk = orc.treecorrection_ref(n, pos, diff, slope, p.rsphere, 0.5, cellw=max(4.0, p.rsphere), targets=ids)
    e32, e64, eref = rel_rms(a[lagpm.PREC_FP32], exact), rel_rms(a[lagpm.PREC_FP64], exact), rel_rms(ref_walk, exact)
    print(f"sampled theta=0.5: fp32 {e32:.2e} fp64 {e64:.2e} reference walk {eref:.2e}", flush=True)
    assert rel_rms(a[lagpm.PREC_FP32], a[lagpm.PREC_FP64]) < 1e-4
    assert e64 < 2.0 * eref and e64 > 0.2 * eref


def test_sampled_walk_rejects_bad_indices(ctx):
    n = 32
    pos = workload(n, "lattice", seed=9)
    with lagpm.Box(ctx, lagpm.make_params(n), len(pos)) as b:
        b.upload_pos(pos)
        with pytest.raises(lagpm.LagpmError, match="repeated"):
            b.tree_accel_sample([3, 5, 3])
        with pytest.raises(lagpm.LagpmError, match="out of range"):
            b.tree_accel_sample([0, len(pos)])
        assert b.tree_accel_sample([]).shape == (0, 3)


def test_walk_census_counts_the_pairs_within_the_cutoff(ctx):
    """With the census on, the FP32 walk counts the evaluated pairs within the cutoff. At theta -> 0
    every pair within the cutoff is evaluated exactly once (itself included), so the count is the
    number of ordered periodic pairs closer than r_cut (scipy's kd-tree); the forces are the same
    with the census on or off."""
    from scipy.spatial import cKDTree
    n = 32
    pos = np.mod(workload(n, "clustered", seed=12), n)
    pos[pos >= n] = 0.0  # np.mod can round up to n itself
    p = lagpm.make_params(n, theta=1e-6)
    a0, _ = tree_accel(ctx, p, pos)
    ctx.set_walk_census(True)
    try:
        a1, st = tree_accel(ctx, p, pos)
    finally:
        ctx.set_walk_census(False)
    exact = cKDTree(pos, boxsize=n).count_neighbors(cKDTree(pos, boxsize=n), p.rsphere)
    print(f"census: evaluated {st.n_pp}, within cutoff {st.n_pp_within}, kd-tree {exact}", flush=True)
    assert np.array_equal(a0, a1)
    assert abs(int(st.n_pp_within) - int(exact)) <= 1e-5 * exact
    assert st.n_pp_within <= st.n_pp
    _, st_off = tree_accel(ctx, p, pos)
    assert st_off.n_pp_within == 0


@pytest.mark.parametrize("mode", [lagpm.GROUPS_RUNS, lagpm.GROUPS_PACKED])
@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_run_groups_theta_zero_is_the_direct_sum(ctx, prec, mode):
    """Groups cut as runs of consecutive particles (of a chaining cell, or packed from the small
    children of a node) give every particle its exact pair sum at theta -> 0, as tree groups do."""
    n = 32
    pos = workload(n, "clustered", seed=14)
    p = lagpm.make_params(n, theta=1e-6, precision=prec)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.set_tuning(16, 64, mode)
        assert b.tuning() == (16, 64, mode)
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(1.0))
        a, st = b.velocities(), b.tree_stats()
    err = rel_rms(a, exact_erfc(n, p, pos))
    print(f"groups mode {mode} theta0 prec={prec}: rms {err:.2e}, groups {st.n_groups} "
          f"({len(pos) / st.n_groups:.1f} targets each)", flush=True)
    assert err < 3e-5  # both precisions interpolate the tabulated law (~2e-6 from the analytic one)


def test_packed_groups_fill_more_lanes_than_tree_groups(ctx):
    """Packing the small children of a node into runs of exactly `group` particles leaves fewer,
    fuller groups than merging whole children; every particle is in exactly one group either way
    (theta -> 0 gives the same exact sums)."""
    n = 48
    pos = workload(n, "clustered", seed=16)
    p = lagpm.make_params(n, theta=1e-6)
    res = {}
    for mode in (lagpm.GROUPS_TREE, lagpm.GROUPS_PACKED):
        with lagpm.Box(ctx, p, len(pos)) as b:
            b.set_tuning(16, 64, mode)
            b.upload_pos(pos)
            b.treecorrection(lagpm.evolfact(1.0))
            res[mode] = (b.velocities(), b.tree_stats())
    (a_t, st_t), (a_p, st_p) = res[lagpm.GROUPS_TREE], res[lagpm.GROUPS_PACKED]
    print(f"targets per group: tree {len(pos) / st_t.n_groups:.1f}, packed {len(pos) / st_p.n_groups:.1f}", flush=True)
    assert st_p.n_groups < st_t.n_groups
    assert rel_rms(a_p, a_t) < 1e-5


@pytest.mark.parametrize("mode", [lagpm.GROUPS_RUNS, lagpm.GROUPS_PACKED])
def test_run_groups_theta_half_fp32_vs_fp64_and_no_less_accurate(ctx, mode):
    """At theta = 0.5 run or packed groups open nodes by the groups' own boxes: FP32 agrees with
    FP64 on the same groups, and the error against the exact sum stays within the reference walk's."""
    n = 48
    pos = workload(n, "clustered", seed=15)
    a = {}
    for prec in (lagpm.PREC_FP32, lagpm.PREC_FP64):
        with lagpm.Box(ctx, lagpm.make_params(n, theta=0.5, precision=prec), len(pos)) as b:
            b.set_tuning(16, 64, mode)
            b.upload_pos(pos)
            b.treecorrection(lagpm.evolfact(1.0))
            a[prec] = b.velocities()
    p = lagpm.make_params(n, theta=0.5)
    exact = exact_erfc(n, p, pos)
    diff, slope = orc.force_spline(1, n, p.epsilon, p.efold / 2)
    ref_walk = orc.treecorrection_ref(n, pos, diff, slope, p.rsphere, 0.5, cellw=max(4.0, p.rsphere))
    e_prec, e_tree, e_ref = rel_rms(a[lagpm.PREC_FP32], a[lagpm.PREC_FP64]), rel_rms(a[lagpm.PREC_FP32], exact), \
        rel_rms(ref_walk, exact)
    print(f"groups mode {mode} theta=0.5: fp32 vs fp64 {e_prec:.2e}; vs exact {e_tree:.2e}; reference walk {e_ref:.2e}",
          flush=True)
    assert e_prec < 1e-4
    assert e_tree <= 1.02 * e_ref


@pytest.mark.parametrize("group,mode", [(24, lagpm.GROUPS_TREE), (48, lagpm.GROUPS_TREE), (96, lagpm.GROUPS_TREE),
                                        (128, lagpm.GROUPS_TREE), (96, lagpm.GROUPS_PACKED), (128, lagpm.GROUPS_RUNS)])
def test_groups_walked_three_targets_per_lane_are_exact(ctx, group, mode):
    """A wave walks a group of 33-48 targets (17-24, 9-12, 5-6) with three targets per lane, and
    of 65-192 targets only that way (packed and run groups fill 96 and 128): at theta -> 0 every
    particle still gets its exact pair sum, and at theta = 0.5 the FP32 walk agrees with the plain
    FP64 walk of the same groups."""
    n = 40
    pos = workload(n, "clustered", seed=21)
    p0 = lagpm.make_params(n, theta=1e-6)
    with lagpm.Box(ctx, p0, len(pos)) as b:
        b.set_tuning(32, group, mode)
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(1.0))
        a0, st = b.velocities(), b.tree_stats()
    err0 = rel_rms(a0, exact_erfc(n, p0, pos))
    a = {}
    for prec in (lagpm.PREC_FP32, lagpm.PREC_FP64):
        with lagpm.Box(ctx, lagpm.make_params(n, theta=0.5, precision=prec), len(pos)) as b:
            b.set_tuning(32, group, mode)
            b.upload_pos(pos)
            b.treecorrection(lagpm.evolfact(1.0))
            a[prec] = b.velocities()
    e_prec = rel_rms(a[lagpm.PREC_FP32], a[lagpm.PREC_FP64])
    print(f"group {group} mode {mode}: {len(pos) / st.n_groups:.1f} targets per group; theta0 rms {err0:.2e}; "
          f"theta 0.5 fp32 vs fp64 {e_prec:.2e}", flush=True)
    assert err0 < 3e-5
    assert e_prec < 1e-4
