"""GPU tests of the whole-step entry points (run with -m gpu): lagpm_kdk_step equals the
reference's step sequence (HubbleDragDM, pmmain KICK, treecorrection, drift; Cosmos/globalT.c:221-306)
made of the single-phase entry points, and lagpm_batch_step gives every box exactly the result
it gets alone, whatever the other boxes in the batch."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from oracle import pyoracle as orc

pytestmark = pytest.mark.gpu


def lattice_box(n, seed, rms=0.3):
    return ics.perturbed_lattice(n, rms=rms, seed=seed)


def run_manual(ctx, p, pos, vel, anow, astep, steps):
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos, vel)
        for s in range(steps):
            ef = lagpm.get_evol_factor(p, anow + s * astep, astep)
            b.hubble_drag(ef.fact1)
            b.pmmain(ef, lagpm.KICK)
            b.treecorrection(ef)
            b.drift(ef.pfact)
        return b.download()


def run_kdk(ctx, p, pos, vel, anow, astep, steps):
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos, vel)
        for s in range(steps):
            b.kdk_step(anow + s * astep, astep)
        st = b.tree_stats()
        return b.download(), st


def test_kdk_step_equals_the_reference_step_sequence(ctx):
    n = 32
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=20, seed=2)
    vel = np.random.default_rng(2).standard_normal(pos.shape) * 0.01
    p = lagpm.make_params(n, omep=0.3, omeplam=0.7)
    print(f"kdk vs manual: N={len(pos)} n={n}", flush=True)
    (x1, v1) = run_manual(ctx, p, pos, vel, 1.0, 0.25, 3)
    (x2, v2), st = run_kdk(ctx, p, pos, vel, 1.0, 0.25, 3)
    assert st.n_pp > 0 and st.n_groups > 0
    assert np.array_equal(x1, x2) and np.array_equal(v1, v2)


def test_kdk_step_moves_particles_by_the_kicked_velocity(ctx):
    """x_new = wrap(x + pfact * v_new), v_new = fact1 v + fact2 (a_PM + a_tree)."""
    n = 32
    pos = lattice_box(n, 4)
    vel = np.random.default_rng(4).standard_normal(pos.shape) * 0.01
    p = lagpm.make_params(n)
    ef = lagpm.get_evol_factor(p, 2.0, 0.5)
    (x1, v1), _ = run_kdk(ctx, p, pos, vel, 2.0, 0.5, 1)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
        b.treecorrection(lagpm.evolfact(1.0))
        acc = b.velocities()
    drag, kick = np.float64(ef.fact1) * vel, np.float64(ef.fact2) * acc
    v_want = drag + kick
    # the kicks are FP32 products of each force with its factor: compare per particle against the
    # size of the terms (a component where drag and kick cancel keeps only their absolute error)
    scale = np.linalg.norm(drag, axis=1) + np.linalg.norm(kick, axis=1)
    assert np.all(np.linalg.norm(v1 - v_want, axis=1) <= 1e-5 * scale + 1e-12)
    x_want = np.mod(pos + np.float64(ef.pfact) * v1, n)
    d = np.abs(x1 - x_want)
    d = np.minimum(d, n - d)
    assert d.max() < 1e-9


@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_batch_equals_each_box_alone(ctx, prec):
    # uneven boxes: different sizes, cosmologies, clustering and scale factors
    specs = [(32, 0.25, 1.0, "clustered"), (40, 0.30, 1.5, "lattice"), (24, 0.35, 3.0, "clustered"),
             (48, 0.28, 1.2, "lattice"), (32, 0.33, 2.0, "lattice")]
    boxes, alone, inputs = [], [], []
    for i, (n, om, a0, kind) in enumerate(specs):
        pos = (ics.clustered_box(n, halo_fraction=0.3, nhalo=max(4, n ** 3 // 2000), seed=10 + i) if kind == "clustered"
               else lattice_box(n, 10 + i))
        vel = np.random.default_rng(i).standard_normal(pos.shape) * 0.01
        p = lagpm.make_params(n, omep=om, omeplam=1 - om, precision=prec)
        inputs.append((p, pos, vel, a0))
    print(f"batch of {len(specs)} boxes, prec={prec}", flush=True)
    for p, pos, vel, a0 in inputs:
        (x, v), _ = run_kdk(ctx, p, pos, vel, a0, 0.1, 2)
        alone.append((x, v))
    try:
        for p, pos, vel, a0 in inputs:
            b = lagpm.Box(ctx, p, len(pos))
            b.upload_pos(pos, vel)
            boxes.append(b)
        anow = np.array([a0 for *_, a0 in inputs])
        for s in range(2):
            lagpm.batch_step(ctx, boxes, anow + 0.1 * s, 0.1)
        for b, (x, v) in zip(boxes, alone):
            xb, vb = b.download()
            assert np.array_equal(xb, x) and np.array_equal(vb, v)
    finally:
        for b in boxes:
            b.close()


@pytest.mark.parametrize("stages", [1, 2, 3])
def test_pipelined_batch_equals_each_box_alone(ctx, stages):
    """A batch run as groups whose walks overlap the next group's fronts (lagpm_set_batch_pipeline)
    gives every box its single-box result bit for bit."""
    specs = [(32, 0.25, 1.0), (40, 0.30, 1.5), (24, 0.35, 3.0), (48, 0.28, 1.2), (32, 0.33, 2.0), (36, 0.27, 1.1),
             (24, 0.31, 1.4)]
    inputs, alone, boxes = [], [], []
    for i, (n, om, a0) in enumerate(specs):
        pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=max(4, n ** 3 // 2000), seed=30 + i)
        vel = np.random.default_rng(40 + i).standard_normal(pos.shape) * 0.01
        inputs.append((lagpm.make_params(n, omep=om, omeplam=1 - om), pos, vel, a0))
    print(f"pipelined batch of {len(specs)} boxes, {stages} stages", flush=True)
    for p, pos, vel, a0 in inputs:
        (x, v), _ = run_kdk(ctx, p, pos, vel, a0, 0.1, 3)
        alone.append((x, v))
    try:
        ctx.set_batch_pipeline(stages, 32)
        for p, pos, vel, a0 in inputs:
            b = lagpm.Box(ctx, p, len(pos))
            b.upload_pos(pos, vel)
            boxes.append(b)
        anow = np.array([a0 for *_, a0 in inputs])
        for s in range(3):  # the third step replays the first step's captured graphs
            lagpm.batch_step(ctx, boxes, anow + 0.1 * s, 0.1)
        for b, (x, v) in zip(boxes, alone):
            xb, vb = b.download()
            assert np.array_equal(xb, x) and np.array_equal(vb, v)
    finally:
        ctx.set_batch_pipeline(3, 32)
        for b in boxes:
            b.close()


def test_mixed_precision_batch_is_not_pipelined_and_equals_each_box_alone(ctx):
    """Seven boxes, FP32 and FP64 mixed: the default pipeline needs one FP32 force table for the
    whole batch, so this batch runs unpipelined (FP64 walks per box, one fused FP32 walk); every box
    still gets its single-box result bit for bit. Bad pipeline settings are refused."""
    with pytest.raises(lagpm.LagpmError):
        ctx.set_batch_pipeline(0, 32)
    with pytest.raises(lagpm.LagpmError):
        ctx.set_batch_pipeline(3, 100000)
    specs = [(24, lagpm.PREC_FP32), (32, lagpm.PREC_FP64), (24, lagpm.PREC_FP64), (36, lagpm.PREC_FP32),
             (24, lagpm.PREC_FP32), (32, lagpm.PREC_FP32), (24, lagpm.PREC_FP64)]
    inputs, alone, boxes = [], [], []
    for i, (n, prec) in enumerate(specs):
        pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=4, seed=50 + i)
        vel = np.random.default_rng(60 + i).standard_normal(pos.shape) * 0.01
        inputs.append((lagpm.make_params(n, precision=prec), pos, vel))
    print(f"mixed batch of {len(specs)} boxes", flush=True)
    for p, pos, vel in inputs:
        (x, v), _ = run_kdk(ctx, p, pos, vel, 1.0, 0.1, 2)
        alone.append((x, v))
    try:
        for p, pos, vel in inputs:
            b = lagpm.Box(ctx, p, len(pos))
            b.upload_pos(pos, vel)
            boxes.append(b)
        for s in range(2):
            lagpm.batch_step(ctx, boxes, 1.0 + 0.1 * s, 0.1)
        for b, (x, v) in zip(boxes, alone):
            xb, vb = b.download()
            assert np.array_equal(xb, x) and np.array_equal(vb, v)
    finally:
        for b in boxes:
            b.close()


@pytest.mark.parametrize("prec", [lagpm.PREC_FP32, lagpm.PREC_FP64])
def test_depth_limit_meets_a_deeper_tree_and_redoes_it(ctx, prec):
    """After a step on a perturbed lattice the build is limited to two levels below that tree's
    deepest level. New particles in tight clumps need the full key depth: the limited build
    flags it, the step redoes the tree at full depth, and the state equals a fresh box's step on
    the same input, bit for bit."""
    n = 32
    lat = lattice_box(n, 3)
    g = np.random.default_rng(21)
    # clumps of 40 (more than a leaf holds) narrower than the finest key cell: full key depth
    centres = g.random((len(lat) // 40, 3)) * n
    tight = np.repeat(centres, 40, axis=0) + g.standard_normal((len(centres) * 40, 3)) * 1e-4
    clumps = np.mod(np.concatenate([tight, g.random((len(lat) - len(tight), 3)) * n]), n)
    vel = g.standard_normal(clumps.shape) * 0.01
    p = lagpm.make_params(n, precision=prec)
    print(f"depth limit: N={len(lat)} prec={prec}", flush=True)
    with lagpm.Box(ctx, p, len(lat)) as a, lagpm.Box(ctx, p, len(lat)) as b:
        a.upload_pos(lat)
        a.kdk_step(1.0, 0.1)
        st = a.tree_stats()
        assert st.depth_limit == min(16, st.deepest + 2) < 16
        a.upload_pos(clumps, vel)
        a.kdk_step(1.1, 0.1)
        st2 = a.tree_stats()
        assert st2.deepest > st.depth_limit and st2.depth_limit == 16
        b.upload_pos(clumps, vel)
        b.kdk_step(1.1, 0.1)
        xa, va = a.download()
        xb, vb = b.download()
    assert np.array_equal(xa, xb) and np.array_equal(va, vb)


@pytest.mark.parametrize("prec,tol", [(lagpm.PREC_FP32, 3e-5), (lagpm.PREC_FP64, 1e-5)])
def test_node_overflow_is_regrown_and_the_kick_is_exact(ctx, prec, tol):
    """Tight 17-particle clumps make about one node per particle, beyond the initial node
    capacity: the build must detect it, grow the arrays and still give the exact pair sum.
    The FP32 walk sees FP32 positions (clump members 2e-4 cells apart differ in the 3rd
    digit at that rounding), so its exact sum is taken over the positions it sees: with the
    erfc law's log-r^2 tables the FP32 tree holds float32(x * sqrt(lam)), lam = float32(32 / rcut^2)."""
    n = 32
    g = np.random.default_rng(12)
    centres = g.random((400, 3)) * n
    pos = (centres[:, None, :] + g.standard_normal((400, 17, 3)) * 2e-4).reshape(-1, 3) % n
    p = lagpm.make_params(n, theta=1e-6, precision=prec)
    with lagpm.Box(ctx, p, len(pos)) as b:
        b.upload_pos(pos)
        b.treecorrection(lagpm.evolfact(1.0))
        a = b.velocities()
        st = b.tree_stats()
        # the regrown arrays hold the tree: a second kick adds exactly the same acceleration
        b.treecorrection(lagpm.evolfact(1.0))
        assert np.array_equal(b.velocities(), 2 * a)
    print(f"clumps prec={prec}: N={len(pos)} nodes {st.n_nodes} groups {st.n_groups}", flush=True)
    assert st.n_nodes > len(pos) // 3
    if prec == lagpm.PREC_FP32:
        sc = np.sqrt(np.float64(np.float32(32.0 / (p.rsphere * p.rsphere))))
        seen = (pos * sc).astype(np.float32).astype(np.float64) / sc
    else:
        seen = pos
    ref = orc.pp_direct_analytic(1, n, p.epsilon, p.efold / 2, seen, p.rsphere)
    err = float(np.sqrt(np.sum((a - ref) ** 2) / np.sum(ref ** 2)))
    print(f"  rms rel {err:.3e}", flush=True)
    assert err < tol
    # and the same input through a whole step (overflow inside the batched path)
    (xk, vk), _ = run_kdk(ctx, p, pos, None, 1.0, 0.1, 1)
    (xm, vm) = run_manual(ctx, p, pos, None, 1.0, 0.1, 1)
    assert np.array_equal(xk, xm) and np.array_equal(vk, vm)


def test_graph_replays_equal_eager_steps_after_every_step(ctx):
    """A box stepped through captured step graphs (replayed, both particle-set parities) equals
    a box stepped with every kernel launched eagerly, compared after each step."""
    n = 32
    pos = ics.clustered_box(n, halo_fraction=0.3, nhalo=16, seed=21)
    vel = np.random.default_rng(21).standard_normal(pos.shape) * 0.01
    p = lagpm.make_params(n)
    with lagpm.Box(ctx, p, len(pos)) as bg, lagpm.Box(ctx, p, len(pos)) as be:
        bg.upload_pos(pos, vel)
        be.upload_pos(pos, vel)
        for s in range(5):
            a = 1.0 + 0.1 * s
            ctx.set_step_graphs(True)
            bg.kdk_step(a, 0.1)
            ctx.set_step_graphs(False)
            be.kdk_step(a, 0.1)
            ctx.synchronize()
            xg, vg = bg.download()
            xe, ve = be.download()
            assert np.array_equal(xg, xe) and np.array_equal(vg, ve), f"step {s}"
        ctx.set_step_graphs(True)
