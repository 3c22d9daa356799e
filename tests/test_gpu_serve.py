"""GPU test of the continuous-batching request loop (run with -m gpu): requests that arrive
while others are running join the running batch as memory frees up, finished boxes are reused
from the pool, and every request's final state equals the state of its box stepped alone."""
import numpy as np
import pytest

from lageunha_amd import ics, lagpm
from lageunha_amd.serve import Request, Server

pytestmark = pytest.mark.gpu


def test_requests_in_the_mix_equal_requests_alone(ctx):
    specs = [(32, 3, 0), (24, 2, 0), (40, 4, 1), (32, 2, 1), (24, 5, 2), (36, 3, 4)]  # (n, steps, arrival)
    reqs = [Request(rid=i, n=n, steps=st, arrival=float(t), omep=0.25 + 0.02 * i, a_start=1.0 + 0.1 * i, da=0.05,
                    seed=i) for i, (n, st, t) in enumerate(specs)]

    def params_for(r):
        return lagpm.make_params(r.n, omep=r.omep, omeplam=1 - r.omep)

    def positions_for(r):
        return ics.perturbed_lattice(r.n, rms=0.3, seed=50 + r.seed)

    biggest = max(lagpm.estimate_box_bytes(params_for(r), r.n ** 3) for r in reqs)
    it = {"t": 0}
    srv = Server(ctx, params_for, positions_for, hbm_budget=1.35 * biggest, keep_results=True,
                 clock=lambda: float(it["t"]))
    for r in reqs:
        srv.submit(r)
    mixes = []
    for _ in range(100):
        if not (srv.queue or srv.active):
            break
        if srv.iterate():
            mixes.append(sorted(srv.active) if srv.active else [])
        it["t"] += 1
    print(f"iterations {srv.iterations}, active sets over time: {mixes}", flush=True)
    assert len(srv.done) == len(reqs)
    # the budget forced some requests to wait, and some iterations held more than one box
    assert any(r.admitted > r.arrival for r in srv.done)
    assert max(it[2] for it in srv.iteration_log) >= 2
    # under this tight budget idle pooled boxes are freed to make room for other sizes
    print(f"boxes created {srv.boxes_created}, reused {srv.boxes_reused}", flush=True)
    assert srv.boxes_created + srv.boxes_reused == len(reqs)
    assert srv.bytes_held <= srv.hbm_budget
    for r in sorted(srv.done, key=lambda r: r.rid):
        assert r.steps_done == r.steps
        with lagpm.Box(ctx, params_for(r), r.n ** 3) as b:
            b.upload_pos(positions_for(r))
            for s in range(r.steps):
                b.kdk_step(r.a_start + s * r.da, r.da)
            x, v = b.download()
        assert np.array_equal(r.result[0], x) and np.array_equal(r.result[1], v), f"request {r.rid}"


def test_pooled_boxes_are_reused_and_step_as_new_ones(ctx):
    """Requests of sizes seen before reuse the finished boxes (their memory, FFT setup and node
    arrays) with their own cosmology and particles; each result equals its box stepped alone."""
    specs = [(32, 2, 0), (24, 3, 0), (32, 3, 3), (24, 2, 4), (32, 2, 6)]  # (n, steps, arrival)
    reqs = [Request(rid=i, n=n, steps=st, arrival=float(t), omep=0.24 + 0.03 * i, a_start=1.0 + 0.2 * i, da=0.1,
                    seed=10 + i) for i, (n, st, t) in enumerate(specs)]

    def params_for(r):
        return lagpm.make_params(r.n, omep=r.omep, omeplam=1 - r.omep)

    def positions_for(r):
        return ics.clustered_box(r.n, halo_fraction=0.3, nhalo=8, seed=r.seed)

    it = {"t": 0}
    srv = Server(ctx, params_for, positions_for, hbm_budget=1e12, keep_results=True, clock=lambda: float(it["t"]))
    for r in reqs:
        srv.submit(r)
    while srv.queue or srv.active:
        srv.iterate()
        it["t"] += 1
    print(f"boxes created {srv.boxes_created}, reused {srv.boxes_reused}", flush=True)
    assert srv.boxes_created == 2 and srv.boxes_reused == 3
    for r in sorted(srv.done, key=lambda r: r.rid):
        with lagpm.Box(ctx, params_for(r), r.n ** 3) as b:
            b.upload_pos(positions_for(r))
            for s in range(r.steps):
                b.kdk_step(r.a_start + s * r.da, r.da)
            x, v = b.download()
        assert np.array_equal(r.result[0], x) and np.array_equal(r.result[1], v), f"request {r.rid}"
    srv.close()
