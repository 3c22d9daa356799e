#!/usr/bin/env python3
"""TreePM gravity-step benchmark on one MI355X.

Workload (the single-box clustered configuration): n^3 particles in a periodic box of n cells
(default n = 256), 70% on seeded-shuffled lattice sites displaced by a Zel'dovich field with
P(k) ~ k^-1 (rms 0.3 cell), 30% in 2000 Plummer spheres (a = 0.5 cell, truncated at 10 a,
counts from dN/dM ~ M^-1.9 on [32, 1e5]); n^3 PM mesh; Gaussian split r_s = 1.25 cells
(efold 2.5) with TSC^2 deconvolution; erfc short-range law, cutoff 4.5 r_s = 5.625 cells;
theta = 0.5; flat LCDM Om = 0.26, AMAX = 48, a = 1, da = 0.25.

One step = one GOTPM step through the C ABI's whole-step entry point (lagpm_kdk_step): Hubble
drag, PM kick, tree kick, drift, in the reference's order (Cosmos/globalT.c:221-306). The sort,
tree build and every kernel run inside every step; the PM kick runs beside the tree walk.
value = particles / step time = TreePM forces/s (one force evaluation per step). Per-phase
device times come from two more steps launched eagerly after the timed ones.

Printed: one JSON line (rank 0) with roofline (dominant kernel, live HIP-event durations),
cpu_baseline (the CPU oracle on a bounded sample), accuracy vs the plain FP64 kernels.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "TreePM forces/s @256^3, rms err<=1e-3 vs FP64; batched p50/p99 step latency"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
PROFILE_TRAFFIC = os.path.join(ROOT, "profiles", "traffic.json")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


class ProfilerRegion:
    """roctxProfilerResume/Pause around the timed steps: under `rocprofv3 --selected-regions` only
    the timed steps are traced, so a kernel summary covers exactly the launches that `value`
    times. Outside a profiler (or without the roctx library) both calls do nothing."""

    def __init__(self):
        import ctypes
        self.lib = None
        for name in ("librocprofiler-sdk-roctx.so.1", "/opt/rocm/lib/librocprofiler-sdk-roctx.so.1"):
            try:
                lib = ctypes.CDLL(name)
                lib.roctxProfilerResume.argtypes = [ctypes.c_uint64]
                lib.roctxProfilerPause.argtypes = [ctypes.c_uint64]
                self.lib = lib
                break
            except (OSError, AttributeError):
                continue

    def __enter__(self):
        if self.lib is not None:
            self.lib.roctxProfilerResume(0)
        return self

    def __exit__(self, *exc):
        if self.lib is not None:
            self.lib.roctxProfilerPause(0)
        return False


def rel_errors(a, ref):
    d = np.linalg.norm(a - ref, axis=1)
    r = np.linalg.norm(ref, axis=1)
    rms = float(np.sqrt(np.sum(d ** 2) / np.sum(r ** 2)))
    per = d / np.maximum(r, 1e-300)
    return rms, float(np.percentile(per, 99)), float(np.median(per))


def force_at(ctx, lagpm, params, pos, tuning):
    """PM + tree acceleration at fixed positions (fact2 = 1, zero velocities). Returns the
    acceleration and two times: the first evaluation on a new box (which also creates the
    precision's FFT plans, possibly compiling their kernels, and finds the tree's depth) and a
    second one from the same positions and zero velocities (the steady-state cost)."""
    with lagpm.Box(ctx, params, len(pos)) as b:
        b.set_tuning(*tuning)
        dts = []
        for rep in range(2):
            b.upload_pos(pos)
            ctx.synchronize()
            t0 = time.perf_counter()
            b.pmmain(lagpm.evolfact(1.0), lagpm.KICK)
            b.treecorrection(lagpm.evolfact(1.0))
            ctx.synchronize()
            dts.append(time.perf_counter() - t0)
            if rep == 0:
                v = b.velocities()
    return v, dts[1], dts[0]


def cpu_info():
    """CPU model and core counts of this host (lscpu's sources: /proc/cpuinfo)."""
    model, cores = "unknown", set()
    try:
        phys = core = None
        for line in open("/proc/cpuinfo"):
            k, _, v = line.partition(":")
            k, v = k.strip(), v.strip()
            if k == "model name" and model == "unknown":
                model = v
            elif k == "physical id":
                phys = v
            elif k == "core id":
                core = v
            elif not k and phys is not None:
                cores.add((phys, core))
                phys = core = None
    except OSError:
        pass
    return model, len(cores) or None, os.cpu_count()


PHASES = ("TSC", "F-FFT", "psolv", "B-FFT", "4FDA")  # the reference's timers, Cosmos/pmmain.c:447-522


def cpu_force_eval(orc, n, pos, prec, theta, targets=None):
    """One TreePM force evaluation by the CPU oracle (the reference algorithm: TSC, FFT, psolver,
    FDA4, then the per-cell reference tree walk of the tree correction) with per-phase wall times.
    prec 1: FP64 mesh chain; 0: the reference's float mesh. The tree walk is the reference's (FP64
    positions, FP32 pair terms). With `targets`, the tree runs for those particles only."""
    diff, slope = orc.force_spline(1, n, 0.1, 1.25)
    pm, ph = orc.pm_accel_timed(n, pos, mass=1.0, efold=2.5, deconvolve=1, prec=prec)
    t0 = time.perf_counter()
    tr = orc.treecorrection_ref(n, pos, diff, slope, 5.625, theta, cellw=5.625, targets=targets)
    ph["Tree"] = time.perf_counter() - t0
    return pm, tr, ph


def cpu_baseline(args, n_cfg, pos_cfg, gpu_force):
    """The CPU leg: the oracle's restatement of the reference's single-rank path, timed on this
    host's cores, and the reference CPU TreePM forces as an accuracy yardstick.

    value: the 64^3 configuration (N = 262,144 on a perturbed lattice, P(k) ~ k^-1, rms 0.3 cell,
    seed 1; FP64 mesh chain; erfc split r_s = 1.25, cutoff 5.625, theta 0.5), one force evaluation
    on every core of this host's share, and again on one thread. cfg1: the bench's own 256^3
    clustered box with the reference's float mesh: the whole PM chain, and the tree for a 1 %
    random subsample of targets (its time scaled to all targets). The same run gives the reference
    CPU TreePM force on that subsample, against which the GPU force of the same particles is
    compared (accuracy.vs_reference_cpu); and a 32^3 clustered box is compared with an FP64 Ewald
    sum (accuracy.vs_ewald)."""
    from lageunha_amd import ics, lagpm
    from oracle import pyoracle as orc
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or min(16, os.cpu_count() or 1)
    model, phys, logical = cpu_info()
    out = {"unit": "forces/s", "kind": "port", "cores": threads, "cpu_model": model,
           "host_physical_cores": phys, "host_logical_cpus": logical}
    # cfg0: 64^3, FP64, all cores then one thread
    n0 = 64
    pos0 = ics.perturbed_lattice(n0, rms=0.3, seed=1)
    rates = {}
    for nt in (threads, 1):
        orc.set_threads(nt)
        t0 = time.perf_counter()
        _, _, ph = cpu_force_eval(orc, n0, pos0, 1, args.theta)
        dt = time.perf_counter() - t0
        rates[nt] = {"forces_per_s": len(pos0) / dt, "seconds": dt, "phases_s": ph}
        log(f"[bench] cpu baseline 64^3 FP64 on {nt} threads: {dt:.2f} s ({ph})")
    out["value"] = rates[threads]["forces_per_s"]
    out["seconds"] = rates[threads]["seconds"]
    out["phases_s"] = rates[threads]["phases_s"]
    out["one_thread"] = rates[1]
    out["sample"] = (f"64^3 perturbed lattice (262,144 particles, P(k)~k^-1, rms 0.3 cell, seed 1), 64^3 mesh, "
                     f"FP64 mesh chain, erfc split r_s 1.25, cutoff 5.625, theta {args.theta}: one force "
                     f"evaluation of the oracle's restatement of the reference path on {threads} OpenMP threads "
                     f"({model}); phases named as the reference's timers (Cosmos/pmmain.c:447-522, Tree: "
                     f"Cosmos/Treemain.c:662)")
    orc.set_threads(threads)
    # cfg1: the 256^3 bench box, float mesh (as the reference), tree on a 1 % subsample
    N = len(pos_cfg)
    sel = np.sort(np.random.default_rng(args.seed + 23).choice(N, max(1, N // 100), replace=False))
    pm, tr, ph = cpu_force_eval(orc, n_cfg, pos_cfg, 0, args.theta, targets=sel)
    t_full = sum(v for k, v in ph.items() if k != "Tree") + ph["Tree"] * N / len(sel)
    out["cfg1"] = {"n": n_cfg, "particles": N, "tree_targets": int(len(sel)), "phases_s": ph,
                   "forces_per_s": N / t_full,
                   "note": "PM chain over all particles; tree walk for the subsample only, its time scaled "
                           "by N / targets in forces_per_s"}
    log(f"[bench] cpu baseline {n_cfg}^3 (tree on {len(sel)} targets): {ph}")
    acc = {}
    if gpu_force is not None:
        ref = pm[sel] + tr
        rms, p99, med = rel_errors(gpu_force[sel], ref)
        acc["vs_reference_cpu"] = {
            "rms_rel": rms, "p99_rel": p99, "median_rel": med, "targets": int(len(sel)),
            "reference": "oracle restatement of the reference CPU TreePM on the same particles: float TSC, FFT, "
                         "psolver (efold 2.5, TSC^2 deconvolution), FDA4 (Cosmos/pmmain.c:390-527) plus the "
                         "per-cell octree walk of the tree correction (Cosmos/treecorrection.c:225-475) with the "
                         f"same erfc law, cutoff and theta {args.theta}"}
    # small N against physics: 32^3 clustered, FP64 Ewald of Plummer-softened particles
    n3 = 32
    pos3 = ics.clustered_box(n3, halo_fraction=0.3, nhalo=16, seed=3)
    ewald = orc.ewald_converged(pos3, n3, eps=0.1)
    ctx = lagpm.Context(0)
    try:
        ve = {"particles": len(pos3), "reference": "FP64 Ewald sum of Plummer-softened (eps 0.1 cell) periodic "
                                                   "point masses, 32^3 clustered (30 % in 16 halos, seed 3)"}
        for th in (args.theta, 1e-6):
            p = lagpm.make_params(n3, theta=th)
            a = lagpm.treepm_accel(ctx, p, pos3).total
            pmr, trr, _ = cpu_force_eval(orc, n3, pos3, 0, th)
            e_gpu, e_ref = rel_errors(a, ewald), rel_errors(pmr + trr, ewald)
            ve[f"theta_{th:g}"] = {"gpu_rms_rel": e_gpu[0], "gpu_p99_rel": e_gpu[1],
                                   "reference_cpu_rms_rel": e_ref[0], "reference_cpu_p99_rel": e_ref[1]}
        acc["vs_ewald"] = ve
    finally:
        ctx.close()
    return out, acc


SMOOTH_SIZES = (32, 36, 48, 54, 64, 72, 96, 108, 128, 144, 162, 192)


MODE_NAMES = {0: "tree", 1: "runs", 2: "packed"}  # lagpm.GROUPS_TREE, GROUPS_RUNS, GROUPS_PACKED


def smooth_round(n):
    """Nearest mesh size with prime factors 2 and 3 only (rocFFT single-kernel lengths)."""
    return min(SMOOTH_SIZES, key=lambda s: (abs(np.log(s / n)), s))


def ensemble_specs(nbox, nmin, nmax, seed):
    """Box sizes log-uniform in [nmin, nmax] rounded to 2,3-smooth meshes; Omega_m ~ U[0.25, 0.35]."""
    g = np.random.default_rng(seed)
    ns = [smooth_round(float(np.exp(g.uniform(np.log(nmin), np.log(nmax))))) for _ in range(nbox)]
    oms = g.uniform(0.25, 0.35, nbox)
    return list(zip(ns, oms))


def ensemble(args, ctx, lagpm, ics):
    """64 independent boxes, N_b = n_b^3 on perturbed lattices (seed b), n_b log-uniform in [32, 128],
    flat LCDM with Omega_m,b ~ U[0.25, 0.35]; 10 GOTPM steps from a = 1 with da = 0.1.
    plain: every box stepped alone (lagpm_kdk_step), one after another;
    batched: all boxes per step in one lagpm_batch_step (fused walk, concurrent streams).
    Both start from the same state; their final states are compared bit for bit."""
    specs = ensemble_specs(args.ens_boxes, 32, 128, args.seed)
    sets = {"plain": [], "batched": []}
    total = 0
    t0 = time.perf_counter()
    for b, (n, om) in enumerate(specs):
        pos = ics.perturbed_lattice(n, rms=0.3, seed=100 + b)
        p = lagpm.make_params(n, law=lagpm.LAW_ERFC, precision=lagpm.PREC_FP32, deconvolve=1, efold=2.5, epsilon=0.1,
                              rsphere=5.625, theta=args.theta, amax=48.0, omep=om, omeplam=1.0 - om)
        for k in sets:
            bx = lagpm.Box(ctx, p, len(pos))
            bx.upload_pos(pos)
            sets[k].append(bx)
        total += len(pos)
    # autotuned tree tuning per box size (on a scratch box with the first box of that size's
    # particles), applied to the boxes of that size in both sets
    tuned = {}
    if not args.no_autotune:
        for b, (n, om) in enumerate(specs):
            if n in tuned:
                continue
            with lagpm.Box(ctx, sets["plain"][b].params, sets["plain"][b].np) as tb:
                tb.upload_pos(ics.perturbed_lattice(n, rms=0.3, seed=100 + b))
                tuned[n] = tb.autotune()
        for k in sets:
            for bx, (n, _) in zip(sets[k], specs):
                bx.set_tuning(*tuned[n])
    log(f"[bench] ensemble: {len(specs)} boxes, {total} particles, sizes {sorted(set(n for n, _ in specs))} "
        f"(setup {time.perf_counter() - t0:.1f} s; tuning {tuned})")
    a0, da, steps = 1.0, 0.1, args.ens_steps
    # warm-up step (plans, tables) in both modes
    for bx in sets["plain"]:
        bx.kdk_step(a0, da)
    lagpm.batch_step(ctx, sets["batched"], a0, da)
    lat = {"plain": [], "batched": []}
    tt = {}
    t0 = time.perf_counter()
    for s in range(steps):
        for bx in sets["plain"]:
            t = time.perf_counter()
            bx.kdk_step(a0 + da * (s + 1), da)
            lat["plain"].append(time.perf_counter() - t)
    tt["plain"] = time.perf_counter() - t0
    log(f"[bench] ensemble plain: {tt['plain']:.2f} s")
    t0 = time.perf_counter()
    for s in range(steps):
        t = time.perf_counter()
        lagpm.batch_step(ctx, sets["batched"], a0 + da * (s + 1), da)
        lat["batched"].append(time.perf_counter() - t)
    tt["batched"] = time.perf_counter() - t0
    log(f"[bench] ensemble batched: {tt['batched']:.2f} s")
    # the same batched steps with every kernel launched (no step graphs), for comparison
    ctx.set_step_graphs(False)
    t0 = time.perf_counter()
    for s in range(2):
        lagpm.batch_step(ctx, sets["batched"], a0 + da * (steps + 1 + s), da)
    t_eager = (time.perf_counter() - t0) / 2
    ctx.set_step_graphs(True)
    for s in range(2):  # keep the plain set in step with the batched one
        for bx in sets["plain"]:
            bx.kdk_step(a0 + da * (steps + 1 + s), da)
    same = True
    for bp, bb in zip(sets["plain"], sets["batched"]):
        xp, vp = bp.download()
        xb, vb = bb.download()
        same = same and np.array_equal(xp, xb) and np.array_equal(vp, vb)
    for k in sets:
        for bx in sets[k]:
            bx.close()
    out = {"boxes": len(specs), "particles": total, "steps": steps,
           "sizes": sorted(n for n, _ in specs), "identical_plain_vs_batched": bool(same),
           "tree_tuning": {str(n): {"nleaf": t[0], "group": t[1], "group_mode": MODE_NAMES[t[2]]}
                           for n, t in sorted(tuned.items())},
           "batch_pipeline": {"stages": 5, "reserve_cu": 32, "share_growth": 1.5,
                              "note": "library default (lagpm_set_batch_pipeline)"}}
    for k in ("plain", "batched"):
        L = np.array(lat[k]) * 1e3
        out[k] = {"particle_steps_per_s": total * steps / tt[k], "seconds": tt[k],
                  "step_latency_p50_ms": float(np.percentile(L, 50)), "step_latency_p99_ms": float(np.percentile(L, 99)),
                  "latency_unit": "one box-step (plain) / one batch step of all boxes (batched)"}
    out["batched_over_plain"] = out["batched"]["particle_steps_per_s"] / out["plain"]["particle_steps_per_s"]
    out["batched_without_graphs"] = {"particle_steps_per_s": total / t_eager, "step_ms": t_eager * 1e3}
    # the plain FP64 kernels, box by box (one step each after a warm-up step)
    t_fp64, ps = 0.0, 0
    for b, (n, om) in enumerate(specs):
        pos = ics.perturbed_lattice(n, rms=0.3, seed=100 + b)
        p = lagpm.make_params(n, law=lagpm.LAW_ERFC, precision=lagpm.PREC_FP64, deconvolve=1, efold=2.5, epsilon=0.1,
                              rsphere=5.625, theta=args.theta, amax=48.0, omep=om, omeplam=1.0 - om)
        with lagpm.Box(ctx, p, len(pos)) as bx:
            bx.upload_pos(pos)
            bx.kdk_step(a0, da)
            t = time.perf_counter()
            bx.kdk_step(a0 + da, da)
            t_fp64 += time.perf_counter() - t
            ps += len(pos)
    out["plain_fp64"] = {"particle_steps_per_s": ps / t_fp64, "seconds": t_fp64, "steps": 1}
    out["batched_over_plain_fp64"] = out["batched"]["particle_steps_per_s"] / out["plain_fp64"]["particle_steps_per_s"]
    log(f"[bench] ensemble plain FP64: {out['plain_fp64']['particle_steps_per_s']:.3e} particle-steps/s")
    return out


def serving(args, ctx, lagpm, ics):
    """Poisson stream of box requests (rate 20/s over a window), n log-uniform in [32, 192]
    rounded to 2,3-smooth meshes, 5..50 steps each, Omega_m ~ U[0.25, 0.35]; admission under a
    200 GB HBM budget and a cap of 32M active particles per iteration (the per-step latency
    bound); continuous batching (one lagpm_batch_step over all active boxes per iteration). Initial conditions: one perturbed lattice per mesh size, prepared before the
    clock starts, shifted by a random whole-cell offset per request."""
    from lageunha_amd.serve import Request, Server
    g = np.random.default_rng(args.seed + 7)
    t, reqs = 0.0, []
    while True:
        t += g.exponential(1.0 / args.serve_rate)
        if t > args.serve_window:
            break
        n = smooth_round(float(np.exp(g.uniform(np.log(32), np.log(192)))))
        reqs.append(Request(rid=len(reqs), n=n, steps=int(g.integers(5, 51)), arrival=t,
                            omep=float(g.uniform(0.25, 0.35)), a_start=1.0, da=0.1, seed=int(g.integers(1 << 30))))
    t0 = time.perf_counter()
    base = {n: ics.perturbed_lattice(n, rms=0.3, seed=n) for n in sorted({r.n for r in reqs})}
    log(f"[bench] serving: {len(reqs)} requests over {args.serve_window:.0f} s, ICs for sizes {sorted(base)} "
        f"({time.perf_counter() - t0:.1f} s)")

    def params_for(r):
        return lagpm.make_params(r.n, omep=r.omep, omeplam=1 - r.omep, theta=args.theta)

    def positions_for(r):
        off = np.random.default_rng(r.seed).integers(0, r.n, 3)
        return np.mod(base[r.n] + off, r.n)

    # one tree tuning for every request's box, autotuned on a 128^3 box (the batched walk is one
    # fused launch over all active boxes; see the ensemble)
    tuning = None
    if not args.no_autotune:
        with lagpm.Box(ctx, params_for(Request(-1, 128, 1, 0.0)), 128 ** 3) as tb:
            tb.upload_pos(ics.perturbed_lattice(128, rms=0.3, seed=128))
            tuning = tb.autotune()
    # warm the FFT plans of every size (not timed)
    for n in base:
        with lagpm.Box(ctx, params_for(Request(-1, n, 1, 0.0)), n ** 3) as b:
            b.upload_pos(base[n])
            b.kdk_step(1.0, 0.1)
    start = time.perf_counter()
    clock = lambda: time.perf_counter() - start  # noqa: E731
    srv = Server(ctx, params_for, positions_for, hbm_budget=args.serve_budget_gb * 1e9,
                 max_active_particles=args.serve_max_particles or None, clock=clock, tuning=tuning)
    for r in reqs:
        srv.submit(r)
    last = 0.0
    while srv.queue or srv.active:
        if not srv.iterate():
            if srv.queue:
                time.sleep(max(0.0, min(0.01, srv.queue[0].arrival - clock())))
        if clock() - last > 5.0:
            last = clock()
            log(f"[bench] serving t={last:.1f} s: done {len(srv.done)}/{len(reqs)}, active {len(srv.active)}, "
                f"queued {len(srv.queue)}")
    wall = clock()
    st = srv.stats()
    st.update({"wall_seconds": wall, "particle_steps_per_s": st["particle_steps"] / wall,
               "arrival_rate_per_s": args.serve_rate, "window_s": args.serve_window,
               "hbm_budget_gb": args.serve_budget_gb, "max_active_particles": args.serve_max_particles or None,
               "tree_tuning": list(tuning) if tuning else None,
               "latency_unit": "one box-step = the lagpm_batch_step iteration that carried it"})
    return st


def copy_calibration(mib):
    """A plain device-to-device copy of known size: FETCH_SIZE / WRITE_SIZE read for it give the
    factor that turns those counters into HBM bytes for the TreePM kernels."""
    import torch
    nbytes = mib << 20
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda:0")
    y = torch.empty_like(x)
    for _ in range(5):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes, "launches": 5}), flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--n", type=int, default=256)
    ap.add_argument("--halo-fraction", type=float, default=0.0, help="default 0.3 (0.4 at n >= 512)")
    ap.add_argument("--nhalo", type=int, default=0, help="default 2000 scaled with volume beyond 256^3")
    ap.add_argument("--acc-sample", type=float, default=1.0, help="fraction of particles in the accuracy statistics")
    ap.add_argument("--theta", type=float, default=0.5)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-accuracy", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--nleaf", type=int, default=0)
    ap.add_argument("--group", type=int, default=0)
    ap.add_argument("--group-mode", choices=["tree", "runs", "packed"], default=None)
    ap.add_argument("--no-ensemble", action="store_true")
    ap.add_argument("--ens-boxes", type=int, default=64)
    ap.add_argument("--ens-steps", type=int, default=10)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--serve-rate", type=float, default=20.0)
    ap.add_argument("--serve-window", type=float, default=8.0)
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    # an iteration over at most 32M particles bounds the per-step latency (about 35 ms); 0: no cap
    ap.add_argument("--serve-max-particles", type=int, default=32_000_000)
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB five times (HBM counter calibration)")
    args = ap.parse_args()
    if args.copy_calibration > 0:
        copy_calibration(args.copy_calibration)
        return

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus > 1 or world > 1:
        print(json.dumps({"error": "this library runs on one GPU: the TreePM step of one box is not sharded; "
                                   "run with --gpus 1 (sharding one box over several GPUs is not built)"}))
        sys.exit(2)

    from lageunha_amd import ics, lagpm

    n = args.n
    t_ic = time.perf_counter()
    hfrac = args.halo_fraction or (0.4 if n >= 512 else 0.3)
    nhalo = args.nhalo or max(2000, int(round(2000 * (n / 256) ** 3)))
    pos = ics.clustered_box(n, halo_fraction=hfrac, nhalo=nhalo, seed=args.seed)
    N = len(pos)
    log(f"[bench] workload: {n}^3 box, {N} particles (IC {time.perf_counter() - t_ic:.1f} s)")
    params = lagpm.make_params(n, law=lagpm.LAW_ERFC, precision=lagpm.PREC_FP32, deconvolve=1, efold=2.5,
                               epsilon=0.1, rsphere=5.625, theta=args.theta, amax=48.0, omep=0.26, omeplam=0.74)
    ef = lagpm.get_evol_factor(params, 1.0, 0.25)

    ctx = lagpm.Context(0)
    box = lagpm.Box(ctx, params, N)
    box.upload_pos(pos)
    if args.nleaf or args.group or args.group_mode:
        box.set_tuning(args.nleaf or 16, args.group or 64,
                       {"tree": lagpm.GROUPS_TREE, "runs": lagpm.GROUPS_RUNS, "packed": lagpm.GROUPS_PACKED}[args.group_mode]
                       if args.group_mode else None)
    elif not args.no_autotune:
        t_at = time.perf_counter()
        nl, gs, gm = box.autotune()
        log(f"[bench] autotune: nleaf {nl}, group {gs}, group mode {gm} ({time.perf_counter() - t_at:.2f} s)")
    tuning = box.tuning()

    phases, stats = [], []

    def step():
        # one GOTPM step through the whole-step entry point: drag, sort and tree build, then the
        # walk beside the PM kick, then drift (step graphs on)
        box.kdk_step(1.0, 0.25)

    # first step traced phase by phase (synchronised), so a slow phase names itself
    tt = time.perf_counter()
    box.hubble_drag(ef.fact1)
    ctx.synchronize()
    log(f"[bench] trace: drag done {time.perf_counter() - tt:.2f} s")
    box.pmmain(ef, lagpm.KICK)
    ctx.synchronize()
    log(f"[bench] trace: pmmain done {time.perf_counter() - tt:.2f} s ({box.timings().as_dict()})")
    box.treecorrection(ef)
    log(f"[bench] trace: treecorrection done {time.perf_counter() - tt:.2f} s ({box.timings().as_dict()})")
    box.drift(ef.pfact)
    ctx.synchronize()
    log(f"[bench] trace: first step done {time.perf_counter() - tt:.2f} s")
    for i in range(args.warmup):
        step()
        log(f"[bench] warmup step {i + 1}/{args.warmup} done")
    ctx.synchronize()
    walk_ms, walk_stats = [], []
    with ProfilerRegion():
        t0 = time.perf_counter()
        for i in range(args.steps):
            step()
            # the walk runs outside the step graphs, between HIP events on the context stream (the
            # stream it is launched on); the step has completed when kdk_step returns
            walk_ms.append(box.timings().ms_walk)
            s = box.tree_stats()
            walk_stats.append((s.n_pp, s.n_pc))
        ctx.synchronize()
        t1 = time.perf_counter()
    ms = (t1 - t0) * 1e3 / args.steps
    log(f"[bench] timed {args.steps} steps: {ms:.2f} ms/step")
    # per-phase device times: two more steps with the kernels launched eagerly (phase events are
    # not recorded inside graphs); the PM phases overlap the walk there as in the timed steps
    ctx.set_step_graphs(False)
    for i in range(2):
        step()
        phases.append(box.timings().as_dict())
        s = box.tree_stats()
        stats.append((s.n_pp, s.n_pc, s.n_nodes, s.n_groups))
    ctx.set_step_graphs(True)
    # one more step with the walk counting the evaluated pairs within the cutoff (slower, untimed)
    ctx.set_walk_census(True)
    step()
    s = box.tree_stats()
    ctx.set_walk_census(False)
    census = {"n_pp": s.n_pp, "n_pp_within_cutoff": s.n_pp_within,
              "useful_pair_fraction": s.n_pp_within / max(1, s.n_pp), "n_groups": s.n_groups,
              "targets_per_group": N / max(1, s.n_groups),
              "note": "the step after the timed and phase steps, counted by the FP32 walk kernel itself"}
    log(f"[bench] census: {census}")

    ph = {k: float(np.mean([p[k] for p in phases])) for k in phases[0]}
    # the walk: its launches in the timed steps themselves, and their interaction counts
    ph["ms_walk"] = float(np.mean(walk_ms))
    npp = float(np.mean([s[0] for s in walk_stats]))
    npc = float(np.mean([s[1] for s in walk_stats]))
    M = n ** 3
    K = (n // 2 + 1) * n * n
    # algorithmic bytes / flops per launch
    kern = {
        "walk": {"bound": "fp32", "work": 21 * npp + 69 * npc, "ms": ph["ms_walk"]},
        "tsc": {"bound": "hbm", "work": 24 * N + 20 * M, "ms": ph["ms_tsc"]},
        "green": {"bound": "hbm", "work": 16 * K, "ms": ph["ms_green"]},
        "grad": {"bound": "hbm", "work": 16 * M, "ms": ph["ms_grad"]},
        "interp": {"bound": "hbm", "work": 72 * N + 12 * M, "ms": ph["ms_interp"]},
    }
    dom = max(kern, key=lambda k: kern[k]["ms"])
    d = kern[dom]
    traffic = None
    if os.path.exists(PROFILE_TRAFFIC):
        try:
            traffic = json.load(open(PROFILE_TRAFFIC)).get(dom)
        except Exception:
            traffic = None
    if d["bound"] == "hbm":
        achieved = d["work"] / (d["ms"] * 1e-3) / 1e9
        roof = {"kernel": dom, "bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    else:
        achieved = d["work"] / (d["ms"] * 1e-3) / 1e12
        roof = {"kernel": "walk_f32_kernel", "bound": "fp32-valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS,
                "engine": "FP32 vector ALU, no MFMA (the chip's FP32 peak, 157.3 TFLOP/s, counts packed FP32, "
                          "which this build does not use: the unpacked FMA rate is 78.6 TFLOP/s; the pair sums "
                          "are not GEMM-shaped)",
                "unit": "TFLOP/s", "frac": achieved / FP32_PEAK_TFLOPS, "traffic": traffic,
                "work": "21 flops per pair-particle + 69 per particle-node interaction evaluated, "
                        "mean over the timed steps",
                "n_pp": npp, "n_pc": npc, "census": census,
                "per_step": [{"walk_ms": w, "n_pp": a, "n_pc": b} for w, (a, b) in zip(walk_ms, walk_stats)]}
    phase_report = {}
    for k, v in kern.items():
        if v["bound"] == "hbm" and v["ms"] > 0:
            phase_report[k] = {"ms": v["ms"], "GB/s": v["work"] / (v["ms"] * 1e-3) / 1e9}
        else:
            phase_report[k] = {"ms": v["ms"], "TFLOP/s": v["work"] / max(v["ms"], 1e-9) / 1e9}
    for k in ("ms_sort", "ms_fft_r2c", "ms_fft_c2r", "ms_tree_sort", "ms_build"):
        phase_report[k[3:]] = {"ms": ph[k]}

    out = {
        "metric": METRIC, "value": N / (ms * 1e-3), "unit": "forces/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "fp32",
        "dtype_detail": "FP32 mesh/FFT/pair arithmetic with FP32 accumulation, int64 fixed-point deposit, "
                        "FP64 positions/velocities and kick-drift",
        "data": "synthetic",
        "config": {"workload": f"single box {n}^3 clustered ({100 - 100 * hfrac:.0f}% perturbed lattice + "
                               f"{100 * hfrac:.0f}% in {nhalo} Plummer halos), "
                               f"{n}^3 mesh, Rsplit 1.25, rcut 5.625, theta {args.theta}, one GOTPM step "
                               f"(drag + PM kick + tree kick + drift) per step",
                   "n": n, "particles": N, "parallelism": "single GPU",
                   "tree_tuning": {"nleaf": tuning[0], "group": tuning[1],
                                   "group_mode": MODE_NAMES[tuning[2]]}},
        "roofline": roof,
        "phases": phase_report,
    }
    box.close()

    a32_full = None
    if not args.no_accuracy:
        # forces at the initial positions: FP32 fast path vs the plain FP64 kernels
        a32, t32, t32_first = force_at(ctx, lagpm, params, pos, tuning)
        a32_full = a32
        acc = {"vs": "plain FP64 kernels, same tree/groups/criteria", "fp32_seconds": t32,
               "fp32_first_call_seconds": t32_first,
               "timing": "a second PM + tree force evaluation on the same box from the same positions; "
                         "the first call also creates the precision's FFT plans and finds the tree depth"}
        if not args.no_fp64:
            p64 = lagpm.make_params(n, law=lagpm.LAW_ERFC, precision=lagpm.PREC_FP64, deconvolve=1, efold=2.5,
                                    epsilon=0.1, rsphere=5.625, theta=args.theta)
            a64, t64, t64_first = force_at(ctx, lagpm, p64, pos, tuning)
            if args.acc_sample < 1.0:  # statistics on a seeded random subsample of the particles
                sel = np.random.default_rng(args.seed + 11).random(N) < args.acc_sample
                a32, a64 = a32[sel], a64[sel]
                acc["sample_fraction"] = args.acc_sample
            rms, p99, med = rel_errors(a32, a64)
            acc.update({"rms_rel": rms, "p99_rel": p99, "median_rel": med, "fp64_plain_seconds": t64,
                        "fp64_plain_first_call_seconds": t64_first,
                        "speedup_vs_fp64_plain": t64 / max(t32, 1e-12)})
            log(f"[bench] accuracy vs FP64 plain: rms {rms:.2e} p99 {p99:.2e}; FP64 plain force eval {t64:.3f} s")
        out["accuracy"] = acc

    if not args.no_ensemble:
        out["ensemble"] = ensemble(args, ctx, lagpm, ics)
        e = out["ensemble"]
        log(f"[bench] ensemble: plain {e['plain']['particle_steps_per_s']:.3e} / batched "
            f"{e['batched']['particle_steps_per_s']:.3e} particle-steps/s, x{e['batched_over_plain']:.2f}, "
            f"identical {e['identical_plain_vs_batched']}")

    if not args.no_serve:
        out["serving"] = serving(args, ctx, lagpm, ics)
        sv = out["serving"]
        log(f"[bench] serving: {sv['requests']} requests, {sv['particle_steps_per_s']:.3e} particle-steps/s, "
            f"step p50 {sv.get('step_latency_p50_ms', 0):.1f} ms p99 {sv.get('step_latency_p99_ms', 0):.1f} ms")

    ctx.close()
    if not args.no_cpu_baseline:
        log("[bench] cpu baseline ...")
        out["cpu_baseline"], cpu_acc = cpu_baseline(args, n, pos, a32_full)
        out.setdefault("accuracy", {}).update(cpu_acc)
        log(f"[bench] cpu baseline: {out['cpu_baseline']['value']:.3e} forces/s; accuracy {cpu_acc}")
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
