"""Continuous-batching request loop over lagpm_batch_step.

A request asks for `steps` GOTPM steps of one periodic box (n^3 particles on an n^3 mesh,
its own Omega_m, scale factor from a_start with step da). The server keeps every admitted
box resident on the device between steps; one loop iteration is one lagpm_batch_step over all
active boxes, so requests that arrive while others run join at the next iteration and share
its launches. Admission is first come first served under two limits: the device bytes held by
the server (active boxes and idle pooled boxes, against the HBM budget) and optionally the
active particle count, which bounds the duration of one iteration and so the per-step latency.

Finished boxes are not freed: they go to a pool keyed by everything that sizes a box (mesh,
particle count, force settings) and are reused, with the next request's cosmology and
particles, by a later request of the same class. The loop then allocates and frees nothing in
the steady state (hipFree synchronises the device, which stalls the running batch), and a
reused box keeps its FFT execution info and grown node arrays. Idle pooled boxes are freed
only when a request of another class needs their memory.

The reference has no counterpart (EUNHA evolves one box per MPI job); the step itself is the
reference's global step, Cosmos/globalT.c:221-306.
"""
from __future__ import annotations

import dataclasses
import time
from collections import deque
from typing import Callable, Deque, Dict, List, Optional

import numpy as np

from . import lagpm


@dataclasses.dataclass
class Request:
    rid: int
    n: int
    steps: int
    arrival: float            # seconds on the server clock
    omep: float = 0.26
    a_start: float = 1.0
    da: float = 0.1
    seed: int = 0
    # filled by the server
    admitted: Optional[float] = None
    finished: Optional[float] = None
    steps_done: int = 0
    step_latency: List[float] = dataclasses.field(default_factory=list)
    result: Optional[tuple] = None  # (pos, vel) when keep_results


def box_class(p: lagpm.Params, npart: int) -> tuple:
    """Everything that sizes a box or its tables (not the cosmology, which a reused box takes)."""
    return (p.n, p.law, p.precision, p.deconvolve, p.mass, p.efold, p.epsilon, p.rsphere, p.theta, int(npart))


class BoxPool:
    """Idle boxes by class, oldest first."""

    def __init__(self):
        self.free: Dict[tuple, List[lagpm.Box]] = {}
        self.order: List[lagpm.Box] = []
        self.bytes = 0

    def take(self, key: tuple) -> Optional[lagpm.Box]:
        lst = self.free.get(key)
        if not lst:
            return None
        box = lst.pop()
        self.order.remove(box)
        self.bytes -= box.device_bytes
        return box

    def give(self, key: tuple, box: lagpm.Box):
        self.free.setdefault(key, []).append(box)
        self.order.append(box)
        self.bytes += box.device_bytes

    def evict_oldest(self) -> bool:
        if not self.order:
            return False
        box = self.order.pop(0)
        for lst in self.free.values():
            if box in lst:
                lst.remove(box)
                break
        self.bytes -= box.device_bytes
        box.close()
        return True

    def close(self):
        while self.evict_oldest():
            pass


class Server:
    def __init__(self, ctx: lagpm.Context, params_for: Callable[[Request], lagpm.Params],
                 positions_for: Callable[[Request], np.ndarray], hbm_budget: float,
                 max_active_particles: Optional[int] = None, keep_results: bool = False,
                 clock: Callable[[], float] = time.perf_counter, pool: bool = True,
                 tuning: Optional[tuple] = None):
        self.ctx = ctx
        self.params_for = params_for
        self.positions_for = positions_for
        self.hbm_budget = float(hbm_budget)
        self.max_active_particles = max_active_particles
        self.keep_results = keep_results
        self.clock = clock
        self.pool = BoxPool() if pool else None
        self.tuning = tuning  # (nleaf, group[, group mode]) for every new box; None: the library's
        self.queue: Deque[Request] = deque()
        self.active: Dict[int, tuple] = {}   # rid -> (request, box, class)
        self.done: List[Request] = []
        self.failed: List[tuple] = []        # (request, error message)
        self.particles_active = 0
        self.iterations = 0
        self.particle_steps = 0
        self.boxes_created = 0
        self.boxes_reused = 0
        self.iteration_log: List[tuple] = []  # (start, seconds, boxes, particles, boxes new to the batch)
        self._joined = 0

    @property
    def bytes_active(self) -> int:
        """Device bytes of the active boxes, as the library reports them now (node arrays may
        have grown during a step)."""
        return sum(v[1].device_bytes for v in self.active.values())

    @property
    def bytes_held(self) -> int:
        return self.bytes_active + (self.pool.bytes if self.pool else 0)

    def submit(self, req: Request):
        self.queue.append(req)

    def _admit(self, now: float):
        while self.queue and self.queue[0].arrival <= now:
            req = self.queue[0]
            p = self.params_for(req)
            npart = req.n ** 3
            key = box_class(p, npart)
            if self.max_active_particles is not None and self.active and \
                    self.particles_active + npart > self.max_active_particles:
                return
            box = self.pool.take(key) if self.pool else None
            if box is None:
                need = lagpm.estimate_box_bytes(p, npart)
                if need > self.hbm_budget:
                    raise ValueError(f"request {req.rid} alone exceeds the HBM budget")
                # idle boxes of other classes make room; active ones never do
                while self.bytes_held + need > self.hbm_budget and self.pool and self.pool.evict_oldest():
                    pass
                if self.bytes_held + need > self.hbm_budget:
                    return
                box = lagpm.Box(self.ctx, p, npart)
                if self.tuning is not None:
                    box.set_tuning(*self.tuning)
                self.boxes_created += 1
            else:
                box.set_cosmology(p.amax, p.omep, p.omeplam)
                self.boxes_reused += 1
            self.queue.popleft()
            box.upload_pos(self.positions_for(req))
            self.active[req.rid] = (req, box, key)
            self.particles_active += npart
            self._joined += 1
            req.admitted = self.clock()

    def _release(self, rid: int, error: Optional[str] = None):
        req, box, key = self.active.pop(rid)
        self.particles_active -= req.n ** 3
        req.finished = self.clock()
        if error is None:
            if self.keep_results:
                req.result = box.download()
            self.done.append(req)
        else:
            self.failed.append((req, error))
        if self.pool is not None and error is None:
            self.pool.give(key, box)
        else:
            box.close()

    def iterate(self) -> bool:
        """Admit what has arrived and fits, then run one step of every active box. Returns
        False when nothing is active (the caller decides whether to wait for arrivals)."""
        self._admit(self.clock())
        if not self.active:
            return False
        reqs = [v[0] for v in self.active.values()]
        boxes = [v[1] for v in self.active.values()]
        anow = np.array([r.a_start + r.steps_done * r.da for r in reqs])
        astep = np.array([r.da for r in reqs])
        joined, self._joined = self._joined, 0
        t0 = self.clock()
        try:
            lagpm.batch_step(self.ctx, boxes, anow, astep)
        except lagpm.LagpmError as e:
            # a failed step leaves these boxes in an undefined state: their requests fail, the
            # boxes are freed (not pooled), and the loop goes on with the others
            for r in reqs:
                self._release(r.rid, error=str(e))
            return True
        dt = self.clock() - t0
        self.iterations += 1
        self.iteration_log.append((t0, dt, len(boxes), self.particles_active, joined))
        for r in reqs:
            r.steps_done += 1
            r.step_latency.append(dt)
            self.particle_steps += r.n ** 3
        for r in reqs:
            if r.steps_done >= r.steps:
                self._release(r.rid)
        return True

    def run(self, until_empty: bool = True, idle_sleep: float = 1e-3):
        while self.queue or self.active:
            if not self.iterate():
                if not self.queue:
                    break
                wait = self.queue[0].arrival - self.clock()
                if wait > 0:
                    time.sleep(min(wait, idle_sleep * 100))
            if not until_empty and not self.active:
                break

    def close(self):
        for rid in list(self.active):
            self._release(rid)
        if self.pool is not None:
            self.pool.close()

    def stats(self) -> dict:
        lat = np.array([x for r in self.done for x in r.step_latency]) * 1e3
        wait = np.array([r.admitted - r.arrival for r in self.done]) * 1e3
        turn = np.array([r.finished - r.arrival for r in self.done]) * 1e3
        busy = sum(it[1] for it in self.iteration_log)
        out = {"requests": len(self.done), "failed": len(self.failed), "iterations": self.iterations,
               "particle_steps": self.particle_steps, "busy_seconds": busy,
               "particle_steps_per_busy_s": self.particle_steps / busy if busy > 0 else 0.0,
               "boxes_created": self.boxes_created, "boxes_reused": self.boxes_reused}
        if len(lat):
            it_ms = np.array([it[1] for it in self.iteration_log]) * 1e3
            parts = np.array([it[3] for it in self.iteration_log], dtype=np.float64)
            out.update({"step_latency_p50_ms": float(np.percentile(lat, 50)),
                        "step_latency_p99_ms": float(np.percentile(lat, 99)),
                        "step_latency_p99_over_p50": float(np.percentile(lat, 99) / np.percentile(lat, 50)),
                        "queue_wait_p50_ms": float(np.percentile(wait, 50)),
                        "queue_wait_p99_ms": float(np.percentile(wait, 99)),
                        "turnaround_p50_ms": float(np.percentile(turn, 50)),
                        "turnaround_p99_ms": float(np.percentile(turn, 99)),
                        "mean_boxes_per_iteration": float(np.mean([it[2] for it in self.iteration_log])),
                        "iteration_ms_p50": float(np.percentile(it_ms, 50)),
                        "iteration_ms_p99": float(np.percentile(it_ms, 99)),
                        "iteration_ns_per_particle_p50": float(np.percentile(it_ms * 1e6 / parts, 50)),
                        "iteration_ns_per_particle_p99": float(np.percentile(it_ms * 1e6 / parts, 99)),
                        "iterations_with_new_boxes": int(sum(1 for it in self.iteration_log if it[4]))})
            for tag, sel in (("with_new_boxes", [it for it in self.iteration_log if it[4]]),
                             ("without_new_boxes", [it for it in self.iteration_log if not it[4]])):
                if sel:
                    ns = np.array([it[1] * 1e9 / it[3] for it in sel])
                    out[f"iteration_ns_per_particle_{tag}_p50"] = float(np.percentile(ns, 50))
                    out[f"iteration_ns_per_particle_{tag}_p99"] = float(np.percentile(ns, 99))
        return out
