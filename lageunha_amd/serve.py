"""Continuous-batching request loop over lagpm_batch_step.

A request asks for `steps` GOTPM steps of one periodic box (n^3 particles on an n^3 mesh,
its own Omega_m, scale factor from a_start with step da). The server keeps every admitted
box resident on the device between steps; one loop iteration is one lagpm_batch_step over all
active boxes, so requests that arrive while others run join at the next iteration and share
its launches. Admission is first come first served under two limits: the device bytes of the
active boxes (lagpm_estimate_box_bytes, the HBM budget) and optionally the active particle
count, which bounds the duration of one iteration and so the per-step latency.

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


class Server:
    def __init__(self, ctx: lagpm.Context, params_for: Callable[[Request], lagpm.Params],
                 positions_for: Callable[[Request], np.ndarray], hbm_budget: float,
                 max_active_particles: Optional[int] = None, keep_results: bool = False,
                 clock: Callable[[], float] = time.perf_counter):
        self.ctx = ctx
        self.params_for = params_for
        self.positions_for = positions_for
        self.hbm_budget = float(hbm_budget)
        self.max_active_particles = max_active_particles
        self.keep_results = keep_results
        self.clock = clock
        self.queue: Deque[Request] = deque()
        self.active: Dict[int, tuple] = {}   # rid -> (request, box, bytes)
        self.done: List[Request] = []
        self.bytes_active = 0
        self.particles_active = 0
        self.iterations = 0
        self.particle_steps = 0
        self.iteration_log: List[tuple] = []  # (start, seconds, boxes, particles)

    def submit(self, req: Request):
        self.queue.append(req)

    def _admit(self, now: float):
        while self.queue and self.queue[0].arrival <= now:
            req = self.queue[0]
            p = self.params_for(req)
            need = lagpm.estimate_box_bytes(p, req.n ** 3)
            npart = req.n ** 3
            if need > self.hbm_budget:
                raise ValueError(f"request {req.rid} alone exceeds the HBM budget")
            fits = self.bytes_active + need <= self.hbm_budget
            if self.max_active_particles is not None and self.active:
                fits = fits and self.particles_active + npart <= self.max_active_particles
            if not fits:
                return
            self.queue.popleft()
            box = lagpm.Box(self.ctx, p, npart)
            box.upload_pos(self.positions_for(req))
            nbytes = box.device_bytes
            self.active[req.rid] = (req, box, nbytes)
            self.bytes_active += nbytes
            self.particles_active += npart
            req.admitted = self.clock()

    def _release(self, rid: int):
        req, box, nbytes = self.active.pop(rid)
        if self.keep_results:
            req.result = box.download()
        box.close()
        self.bytes_active -= nbytes
        self.particles_active -= req.n ** 3
        req.finished = self.clock()
        self.done.append(req)

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
        t0 = self.clock()
        lagpm.batch_step(self.ctx, boxes, anow, astep)
        dt = self.clock() - t0
        self.iterations += 1
        self.iteration_log.append((t0, dt, len(boxes), self.particles_active))
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

    def stats(self) -> dict:
        lat = np.array([x for r in self.done for x in r.step_latency]) * 1e3
        wait = np.array([r.admitted - r.arrival for r in self.done]) * 1e3
        turn = np.array([r.finished - r.arrival for r in self.done]) * 1e3
        busy = sum(dt for _, dt, _, _ in self.iteration_log)
        out = {"requests": len(self.done), "iterations": self.iterations, "particle_steps": self.particle_steps,
               "busy_seconds": busy,
               "particle_steps_per_busy_s": self.particle_steps / busy if busy > 0 else 0.0}
        if len(lat):
            out.update({"step_latency_p50_ms": float(np.percentile(lat, 50)),
                        "step_latency_p99_ms": float(np.percentile(lat, 99)),
                        "queue_wait_p50_ms": float(np.percentile(wait, 50)),
                        "queue_wait_p99_ms": float(np.percentile(wait, 99)),
                        "turnaround_p50_ms": float(np.percentile(turn, 50)),
                        "turnaround_p99_ms": float(np.percentile(turn, 99)),
                        "mean_boxes_per_iteration": float(np.mean([b for _, _, b, _ in self.iteration_log]))})
        return out
