"""Continuous-batching serving of independent MPM scenes on one GPU.

Requests (a world, its substep dt, a horizon in substeps, optional rigid
bodies driving some of its obstacles) arrive over time.  One
:class:`~tessera_amd.mpm.HipMpmBatch` advances every admitted scene together,
one kernel launch per stage per substep.  Between ticks the server:

* admits waiting requests first-come first-served while the estimated device
  bytes of the live scenes stay within an HBM budget;
* retires scenes that reached their horizon, reading their final world (and
  rigid bodies) as the response.

A tick is as many substeps as the shortest remaining horizon among the live
scenes, at most ``steps_per_tick``, so every scene stops exactly at its
horizon.  Because blocks never mix scenes and the transfers are
order-independent, a scene's result does not depend on what it was batched
with: it equals the same world stepped alone (tests/test_gpu_serving.py).

Reported: aggregate particle-updates/s, per-step latency (tick wall time per
substep, one sample per live scene per tick) p50/p99, and queueing delay.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from .mpm import HipMpmBatch, HipMpmTransfers, MpmWorld, RigidBody, RigidMapping

# Device bytes per particle of a batch: three particle buffers (live, reorder
# target, last synchronized copy) of 112 B, the one-shot force, the binning
# per-particle arrays and lists, plus ~1/100 of a block's tiles and grid.
BYTES_PER_PARTICLE = 3 * 112 + 16 + 12 + 32 + 64


@dataclass
class Request:
    name: str
    world: MpmWorld
    dt: float
    horizon: int
    arrival: float = 0.0
    bodies: Sequence[RigidBody] = ()
    mappings: Sequence[RigidMapping] = ()
    gravity: tuple = (0.0, -9.8, 0.0)

    @property
    def particles(self) -> int:
        return len(self.world.particles)

    def device_bytes(self) -> int:
        return self.particles * BYTES_PER_PARTICLE


@dataclass
class Completed:
    name: str
    particles: int
    horizon: int
    arrival: float
    admitted: float
    finished: float
    world: MpmWorld | None = None
    bodies: list = field(default_factory=list)
    error: Exception | None = None


@dataclass
class _Live:
    request: Request
    slot: int
    admitted: float
    done: int = 0


class Server:
    def __init__(self, transfers: HipMpmTransfers, hbm_budget: int, steps_per_tick: int = 16, mode: str = "fp32",
                 keep_worlds: bool = True):
        self.transfers = transfers
        self.hbm_budget = int(hbm_budget)
        self.steps_per_tick = int(steps_per_tick)
        self.mode = mode
        self.keep_worlds = keep_worlds
        self.batch: HipMpmBatch | None = None
        self.live: dict[int, _Live] = {}
        self.latency: list[float] = []   # seconds per substep, one sample per live scene per tick
        self.completed: list[Completed] = []
        self.updates = 0

    # -- admission ------------------------------------------------------------
    def _bytes_live(self) -> int:
        return sum(l.request.device_bytes() for l in self.live.values())

    def _admit(self, req: Request, now: float) -> None:
        if self.batch is None:
            self.batch = HipMpmBatch(self.transfers, [req.world], [req.dt], self.mode)
            slot = 0
        else:
            slot = self.batch.add_scene(req.world, req.dt)
        if req.bodies:
            self.batch.set_rigid_bodies(slot, list(req.bodies), list(req.mappings), gravity=req.gravity)
        self.live[slot] = _Live(req, slot, now)

    def _retire(self, live: _Live, now: float) -> None:
        b = self.batch
        err = b.scene_error(live.slot)
        done = Completed(live.request.name, live.request.particles, live.done, live.request.arrival, live.admitted,
                         now, error=err)
        if self.keep_worlds:
            done.world = b.world(live.slot)
            if live.request.bodies:
                done.bodies = b.rigid_bodies(live.slot)
        b.remove_scene(live.slot)
        del self.live[live.slot]
        self.completed.append(done)

    # -- serving loop -----------------------------------------------------------
    def run(self, requests: Sequence[Request], time_limit: float = 600.0) -> dict:
        """Serves `requests` (sorted by arrival, seconds from the start) to completion."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        start = time.perf_counter()
        nxt = 0
        while nxt < len(pending) or waiting or self.live:
            now = time.perf_counter() - start
            if now > time_limit:
                break
            while nxt < len(pending) and pending[nxt].arrival <= now:
                waiting.append(pending[nxt])
                nxt += 1
            while waiting and (not self.live or
                               self._bytes_live() + waiting[0].device_bytes() <= self.hbm_budget):
                self._admit(waiting.pop(0), now)
            if not self.live:
                if nxt < len(pending):
                    time.sleep(max(0.0, min(pending[nxt].arrival - now, 0.05)))
                continue
            steps = min(self.steps_per_tick, min(l.request.horizon - l.done for l in self.live.values()))
            t0 = time.perf_counter()
            self.batch.step(steps)
            dt = time.perf_counter() - t0
            for l in list(self.live.values()):
                l.done += steps
                self.updates += l.request.particles * steps
                self.latency.append(dt / steps)
            now = time.perf_counter() - start
            for l in list(self.live.values()):
                if l.done >= l.request.horizon or self.batch.scene_error(l.slot) is not None:
                    self._retire(l, now)
        wall = time.perf_counter() - start
        lat = np.array(self.latency) if self.latency else np.zeros(1)
        waits = np.array([c.admitted - c.arrival for c in self.completed]) if self.completed else np.zeros(1)
        return {
            "requests": len(requests),
            "completed": len(self.completed),
            "failed": sum(1 for c in self.completed if c.error is not None),
            "wall_s": wall,
            "particle_updates": self.updates,
            "value": self.updates / wall if wall > 0 else 0.0,
            "unit": "particle-updates/s",
            "step_latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
            "step_latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
            "queue_wait_s_p50": float(np.percentile(waits, 50)),
            "queue_wait_s_p99": float(np.percentile(waits, 99)),
        }

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
            self.batch = None
