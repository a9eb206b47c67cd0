"""Continuous-batching serving of independent MPM scenes on one GPU.

Requests (a world, its substep dt, a horizon in substeps, optional rigid
bodies driving some of its obstacles) arrive over time.  One
:class:`~tessera_amd.mpm.HipMpmBatch` advances every admitted scene together,
one kernel launch per stage per substep.  The loop overlaps host and device
work:

* a tick submits as many substeps as the shortest remaining horizon among the
  live scenes (at most ``steps_per_tick``, so every scene stops exactly at its
  horizon) and polls ``synchronize_async`` until the device is done;
* meanwhile a host thread stages the next waiting requests: validation, the
  f64 copy of the world, the f32 packing and the block count
  (``StagedScene``), which is most of an admission's cost;
* between ticks, scenes that reached their horizon (or failed) retire, with
  their final world (and rigid bodies) read as the response, and staged scenes
  are admitted (the device half: uploads) first-come first-served while the
  live scenes' device bytes stay within an HBM budget.  The bytes per particle
  are the batch's own: the device bytes it reports over its particle capacity.

Because blocks never mix scenes and the transfers are order-independent, a
scene's result does not depend on what it was batched with: it equals the
same world stepped alone (tests/test_gpu_serving.py).

:class:`PlainServer` serves the same stream the way the reference would: one
resident session per request (``GpuMpmResidentSession``, gpu.rs:206-662),
each on its own stream, all submitted before any is synchronized.

Reported: aggregate particle-updates/s over the stream's wall time, per-step
latency and queueing delay.  A latency sample is the wall time between two
consecutive completions of a scene's substeps divided by the substeps between
them (the first from its admission), so the host work between ticks that
every live scene waits through is part of it.
"""

from __future__ import annotations

import time
from concurrent.futures import Future, ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from .mpm import GpuMpmError, HipMpmBatch, HipMpmResidentSession, HipMpmTransfers, InvalidInput, MpmWorld, \
    RigidBody, RigidMapping, StagedScene

# Device bytes per particle assumed before a batch exists to report its own:
# three particle buffers (live, reorder target, last synchronized copy) of
# 124 B, the one-shot force, the binning per-particle arrays, plus about a
# hundredth of a block's tiles and grid.
BYTES_PER_PARTICLE = 3 * 124 + 16 + 16 + 64


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
    # one-shot forces (particle index, force xyz) for the first substep, as
    # set_particle_force (gpu.rs:457-485) applies them
    forces: Sequence = ()

    @property
    def particles(self) -> int:
        return len(self.world.particles)

    def device_bytes(self) -> int:
        """Device bytes assumed for this request before a batch reports its own bytes per particle."""
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
    slot: object          # batch slot (int) or the request's own session
    admitted: float
    last: float           # wall time of the last completed tick (or of admission)
    done: int = 0


def _percentiles(samples: list, scale: float = 1.0) -> tuple[float, float]:
    a = np.asarray(samples, dtype=np.float64) if samples else np.zeros(1)
    return float(np.percentile(a, 50) * scale), float(np.percentile(a, 99) * scale)


class _Loop:
    """What both servers share: arrival bookkeeping, latency samples, the report."""

    def __init__(self, transfers: HipMpmTransfers, hbm_budget: int, steps_per_tick: int = 16, mode: str = "fp32",
                 keep_worlds: bool = True):
        self.transfers = transfers
        self.hbm_budget = int(hbm_budget)
        self.steps_per_tick = int(steps_per_tick)
        self.mode = mode
        self.keep_worlds = keep_worlds
        self.live: list[_Live] = []
        self.latency: list[float] = []  # seconds per substep (see the module docstring)
        self.completed: list[Completed] = []
        self.updates = 0
        self.ticks = 0
        self._start = 0.0

    def _now(self) -> float:
        return time.perf_counter() - self._start

    def _reject(self, req: Request, err: Exception) -> None:
        now = self._now()
        self.completed.append(Completed(req.name, req.particles, 0, req.arrival, now, now, error=err))

    def _finish_tick(self, steps: int, failed: set) -> None:
        """Counts a completed tick of `steps` substeps: updates and latency of the scenes that ran it
        (a scene that failed during the tick counts neither)."""
        t = self._now()
        for l in self.live:
            if id(l) in failed:
                continue
            l.done += steps
            self.updates += l.request.particles * steps
            self.latency.append((t - l.last) / steps)
            l.last = t
        self.ticks += 1

    def _report(self, requests: Sequence[Request], wall: float) -> dict:
        p50, p99 = _percentiles(self.latency, 1e3)
        w50, w99 = _percentiles([c.admitted - c.arrival for c in self.completed])
        return {
            "requests": len(requests),
            "completed": len(self.completed),
            "failed": sum(1 for c in self.completed if c.error is not None),
            "wall_s": wall,
            "ticks": self.ticks,
            "particle_updates": self.updates,
            "value": self.updates / wall if wall > 0 else 0.0,
            "unit": "particle-updates/s",
            "step_latency_ms_p50": p50,
            "step_latency_ms_p99": p99,
            "queue_wait_s_p50": w50,
            "queue_wait_s_p99": w99,
        }


class Server(_Loop):
    """Continuous batching: every admitted scene in one batch, admissions staged while the device
    runs the previous tick."""

    def __init__(self, transfers: HipMpmTransfers, hbm_budget: int, steps_per_tick: int = 16, mode: str = "fp32",
                 keep_worlds: bool = True, stage_ahead: int = 4, poll_s: float = 5.0e-5):
        super().__init__(transfers, hbm_budget, steps_per_tick, mode, keep_worlds)
        self.stage_ahead = int(stage_ahead)
        self.poll_s = float(poll_s)
        self.batch: HipMpmBatch | None = None
        self._pool = ThreadPoolExecutor(max_workers=1, thread_name_prefix="tessera-stage")

    # -- memory ---------------------------------------------------------------
    def _bytes_per_particle(self) -> float:
        if self.batch is None:
            return float(BYTES_PER_PARTICLE)
        mem = self.batch.memory()
        return mem["device_bytes"] / max(1, mem["particle_capacity"])

    def _fits(self, req: Request) -> bool:
        if not self.live:
            return True
        bpp = self._bytes_per_particle()
        live = sum(l.request.particles for l in self.live)
        return (live + req.particles) * bpp <= self.hbm_budget

    # -- admission and retirement (no substeps pending) ---------------------------
    def _admit(self, req: Request, staged: StagedScene) -> None:
        try:
            if self.batch is None:
                staged.close()
                self.batch = HipMpmBatch(self.transfers, [req.world], [req.dt], self.mode)
                slot = 0
            else:
                slot = self.batch.add_staged(staged)
            if req.bodies:
                self.batch.set_rigid_bodies(slot, list(req.bodies), list(req.mappings), gravity=req.gravity)
            for i, f in req.forces:
                self.batch.set_particle_force(slot, i, f)
        except GpuMpmError as e:
            staged.close()
            if self.batch is not None and not self.batch.valid:
                self._fail_all(e)
            self._reject(req, e)
            return
        now = self._now()
        self.live.append(_Live(req, slot, now, now))

    def _retire(self, live: _Live) -> None:
        b = self.batch
        err = b.scene_error(live.slot)
        done = Completed(live.request.name, live.request.particles, live.done, live.request.arrival, live.admitted,
                         self._now(), error=err)
        if self.keep_worlds:
            done.world = b.world(live.slot)
            if live.request.bodies:
                done.bodies = b.rigid_bodies(live.slot)
        b.remove_scene(live.slot)
        self.live.remove(live)
        self.completed.append(done)

    def _fail_all(self, err: Exception) -> None:
        """A failure of the whole batch (device, grid capacity): every live request completes with it."""
        now = self._now()
        for l in self.live:
            self.completed.append(Completed(l.request.name, l.request.particles, l.done, l.request.arrival,
                                            l.admitted, now, error=err))
        self.live = []
        if self.batch is not None:
            self.batch.close()
            self.batch = None

    # -- serving loop -----------------------------------------------------------
    def run(self, requests: Sequence[Request], time_limit: float = 600.0) -> dict:
        """Serves `requests` (arrival = seconds from the start) to completion."""
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        staging: list[tuple[Request, Future]] = []
        self._start = time.perf_counter()
        nxt = 0

        def stage(req: Request) -> StagedScene:
            return StagedScene(req.world, req.dt)

        while nxt < len(pending) or waiting or staging or self.live:
            now = self._now()
            if now > time_limit:
                break
            while nxt < len(pending) and pending[nxt].arrival <= now:
                req = pending[nxt]
                nxt += 1
                if req.horizon < 1:
                    self._reject(req, InvalidInput(f"horizon {req.horizon} < 1"))
                else:
                    waiting.append(req)
            while waiting and len(staging) < self.stage_ahead:
                req = waiting.pop(0)
                staging.append((req, self._pool.submit(stage, req)))
            # admissions, first come first served, while the budget allows
            while staging and self._fits(staging[0][0]):
                req, fut = staging[0]
                if not fut.done() and self.live:
                    break  # still staging: admit it after the next tick
                staging.pop(0)
                try:
                    staged = fut.result()
                except GpuMpmError as e:
                    self._reject(req, e)
                    continue
                self._admit(req, staged)
            if not self.live:
                if nxt < len(pending) and not staging:
                    time.sleep(max(0.0, min(pending[nxt].arrival - self._now(), 0.05)))
                continue
            steps = min(self.steps_per_tick, min(l.request.horizon - l.done for l in self.live))
            b = self.batch
            try:
                b.submit_steps(steps)
                while not b.synchronize_async():
                    time.sleep(self.poll_s)
            except GpuMpmError as e:
                if b.valid:
                    raise
                self._fail_all(e)
                continue
            failed = {id(l) for l in self.live if b.scene_error(l.slot) is not None}
            self._finish_tick(steps, failed)
            for l in list(self.live):
                if id(l) in failed or l.done >= l.request.horizon:
                    self._retire(l)
        for _, fut in staging:
            try:
                fut.result().close()
            except GpuMpmError:
                pass
        return self._report(requests, self._now())

    def close(self) -> None:
        if self.batch is not None:
            self.batch.close()
            self.batch = None
        self._pool.shutdown(wait=True)


class PlainServer(_Loop):
    """The plain per-scene path on the same GPU: one resident session per admitted request, each on
    its own stream.  A tick submits to every live session before synchronizing any, so their
    substeps overlap on the device as far as the streams allow."""

    def _session_bytes(self) -> int:
        return sum(l.slot.memory()["device_bytes"] for l in self.live)

    def _fits(self, req: Request) -> bool:
        if not self.live:
            return True
        live = sum(l.request.particles for l in self.live)
        bpp = self._session_bytes() / max(1, live)
        return (live + req.particles) * bpp <= self.hbm_budget

    def _admit(self, req: Request) -> None:
        try:
            s = HipMpmResidentSession(self.transfers, req.world, req.dt, self.mode)
            if req.bodies:
                s.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
            for i, f in req.forces:
                s.set_particle_force(i, f)
        except GpuMpmError as e:
            self._reject(req, e)
            return
        now = self._now()
        self.live.append(_Live(req, s, now, now))

    def _retire(self, live: _Live, err: Exception | None) -> None:
        s = live.slot
        done = Completed(live.request.name, live.request.particles, live.done, live.request.arrival, live.admitted,
                         self._now(), error=err)
        if self.keep_worlds:
            done.world = s.world()
            if live.request.bodies:
                done.bodies = s.rigid_bodies()
        s.close()
        self.live.remove(live)
        self.completed.append(done)

    def run(self, requests: Sequence[Request], time_limit: float = 600.0) -> dict:
        pending = sorted(requests, key=lambda r: r.arrival)
        waiting: list[Request] = []
        self._start = time.perf_counter()
        nxt = 0
        while nxt < len(pending) or waiting or self.live:
            now = self._now()
            if now > time_limit:
                break
            while nxt < len(pending) and pending[nxt].arrival <= now:
                req = pending[nxt]
                nxt += 1
                if req.horizon < 1:
                    self._reject(req, InvalidInput(f"horizon {req.horizon} < 1"))
                else:
                    waiting.append(req)
            while waiting and self._fits(waiting[0]):
                self._admit(waiting.pop(0))
            if not self.live:
                if nxt < len(pending):
                    time.sleep(max(0.0, min(pending[nxt].arrival - self._now(), 0.05)))
                continue
            steps = min(self.steps_per_tick, min(l.request.horizon - l.done for l in self.live))
            for l in self.live:
                l.slot.submit_steps(steps)
            errors = {}
            for l in self.live:
                try:
                    l.slot.synchronize()
                except GpuMpmError as e:
                    errors[id(l)] = e
            self._finish_tick(steps, set(errors))
            for l in list(self.live):
                if id(l) in errors or l.done >= l.request.horizon:
                    self._retire(l, errors.get(id(l)))
        return self._report(requests, self._now())

    def close(self) -> None:
        for l in self.live:
            l.slot.close()
        self.live = []
