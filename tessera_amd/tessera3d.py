"""The Python-facing ``MpmWorld`` of the reference's bindings
(crates/tessera-python3d/src/mpm.rs:630-1283), on the HIP backend.

The reference object owns a CPU ``MpmWorld`` plus an optional resident GPU
session, and flattens every error into ``TesseraError``.  Here the world lives
in a :class:`~tessera_amd.mpm.HipMpmResidentSession` and every step runs on
the device: ``step`` and ``step_gpu`` both take CFL-limited device substeps
(the reference's ``step`` runs its CPU path, which this backend does not
ship), and ``step_gpu_fixed`` keeps one resident session while its substep dt
is unchanged.  Configuration changes (gravity, parameters, CPIC, transfer
colours) need no pending substeps and rebuild the session from the
synchronized world, keeping particles, forces, chunks and obstacles.

Rigid-world coupling (``*_with_primitive_world``, ``*_with_sphere_world``)
belongs to the reference's rigid physics crate and is not mirrored; the
MPM-owned coupling is :meth:`HipMpmResidentSession.set_rigid_bodies`.
"""

from __future__ import annotations

import copy
import math
from typing import Optional, Sequence

import numpy as np

from .emitters import MeshEmitter
from .mpm import (GpuMpmError, HipMpmResidentSession, HipMpmTransfers, MaterialModel, MpmParams, MpmParticle,
                  MpmWorld as _World, ObstacleShape, ParticleArrays, RigidObstacle, WorldBounds)


# Substep dt of a session that only takes CFL-limited (adaptive) substeps: a
# resident session validates its fixed dt against the stable timestep, and
# this one is never used for a step.
_ADAPTIVE_DT = 1e-12


class TesseraError(Exception):
    """``TesseraError::Failed { reason }`` (py/src/lib.rs:82-89)."""


def _vec(v) -> np.ndarray:
    return np.asarray(v, dtype=np.float64).reshape(3)


def _bounds(bounds_min, bounds_max) -> Optional[WorldBounds]:
    if bounds_min is None and bounds_max is None:
        return None
    if bounds_min is None or bounds_max is None:
        raise TesseraError("both MPM bounds corners must be supplied")
    return WorldBounds(_vec(bounds_min), _vec(bounds_max))


class MpmWorld:
    """``tessera3d.MpmWorld`` (mpm.rs:630-1283)."""

    def __init__(self, particles: Sequence[MpmParticle] | ParticleArrays, gravity, cell_width: float,
                 max_substep: float, bounds_min=None, bounds_max=None, transfers: HipMpmTransfers | None = None):
        pa = particles if isinstance(particles, ParticleArrays) else ParticleArrays.from_particles(list(particles))
        params = MpmParams(cell_width=float(cell_width), gravity=_vec(gravity), max_substep=float(max_substep),
                           bounds=_bounds(bounds_min, bounds_max))
        self._transfers = transfers or HipMpmTransfers(0)
        self._frame_substeps = 1
        self._session: HipMpmResidentSession | None = None
        self._fixed_dt: float | None = None
        self._next_chunk = int(pa.chunk_id.max()) + 1 if len(pa) else 1
        self._build(_World(pa, params), _ADAPTIVE_DT)

    # -- session management ----------------------------------------------------
    def _build(self, world: _World, dt: float) -> None:
        world.next_chunk_id = self._next_chunk
        try:
            session = HipMpmResidentSession(self._transfers, world, dt)
        except GpuMpmError as e:
            raise TesseraError(str(e)) from e
        if self._session is not None:
            self._session.close()
        self._session = session
        self._dt = dt

    def _reject_pending(self) -> None:
        if self._session.pending_substeps() > 0:
            raise TesseraError("synchronize pending GPU substeps first")

    def _rebuild(self, edit) -> None:
        """Applies `edit` to a copy of the synchronized world and rebuilds the session;
        on failure the world is left as it was."""
        self._reject_pending()
        world = self._session.world()
        trial = world.clone()
        trial.params = copy.deepcopy(world.params)
        edit(trial)
        self._build(trial, _ADAPTIVE_DT)
        self._fixed_dt = None

    def _call(self, fn, *args):
        try:
            return fn(*args)
        except GpuMpmError as e:
            raise TesseraError(str(e)) from e

    # -- configuration (mpm.rs:684-756) ------------------------------------------
    def set_gravity(self, gravity) -> None:
        g = _vec(gravity)

        def edit(w):
            w.params.gravity = g

        self._rebuild(edit)

    def gravity(self) -> np.ndarray:
        return self._session.world().params.gravity.copy()

    def set_params(self, gravity, cell_width: float, max_substep: float, bounds_min=None, bounds_max=None) -> None:
        bounds = _bounds(bounds_min, bounds_max)
        if not (math.isfinite(cell_width) and cell_width > 0 and math.isfinite(max_substep) and max_substep > 0):
            raise TesseraError("invalid MPM parameters")

        def edit(w):
            w.params = MpmParams(cell_width=float(cell_width), gravity=_vec(gravity), max_substep=float(max_substep),
                                 bounds=bounds)

        self._rebuild(edit)

    def set_substeps(self, substeps: int) -> None:
        self._reject_pending()
        self._frame_substeps = max(1, int(substeps))

    def frame_substeps(self) -> int:
        return self._frame_substeps

    def set_use_cpic(self, enabled: bool) -> None:
        def edit(w):
            w.use_cpic = bool(enabled)

        self._rebuild(edit)

    def use_cpic(self) -> bool:
        return bool(self._session.world().use_cpic)

    # -- stepping (mpm.rs:758-891) -------------------------------------------------
    def step(self, dt: float) -> None:
        """Frame of `frame_substeps` subdivisions, each with CFL substeps, on the device."""
        self._reject_pending()
        sub = dt / self._frame_substeps
        if not math.isfinite(dt) or not sub > 0.0:
            raise TesseraError("invalid MPM frame timestep")
        for _ in range(self._frame_substeps):
            self._call(self._session.step_adaptive, sub)

    step_gpu = step

    def step_gpu_fixed(self, substep_dt: float, steps: int) -> None:
        if not (math.isfinite(substep_dt) and substep_dt > 0.0 and 1 <= steps <= 64):
            raise TesseraError("invalid fixed GPU substep input")
        self._reject_pending()
        if self._fixed_dt != substep_dt or self._dt != substep_dt:
            self._build(self._session.world(), float(substep_dt))
            self._fixed_dt = float(substep_dt)
        self._call(self._session.step, int(steps))

    def submit_gpu_fixed(self, substep_dt: float, steps: int) -> None:
        """Queues fixed substeps without waiting (``synchronize_gpu`` reads them back)."""
        if not (math.isfinite(substep_dt) and substep_dt > 0.0 and 1 <= steps <= 64):
            raise TesseraError("invalid fixed GPU substep input")
        if self._fixed_dt != substep_dt or self._dt != substep_dt:
            self._reject_pending()
            self._build(self._session.world(), float(substep_dt))
            self._fixed_dt = float(substep_dt)
        self._call(self._session.submit_steps, int(steps))

    def particles(self) -> list:
        pa = self._session.read_particles()
        return [pa.particle(i) for i in range(len(pa))]

    def substeps(self) -> int:
        return int(self._session.substeps)

    def pending_gpu_substeps(self) -> int:
        return int(self._session.pending_substeps())

    def synchronize_gpu(self) -> None:
        self._call(self._session.synchronize)

    # -- edits (mpm.rs:893-1283) ----------------------------------------------------
    def add_particles(self, particles) -> int:
        chunk = int(self._call(self._session.add_particles, particles))
        self._next_chunk = max(self._next_chunk, chunk + 1)
        return chunk

    def add_closed_mesh_particles(self, vertices, triangles, spacing: float, density: float,
                                  material: MaterialModel, max_candidates: int, max_particles: int,
                                  velocity=(0.0, 0.0, 0.0), damping: float = 0.0) -> int:
        emitter = MeshEmitter(np.asarray(vertices, dtype=np.float64), np.asarray(triangles, dtype=np.uint32),
                              float(spacing), float(density), material, int(max_candidates), int(max_particles),
                              velocity=_vec(velocity), damping=float(damping))
        try:
            batch = emitter.sample(self._transfers)
        except GpuMpmError as e:
            raise TesseraError(str(e)) from e
        return self.add_particles(batch)

    def remove_chunk(self, chunk_id: int) -> int:
        return int(self._call(self._session.remove_chunk, int(chunk_id)))

    def set_obstacles(self, obstacles: Sequence[RigidObstacle]) -> None:
        obstacles = [copy.deepcopy(o) for o in obstacles]
        for o in obstacles:
            # the binding's unit_orientation (mpm.rs:408-417): normalized when within 1e-5
            q = np.asarray(o.orientation, dtype=np.float64)
            n2 = float(q @ q)
            if not math.isfinite(n2) or abs(n2 - 1.0) > 1e-5:
                raise TesseraError("obstacle orientation must be a unit quaternion")
            o.orientation = q / math.sqrt(n2)
        self._call(self._session.set_obstacles, obstacles)

    def obstacles(self) -> list:
        return list(self._session.world().obstacles)

    def obstacle_count(self) -> int:
        return len(self._session.world().obstacles)

    def set_obstacle_boundaries(self, boundaries: Sequence[int]) -> None:
        self._reject_pending()
        obstacles = [copy.deepcopy(o) for o in self._session.world().obstacles]
        if len(boundaries) != len(obstacles):
            raise TesseraError("MPM boundary count must match obstacle count")
        for o, b in zip(obstacles, boundaries):
            o.boundary = int(b)
        self.set_obstacles(obstacles)

    def obstacle_boundaries(self) -> list:
        return [int(o.boundary) for o in self._session.world().obstacles]

    def set_obstacle_cpic_groups(self, groups: Sequence[Optional[int]]) -> None:
        self._reject_pending()
        obstacles = [copy.deepcopy(o) for o in self._session.world().obstacles]
        if len(groups) != len(obstacles):
            raise TesseraError("MPM CPIC group count must match obstacle count")
        for o, g in zip(obstacles, groups):
            if g is not None and not (0 <= int(g) < 32 and o.shape == ObstacleShape.TRIANGLE_PRISM):
                raise TesseraError("invalid MPM CPIC group or obstacle shape")
            o.cpic_group = None if g is None else int(g)
        self.set_obstacles(obstacles)

    def obstacle_cpic_groups(self) -> list:
        return [o.cpic_group for o in self._session.world().obstacles]

    def set_particle_force(self, index: int, force) -> None:
        f = _vec(force)
        if not np.all(np.isfinite(f)):
            raise TesseraError("MPM force must be finite")
        if not 0 <= int(index) < len(self._session.world().particles):
            raise TesseraError("MPM particle index out of range")
        self._call(self._session.set_particle_force, int(index), f)

    def set_particle_transfer_color(self, index: int, color: int) -> None:
        if not 0 <= int(color) <= 0xFFFF:
            raise TesseraError("MPM transfer color out of range")
        if not 0 <= int(index) < len(self._session.world().particles):
            raise TesseraError("MPM particle index out of range")

        def edit(w):
            w.particles.transfer_color[int(index)] = int(color)

        self._rebuild(edit)

    def close(self) -> None:
        if self._session is not None:
            self._session.close()
            self._session = None
