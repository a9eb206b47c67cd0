"""Particle emitters (crates/tessera-mpm/src/emitter.rs) over the HIP backend.

* :class:`BoxEmitter` — a centred lattice patch (emitter.rs:116-229).
* :class:`MeshEmitter` — a closed triangle mesh filled with cubic material
  points (emitter.rs:20-113), sampled by :func:`sample_closed_mesh_volume`
  (sampling.rs:47-146), whose solid-angle test runs on the device.

``sample()`` returns the particles as a :class:`~tessera_amd.mpm.ParticleArrays`
without touching a session; ``emit(session)`` adds them as one removable chunk
and returns its id (``MpmWorld::add_particles``, world.rs:244-265), which
``session.remove_chunk`` takes back.
"""

from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from .mpm import (Capacity, InvalidInput, MaterialModel, MpmParticle, ParticleArrays, _check, _dp, _vec,
                  load_library)


def sample_closed_mesh_volume(transfers, vertices, triangles, spacing: float, max_candidates: int,
                              max_samples: int) -> np.ndarray:
    """Lattice centres inside a closed, consistently oriented mesh, in lattice order
    (``sample_closed_mesh_volume``, sampling.rs:47-146).  Raises InvalidInput for bad
    geometry and Capacity when a limit is exceeded."""
    lib = load_library()
    v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, dtype=np.uint32).reshape(-1, 3))
    if int(max_samples) <= 0:
        raise InvalidInput("max_samples must be positive")
    out = np.zeros((int(max_samples), 3))
    count = C.c_uint64()
    _check(lib.tmpm_sample_closed_mesh_volume(transfers.handle, v.ctypes.data_as(_dp), len(v),
                                              t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t), float(spacing),
                                              int(max_candidates), int(max_samples), out.ctypes.data_as(_dp),
                                              C.byref(count)), lib)
    return out[: count.value].copy()


def _material_valid(material: MaterialModel) -> bool:
    m = material.to_c()
    return bool(load_library().tmpm_material_valid(C.byref(m)))


def _finite(*values) -> bool:
    return all(np.all(np.isfinite(np.asarray(v, dtype=np.float64))) for v in values)


@dataclass
class MeshEmitter:
    """``MeshEmitter`` (emitter.rs:20-113): particles of spacing^3 volume at the lattice
    centres inside a closed mesh (world-space vertices)."""

    vertices: np.ndarray
    triangles: np.ndarray
    spacing: float
    density: float
    material: MaterialModel
    max_candidates: int
    max_particles: int
    velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    damping: float = 0.0
    group_id: int = 0

    def sample(self, transfers) -> ParticleArrays:
        radius = self.spacing * 0.5
        volume = self.spacing ** 3
        mass = self.density * volume
        if not (_finite(radius, volume, mass, self.velocity, self.damping) and radius > 0.0 and volume > 0.0
                and mass > 0.0 and self.damping >= 0.0 and _material_valid(self.material)):
            raise InvalidInput("invalid emitter parameters")
        positions = sample_closed_mesh_volume(transfers, self.vertices, self.triangles, self.spacing,
                                              self.max_candidates, self.max_particles)
        if len(positions) == 0:
            raise InvalidInput("the mesh contains no lattice centre")
        return _particles_at(positions, radius, self.density, self.material, self.velocity, self.damping,
                             self.group_id)

    def emit(self, session) -> int:
        return session.add_particles(self.sample(session.transfers))


@dataclass
class BoxEmitter:
    """``BoxEmitter`` (emitter.rs:116-229): counts[0] x counts[1] x counts[2] particles
    centred on `center`, `spacing` apart per axis, in x, y, z loop order."""

    center: np.ndarray
    counts: tuple
    spacing: np.ndarray
    radius: float
    density: float
    material: MaterialModel
    velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    damping: float = 0.0
    group_id: int = 0

    def sample(self) -> ParticleArrays:
        counts = [int(c) for c in self.counts]
        center, spacing = _vec(self.center), _vec(self.spacing)
        volume = (2.0 * self.radius) ** 3
        ok = (all(c > 0 for c in counts) and _finite(center, spacing, self.radius, self.density, volume,
                                                      self.velocity, self.damping)
              and np.all(spacing > 0.0) and self.radius > 0.0 and self.density > 0.0 and volume > 0.0
              and np.isfinite(self.density * volume) and self.density * volume > 0.0 and self.damping >= 0.0
              and _material_valid(self.material))
        if ok:
            half = (np.array(counts, dtype=np.float64) - 1.0) * spacing * 0.5
            ok = _finite(half, center - half, center + half)
        if not ok:
            raise InvalidInput("invalid emitter parameters")
        ci = (np.array(counts, dtype=np.float64) - 1.0) * 0.5
        idx = np.stack(np.meshgrid(*[np.arange(c, dtype=np.float64) for c in counts], indexing="ij"),
                       axis=-1).reshape(-1, 3)
        positions = center + (idx - ci) * spacing
        return _particles_at(positions, self.radius, self.density, self.material, self.velocity, self.damping,
                             self.group_id)

    def emit(self, session) -> int:
        return session.add_particles(self.sample())


def _particles_at(positions, radius, density, material, velocity, damping, group_id) -> ParticleArrays:
    """``MpmParticle::new`` (world.rs:120-150) at every position, with the emitter's
    velocity, damping and group."""
    n = len(positions)
    template = MpmParticle.new(np.zeros(3), radius, density, material)
    pa = ParticleArrays(n, [material])
    pa.position[:] = positions
    pa.velocity[:] = _vec(velocity)
    pa.mass[:] = template.mass
    pa.rest_volume[:] = template.rest_volume
    pa.radius[:] = template.radius
    pa.damping[:] = float(damping)
    pa.plastic[:] = (template.plastic.plastic_det, template.plastic.hardening, template.plastic.log_volume_gain)
    pa.group_id[:] = int(group_id)
    return pa
