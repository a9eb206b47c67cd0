"""Python binding of the HIP MLS-MPM backend (ctypes over include/tessera_mpm.h).

The classes mirror the reference crate's public types so that code written
against tessera-mpm reads the same:

=============================  ==================================================
reference (crates/tessera-mpm)  this module
=============================  ==================================================
``MaterialModel``               :class:`MaterialModel` (material.rs:28-169)
``PlasticState``                :class:`PlasticState` (material.rs:5-24)
``MpmParticle::new``            :meth:`MpmParticle.new` (world.rs:118-150)
``MpmParams`` / ``WorldBounds``  :class:`MpmParams`, :class:`WorldBounds` (world.rs:10-49)
``RigidObstacle``               :class:`RigidObstacle` (obstacle.rs:79-262)
``MpmWorld``                    :class:`MpmWorld` (world.rs:206-303), a container
``GpuMpmTransfers``             :class:`HipMpmTransfers` (gpu.rs:160-167, 927-992)
``GpuMpmResidentSession``       :class:`HipMpmResidentSession` (gpu.rs:206-662)
``GpuMpmError``                 :class:`GpuMpmError` and subclasses (gpu.rs:135-154)
=============================  ==================================================

Bulk particle data lives in :class:`ParticleArrays` (structure of arrays); the
per-object :class:`MpmParticle` exists for small scenes and tests.  Matrices
are numpy ``(3, 3)`` arrays indexed ``[row, column]`` like nalgebra's
``Matrix3``.

This module performs no simulation on the CPU: stepping happens in the HIP
library, which must be built (``tessera_amd.build``) and fails loudly when it
is missing.
"""

from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass, field
from pathlib import Path
from typing import Iterable, Sequence

import numpy as np

LIBRARY_PATH = Path(__file__).resolve().parent / "libtessera_mpm_hip.so"

# ---------------------------------------------------------------------------
# Status codes and errors (gpu.rs:135-154, world.rs:165-181)
# ---------------------------------------------------------------------------
OK = 0
STATUS_NAMES = {
    1: "InvalidInput",
    2: "Capacity",
    3: "PendingSteps",
    4: "Readback",
    5: "ResidentUnavailable",
    6: "ExcessiveSubsteps",
    7: "NonFiniteState",
    8: "UnknownChunk",
    9: "ParticleCapacity",
    10: "NoDevice",
}


class GpuMpmError(Exception):
    """Base of all backend errors; ``status`` is the C status code."""

    status = -1


def _error_class(name: str, code: int):
    return type(name, (GpuMpmError,), {"status": code})


InvalidInput = _error_class("InvalidInput", 1)
Capacity = _error_class("Capacity", 2)
PendingSteps = _error_class("PendingSteps", 3)
Readback = _error_class("Readback", 4)
ResidentUnavailable = _error_class("ResidentUnavailable", 5)
ExcessiveSubsteps = _error_class("ExcessiveSubsteps", 6)
NonFiniteState = _error_class("NonFiniteState", 7)
UnknownChunk = _error_class("UnknownChunk", 8)
ParticleCapacity = _error_class("ParticleCapacity", 9)
NoDevice = _error_class("NoDevice", 10)
_ERRORS = {cls.status: cls for cls in (InvalidInput, Capacity, PendingSteps, Readback, ResidentUnavailable,
                                        ExcessiveSubsteps, NonFiniteState, UnknownChunk, ParticleCapacity, NoDevice)}

# ---------------------------------------------------------------------------
# C structures (include/tessera_mpm.h)
# ---------------------------------------------------------------------------
_dp = C.POINTER(C.c_double)


class CMaterial(C.Structure):
    _fields_ = [("kind", C.c_int32), ("reserved", C.c_int32), ("p", C.c_double * 5)]


class CParams(C.Structure):
    _fields_ = [
        ("cell_width", C.c_double),
        ("gravity", C.c_double * 3),
        ("max_substep", C.c_double),
        ("has_bounds", C.c_int32),
        ("use_cpic", C.c_int32),
        ("bound_min", C.c_double * 3),
        ("bound_max", C.c_double * 3),
    ]


class CObstacle(C.Structure):
    _fields_ = [
        ("shape", C.c_int32),
        ("boundary", C.c_int32),
        ("cpic_group", C.c_int32),
        ("plane_count", C.c_uint32),
        ("center", C.c_double * 3),
        ("orientation", C.c_double * 4),
        ("linear_velocity", C.c_double * 3),
        ("angular_velocity", C.c_double * 3),
        ("friction", C.c_double),
        ("dims", C.c_double * 4),
        ("vertices", C.c_double * 9),
        ("planes", _dp),
    ]


class CParticles(C.Structure):
    _fields_ = [
        ("count", C.c_uint64),
        ("position", _dp),
        ("velocity", _dp),
        ("deformation", _dp),
        ("affine", _dp),
        ("force", _dp),
        ("mass", _dp),
        ("rest_volume", _dp),
        ("radius", _dp),
        ("damping", _dp),
        ("plastic", _dp),
        ("material_index", C.POINTER(C.c_uint32)),
        ("group_id", C.POINTER(C.c_uint32)),
        ("transfer_color", C.POINTER(C.c_uint16)),
        ("enabled", C.POINTER(C.c_uint8)),
        ("fixed", C.POINTER(C.c_uint8)),
        ("chunk_id", C.POINTER(C.c_uint64)),
    ]


class CWorld(C.Structure):
    _fields_ = [
        ("params", CParams),
        ("particles", CParticles),
        ("materials", C.POINTER(CMaterial)),
        ("material_count", C.c_uint32),
        ("obstacle_count", C.c_uint32),
        ("obstacles", C.POINTER(CObstacle)),
        ("substeps", C.c_uint64),
        ("next_chunk_id", C.c_uint64),
    ]


class CTopologyStats(C.Structure):
    _fields_ = [
        ("live_nodes", C.c_uint64),
        ("grid_blocks", C.c_uint64),
        ("particle_blocks", C.c_uint64),
        ("binned_particles", C.c_uint64),
    ]


class CRigidBody(C.Structure):
    _fields_ = [
        ("position", C.c_double * 3),
        ("inverse_mass", C.c_double),
        ("orientation", C.c_double * 4),
        ("linear_velocity", C.c_double * 3),
        ("angular_velocity", C.c_double * 3),
        ("inverse_inertia", C.c_double * 3),
    ]


class CRigidMapping(C.Structure):
    _fields_ = [
        ("body", C.c_uint32),
        ("reserved", C.c_uint32),
        ("local_center", C.c_double * 3),
        ("local_orientation", C.c_double * 4),
    ]


# ---------------------------------------------------------------------------
# Library loading
# ---------------------------------------------------------------------------
_LIB = None


def load_library(path: Path | str | None = None) -> C.CDLL:
    """Loads the HIP library (built in-tree); raises when it is missing."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    lib_path = Path(path) if path else LIBRARY_PATH
    if not lib_path.exists():
        raise RuntimeError(
            f"HIP library {lib_path} is missing; build it with `python -m tessera_amd.build` "
            "(there is no CPU fallback)"
        )
    lib = C.CDLL(str(lib_path))
    vp = C.c_void_p
    sig = {
        "tmpm_last_error": (C.c_char_p, []),
        "tmpm_version": (C.c_char_p, []),
        "tmpm_context_create": (C.c_int, [C.c_int32, C.POINTER(vp)]),
        "tmpm_context_destroy": (None, [vp]),
        "tmpm_session_create": (C.c_int, [vp, C.POINTER(CWorld), C.c_double, C.c_int, C.POINTER(vp)]),
        "tmpm_session_destroy": (None, [vp]),
        "tmpm_session_submit_steps": (C.c_int, [vp, C.c_uint32]),
        "tmpm_session_wait": (C.c_int, [vp]),
        "tmpm_session_synchronize": (C.c_int, [vp]),
        "tmpm_session_step": (C.c_int, [vp, C.c_uint32]),
        "tmpm_session_synchronize_async": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "tmpm_session_memory": (C.c_int, [vp, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
        "tmpm_session_reaction_saturations": (C.c_uint64, [vp]),
        "tmpm_batch_stage_scene": (C.c_int, [C.POINTER(CWorld), C.c_double, C.POINTER(vp)]),
        "tmpm_batch_add_staged": (C.c_int, [vp, vp, C.POINTER(C.c_uint32)]),
        "tmpm_staged_scene_destroy": (None, [vp]),
        "tmpm_session_set_audit": (C.c_int, [vp, C.c_int32]),
        "tmpm_session_read_audit": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_double), C.c_uint64,
                                              C.POINTER(C.c_uint64)]),
        "tmpm_tune_cache_clear": (None, []),
        "tmpm_tune_cache_size": (C.c_uint64, []),
        "tmpm_session_step_adaptive": (C.c_int, [vp, C.c_double]),
        "tmpm_session_pending_substeps": (C.c_uint64, [vp]),
        "tmpm_session_valid": (C.c_int32, [vp]),
        "tmpm_session_substeps": (C.c_uint64, [vp]),
        "tmpm_session_particle_count": (C.c_uint64, [vp]),
        "tmpm_session_read_particles": (C.c_int, [vp, C.POINTER(CParticles)]),
        "tmpm_session_set_particle_force": (C.c_int, [vp, C.c_uint64, C.c_double * 3]),
        "tmpm_session_set_obstacles": (C.c_int, [vp, C.POINTER(CObstacle), C.c_uint32]),
        "tmpm_session_add_particles": (
            C.c_int, [vp, C.POINTER(CParticles), C.POINTER(CMaterial), C.c_uint32, C.POINTER(C.c_uint64)]),
        "tmpm_session_remove_chunk": (C.c_int, [vp, C.c_uint64, C.POINTER(C.c_uint64)]),
        "tmpm_session_encode_snapshot": (C.c_int, [vp, vp, C.c_uint64, vp, C.POINTER(C.c_uint32)]),
        "tmpm_session_stable_timestep": (C.c_double, [vp]),
        "tmpm_session_topology_stats": (C.c_int, [vp, C.POINTER(CTopologyStats)]),
        "tmpm_session_export_nodes": (C.c_int, [vp, C.POINTER(C.c_int32), C.c_uint64, C.POINTER(C.c_uint64)]),
        "tmpm_session_export_blocks": (
            C.c_int, [vp, C.POINTER(C.c_int32), C.c_uint64, C.POINTER(C.c_int64), C.POINTER(C.c_uint64)]),
        "tmpm_session_set_timing": (C.c_int, [vp, C.c_int32]),
        "tmpm_session_set_graphs": (C.c_int, [vp, C.c_int32]),
        "tmpm_session_autotune": (C.c_int, [vp, C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_session_launch_config": (C.c_int, [vp, C.POINTER(C.c_int32)]),
        "tmpm_session_kernel_times": (
            C.c_int,
            [vp, C.c_char_p, C.c_uint64, C.POINTER(C.c_double), C.POINTER(C.c_uint64), C.c_uint32,
             C.POINTER(C.c_uint32)]),
        "tmpm_batch_create": (
            C.c_int, [vp, C.POINTER(CWorld), C.POINTER(C.c_double), C.c_uint32, C.c_int, C.POINTER(vp)]),
        "tmpm_batch_scene_count": (C.c_uint32, [vp]),
        "tmpm_batch_scene_status": (C.c_int, [vp, C.c_uint32]),
        "tmpm_batch_scene_particle_count": (C.c_uint64, [vp, C.c_uint32]),
        "tmpm_batch_scene_substeps": (C.c_uint64, [vp, C.c_uint32]),
        "tmpm_batch_read_particles": (C.c_int, [vp, C.c_uint32, C.POINTER(CParticles)]),
        "tmpm_batch_set_particle_force": (C.c_int, [vp, C.c_uint32, C.c_uint64, C.c_double * 3]),
        "tmpm_batch_set_obstacles": (C.c_int, [vp, C.c_uint32, C.POINTER(CObstacle), C.c_uint32]),
        "tmpm_batch_add_scene": (C.c_int, [vp, C.POINTER(CWorld), C.c_double, C.POINTER(C.c_uint32)]),
        "tmpm_batch_remove_scene": (C.c_int, [vp, C.c_uint32]),
        "tmpm_session_set_reactions": (C.c_int, [vp, C.c_int32]),
        "tmpm_session_read_reactions": (C.c_int, [vp, C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_session_set_rigid_bodies": (
            C.c_int, [vp, C.POINTER(CRigidBody), C.c_uint32, C.POINTER(CRigidMapping), C.c_uint32, C.c_double * 3,
                      C.c_double, C.c_int32]),
        "tmpm_session_read_rigid_bodies": (C.c_int, [vp, C.POINTER(CRigidBody), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_session_read_obstacles": (C.c_int, [vp, C.POINTER(CObstacle), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_batch_read_reactions": (
            C.c_int, [vp, C.c_uint32, C.POINTER(C.c_double), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_batch_set_rigid_bodies": (
            C.c_int, [vp, C.c_uint32, C.POINTER(CRigidBody), C.c_uint32, C.POINTER(CRigidMapping), C.c_uint32,
                      C.c_double * 3, C.c_double, C.c_int32]),
        "tmpm_batch_read_rigid_bodies": (
            C.c_int, [vp, C.c_uint32, C.POINTER(CRigidBody), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_batch_read_obstacles": (
            C.c_int, [vp, C.c_uint32, C.POINTER(CObstacle), C.c_uint32, C.POINTER(C.c_uint32)]),
        "tmpm_bandwidth_probe": (C.c_int, [C.c_int32, C.c_uint64, C.c_uint32, C.POINTER(C.c_double)]),
        "tmpm_timestep_bound": (C.c_double, [C.POINTER(CMaterial), C.c_double, C.c_double * 3, C.c_double, C.c_double]),
        "tmpm_material_valid": (C.c_int32, [C.POINTER(CMaterial)]),
        "tmpm_sample_closed_mesh_volume": (C.c_int, [vp, _dp, C.c_uint64, C.POINTER(C.c_uint32), C.c_uint64,
                                                     C.c_double, C.c_uint64, C.c_uint64, _dp,
                                                     C.POINTER(C.c_uint64)]),
        "tmpm_sample_closed_mesh_volume_noalloc": (C.c_int, [vp, _dp, C.c_uint64, C.POINTER(C.c_uint32),
                                                             C.c_uint64, C.c_double, C.c_uint64, C.c_uint64, _dp,
                                                             C.POINTER(C.c_uint64)]),
        "tmpm_obstacle_valid": (C.c_int32, [C.POINTER(CObstacle)]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    """Names of every function the C header declares (checked against the .so)."""
    header = Path(__file__).resolve().parent.parent / "include" / "tessera_mpm.h"
    import re

    text = header.read_text()
    return sorted(set(re.findall(r"\b(tmpm_[a-z_0-9]+)\s*\(", text)))


def _check(status: int, lib=None) -> None:
    if status == OK:
        return
    lib = lib or load_library()
    msg = lib.tmpm_last_error().decode(errors="replace")
    cls = _ERRORS.get(status, GpuMpmError)
    raise cls(msg or STATUS_NAMES.get(status, str(status)))


# ---------------------------------------------------------------------------
# Domain types (mirrors of the reference's Rust types)
# ---------------------------------------------------------------------------
class MaterialModel:
    """Constitutive model (material.rs:28-169); ``kind`` follows variant order."""

    LINEAR_ELASTIC = 0
    NEO_HOOKEAN = 1
    SAND = 2
    FLUID = 3
    SNOW = 4
    SAND_NEO_HOOKEAN = 5

    __slots__ = ("kind", "params")

    def __init__(self, kind: int, params: Sequence[float]):
        self.kind = int(kind)
        p = [float(v) for v in params] + [0.0] * (5 - len(params))
        self.params = tuple(p[:5])

    @classmethod
    def elastic(cls, young_modulus: float, poisson_ratio: float) -> "MaterialModel":
        return cls(cls.LINEAR_ELASTIC, (young_modulus, poisson_ratio))

    @classmethod
    def neo_hookean(cls, young_modulus: float, poisson_ratio: float) -> "MaterialModel":
        return cls(cls.NEO_HOOKEAN, (young_modulus, poisson_ratio))

    @classmethod
    def sand(cls, young_modulus, poisson_ratio, friction_angle, cohesion) -> "MaterialModel":
        return cls(cls.SAND, (young_modulus, poisson_ratio, friction_angle, cohesion))

    @classmethod
    def sand_neo_hookean(cls, young_modulus, poisson_ratio, friction_angle, cohesion) -> "MaterialModel":
        return cls(cls.SAND_NEO_HOOKEAN, (young_modulus, poisson_ratio, friction_angle, cohesion))

    @classmethod
    def fluid(cls, bulk_modulus: float, gamma: float, viscosity: float, tensile_stiffness: float = 0.25):
        """``MaterialModel::fluid`` (material.rs:151-158) fixes tensile stiffness 0.25."""
        return cls(cls.FLUID, (bulk_modulus, gamma, viscosity, tensile_stiffness))

    @classmethod
    def snow(cls, young_modulus: float, poisson_ratio: float) -> "MaterialModel":
        """Stomakhin yield box defaults (material.rs:161-169)."""
        return cls(cls.SNOW, (young_modulus, poisson_ratio, 2.5e-2, 7.5e-3, 10.0))

    def to_c(self) -> CMaterial:
        m = CMaterial()
        m.kind = self.kind
        for i, v in enumerate(self.params):
            m.p[i] = v
        return m

    def key(self) -> tuple:
        return (self.kind, self.params)

    def __eq__(self, other) -> bool:
        return isinstance(other, MaterialModel) and self.key() == other.key()

    def __hash__(self) -> int:
        return hash(self.key())

    def __repr__(self) -> str:
        return f"MaterialModel(kind={self.kind}, params={self.params})"


@dataclass
class PlasticState:
    """material.rs:5-24."""

    plastic_det: float = 1.0
    hardening: float = 0.0
    log_volume_gain: float = 0.0


def _vec(v) -> np.ndarray:
    return np.array(v, dtype=np.float64).reshape(3)


@dataclass
class MpmParticle:
    """One material point (world.rs:67-102)."""

    position: np.ndarray
    velocity: np.ndarray
    deformation: np.ndarray
    affine: np.ndarray
    force: np.ndarray
    mass: float
    rest_volume: float
    radius: float
    material: MaterialModel
    plastic: PlasticState
    damping: float = 0.0
    group_id: int = 0
    transfer_color: int = 0
    enabled: bool = True
    fixed: bool = False
    chunk_id: int = 0  # read-out only (MpmParticle::chunk_id, world.rs:153-156)

    @classmethod
    def new(cls, position, radius: float, density: float, material: MaterialModel) -> "MpmParticle":
        """``MpmParticle::new`` (world.rs:120-150): rest volume (2r)^3."""
        rest_volume = (2.0 * radius) ** 3
        hardening = 1.0 if material.kind in (MaterialModel.SAND, MaterialModel.SAND_NEO_HOOKEAN) else 0.0
        return cls(
            position=_vec(position),
            velocity=np.zeros(3),
            deformation=np.eye(3),
            affine=np.zeros((3, 3)),
            force=np.zeros(3),
            mass=density * rest_volume,
            rest_volume=rest_volume,
            radius=float(radius),
            material=material,
            plastic=PlasticState(hardening=hardening),
        )

    def clone(self) -> "MpmParticle":
        return MpmParticle(
            position=self.position.copy(), velocity=self.velocity.copy(), deformation=self.deformation.copy(),
            affine=self.affine.copy(), force=self.force.copy(), mass=self.mass, rest_volume=self.rest_volume,
            radius=self.radius, material=self.material, plastic=PlasticState(**vars(self.plastic)),
            damping=self.damping, group_id=self.group_id, transfer_color=self.transfer_color,
            enabled=self.enabled, fixed=self.fixed)


def _column_major(m: np.ndarray) -> np.ndarray:
    """(n, 3, 3) row-indexed view of a copy of `m` whose memory is column-major per matrix."""
    buf = np.ascontiguousarray(np.transpose(m, (0, 2, 1)))
    return buf.transpose(0, 2, 1)


class ParticleArrays:
    """Structure-of-arrays particle storage in insertion order."""

    FIELDS = ("position", "velocity", "deformation", "affine", "force", "mass", "rest_volume", "radius", "damping",
              "plastic", "material_index", "group_id", "transfer_color", "enabled", "fixed", "chunk_id")

    def __init__(self, n: int, materials: Sequence[MaterialModel] | None = None):
        self.position = np.zeros((n, 3))
        self.velocity = np.zeros((n, 3))
        # matrices are (n, 3, 3) views of column-major storage, the layout of
        # the C ABI (nalgebra order), so handing them to the library copies nothing
        self.deformation = _column_major(np.tile(np.eye(3), (n, 1, 1)))
        self.affine = _column_major(np.zeros((n, 3, 3)))
        self.force = np.zeros((n, 3))
        self.mass = np.zeros(n)
        self.rest_volume = np.zeros(n)
        self.radius = np.zeros(n)
        self.damping = np.zeros(n)
        self.plastic = np.tile(np.array([1.0, 0.0, 0.0]), (n, 1))
        self.material_index = np.zeros(n, dtype=np.uint32)
        self.group_id = np.zeros(n, dtype=np.uint32)
        self.transfer_color = np.zeros(n, dtype=np.uint16)
        self.enabled = np.ones(n, dtype=bool)
        self.fixed = np.zeros(n, dtype=bool)
        self.chunk_id = np.zeros(n, dtype=np.uint64)
        self.materials: list[MaterialModel] = list(materials or [])

    def __len__(self) -> int:
        return len(self.mass)

    @classmethod
    def from_particles(cls, particles: Iterable[MpmParticle]) -> "ParticleArrays":
        plist = list(particles)
        out = cls(len(plist))
        index: dict[MaterialModel, int] = {}
        for i, p in enumerate(plist):
            out.position[i] = p.position
            out.velocity[i] = p.velocity
            out.deformation[i] = p.deformation
            out.affine[i] = p.affine
            out.force[i] = p.force
            out.mass[i] = p.mass
            out.rest_volume[i] = p.rest_volume
            out.radius[i] = p.radius
            out.damping[i] = p.damping
            out.plastic[i] = (p.plastic.plastic_det, p.plastic.hardening, p.plastic.log_volume_gain)
            if p.material not in index:
                index[p.material] = len(out.materials)
                out.materials.append(p.material)
            out.material_index[i] = index[p.material]
            out.group_id[i] = p.group_id
            out.transfer_color[i] = p.transfer_color
            out.enabled[i] = p.enabled
            out.fixed[i] = p.fixed
        return out

    def particle(self, i: int) -> MpmParticle:
        return MpmParticle(
            position=self.position[i].copy(), velocity=self.velocity[i].copy(),
            deformation=self.deformation[i].copy(), affine=self.affine[i].copy(), force=self.force[i].copy(),
            mass=float(self.mass[i]), rest_volume=float(self.rest_volume[i]), radius=float(self.radius[i]),
            material=self.materials[int(self.material_index[i])],
            plastic=PlasticState(*map(float, self.plastic[i])), damping=float(self.damping[i]),
            group_id=int(self.group_id[i]), transfer_color=int(self.transfer_color[i]),
            enabled=bool(self.enabled[i]), fixed=bool(self.fixed[i]), chunk_id=int(self.chunk_id[i]))

    def to_particles(self) -> list[MpmParticle]:
        return [self.particle(i) for i in range(len(self))]

    def copy(self) -> "ParticleArrays":
        out = ParticleArrays(0, self.materials)
        for f in self.FIELDS:
            setattr(out, f, getattr(self, f).copy(order="K"))
        return out

    def concat(self, other: "ParticleArrays") -> "ParticleArrays":
        out = ParticleArrays(0)
        out.materials = list(self.materials)
        remap = []
        for m in other.materials:
            if m in out.materials:
                remap.append(out.materials.index(m))
            else:
                remap.append(len(out.materials))
                out.materials.append(m)
        for f in self.FIELDS:
            a, b = getattr(self, f), getattr(other, f)
            if f == "material_index":
                b = np.array(remap, dtype=np.uint32)[b] if len(b) else b
            setattr(out, f, np.concatenate([a, b.astype(a.dtype)]))
        return out

    def select(self, mask: np.ndarray) -> "ParticleArrays":
        out = ParticleArrays(0, self.materials)
        for f in self.FIELDS:
            setattr(out, f, getattr(self, f)[mask].copy())
        return out

    # -- C view -----------------------------------------------------------
    def c_view(self):
        """Returns (CParticles, CMaterial array, keep-alive list)."""
        n = len(self)
        keep = []

        def arr(a, dtype):
            b = np.ascontiguousarray(a, dtype=dtype)
            keep.append(b)
            return b

        cp = CParticles()
        cp.count = n
        cp.position = arr(self.position.reshape(n, 3), np.float64).ctypes.data_as(_dp)
        cp.velocity = arr(self.velocity.reshape(n, 3), np.float64).ctypes.data_as(_dp)
        # nalgebra column-major: element (r, c) at c*3 + r == row-major transpose
        cp.deformation = arr(np.transpose(self.deformation, (0, 2, 1)).reshape(n, 9), np.float64).ctypes.data_as(_dp)
        cp.affine = arr(np.transpose(self.affine, (0, 2, 1)).reshape(n, 9), np.float64).ctypes.data_as(_dp)
        cp.force = arr(self.force.reshape(n, 3), np.float64).ctypes.data_as(_dp)
        cp.mass = arr(self.mass, np.float64).ctypes.data_as(_dp)
        cp.rest_volume = arr(self.rest_volume, np.float64).ctypes.data_as(_dp)
        cp.radius = arr(self.radius, np.float64).ctypes.data_as(_dp)
        cp.damping = arr(self.damping, np.float64).ctypes.data_as(_dp)
        cp.plastic = arr(self.plastic.reshape(n, 3), np.float64).ctypes.data_as(_dp)
        cp.material_index = arr(self.material_index, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32))
        cp.group_id = arr(self.group_id, np.uint32).ctypes.data_as(C.POINTER(C.c_uint32))
        cp.transfer_color = arr(self.transfer_color, np.uint16).ctypes.data_as(C.POINTER(C.c_uint16))
        cp.enabled = arr(self.enabled, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
        cp.fixed = arr(self.fixed, np.uint8).ctypes.data_as(C.POINTER(C.c_uint8))
        cp.chunk_id = arr(self.chunk_id, np.uint64).ctypes.data_as(C.POINTER(C.c_uint64))
        mats = (CMaterial * max(1, len(self.materials)))()
        for i, m in enumerate(self.materials):
            mats[i] = m.to_c()
        keep.append(mats)
        return cp, mats, keep


@dataclass
class WorldBounds:
    """world.rs:10-25."""

    min: np.ndarray
    max: np.ndarray


@dataclass
class MpmParams:
    """world.rs:27-49 defaults: h=0.1, g=(0,0,-9.81), max_substep=1e-3, no bounds."""

    cell_width: float = 0.1
    gravity: np.ndarray = field(default_factory=lambda: np.array([0.0, 0.0, -9.81]))
    max_substep: float = 1.0 / 1000.0
    bounds: WorldBounds | None = None

    def to_c(self, use_cpic: bool = True) -> CParams:
        p = CParams()
        p.cell_width = self.cell_width
        for i in range(3):
            p.gravity[i] = float(self.gravity[i])
        p.max_substep = self.max_substep
        p.has_bounds = 1 if self.bounds is not None else 0
        p.use_cpic = 1 if use_cpic else 0
        if self.bounds is not None:
            for i in range(3):
                p.bound_min[i] = float(self.bounds.min[i])
                p.bound_max[i] = float(self.bounds.max[i])
        return p


class ObstacleShape:
    SPHERE, BOX, CAPSULE, CYLINDER, CONE, GROUND, TRIANGLE_PRISM, CONVEX = range(8)


class ObstacleBoundary:
    """obstacle.rs:65-77 (Slip is the default)."""

    SLIP, STICK, SEPARATE, NON_REFLECTING, FREE_SLIP = range(5)


def quat_from_axis_angle(axis, angle: float) -> np.ndarray:
    """Unit quaternion (i, j, k, w) of a rotation about ``axis``."""
    a = np.asarray(axis, dtype=np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(angle / 2.0)
    return np.array([a[0] * s, a[1] * s, a[2] * s, math.cos(angle / 2.0)])


IDENTITY_QUAT = np.array([0.0, 0.0, 0.0, 1.0])


@dataclass
class RigidObstacle:
    """A prescribed rigid obstacle (obstacle.rs:79-262)."""

    shape: int
    center: np.ndarray
    orientation: np.ndarray = field(default_factory=lambda: IDENTITY_QUAT.copy())
    dims: tuple = (0.0, 0.0, 0.0, 0.0)
    vertices: np.ndarray | None = None
    planes: np.ndarray | None = None
    linear_velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    angular_velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    friction: float = 0.0
    boundary: int = ObstacleBoundary.SLIP
    cpic_group: int | None = None

    @classmethod
    def sphere(cls, center, radius: float) -> "RigidObstacle":
        return cls(ObstacleShape.SPHERE, _vec(center), dims=(radius, 0.0, 0.0, 0.0))

    @classmethod
    def cuboid(cls, center, half_extents, orientation=None) -> "RigidObstacle":
        h = _vec(half_extents)
        return cls(ObstacleShape.BOX, _vec(center), _q(orientation), dims=(h[0], h[1], h[2], 0.0))

    @classmethod
    def capsule(cls, center, half_height, radius, orientation=None) -> "RigidObstacle":
        return cls(ObstacleShape.CAPSULE, _vec(center), _q(orientation), dims=(half_height, radius, 0.0, 0.0))

    @classmethod
    def cylinder(cls, center, half_height, radius, orientation=None) -> "RigidObstacle":
        return cls(ObstacleShape.CYLINDER, _vec(center), _q(orientation), dims=(half_height, radius, 0.0, 0.0))

    @classmethod
    def cone(cls, center, half_height, radius, orientation=None) -> "RigidObstacle":
        return cls(ObstacleShape.CONE, _vec(center), _q(orientation), dims=(half_height, radius, 0.0, 0.0))

    @classmethod
    def ground(cls, center, half_extents, orientation=None) -> "RigidObstacle":
        return cls(ObstacleShape.GROUND, _vec(center), _q(orientation),
                   dims=(float(half_extents[0]), float(half_extents[1]), 0.0, 0.0))

    @classmethod
    def triangle_prism(cls, center, vertices, half_thickness, orientation=None) -> "RigidObstacle":
        return cls(ObstacleShape.TRIANGLE_PRISM, _vec(center), _q(orientation), dims=(half_thickness, 0, 0, 0),
                   vertices=np.asarray(vertices, dtype=np.float64).reshape(3, 3))

    @classmethod
    def convex(cls, center, vertices, face_normals, orientation=None) -> "RigidObstacle":
        """obstacle.rs:211-245: planes from face normals and support offsets."""
        verts = np.asarray(vertices, dtype=np.float64).reshape(-1, 3)
        normals = np.asarray(face_normals, dtype=np.float64).reshape(-1, 3)
        planes = np.array([[n[0], n[1], n[2], max(float(n @ v) for v in verts)] for n in normals])
        if len(verts) >= 4 and np.all(np.isfinite(verts)):
            bound = float(max(np.linalg.norm(v) for v in verts))
        else:
            bound = float("nan")
        return cls(ObstacleShape.CONVEX, _vec(center), _q(orientation), dims=(bound, 0, 0, 0), planes=planes)

    def to_c(self, keep: list) -> CObstacle:
        o = CObstacle()
        o.shape = self.shape
        o.boundary = self.boundary
        o.cpic_group = -1 if self.cpic_group is None else int(self.cpic_group)
        for i in range(3):
            o.center[i] = float(self.center[i])
            o.linear_velocity[i] = float(self.linear_velocity[i])
            o.angular_velocity[i] = float(self.angular_velocity[i])
        for i in range(4):
            o.orientation[i] = float(self.orientation[i])
            o.dims[i] = float(self.dims[i])
        o.friction = float(self.friction)
        if self.vertices is not None:
            for i, v in enumerate(np.asarray(self.vertices).reshape(9)):
                o.vertices[i] = float(v)
        if self.planes is not None:
            pl = np.ascontiguousarray(self.planes, dtype=np.float64).reshape(-1)
            keep.append(pl)
            o.planes = pl.ctypes.data_as(_dp)
            o.plane_count = len(pl) // 4
        return o


def _q(orientation) -> np.ndarray:
    return IDENTITY_QUAT.copy() if orientation is None else np.asarray(orientation, dtype=np.float64).reshape(4)


def obstacles_to_c(obstacles: Sequence[RigidObstacle]):
    keep: list = []
    arr = (CObstacle * max(1, len(obstacles)))()
    for i, o in enumerate(obstacles):
        arr[i] = o.to_c(keep)
    keep.append(arr)
    return arr, keep


class MpmWorld:
    """A world container (world.rs:206-221): particles in insertion order,
    parameters, obstacles and the substep count.  Stepping is done by a
    :class:`HipMpmResidentSession`."""

    def __init__(self, particles, params: MpmParams | None = None):
        if not isinstance(particles, ParticleArrays):
            particles = ParticleArrays.from_particles(particles)
        self.particles = particles
        self.params = params or MpmParams()
        self.obstacles: list[RigidObstacle] = []
        self.substeps = 0
        self.use_cpic = True

    def set_obstacles(self, obstacles: Sequence[RigidObstacle]) -> None:
        self.obstacles = list(obstacles)

    def clone(self) -> "MpmWorld":
        w = MpmWorld(self.particles.copy(), self.params)
        w.obstacles = list(self.obstacles)
        w.substeps = self.substeps
        w.use_cpic = self.use_cpic
        return w

    def stable_timestep(self) -> float:
        """``MpmWorld::stable_timestep`` (world.rs:289-303), evaluated by the library."""
        lib = load_library()
        p = self.particles
        s = self.params.max_substep
        for i in range(len(p)):
            if not p.enabled[i] or p.fixed[i]:
                continue
            m = p.materials[int(p.material_index[i])].to_c()
            v = (C.c_double * 3)(*map(float, p.velocity[i]))
            b = lib.tmpm_timestep_bound(C.byref(m), float(p.mass[i] / p.rest_volume[i]), v, self.params.cell_width,
                                        float(np.linalg.det(p.deformation[i])))
            s = min(s, b)
        return min(s, self.params.max_substep)

    def c_desc(self):
        cp, mats, keep = self.particles.c_view()
        obs, okeep = obstacles_to_c(self.obstacles)
        keep += okeep
        d = CWorld()
        d.params = self.params.to_c(self.use_cpic)
        d.particles = cp
        d.materials = C.cast(mats, C.POINTER(CMaterial))
        d.material_count = len(self.particles.materials)
        d.obstacle_count = len(self.obstacles)
        d.obstacles = C.cast(obs, C.POINTER(CObstacle))
        d.substeps = self.substeps
        d.next_chunk_id = int(getattr(self, "next_chunk_id", 0) or 0)
        keep.append(d)
        return d, keep


# ---------------------------------------------------------------------------
# Device context and resident session
# ---------------------------------------------------------------------------
@dataclass
class RigidBody:
    """Rigid body state (GpuRigidBodyState, phys/src/gpu_rigid_state.rs:17-28); zero inverse
    mass is static or kinematic, orientation is XYZW, inertia the body-frame inverse principal."""

    position: np.ndarray
    inverse_mass: float
    orientation: np.ndarray = field(default_factory=lambda: IDENTITY_QUAT.copy())
    linear_velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    angular_velocity: np.ndarray = field(default_factory=lambda: np.zeros(3))
    inverse_inertia: np.ndarray = field(default_factory=lambda: np.zeros(3))

    @classmethod
    def box(cls, center, half_extents, density: float, orientation=None) -> "RigidBody":
        """Solid box: mass and principal inertia from the box formulas."""
        h = _vec(half_extents)
        mass = density * 8.0 * h[0] * h[1] * h[2]
        inertia = mass / 3.0 * np.array([h[1] ** 2 + h[2] ** 2, h[0] ** 2 + h[2] ** 2, h[0] ** 2 + h[1] ** 2])
        return cls(_vec(center), 1.0 / mass, _q(orientation), inverse_inertia=1.0 / inertia)

    def to_c(self) -> "CRigidBody":
        c = CRigidBody()
        c.position[:] = list(map(float, self.position))
        c.inverse_mass = float(self.inverse_mass)
        c.orientation[:] = list(map(float, self.orientation))
        c.linear_velocity[:] = list(map(float, self.linear_velocity))
        c.angular_velocity[:] = list(map(float, self.angular_velocity))
        c.inverse_inertia[:] = list(map(float, self.inverse_inertia))
        return c

    @classmethod
    def from_c(cls, c: "CRigidBody") -> "RigidBody":
        return cls(np.array(c.position[:]), float(c.inverse_mass), np.array(c.orientation[:]),
                   np.array(c.linear_velocity[:]), np.array(c.angular_velocity[:]), np.array(c.inverse_inertia[:]))

    def as_dict(self) -> dict:
        return {"position": np.array(self.position, float), "inverse_mass": float(self.inverse_mass),
                "orientation": np.array(self.orientation, float),
                "linear_velocity": np.array(self.linear_velocity, float),
                "angular_velocity": np.array(self.angular_velocity, float),
                "inverse_inertia": np.array(self.inverse_inertia, float)}


@dataclass
class RigidMapping:
    """Obstacle -> body mapping (gpu_rigid_sync.rs:15-30); body None: the obstacle is not driven."""

    body: int | None
    local_center: np.ndarray = field(default_factory=lambda: np.zeros(3))
    local_orientation: np.ndarray = field(default_factory=lambda: IDENTITY_QUAT.copy())

    def to_c(self) -> "CRigidMapping":
        c = CRigidMapping()
        c.body = 0xFFFFFFFF if self.body is None else int(self.body)
        c.local_center[:] = list(map(float, self.local_center))
        c.local_orientation[:] = list(map(float, self.local_orientation))
        return c

    def as_dict(self) -> dict:
        return {"body": self.body, "local_center": np.array(self.local_center, float),
                "local_orientation": np.array(self.local_orientation, float)}


def _rigid_args(bodies, mappings):
    cb = (CRigidBody * max(1, len(bodies)))(*[b.to_c() for b in bodies])
    cm = (CRigidMapping * max(1, len(mappings)))(*[m.to_c() for m in mappings])
    return cb, cm


def _read_n(fn, ctype, cap_hint: int):
    count = C.c_uint32()
    cap = max(1, cap_hint)
    while True:
        arr = (ctype * cap)()
        st = fn(arr, cap, C.byref(count))
        if st == OK:
            return arr, int(count.value)
        if count.value > cap:
            cap = int(count.value)
            continue
        _check(st)


def _obstacle_from_c(c: "CObstacle") -> "RigidObstacle":
    o = RigidObstacle(int(c.shape), np.array(c.center[:]), np.array(c.orientation[:]), tuple(c.dims[:]))
    o.vertices = np.array(c.vertices[:]).reshape(3, 3) if c.shape == ObstacleShape.TRIANGLE_PRISM else None
    o.linear_velocity = np.array(c.linear_velocity[:])
    o.angular_velocity = np.array(c.angular_velocity[:])
    o.friction = float(c.friction)
    o.boundary = int(c.boundary)
    o.cpic_group = None if c.cpic_group < 0 else int(c.cpic_group)
    return o


class HipMpmTransfers:
    """Device context (the reference's ``GpuMpmTransfers``)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = C.c_void_p()
        _check(self.lib.tmpm_context_create(int(device), C.byref(h)), self.lib)
        self._h = h

    @property
    def handle(self):
        return self._h

    def close(self) -> None:
        if self._h:
            self.lib.tmpm_context_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


P2G_MODES = {"fp32": 0, "fixed64": 1}


def _read_particle_arrays(n: int, read, lib) -> ParticleArrays:
    """Reads n particles through `read(CParticles*)` into a ParticleArrays (row-major matrices)."""
    out = ParticleArrays(n)
    flat = {
        "position": np.zeros((n, 3)), "velocity": np.zeros((n, 3)), "deformation": np.zeros((n, 9)),
        "affine": np.zeros((n, 9)), "force": np.zeros((n, 3)), "mass": np.zeros(n), "rest_volume": np.zeros(n),
        "radius": np.zeros(n), "damping": np.zeros(n), "plastic": np.zeros((n, 3)),
    }
    ints = {
        "group_id": np.zeros(n, np.uint32), "transfer_color": np.zeros(n, np.uint16),
        "enabled": np.zeros(n, np.uint8), "fixed": np.zeros(n, np.uint8), "chunk_id": np.zeros(n, np.uint64),
    }
    cp = CParticles()
    cp.count = n
    for k, a in flat.items():
        setattr(cp, k, a.ctypes.data_as(_dp))
    cp.group_id = ints["group_id"].ctypes.data_as(C.POINTER(C.c_uint32))
    cp.transfer_color = ints["transfer_color"].ctypes.data_as(C.POINTER(C.c_uint16))
    cp.enabled = ints["enabled"].ctypes.data_as(C.POINTER(C.c_uint8))
    cp.fixed = ints["fixed"].ctypes.data_as(C.POINTER(C.c_uint8))
    cp.chunk_id = ints["chunk_id"].ctypes.data_as(C.POINTER(C.c_uint64))
    _check(read(C.byref(cp)), lib)
    out.position = flat["position"]
    out.velocity = flat["velocity"]
    out.deformation = flat["deformation"].reshape(n, 3, 3).transpose(0, 2, 1)
    out.affine = flat["affine"].reshape(n, 3, 3).transpose(0, 2, 1)
    out.force = flat["force"]
    out.mass = flat["mass"]
    out.rest_volume = flat["rest_volume"]
    out.radius = flat["radius"]
    out.damping = flat["damping"]
    out.plastic = flat["plastic"]
    out.group_id = ints["group_id"]
    out.transfer_color = ints["transfer_color"]
    out.enabled = ints["enabled"].astype(bool)
    out.fixed = ints["fixed"].astype(bool)
    out.chunk_id = ints["chunk_id"]
    return out


class HipMpmResidentSession:
    """Resident HIP session (the reference's ``GpuMpmResidentSession``).

    ``p2g_mode`` selects FP32 atomics ("fp32") or 64-bit fixed-point
    accumulation ("fixed64", bit-reproducible)."""

    def __init__(self, transfers: HipMpmTransfers, world: MpmWorld, substep_dt: float, p2g_mode: str = "fp32"):
        self.lib = transfers.lib
        self.transfers = transfers
        self._world = world.clone()
        desc, keep = world.c_desc()
        h = C.c_void_p()
        _check(self.lib.tmpm_session_create(transfers.handle, C.byref(desc), float(substep_dt),
                                            P2G_MODES[p2g_mode], C.byref(h)), self.lib)
        del keep
        self._h = h
        self.substep_dt = substep_dt
        self.p2g_mode = p2g_mode

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.tmpm_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping (gpu.rs:286-450) -----------------------------------------
    def step(self, steps: int) -> None:
        _check(self.lib.tmpm_session_step(self._h, int(steps)), self.lib)

    def submit_steps(self, steps: int) -> None:
        _check(self.lib.tmpm_session_submit_steps(self._h, int(steps)), self.lib)

    def synchronize(self) -> None:
        _check(self.lib.tmpm_session_synchronize(self._h), self.lib)

    def wait(self) -> None:
        """Device completion only (no readback, no validation)."""
        _check(self.lib.tmpm_session_wait(self._h), self.lib)

    def synchronize_async(self) -> bool:
        """``synchronize_async`` (gpu.rs:413-450) as a poll: False while the device still runs the
        submitted substeps (nothing changes; they stay pending), True once they were validated and
        committed as by :meth:`synchronize`, whose errors it raises."""
        return _poll_synchronize(self.lib, self._h)

    def memory(self) -> dict:
        """Device bytes the session holds and the particle capacity of its state buffers."""
        return _memory(self.lib, self._h)

    def set_audit(self, enabled: bool) -> None:
        """FP64 per-substep audit of mass, momentum, kinetic energy, first moment and angular momentum."""
        _check(self.lib.tmpm_session_set_audit(self._h, 1 if enabled else 0), self.lib)

    def audit(self) -> np.ndarray:
        """Audit records of the substeps of the last synchronize, one row of AUDIT_FIELDS each."""
        return _read_audit(self.lib, self._h, 0)

    def step_adaptive(self, dt: float) -> None:
        """CFL-limited substeps (``MpmWorld::step_with_gpu_transfers``)."""
        _check(self.lib.tmpm_session_step_adaptive(self._h, float(dt)), self.lib)

    def has_pending_steps(self) -> bool:
        return self.pending_substeps() > 0

    def pending_substeps(self) -> int:
        return int(self.lib.tmpm_session_pending_substeps(self._h))

    @property
    def valid(self) -> bool:
        return bool(self.lib.tmpm_session_valid(self._h))

    @property
    def substeps(self) -> int:
        return int(self.lib.tmpm_session_substeps(self._h))

    def stable_timestep(self) -> float:
        return float(self.lib.tmpm_session_stable_timestep(self._h))

    # -- state ---------------------------------------------------------------
    def read_particles(self) -> ParticleArrays:
        n = int(self.lib.tmpm_session_particle_count(self._h))
        out = _read_particle_arrays(n, lambda cp: self.lib.tmpm_session_read_particles(self._h, cp), self.lib)
        out.materials = list(self._world.particles.materials)
        out.material_index = self._material_index(n)
        return out

    def _material_index(self, n: int) -> np.ndarray:
        idx = self._world.particles.material_index
        if len(idx) == n:
            return idx.copy()
        return np.zeros(n, dtype=np.uint32)

    def world(self) -> MpmWorld:
        """Last synchronized world (``world()``, gpu.rs:653-656)."""
        w = MpmWorld(self.read_particles(), self._world.params)
        w.obstacles = list(self._world.obstacles)
        w.use_cpic = self._world.use_cpic
        w.substeps = self.substeps
        return w

    # -- edits ----------------------------------------------------------------
    def set_particle_force(self, index: int, force) -> None:
        f = (C.c_double * 3)(*map(float, force))
        _check(self.lib.tmpm_session_set_particle_force(self._h, int(index), f), self.lib)

    def set_obstacles(self, obstacles: Sequence[RigidObstacle]) -> None:
        arr, keep = obstacles_to_c(obstacles)
        _check(self.lib.tmpm_session_set_obstacles(self._h, C.cast(arr, C.POINTER(CObstacle)), len(obstacles)),
               self.lib)
        self._world.obstacles = list(obstacles)

    def add_particles(self, particles) -> int:
        if not isinstance(particles, ParticleArrays):
            particles = ParticleArrays.from_particles(particles)
        cp, mats, keep = particles.c_view()
        chunk = C.c_uint64()
        _check(self.lib.tmpm_session_add_particles(self._h, C.byref(cp), C.cast(mats, C.POINTER(CMaterial)),
                                                   len(particles.materials), C.byref(chunk)), self.lib)
        added = particles.copy()
        added.chunk_id[:] = chunk.value
        self._world.particles = self._world.particles.concat(added)
        return int(chunk.value)

    def remove_chunk(self, chunk: int) -> int:
        removed = C.c_uint64()
        _check(self.lib.tmpm_session_remove_chunk(self._h, int(chunk), C.byref(removed)), self.lib)
        # keep the material table aligned with the remaining particles
        self._world.particles = self._world.particles.select(self._world.particles.chunk_id != int(chunk))
        return int(removed.value)

    def encode_particle_snapshot(self, device_ptr: int, nbytes: int, stream: int | None = None) -> int:
        """Writes 32-byte {position, mass; velocity, volume} records (gpu.rs:595-651)."""
        count = C.c_uint32()
        _check(self.lib.tmpm_session_encode_snapshot(self._h, C.c_void_p(device_ptr), int(nbytes),
                                                     C.c_void_p(stream) if stream else None, C.byref(count)),
               self.lib)
        return int(count.value)

    # -- reactions and rigid coupling ----------------------------------------------
    def set_reactions(self, enabled: bool) -> None:
        """Compute obstacle reactions (step_with_obstacle_reactions, world.rs:309-330)."""
        _check(self.lib.tmpm_session_set_reactions(self._h, 1 if enabled else 0), self.lib)

    def reaction_saturations(self) -> int:
        """Particle-side reaction impulses clamped to the int64 bins' range since creation (0: exact sums)."""
        return int(self.lib.tmpm_session_reaction_saturations(self._h))

    def reactions(self) -> np.ndarray:
        """(obstacles, 6): linear and angular impulse of the last synchronized substeps."""
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_session_read_reactions(self._h, a, c // 6, k), C.c_double,
                         6 * max(1, len(self._world.obstacles)))
        return np.array(arr[: 6 * n]).reshape(n, 6)

    def set_rigid_bodies(self, bodies, mappings, gravity=(0.0, -9.8, 0.0), max_linear_speed: float = 0.0,
                         integrate: bool = True) -> None:
        """Device-resident MPM-owned rigid coupling (gpu_rigid_sync.rs:187-213)."""
        cb, cm = _rigid_args(bodies, mappings)
        g = (C.c_double * 3)(*map(float, gravity))
        _check(self.lib.tmpm_session_set_rigid_bodies(self._h, cb, len(bodies), cm, len(mappings), g,
                                                      float(max_linear_speed), 1 if integrate else 0), self.lib)

    def rigid_bodies(self) -> list:
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_session_read_rigid_bodies(self._h, a, c, k), CRigidBody, 8)
        return [RigidBody.from_c(arr[i]) for i in range(n)]

    def obstacles(self) -> list:
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_session_read_obstacles(self._h, a, c, k), CObstacle,
                         len(self._world.obstacles))
        return [_obstacle_from_c(arr[i]) for i in range(n)]

    # -- diagnostics ------------------------------------------------------------
    def topology_stats(self) -> dict:
        st = CTopologyStats()
        _check(self.lib.tmpm_session_topology_stats(self._h, C.byref(st)), self.lib)
        return {k: int(getattr(st, k)) for k, _ in CTopologyStats._fields_}

    def export_nodes(self) -> np.ndarray:
        count = C.c_uint64()
        _check(self.lib.tmpm_session_export_nodes(self._h, None, 0, C.byref(count)), self.lib)
        keys = np.zeros((max(1, count.value), 4), dtype=np.int32)
        _check(self.lib.tmpm_session_export_nodes(self._h, keys.ctypes.data_as(C.POINTER(C.c_int32)), count.value,
                                                  C.byref(count)), self.lib)
        return keys[: count.value]

    def export_blocks(self) -> tuple[np.ndarray, np.ndarray]:
        n = int(self.lib.tmpm_session_particle_count(self._h))
        count = C.c_uint64()
        part = np.zeros(max(1, n), dtype=np.int64)
        _check(self.lib.tmpm_session_export_blocks(self._h, None, 0, part.ctypes.data_as(C.POINTER(C.c_int64)),
                                                   C.byref(count)), self.lib)
        keys = np.zeros((max(1, count.value), 4), dtype=np.int32)
        _check(self.lib.tmpm_session_export_blocks(self._h, keys.ctypes.data_as(C.POINTER(C.c_int32)), count.value,
                                                   part.ctypes.data_as(C.POINTER(C.c_int64)), C.byref(count)),
               self.lib)
        return keys[: count.value], part[:n]

    def set_timing(self, enabled: bool) -> None:
        _check(self.lib.tmpm_session_set_timing(self._h, 1 if enabled else 0), self.lib)

    def set_graphs(self, enabled: bool) -> None:
        """Replay each substep as a captured HIP graph (default off); results are identical."""
        _check(self.lib.tmpm_session_set_graphs(self._h, 1 if enabled else 0), self.lib)

    def autotune(self, trial: int = 32) -> dict:
        """Times the launch shapes over real substeps (which stay pending) and keeps the fastest."""
        ran = C.c_uint32()
        _check(self.lib.tmpm_session_autotune(self._h, int(trial), C.byref(ran)), self.lib)
        cfg = self.launch_config()
        cfg["substeps_run"] = int(ran.value)
        return cfg

    def launch_config(self) -> dict:
        out = (C.c_int32 * 5)()
        _check(self.lib.tmpm_session_launch_config(self._h, out), self.lib)
        return {"p2g_threads": out[0], "p2g_copies": out[1], "g2p_threads": out[2], "reorder_interval": out[3],
                "graphs": bool(out[4])}

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        cap = 64
        names = C.create_string_buffer(4096)
        avg = (C.c_double * cap)()
        calls = (C.c_uint64 * cap)()
        count = C.c_uint32()
        _check(self.lib.tmpm_session_kernel_times(self._h, names, 4096, avg, calls, cap, C.byref(count)), self.lib)
        keys = names.value.decode().split(",") if names.value else []
        return {k: (avg[i], int(calls[i])) for i, k in enumerate(keys)}


class _SceneHeader:
    """What a batch keeps of a scene's world on the host: everything but the
    particle state, which lives on the device (and in the library's host
    copy) and is read back with the scene."""

    __slots__ = ("params", "obstacles", "use_cpic", "materials", "material_index")

    def __init__(self, world: MpmWorld):
        self.params = world.params
        self.obstacles = list(world.obstacles)
        self.use_cpic = world.use_cpic
        self.materials = list(world.particles.materials)
        self.material_index = np.array(world.particles.material_index, dtype=np.uint32)


class HipMpmBatch:
    """Independent worlds advanced together, one kernel launch per stage per
    substep for all of them (the batch/serving boundary; the reference steps
    one ``GpuMpmResidentSession`` per world).  Each scene keeps the resident
    session's semantics; a failure after submission stops only its scene
    (``scene_error``), whose last synchronized world stays readable."""

    def __init__(self, transfers: HipMpmTransfers, worlds: Sequence[MpmWorld], substep_dts: Sequence[float],
                 p2g_mode: str = "fp32"):
        if len(worlds) != len(substep_dts):
            raise InvalidInput("one substep dt per world")
        self.lib = transfers.lib
        self.transfers = transfers
        descs = (CWorld * max(1, len(worlds)))()
        keep = []
        for k, w in enumerate(worlds):
            d, kk = w.c_desc()
            descs[k] = d
            keep.append(kk)
        dts = (C.c_double * max(1, len(worlds)))(*map(float, substep_dts))
        h = C.c_void_p()
        _check(self.lib.tmpm_batch_create(transfers.handle, descs, dts, len(worlds), P2G_MODES[p2g_mode],
                                          C.byref(h)), self.lib)
        del keep
        self._h = h
        self._worlds = {k: _SceneHeader(w) for k, w in enumerate(worlds)}
        self.p2g_mode = p2g_mode

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.tmpm_session_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- stepping (all scenes) ----------------------------------------------
    def submit_steps(self, steps: int) -> None:
        _check(self.lib.tmpm_session_submit_steps(self._h, int(steps)), self.lib)

    def synchronize(self) -> None:
        _check(self.lib.tmpm_session_synchronize(self._h), self.lib)

    def wait(self) -> None:
        _check(self.lib.tmpm_session_wait(self._h), self.lib)

    def step(self, steps: int) -> None:
        _check(self.lib.tmpm_session_step(self._h, int(steps)), self.lib)

    def synchronize_async(self) -> bool:
        """A poll: False while the submitted substeps still run, True once synchronized."""
        return _poll_synchronize(self.lib, self._h)

    def pending_substeps(self) -> int:
        return int(self.lib.tmpm_session_pending_substeps(self._h))

    @property
    def valid(self) -> bool:
        return bool(self.lib.tmpm_session_valid(self._h))

    def memory(self) -> dict:
        """Device bytes the batch holds and the particle capacity of its state buffers (kept when
        scenes retire, reused by later admissions)."""
        return _memory(self.lib, self._h)

    def set_audit(self, enabled: bool) -> None:
        _check(self.lib.tmpm_session_set_audit(self._h, 1 if enabled else 0), self.lib)

    def audit(self, scene: int) -> np.ndarray:
        """Audit records of one scene for the substeps of the last synchronize."""
        return _read_audit(self.lib, self._h, scene)

    # -- scenes -------------------------------------------------------------
    def scene_count(self) -> int:
        return int(self.lib.tmpm_batch_scene_count(self._h))

    def scenes(self) -> list[int]:
        """Live scene slots."""
        return sorted(self._worlds)

    def scene_error(self, scene: int):
        """None while the scene steps, else the GpuMpmError that stopped it."""
        st = int(self.lib.tmpm_batch_scene_status(self._h, int(scene)))
        if st == OK:
            return None
        return _ERRORS.get(st, GpuMpmError)(STATUS_NAMES.get(st, str(st)))

    def substeps(self, scene: int) -> int:
        return int(self.lib.tmpm_batch_scene_substeps(self._h, int(scene)))

    def read_particles(self, scene: int) -> ParticleArrays:
        n = int(self.lib.tmpm_batch_scene_particle_count(self._h, int(scene)))
        out = _read_particle_arrays(n, lambda cp: self.lib.tmpm_batch_read_particles(self._h, int(scene), cp),
                                    self.lib)
        src = self._worlds[int(scene)]
        out.materials = list(src.materials)
        out.material_index = src.material_index.copy() if len(src.material_index) == n else np.zeros(n, np.uint32)
        return out

    def world(self, scene: int) -> MpmWorld:
        base = self._worlds[int(scene)]
        w = MpmWorld(self.read_particles(scene), base.params)
        w.obstacles = list(base.obstacles)
        w.use_cpic = base.use_cpic
        w.substeps = self.substeps(scene)
        return w

    def set_particle_force(self, scene: int, index: int, force) -> None:
        f = (C.c_double * 3)(*map(float, force))
        _check(self.lib.tmpm_batch_set_particle_force(self._h, int(scene), int(index), f), self.lib)

    def set_obstacles(self, scene: int, obstacles: Sequence[RigidObstacle]) -> None:
        arr, keep = obstacles_to_c(obstacles)
        _check(self.lib.tmpm_batch_set_obstacles(self._h, int(scene), C.cast(arr, C.POINTER(CObstacle)),
                                                 len(obstacles)), self.lib)
        self._worlds[int(scene)].obstacles = list(obstacles)

    def add_scene(self, world: MpmWorld, substep_dt: float) -> int:
        desc, keep = world.c_desc()
        slot = C.c_uint32()
        _check(self.lib.tmpm_batch_add_scene(self._h, C.byref(desc), float(substep_dt), C.byref(slot)), self.lib)
        del keep
        self._worlds[int(slot.value)] = _SceneHeader(world)
        return int(slot.value)

    def add_staged(self, staged: "StagedScene") -> int:
        """Admits a scene staged by :class:`StagedScene` (needs no pending substeps); consumes it."""
        h = staged.take()
        slot = C.c_uint32()
        _check(self.lib.tmpm_batch_add_staged(self._h, h, C.byref(slot)), self.lib)
        self._worlds[int(slot.value)] = _SceneHeader(staged.world)
        return int(slot.value)

    def set_reactions(self, enabled: bool) -> None:
        _check(self.lib.tmpm_session_set_reactions(self._h, 1 if enabled else 0), self.lib)

    def reactions(self, scene: int) -> np.ndarray:
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_batch_read_reactions(self._h, int(scene), a, c // 6, k),
                         C.c_double, 6 * max(1, len(self._worlds[int(scene)].obstacles)))
        return np.array(arr[: 6 * n]).reshape(n, 6)

    def set_rigid_bodies(self, scene: int, bodies, mappings, gravity=(0.0, -9.8, 0.0),
                         max_linear_speed: float = 0.0, integrate: bool = True) -> None:
        cb, cm = _rigid_args(bodies, mappings)
        g = (C.c_double * 3)(*map(float, gravity))
        _check(self.lib.tmpm_batch_set_rigid_bodies(self._h, int(scene), cb, len(bodies), cm, len(mappings), g,
                                                    float(max_linear_speed), 1 if integrate else 0), self.lib)

    def rigid_bodies(self, scene: int) -> list:
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_batch_read_rigid_bodies(self._h, int(scene), a, c, k),
                         CRigidBody, 8)
        return [RigidBody.from_c(arr[i]) for i in range(n)]

    def obstacles(self, scene: int) -> list:
        arr, n = _read_n(lambda a, c, k: self.lib.tmpm_batch_read_obstacles(self._h, int(scene), a, c, k), CObstacle,
                         len(self._worlds[int(scene)].obstacles))
        return [_obstacle_from_c(arr[i]) for i in range(n)]

    def remove_scene(self, scene: int) -> None:
        _check(self.lib.tmpm_batch_remove_scene(self._h, int(scene)), self.lib)
        self._worlds.pop(int(scene), None)

    # -- diagnostics ----------------------------------------------------------
    def set_timing(self, enabled: bool) -> None:
        _check(self.lib.tmpm_session_set_timing(self._h, 1 if enabled else 0), self.lib)

    def set_graphs(self, enabled: bool) -> None:
        """Replay each substep as a captured HIP graph (default off); results are identical."""
        _check(self.lib.tmpm_session_set_graphs(self._h, 1 if enabled else 0), self.lib)

    def autotune(self, trial: int = 32) -> dict:
        """Times the launch shapes over real substeps (which stay pending) and keeps the fastest."""
        ran = C.c_uint32()
        _check(self.lib.tmpm_session_autotune(self._h, int(trial), C.byref(ran)), self.lib)
        cfg = self.launch_config()
        cfg["substeps_run"] = int(ran.value)
        return cfg

    def launch_config(self) -> dict:
        out = (C.c_int32 * 5)()
        _check(self.lib.tmpm_session_launch_config(self._h, out), self.lib)
        return {"p2g_threads": out[0], "p2g_copies": out[1], "g2p_threads": out[2], "reorder_interval": out[3],
                "graphs": bool(out[4])}

    def kernel_times(self) -> dict[str, tuple[float, int]]:
        return HipMpmResidentSession.kernel_times(self)  # same session-level entry point


AUDIT_FIELDS = ("mass", "px", "py", "pz", "kinetic", "mx", "my", "mz", "lx", "ly", "lz", "count")


def _memory(lib, h) -> dict:
    nbytes, cap = C.c_uint64(), C.c_uint64()
    _check(lib.tmpm_session_memory(h, C.byref(nbytes), C.byref(cap)), lib)
    return {"device_bytes": int(nbytes.value), "particle_capacity": int(cap.value)}


def _poll_synchronize(lib, h) -> bool:
    done = C.c_int32()
    _check(lib.tmpm_session_synchronize_async(h, C.byref(done)), lib)
    return bool(done.value)


def _read_audit(lib, h, scene: int) -> np.ndarray:
    count = C.c_uint64()
    _check(lib.tmpm_session_read_audit(h, int(scene), None, 0, C.byref(count)), lib)
    n = int(count.value)
    buf = (C.c_double * max(1, n * len(AUDIT_FIELDS)))()
    _check(lib.tmpm_session_read_audit(h, int(scene), buf, n, C.byref(count)), lib)
    return np.frombuffer(buf, dtype=np.float64, count=n * len(AUDIT_FIELDS)).reshape(n, len(AUDIT_FIELDS)).copy()


class StagedScene:
    """The host half of a batch admission (validation, f64 copy, f32 packing, block count),
    done without a session or the device, so it can run while a batch has substeps in flight;
    :meth:`HipMpmBatch.add_staged` admits it."""

    def __init__(self, world: MpmWorld, substep_dt: float):
        self.lib = load_library()
        self.world = world
        self.dt = float(substep_dt)
        desc, keep = world.c_desc()
        h = C.c_void_p()
        _check(self.lib.tmpm_batch_stage_scene(C.byref(desc), self.dt, C.byref(h)), self.lib)
        del keep
        self._h = h

    def take(self):
        if not getattr(self, "_h", None):
            raise InvalidInput("staged scene already admitted or released")
        h, self._h = self._h, None
        return h

    def close(self) -> None:
        if getattr(self, "_h", None):
            self.lib.tmpm_staged_scene_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def tune_cache_clear() -> None:
    """Forgets every tuned launch shape of this process (the next autotune runs its trials)."""
    load_library().tmpm_tune_cache_clear()


def tune_cache_size() -> int:
    return int(load_library().tmpm_tune_cache_size())


def bandwidth_probe(device: int = 0, nbytes: int = 1 << 30, iterations: int = 20) -> float:
    """Achieved copy bandwidth (read + write GB/s) of a plain float4 copy."""
    lib = load_library()
    out = C.c_double()
    _check(lib.tmpm_bandwidth_probe(int(device), int(nbytes), int(iterations), C.byref(out)), lib)
    return float(out.value)
