"""TEST INFRASTRUCTURE ONLY — ctypes wrapper of the FP64 oracle (liboracle.so).

The oracle restates the reference CPU pipeline of tessera-mpm
(crates/tessera-mpm/src/world.rs, material.rs, obstacle.rs) in C++
(oracle/mpm_oracle.cpp).  Only tests, the smoke check and the benchmark's CPU
baseline may use it, always as the checker, never as the thing measured or
shipped.  Parity pinning of every function is described in the header of
mpm_oracle.cpp.
"""

from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from tessera_amd.mpm import CObstacle, MpmParams, MpmWorld, ParticleArrays, obstacles_to_c

ORACLE_PATH = Path(__file__).resolve().parent / "liboracle.so"
_dp = C.POINTER(C.c_double)
_LIB = None


class OracleError(Exception):
    def __init__(self, status: int):
        names = {1: "InvalidInput", 6: "ExcessiveSubsteps", 7: "NonFiniteState", 8: "UnknownChunk",
                 9: "ParticleCapacity"}
        super().__init__(names.get(status, str(status)))
        self.status = status
        self.name = names.get(status, str(status))


class OParams(C.Structure):
    _fields_ = [
        ("cell_width", C.c_double),
        ("gravity", C.c_double * 3),
        ("max_substep", C.c_double),
        ("has_bounds", C.c_int32),
        ("use_cpic", C.c_int32),
        ("bound_min", C.c_double * 3),
        ("bound_max", C.c_double * 3),
    ]


def lib() -> C.CDLL:
    global _LIB
    if _LIB is not None:
        return _LIB
    if not ORACLE_PATH.exists():
        raise RuntimeError(f"{ORACLE_PATH} missing; run oracle/build_oracle.sh")
    L = C.CDLL(str(ORACLE_PATH))
    vp = C.c_void_p
    u8p, u16p, i32p = C.POINTER(C.c_uint8), C.POINTER(C.c_uint16), C.POINTER(C.c_int32)
    L.orc_world_new.restype = vp
    L.orc_world_new.argtypes = [C.POINTER(OParams)]
    L.orc_world_free.argtypes = [vp]
    L.orc_world_add.restype = C.c_int
    L.orc_world_add.argtypes = [vp, C.c_uint64, C.c_int32, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, _dp, i32p, _dp, _dp,
                                u16p, u8p, u8p, C.POINTER(C.c_uint64)]
    L.orc_world_remove_chunk.restype = C.c_int
    L.orc_world_remove_chunk.argtypes = [vp, C.c_uint64, C.POINTER(C.c_uint64)]
    L.orc_world_set_obstacles.restype = C.c_int
    L.orc_world_set_obstacles.argtypes = [vp, C.c_uint64, C.POINTER(CObstacle)]
    L.orc_world_set_use_cpic.argtypes = [vp, C.c_int32]
    L.orc_world_step.restype = C.c_int
    L.orc_world_step.argtypes = [vp, C.c_double]
    L.orc_world_step_with_reactions.restype = C.c_int
    L.orc_world_step_with_reactions.argtypes = [vp, C.c_double, _dp]
    L.orc_world_substeps_fixed.restype = C.c_int
    L.orc_world_substeps_fixed.argtypes = [vp, C.c_double, C.c_uint32]
    L.orc_world_substeps_fixed_threads.restype = C.c_int
    L.orc_world_substeps_fixed_threads.argtypes = [vp, C.c_double, C.c_uint32, C.c_int32]
    L.orc_world_fp32_substeps.restype = C.c_int
    L.orc_world_fp32_substeps.argtypes = [vp, C.c_double, C.c_uint32, _dp, _dp, _dp, _dp]
    L.orc_world_stable_timestep.restype = C.c_double
    L.orc_world_stable_timestep.argtypes = [vp]
    L.orc_world_count.restype = C.c_uint64
    L.orc_world_count.argtypes = [vp]
    L.orc_world_substeps.restype = C.c_uint64
    L.orc_world_substeps.argtypes = [vp]
    L.orc_world_read.argtypes = [vp, _dp, _dp, _dp, _dp, _dp, _dp, C.POINTER(C.c_uint64)]
    L.orc_world_write.argtypes = [vp, _dp, _dp, _dp, _dp, _dp]
    L.orc_world_p2g_totals.restype = C.c_uint64
    L.orc_world_p2g_totals.argtypes = [vp, C.c_double, _dp, _dp]
    L.orc_kirchhoff_stress.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, _dp]
    L.orc_project_deformation.argtypes = [C.c_int32, _dp, _dp, _dp, _dp, _dp]
    L.orc_timestep_bound.restype = C.c_double
    L.orc_timestep_bound.argtypes = [C.c_int32, _dp, C.c_double, _dp, C.c_double, C.c_double]
    L.orc_material_valid.restype = C.c_int32
    L.orc_material_valid.argtypes = [C.c_int32, _dp]
    L.orc_svd.argtypes = [_dp, _dp, _dp, _dp]
    L.orc_determinant.restype = C.c_double
    L.orc_determinant.argtypes = [_dp]
    L.orc_obstacle_surface.argtypes = [C.POINTER(CObstacle), _dp, _dp, _dp]
    L.orc_obstacle_contact_velocity.argtypes = [C.POINTER(CObstacle), _dp, _dp, _dp, _dp]
    L.orc_obstacle_may_contact.restype = C.c_int32
    L.orc_obstacle_may_contact.argtypes = [C.POINTER(CObstacle), _dp, C.c_double]
    L.orc_obstacle_valid.restype = C.c_int32
    L.orc_obstacle_valid.argtypes = [C.POINTER(CObstacle)]
    L.orc_cell_f32.restype = C.c_int32
    L.orc_cell_f32.argtypes = [C.POINTER(C.c_float), C.c_float, i32p]
    L.orc_node_set.restype = C.c_uint64
    L.orc_node_set.argtypes = [C.c_uint64, C.POINTER(C.c_float), u8p, C.POINTER(C.c_uint32), C.c_float, i32p,
                               C.c_uint64]
    L.orc_block_partition.restype = C.c_uint64
    L.orc_block_partition.argtypes = [C.c_uint64, C.POINTER(C.c_float), u8p, C.POINTER(C.c_uint32), C.c_float, i32p,
                                      C.c_uint64, C.POINTER(C.c_int64)]
    _LIB = L
    return L


def _ptr(a: np.ndarray, ctype=C.c_double):
    return a.ctypes.data_as(C.POINTER(ctype))


def _cm(m: np.ndarray) -> np.ndarray:
    """(n,3,3) row/col -> (n,9) nalgebra column-major."""
    return np.ascontiguousarray(np.transpose(np.asarray(m, dtype=np.float64).reshape(-1, 3, 3), (0, 2, 1)).reshape(-1, 9))


def _from_cm(flat: np.ndarray) -> np.ndarray:
    return flat.reshape(-1, 3, 3).transpose(0, 2, 1).copy()


class OracleWorld:
    """FP64 CPU world; the same API shape as the reference's ``MpmWorld``."""

    def __init__(self, world: MpmWorld):
        L = lib()
        p = world.params
        op = OParams()
        op.cell_width = p.cell_width
        for i in range(3):
            op.gravity[i] = float(p.gravity[i])
        op.max_substep = p.max_substep
        op.has_bounds = 1 if p.bounds is not None else 0
        op.use_cpic = 1 if world.use_cpic else 0
        if p.bounds is not None:
            for i in range(3):
                op.bound_min[i] = float(p.bounds.min[i])
                op.bound_max[i] = float(p.bounds.max[i])
        h = L.orc_world_new(C.byref(op))
        if not h:
            raise OracleError(1)
        self._h = h
        self.params = p
        self.materials = list(world.particles.materials)
        self.material_index = world.particles.material_index.copy()
        self.meta = world.particles.copy()
        self._add(world.particles, initial=True)
        self.set_obstacles(world.obstacles)
        self._substeps0 = world.substeps

    def __del__(self):
        try:
            lib().orc_world_free(self._h)
        except Exception:
            pass

    def _add(self, pa: ParticleArrays, initial: bool) -> int:
        L = lib()
        n = len(pa)
        kinds = np.array([pa.materials[int(i)].kind for i in pa.material_index], dtype=np.int32)
        mp = np.array([pa.materials[int(i)].params for i in pa.material_index], dtype=np.float64).reshape(n, 5)
        arrays = dict(
            pos=np.ascontiguousarray(pa.position, dtype=np.float64), vel=np.ascontiguousarray(pa.velocity, np.float64),
            F=_cm(pa.deformation), Cm=_cm(pa.affine), force=np.ascontiguousarray(pa.force, np.float64),
            mass=np.ascontiguousarray(pa.mass, np.float64), vol=np.ascontiguousarray(pa.rest_volume, np.float64),
            rad=np.ascontiguousarray(pa.radius, np.float64), damp=np.ascontiguousarray(pa.damping, np.float64),
            plastic=np.ascontiguousarray(pa.plastic, np.float64), color=np.ascontiguousarray(pa.transfer_color, np.uint16),
            en=np.ascontiguousarray(pa.enabled, np.uint8), fx=np.ascontiguousarray(pa.fixed, np.uint8))
        chunk = C.c_uint64()
        st = L.orc_world_add(self._h, n, 1 if initial else 0, _ptr(arrays["pos"]), _ptr(arrays["vel"]),
                             _ptr(arrays["F"]), _ptr(arrays["Cm"]), _ptr(arrays["force"]), _ptr(arrays["mass"]),
                             _ptr(arrays["vol"]), _ptr(arrays["rad"]), _ptr(arrays["damp"]), _ptr(kinds, C.c_int32),
                             _ptr(mp), _ptr(arrays["plastic"]), _ptr(arrays["color"], C.c_uint16),
                             _ptr(arrays["en"], C.c_uint8), _ptr(arrays["fx"], C.c_uint8), C.byref(chunk))
        if st:
            raise OracleError(st)
        return int(chunk.value)

    def add_particles(self, pa: ParticleArrays) -> int:
        chunk = self._add(pa, initial=False)
        self.meta = self.meta.concat(pa)
        return chunk

    def remove_chunk(self, chunk: int) -> int:
        removed = C.c_uint64()
        st = lib().orc_world_remove_chunk(self._h, int(chunk), C.byref(removed))
        if st:
            raise OracleError(st)
        return int(removed.value)

    def set_obstacles(self, obstacles) -> None:
        arr, keep = obstacles_to_c(obstacles)
        st = lib().orc_world_set_obstacles(self._h, len(obstacles), C.cast(arr, C.POINTER(CObstacle)))
        if st:
            raise OracleError(st)

    def set_use_cpic(self, on: bool) -> None:
        lib().orc_world_set_use_cpic(self._h, 1 if on else 0)

    def step(self, dt: float) -> None:
        st = lib().orc_world_step(self._h, float(dt))
        if st:
            raise OracleError(st)

    def step_with_obstacle_reactions(self, dt: float, n_obstacles: int) -> np.ndarray:
        out = np.zeros((max(1, n_obstacles), 6))
        st = lib().orc_world_step_with_reactions(self._h, float(dt), _ptr(out))
        if st:
            raise OracleError(st)
        return out[:n_obstacles]

    def substeps_fixed(self, dt: float, steps: int) -> None:
        st = lib().orc_world_substeps_fixed(self._h, float(dt), int(steps))
        if st:
            raise OracleError(st)

    def substeps_fixed_threads(self, dt: float, steps: int, threads: int) -> None:
        """The same substeps on `threads` host threads (per-chunk P2G sums merged in
        thread order; deterministic per thread count, no reactions)."""
        st = lib().orc_world_substeps_fixed_threads(self._h, float(dt), int(steps), int(threads))
        if st:
            raise OracleError(st)

    def fp32_substeps(self, dt: float, steps: int) -> dict:
        """The FP32 restatement of the reference's resident GPU substep (gpu.wgsl:428-1019):
        `steps` substeps of dt from this world's current state, which stays unchanged.  The
        error floor f32 arithmetic leaves against the FP64 substeps."""
        n = len(self)
        pos, vel, F, Cm = np.zeros((n, 3)), np.zeros((n, 3)), np.zeros((n, 9)), np.zeros((n, 9))
        st = lib().orc_world_fp32_substeps(self._h, float(dt), int(steps), _ptr(pos), _ptr(vel), _ptr(F), _ptr(Cm))
        if st:
            raise OracleError(st)
        return dict(position=pos, velocity=vel, deformation=_from_cm(F), affine=_from_cm(Cm))

    def stable_timestep(self) -> float:
        return float(lib().orc_world_stable_timestep(self._h))

    @property
    def substeps(self) -> int:
        return int(lib().orc_world_substeps(self._h)) + self._substeps0

    def __len__(self) -> int:
        return int(lib().orc_world_count(self._h))

    def read(self) -> dict:
        n = len(self)
        pos, vel, force, plastic = (np.zeros((n, 3)) for _ in range(4))
        F, Cm = np.zeros((n, 9)), np.zeros((n, 9))
        chunk = np.zeros(n, dtype=np.uint64)
        lib().orc_world_read(self._h, _ptr(pos), _ptr(vel), _ptr(F), _ptr(Cm), _ptr(force), _ptr(plastic),
                             _ptr(chunk, C.c_uint64))
        return dict(position=pos, velocity=vel, deformation=_from_cm(F), affine=_from_cm(Cm), force=force,
                    plastic=plastic, chunk_id=chunk)

    def write(self, position, velocity, deformation, affine, force) -> None:
        lib().orc_world_write(self._h, _ptr(np.ascontiguousarray(position, np.float64)),
                              _ptr(np.ascontiguousarray(velocity, np.float64)), _ptr(_cm(deformation)),
                              _ptr(_cm(affine)), _ptr(np.ascontiguousarray(force, np.float64)))

    def p2g_totals(self, dt: float) -> tuple[int, float, np.ndarray]:
        m = C.c_double()
        p = np.zeros(3)
        nodes = lib().orc_world_p2g_totals(self._h, float(dt), C.byref(m), _ptr(p))
        return int(nodes), float(m.value), p


# --- material / obstacle / SVD entry points ----------------------------------------
def kirchhoff_stress(material, F, Cmat=None, plastic=(1.0, 0.0, 0.0)) -> np.ndarray:
    out = np.zeros(9)
    mp = np.array(material.params, dtype=np.float64)
    Cmat = np.zeros((3, 3)) if Cmat is None else Cmat
    lib().orc_kirchhoff_stress(material.kind, _ptr(mp), _ptr(_cm(F)), _ptr(_cm(Cmat)),
                               _ptr(np.array(plastic, dtype=np.float64)), _ptr(out))
    return _from_cm(out)[0]


def project_deformation(material, F, plastic=(1.0, 0.0, 0.0)):
    Fo = np.zeros(9)
    po = np.zeros(3)
    mp = np.array(material.params, dtype=np.float64)
    lib().orc_project_deformation(material.kind, _ptr(mp), _ptr(_cm(F)), _ptr(np.array(plastic, dtype=np.float64)),
                                  _ptr(Fo), _ptr(po))
    return _from_cm(Fo)[0], po


def timestep_bound(material, density, velocity, cell_width, det=1.0) -> float:
    mp = np.array(material.params, dtype=np.float64)
    return float(lib().orc_timestep_bound(material.kind, _ptr(mp), float(density),
                                          _ptr(np.array(velocity, dtype=np.float64)), float(cell_width), float(det)))


def material_valid(material) -> bool:
    return bool(lib().orc_material_valid(material.kind, _ptr(np.array(material.params, dtype=np.float64))))


def svd(F):
    U, V = np.zeros(9), np.zeros(9)
    s = np.zeros(3)
    lib().orc_svd(_ptr(_cm(F)), _ptr(U), _ptr(s), _ptr(V))
    return _from_cm(U)[0], s, _from_cm(V)[0]


def determinant(F) -> float:
    return float(lib().orc_determinant(_ptr(_cm(F))))


def _obs(o):
    keep = []
    c = o.to_c(keep)
    return c, keep


def obstacle_surface(obstacle, point):
    c, keep = _obs(obstacle)
    d = C.c_double()
    n = np.zeros(3)
    lib().orc_obstacle_surface(C.byref(c), _ptr(np.array(point, dtype=np.float64)), C.byref(d), _ptr(n))
    return float(d.value), n


def contact_velocity(obstacle, point, velocity, normal) -> np.ndarray:
    c, keep = _obs(obstacle)
    out = np.zeros(3)
    lib().orc_obstacle_contact_velocity(C.byref(c), _ptr(np.array(point, dtype=np.float64)),
                                        _ptr(np.array(velocity, dtype=np.float64)),
                                        _ptr(np.array(normal, dtype=np.float64)), _ptr(out))
    return out


def may_contact(obstacle, point, margin) -> bool:
    c, keep = _obs(obstacle)
    return bool(lib().orc_obstacle_may_contact(C.byref(c), _ptr(np.array(point, dtype=np.float64)), float(margin)))


def obstacle_valid(obstacle) -> bool:
    c, keep = _obs(obstacle)
    return bool(lib().orc_obstacle_valid(C.byref(c)))


# --- binning (f32 rule) -------------------------------------------------------------------
def cell_f32(x, h: float):
    xs = np.asarray(x, dtype=np.float32).reshape(3)
    out = np.zeros(3, dtype=np.int32)
    ok = lib().orc_cell_f32(_ptr(xs, C.c_float), np.float32(h), _ptr(out, C.c_int32))
    return (out if ok else None)


def _bin_inputs(positions, active, region):
    pos = np.ascontiguousarray(np.asarray(positions, dtype=np.float64).astype(np.float32).reshape(-1, 3))
    n = len(pos)
    act = np.ascontiguousarray(np.ones(n, np.uint8) if active is None else np.asarray(active, np.uint8))
    reg = np.ascontiguousarray(np.zeros(n, np.uint32) if region is None else np.asarray(region, np.uint32))
    return pos, act, reg


def node_set(positions, h: float, active=None, region=None) -> np.ndarray:
    """Sorted live node keys (x, y, z, region) of the 27-node stencils."""
    pos, act, reg = _bin_inputs(positions, active, region)
    n = len(pos)
    cnt = lib().orc_node_set(n, _ptr(pos, C.c_float), _ptr(act, C.c_uint8), _ptr(reg, C.c_uint32), np.float32(h),
                             None, 0)
    if cnt == 2**64 - 1:
        raise OracleError(1)
    out = np.zeros((max(1, cnt), 4), dtype=np.int32)
    lib().orc_node_set(n, _ptr(pos, C.c_float), _ptr(act, C.c_uint8), _ptr(reg, C.c_uint32), np.float32(h),
                       _ptr(out, C.c_int32), cnt)
    return out[:cnt]


def block_partition(positions, h: float, active=None, region=None):
    """(sorted unique block keys, per-particle block index or -1)."""
    pos, act, reg = _bin_inputs(positions, active, region)
    n = len(pos)
    part = np.zeros(max(1, n), dtype=np.int64)
    cnt = lib().orc_block_partition(n, _ptr(pos, C.c_float), _ptr(act, C.c_uint8), _ptr(reg, C.c_uint32),
                                    np.float32(h), None, 0, _ptr(part, C.c_int64))
    if cnt == 2**64 - 1:
        raise OracleError(1)
    out = np.zeros((max(1, cnt), 4), dtype=np.int32)
    lib().orc_block_partition(n, _ptr(pos, C.c_float), _ptr(act, C.c_uint8), _ptr(reg, C.c_uint32), np.float32(h),
                              _ptr(out, C.c_int32), cnt, _ptr(part, C.c_int64))
    return out[:cnt], part[:n]
