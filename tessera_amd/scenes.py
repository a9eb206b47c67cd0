"""Seeded synthetic scenes for the benchmark and the tests.

Every generator is deterministic (numpy PCG64 via ``default_rng(seed)``).
Units are SI and gravity points along -y.  Particle volume and mass follow
``MpmParticle::new`` (crates/tessera-mpm/src/world.rs:120-150): rest volume
(2r)^3 and mass density * volume.

* :func:`uniform_cube`       8,192 particles uniform in [0.4, 0.6]^3, elastic
                             E=1e4, nu=0.3, rho=1e3, h=1/64, dt=1e-4, sticky
                             ground at y=0.05 (200 steps).
* :func:`four_cubes`         1,048,576 particles: four jittered 64^3 lattices of
                             edge 0.2, velocities N(0, 0.5) per component,
                             elastic, h=1/256 (4^3-cell blocks over a 256^3
                             domain), dt=5e-5 (1000 steps).
* :func:`dam_break`          4,194,304 fluid particles (K=1e5, gamma=7) filling
                             [0,0.4]x[0,0.6]x[0,1] at 2 per cell; h=0.0048559
                             so that N, the particles per cell and a 512^3 cell
                             domain agree; dt=2e-5 (2000 steps).
* :func:`serving_requests`   a Poisson stream of requests: N log-uniform in
                             [4096, 524288] at 8 particles per cell (h the
                             largest of 1/64, 1/96, 1/128 that keeps the cube
                             within 0.4 m) on a sticky floor, elastic or
                             fluid, 1-4 falling rigid boxes (edge U[0.05,
                             0.15], density U[500, 5000]) driven by the MPM's
                             reactions, horizons uniform in {50..2000}
                             substeps of dt=5e-5.
* :func:`batch_scenes`       256 independent scenes, N_i log-uniform in
                             [2048, 65536] (sum ~4.7M), h_i in {1/64 .. 1/128},
                             a jittered cube of 8 particles per cell with
                             N(0, 0.3) velocities, elastic (corotated or
                             neo-Hookean) or fluid, dt=5e-5 each.
"""

from __future__ import annotations

import hashlib
import math
from dataclasses import dataclass

import numpy as np

from .mpm import (MaterialModel, MpmParams, MpmWorld, ObstacleBoundary, ParticleArrays, RigidObstacle, WorldBounds,
                  quat_from_axis_angle)

GRAVITY_Y = np.array([0.0, -9.8, 0.0])


@dataclass
class Scene:
    name: str
    world: MpmWorld
    dt: float
    steps: int


def _particles(pos: np.ndarray, vel: np.ndarray, radius: float, density: float, material: MaterialModel) -> ParticleArrays:
    n = len(pos)
    pa = ParticleArrays(n, [material])
    pa.position[:] = pos
    pa.velocity[:] = vel
    vol = (2.0 * radius) ** 3
    pa.rest_volume[:] = vol
    pa.mass[:] = density * vol
    pa.radius[:] = radius
    return pa


def sticky_floor(y: float = 0.05) -> RigidObstacle:
    """Finite ground with local +z mapped to world +y, Stick response
    (obstacle.rs:185-191, 517-519)."""
    q = quat_from_axis_angle([1.0, 0.0, 0.0], -math.pi / 2)
    g = RigidObstacle.ground([0.5, y, 0.5], [1.0, 1.0], q)
    g.boundary = ObstacleBoundary.STICK
    return g


def uniform_cube(n: int = 8192, seed: int = 0) -> Scene:
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0.4, 0.6, size=(n, 3))
    vol = 0.008 / n
    radius = 0.5 * vol ** (1.0 / 3.0)
    pa = _particles(pos, np.zeros((n, 3)), radius, 1.0e3, MaterialModel.elastic(1.0e4, 0.3))
    params = MpmParams(cell_width=1.0 / 64.0, gravity=GRAVITY_Y.copy(), max_substep=1.0e-4,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    w.set_obstacles([sticky_floor(0.05)])
    return Scene("uniform_cube_8k", w, 1.0e-4, 200)


def jittered_lattice(rng, counts, lo, hi) -> np.ndarray:
    counts = np.asarray(counts)
    lo = np.asarray(lo, dtype=np.float64)
    hi = np.asarray(hi, dtype=np.float64)
    spacing = (hi - lo) / counts
    idx = np.stack(np.meshgrid(*[np.arange(c) for c in counts], indexing="ij"), axis=-1).reshape(-1, 3)
    jitter = rng.uniform(-0.5, 0.5, size=idx.shape)
    return lo + (idx + 0.5 + jitter) * spacing


FOUR_CUBE_CENTERS = np.array([[0.3, 0.3, 0.3], [0.7, 0.3, 0.7], [0.3, 0.7, 0.7], [0.7, 0.7, 0.3]])


def four_cubes(per_axis: int = 64, seed: int = 0, edge: float = 0.2) -> Scene:
    rng = np.random.default_rng(seed)
    parts = []
    for c in FOUR_CUBE_CENTERS:
        parts.append(jittered_lattice(rng, [per_axis] * 3, c - edge / 2, c + edge / 2))
    pos = np.concatenate(parts)
    n = len(pos)
    vel = rng.normal(0.0, 0.5, size=(n, 3))
    spacing = edge / per_axis
    pa = _particles(pos, vel, spacing / 2, 1.0e3, MaterialModel.elastic(1.0e4, 0.3))
    params = MpmParams(cell_width=1.0 / 256.0, gravity=GRAVITY_Y.copy(), max_substep=5.0e-5,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    return Scene(f"four_cubes_{n}", MpmWorld(pa, params), 5.0e-5, 1000)


DAM_CELL = 0.0048559


def dam_break(counts=(128, 128, 256), seed: int = 0) -> Scene:
    rng = np.random.default_rng(seed)
    extent = np.array([0.4, 0.6, 1.0])
    pos = jittered_lattice(rng, counts, np.zeros(3), extent)
    n = len(pos)
    vol = float(np.prod(extent)) / n
    radius = 0.5 * vol ** (1.0 / 3.0)
    pa = _particles(pos, np.zeros((n, 3)), radius, 1.0e3, MaterialModel.fluid(1.0e5, 7.0, 0.0))
    h = DAM_CELL
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y.copy(), max_substep=2.0e-5,
                       bounds=WorldBounds(np.zeros(3), np.full(3, 512 * h)))
    return Scene(f"dam_break_{n}", MpmWorld(pa, params), 2.0e-5, 2000)


BATCH_CELLS = (1.0 / 64, 1.0 / 80, 1.0 / 96, 1.0 / 112, 1.0 / 128)


def batch_scene(rng, n_target: int, index: int) -> Scene:
    """One scene of a batch: a jittered cube of ~n_target particles at 8 per cell."""
    h = BATCH_CELLS[int(rng.integers(len(BATCH_CELLS)))]
    per_axis = max(2, int(round((n_target) ** (1.0 / 3.0))))
    spacing = h / 2.0
    edge = per_axis * spacing
    lo = rng.uniform(0.2, 0.8 - edge, size=3)
    pos = jittered_lattice(rng, [per_axis] * 3, lo, lo + edge)
    n = len(pos)
    vel = rng.normal(0.0, 0.3, size=(n, 3))
    kind = int(rng.integers(3))
    if kind == 0:
        material = MaterialModel.elastic(float(rng.uniform(5e3, 5e4)), 0.3)
    elif kind == 1:
        material = MaterialModel.neo_hookean(float(rng.uniform(5e3, 5e4)), 0.3)
    else:
        material = MaterialModel.fluid(1.0e5, 7.0, 0.0)
    pa = _particles(pos, vel, spacing / 2, 1.0e3, material)
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y.copy(), max_substep=5.0e-5,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    return Scene(f"batch_{index}_{n}", MpmWorld(pa, params), 5.0e-5, 100)


def batch_scenes(count: int = 256, seed: int = 0, n_min: int = 2048, n_max: int = 65536) -> list[Scene]:
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=count))
    return [batch_scene(rng, int(n), k) for k, n in enumerate(sizes)]


def serving_request(rng, n_target: int, name: str, horizon: int, arrival: float, max_boxes: int = 4):
    """One serving request: particle cube on a sticky floor with falling rigid boxes."""
    from .mpm import RigidBody, RigidMapping
    from .serving import Request

    cells = (n_target / 8.0) ** (1.0 / 3.0)
    h = next((c for c in (1 / 64, 1 / 96, 1 / 128) if cells * c <= 0.4), 1 / 128)
    per_axis = max(2, int(round(n_target ** (1.0 / 3.0))))
    spacing = h / 2
    edge = per_axis * spacing
    lo = np.array([0.5 - edge / 2, 0.06, 0.5 - edge / 2])
    pos = jittered_lattice(rng, [per_axis] * 3, lo, lo + edge)
    vel = rng.normal(0.0, 0.1, size=pos.shape)
    material = MaterialModel.elastic(float(rng.uniform(1e4, 5e4)), 0.3) if rng.integers(2) == 0 else \
        MaterialModel.fluid(1.0e5, 7.0, 0.0)
    pa = _particles(pos, vel, spacing / 2, 1.0e3, material)
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y.copy(), max_substep=5.0e-5,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    obstacles = [sticky_floor(0.05)]
    bodies, maps = [], [RigidMapping(None)]
    top = float(pos[:, 1].max())
    for b in range(int(rng.integers(1, max_boxes + 1))):
        half = 0.5 * float(rng.uniform(0.05, 0.15))
        center = [0.5 + rng.uniform(-0.3, 0.3) * edge, top + half + 0.5 * h + 0.02 * b,
                  0.5 + rng.uniform(-0.3, 0.3) * edge]
        body = RigidBody.box(center, [half, half, half], float(rng.uniform(500.0, 5000.0)))
        body.linear_velocity = np.array([0.0, -0.5, 0.0])
        bodies.append(body)
        obstacles.append(RigidObstacle.cuboid(center, [half, half, half]))
        maps.append(RigidMapping(len(bodies) - 1))
    w.set_obstacles(obstacles)
    return Request(name, w, 5.0e-5, int(horizon), float(arrival), bodies, maps)


def serving_requests(count: int, rate: float, seed: int = 0, n_min: int = 4096, n_max: int = 524288,
                     horizons=(50, 2000), max_boxes: int = 4):
    """`count` requests with Poisson arrivals at `rate` per second."""
    rng = np.random.default_rng(seed)
    arrivals = np.cumsum(rng.exponential(1.0 / rate, size=count))
    sizes = np.exp(rng.uniform(math.log(n_min), math.log(n_max), size=count))
    hz = rng.integers(horizons[0], horizons[1] + 1, size=count)
    return [serving_request(rng, int(n), f"req{k}", int(hz[k]), float(arrivals[k]), max_boxes)
            for k, n in enumerate(sizes)]


def record_hash(pa: ParticleArrays, count: int = 4096) -> str:
    """SHA-256 of the first ``count`` position/velocity records (f64 bytes)."""
    m = min(count, len(pa))
    data = np.concatenate([pa.position[:m].ravel(), pa.velocity[:m].ravel(), pa.mass[:m], pa.rest_volume[:m]])
    return hashlib.sha256(np.ascontiguousarray(data, dtype=np.float64).tobytes()).hexdigest()


SCENES = {"c1": uniform_cube, "c2": four_cubes, "c3": dam_break, "c4": batch_scenes}
