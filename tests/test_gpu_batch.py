"""Batched scenes on the GPU.

A batch advances independent worlds with one launch per stage per substep.
The block key carries the scene index, so no block, tile or grid node mixes
scenes, and P2G's fixed-point sums do not depend on which other particles are
binned alongside.  Each scene of a batch must therefore reproduce, bit for
bit, the same world stepped alone in its own resident session, in both P2G
modes; single sessions are checked against the FP64 oracle elsewhere, and one
scene here is checked against it directly.

Failure semantics per scene follow the resident session (gpu.rs:377-381): a
failure after submission stops the scene and keeps its last synchronized
world, without disturbing the others."""

import numpy as np
import pytest

from oracle import oracle as O
from tessera_amd.mpm import (HipMpmBatch, HipMpmResidentSession, InvalidInput, MaterialModel, MpmParams, MpmWorld,
                             ResidentUnavailable, RigidObstacle, WorldBounds)
from tessera_amd.scenes import GRAVITY_Y, _particles, batch_scenes, jittered_lattice

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "affine", "deformation")


def small_scene(seed: int, kind: int, h: float, n_axis: int, obstacle: bool = False) -> MpmWorld:
    rng = np.random.default_rng(seed)
    spacing = h / 2
    edge = n_axis * spacing
    lo = rng.uniform(0.3, 0.7 - edge, size=3)
    pos = jittered_lattice(rng, [n_axis] * 3, lo, lo + edge)
    vel = rng.normal(0.0, 0.4, size=pos.shape)
    mat = [MaterialModel.elastic(2.0e4, 0.3), MaterialModel.neo_hookean(1.5e4, 0.25),
           MaterialModel.fluid(5.0e4, 7.0, 0.02)][kind]
    pa = _particles(pos, vel, spacing / 2, 1.0e3, mat)
    pa.transfer_color[::7] = 3
    pa.fixed[5] = True
    pa.enabled[11] = False
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y * (1.0 + 0.1 * seed), max_substep=1.0e-4,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    if obstacle:
        # just below the cube: grid nodes within one cell of its surface take its contact law
        w.set_obstacles([RigidObstacle.sphere(lo + np.array([edge / 2, -0.06, edge / 2]), 0.05)])
    return w


def mixed_worlds():
    return [small_scene(1, 0, 1 / 64, 12, obstacle=True), small_scene(2, 1, 1 / 96, 10),
            small_scene(3, 2, 1 / 80, 14), small_scene(4, 0, 1 / 128, 9), small_scene(5, 2, 1 / 64, 8, True)]


def alone(transfers, world, dt, steps, mode):
    s = HipMpmResidentSession(transfers, world, dt, mode)
    left = steps
    while left:
        k = min(left, 64)
        s.submit_steps(k)
        left -= k
    s.synchronize()
    return s.read_particles()


def assert_same(a, b):
    for f in FIELDS:
        x, y = getattr(a, f), getattr(b, f)
        assert np.array_equal(x, y), (f, float(np.max(np.abs(x - y))))


def dts_for(worlds):
    return [min(5.0e-5, 0.5 * w.stable_timestep()) for w in worlds]


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_batch_scenes_match_the_same_worlds_stepped_alone_bit_for_bit(transfers, mode):
    worlds = mixed_worlds()
    dts = dts_for(worlds)
    b = HipMpmBatch(transfers, worlds, dts, mode)
    b.submit_steps(17)  # crosses the periodic reorder of the particle state
    b.submit_steps(7)
    b.synchronize()
    for k, w in enumerate(worlds):
        assert b.scene_error(k) is None
        assert b.substeps(k) == 24
        assert_same(b.read_particles(k), alone(transfers, w, dts[k], 24, mode))


def test_batch_scene_tracks_the_fp64_oracle(transfers):
    worlds = mixed_worlds()[:2]
    dts = dts_for(worlds)
    b = HipMpmBatch(transfers, worlds, dts)
    b.step(10)
    cpu = O.OracleWorld(worlds[1])
    cpu.substeps_fixed(dts[1], 10)
    ref, got = cpu.read(), b.read_particles(1)
    live = np.asarray(got.enabled) & ~np.asarray(got.fixed)
    assert np.max(np.linalg.norm(ref["position"][live] - got.position[live], axis=1)) < 1e-5
    assert np.max(np.linalg.norm(ref["velocity"][live] - got.velocity[live], axis=1)) < 1e-3


def test_failure_stops_only_its_scene_and_keeps_its_last_world(transfers):
    worlds = mixed_worlds()[:3]
    dts = dts_for(worlds)
    b = HipMpmBatch(transfers, worlds, dts, "fixed64")
    b.step(4)
    before = b.read_particles(1)
    # a one-shot force that takes one particle of scene 1 to 200 m/s, beyond
    # the CFL bound, while it stays well inside the walls
    m = float(worlds[1].particles.mass[0])
    b.set_particle_force(1, 0, [200.0 * m / dts[1], 0.0, 0.0])
    b.submit_steps(6)
    b.synchronize()
    assert isinstance(b.scene_error(1), ResidentUnavailable)
    assert b.scene_error(0) is None and b.scene_error(2) is None
    assert b.substeps(1) == 4
    assert_same(b.read_particles(1), before)
    with pytest.raises((InvalidInput, ResidentUnavailable)):
        b.set_particle_force(1, 0, [0.0, 0.0, 0.0])
    b.step(3)  # the stopped scene no longer runs; the others continue
    for k in (0, 2):
        assert b.substeps(k) == 13
        assert_same(b.read_particles(k), alone(transfers, worlds[k], dts[k], 13, "fixed64"))


def test_scenes_join_and_leave_between_submissions(transfers):
    worlds = mixed_worlds()
    dts = dts_for(worlds)
    b = HipMpmBatch(transfers, worlds[:2], dts[:2])
    b.step(3)
    k2 = b.add_scene(worlds[2], dts[2])
    assert k2 == 2
    b.step(5)
    b.remove_scene(0)
    with pytest.raises(InvalidInput):
        b.read_particles(0)
    b.step(2)
    k3 = b.add_scene(worlds[3], dts[3])
    assert k3 == 0  # the freed slot is reused
    b.step(4)
    assert_same(b.read_particles(1), alone(transfers, worlds[1], dts[1], 14, "fp32"))
    assert_same(b.read_particles(2), alone(transfers, worlds[2], dts[2], 11, "fp32"))
    assert_same(b.read_particles(0), alone(transfers, worlds[3], dts[3], 4, "fp32"))
    assert b.substeps(2) == 11 and b.substeps(0) == 4


def test_256_scene_batch_conserves_each_scenes_momentum_under_gravity(transfers):
    scenes = batch_scenes(256, seed=0)
    worlds = [s.world for s in scenes]
    b = HipMpmBatch(transfers, worlds, [s.dt for s in scenes], "fixed64")
    steps = 8
    b.step(steps)
    for k in range(0, 256, 17):
        w = worlds[k]
        got = b.read_particles(k)
        m = np.asarray(w.particles.mass)
        p0 = (m[:, None] * np.asarray(w.particles.velocity)).sum(0)
        p1 = (m[:, None] * got.velocity).sum(0)
        impulse = m.sum() * np.asarray(w.params.gravity) * scenes[k].dt * steps
        scale = float(np.abs(m[:, None] * np.asarray(w.particles.velocity)).sum())
        assert b.scene_error(k) is None
        assert np.max(np.abs(p1 - p0 - impulse)) < 1e-5 * scale, k


def test_a_diverging_unbounded_scene_cannot_exhaust_the_shared_capacity(transfers):
    """A scene without walls whose particles are all thrown past the CFL bound
    would, binned again, need a block per particle and overflow the hash and
    grid capacity that the batch's scenes share (a session-wide Capacity
    failure).  Once a scene has failed it is no longer binned: it alone stops,
    keeps its last synchronized world, and the other scenes equal the same
    worlds stepped alone."""
    import dataclasses

    worlds = mixed_worlds()[:3]
    free = worlds[1].clone()
    free.params = dataclasses.replace(free.params, bounds=None)
    worlds[1] = free
    dts = dts_for(worlds)
    b = HipMpmBatch(transfers, worlds, dts, "fp32")
    b.step(2)
    before = b.read_particles(1)
    n = len(free.particles)
    rng = np.random.default_rng(3)
    dirs = rng.normal(size=(n, 3))
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    m = np.asarray(free.particles.mass)
    # 1e4 m/s in random directions: ~0.5 m (dozens of cells) in one substep
    for i in range(n):
        b.set_particle_force(1, i, dirs[i] * 1.0e4 * m[i] / dts[1])
    b.submit_steps(8)
    b.synchronize()
    assert b.valid
    assert isinstance(b.scene_error(1), ResidentUnavailable)
    assert b.substeps(1) == 2
    assert_same(b.read_particles(1), before)
    b.step(3)
    for k in (0, 2):
        assert b.scene_error(k) is None
        assert b.substeps(k) == 13
        assert_same(b.read_particles(k), alone(transfers, worlds[k], dts[k], 13, "fp32"))
