"""Obstacle reactions and device-resident rigid coupling on the GPU.

Reactions (ObstacleReaction, world.rs:190-204) are checked substep by
substep against the FP64 oracle's step_with_obstacle_reactions and against
the momentum balance the reference's own CPU test asserts (world.rs tests:
particle momentum change + reactions = gravity impulse).  The MPM-owned rigid
coupling (gpu_rigid_sync.wgsl:43-62, gpu_rigid_reaction.wgsl:72-150) is
checked against a host loop of the oracle world and oracle/rigid.py.  Both
are reproducible run to run and identical in a batch."""

import numpy as np
import pytest

from oracle import oracle as O
from oracle import rigid as R
from tessera_amd.mpm import (HipMpmBatch, HipMpmResidentSession, MaterialModel, MpmParams, MpmParticle, MpmWorld,
                             ObstacleBoundary, RigidBody, RigidMapping, RigidObstacle, WorldBounds)
from tessera_amd.scenes import GRAVITY_Y, _particles, jittered_lattice

pytestmark = pytest.mark.gpu


def cube_on_box(seed=0, boundary=ObstacleBoundary.SLIP, friction=0.0, n_axis=10):
    rng = np.random.default_rng(seed)
    h = 1.0 / 64
    spacing = h / 2
    edge = n_axis * spacing
    lo = np.array([0.5 - edge / 2, 0.4 + 0.25 * h, 0.5 - edge / 2])
    pos = jittered_lattice(rng, [n_axis] * 3, lo, lo + edge)
    vel = np.tile([0.05, -0.5, 0.02], (len(pos), 1)) + rng.normal(0.0, 0.05, size=pos.shape)
    pa = _particles(pos, vel, spacing / 2, 1.0e3, MaterialModel.elastic(2.0e4, 0.3))
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y.copy(), max_substep=1.0e-4,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    box = RigidObstacle.cuboid([0.5, 0.35, 0.5], [0.15, 0.05, 0.15])
    box.boundary = boundary
    box.friction = friction
    far = RigidObstacle.sphere([0.15, 0.8, 0.15], 0.05)
    w.set_obstacles([box, far])
    return w


def momentum(pa):
    return (np.asarray(pa.mass)[:, None] * np.asarray(pa.velocity)).sum(0)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
@pytest.mark.parametrize("boundary,friction", [(ObstacleBoundary.SLIP, 0.3), (ObstacleBoundary.STICK, 0.0)])
def test_reactions_match_the_oracle_substep_by_substep(transfers, mode, boundary, friction):
    w = cube_on_box(boundary=boundary, friction=friction)
    dt = min(5.0e-5, 0.5 * w.stable_timestep())
    s = HipMpmResidentSession(transfers, w, dt, mode)
    s.set_reactions(True)
    cpu = O.OracleWorld(w)
    total_gpu = np.zeros((2, 6))
    total_cpu = np.zeros((2, 6))
    p_prev = momentum(w.particles)
    mass = float(np.sum(w.particles.mass))
    for k in range(24):
        s.step(1)
        r_gpu = s.reactions()
        r_cpu = cpu.step_with_obstacle_reactions(dt, 2)
        scale = max(float(np.abs(r_cpu).max()), 1e-12)
        assert np.abs(r_gpu - r_cpu).max() <= 2e-3 * scale + 1e-10, (k, r_gpu, r_cpu)
        assert np.all(r_gpu[1] == 0.0)  # the far sphere deflects nothing
        # momentum balance of this substep (gravity impulse = particle change + reactions)
        p = momentum(s.read_particles())
        balance = p - p_prev + r_gpu[0, :3] - mass * GRAVITY_Y * dt
        assert np.abs(balance).max() <= 1e-5 * mass * 0.5
        p_prev = p
        total_gpu += r_gpu
        total_cpu += r_cpu
    assert np.abs(total_cpu[0, 1]) > 0.0  # contact happened
    assert np.abs(total_gpu - total_cpu).max() <= 2e-3 * np.abs(total_cpu).max()
    assert s.reaction_saturations() == 0  # every particle impulse fit the int64 bins


def test_reactions_are_summed_over_the_synchronized_substeps_and_reproducible(transfers):
    w = cube_on_box(seed=2)
    dt = min(5.0e-5, 0.5 * w.stable_timestep())
    runs = []
    for _ in range(2):
        s = HipMpmResidentSession(transfers, w, dt, "fixed64")
        s.set_reactions(True)
        s.submit_steps(7)
        s.submit_steps(5)
        s.synchronize()
        runs.append(s.reactions())
    assert np.array_equal(runs[0], runs[1])
    single = HipMpmResidentSession(transfers, w, dt, "fixed64")
    single.set_reactions(True)
    acc = np.zeros((2, 6))
    for _ in range(12):
        single.step(1)
        acc += single.reactions()
    assert np.allclose(acc, runs[0], rtol=1e-6, atol=1e-12)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_reactions_do_not_depend_on_the_g2p_workgroup_or_storage_order(transfers, mode, monkeypatch):
    # Each particle's impulse is rounded to the bins' integer units before it is
    # summed, so G2P's workgroup size (which lane takes which particle) and the
    # reorder interval (the storage order) change no reaction bit.
    w = cube_on_box(seed=5, boundary=ObstacleBoundary.STICK, n_axis=16)
    dt = min(5.0e-5, 0.5 * w.stable_timestep())
    runs = []
    for threads, interval in [(64, 16), (256, 16), (128, 3)]:
        monkeypatch.setenv("TMPM_PBLOCK_THREADS", str(threads))
        monkeypatch.setenv("TMPM_REORDER_INTERVAL", str(interval))
        s = HipMpmResidentSession(transfers, w, dt, mode)
        assert s.launch_config()["g2p_threads"] == threads
        s.set_reactions(True)
        got = []
        for _ in range(20):
            s.step(1)
            got.append(s.reactions())
        runs.append((np.stack(got), s.read_particles()))
    assert np.abs(runs[0][0][:, 0, 1]).max() > 0.0  # the cube hit the box
    for reactions, pa in runs[1:]:
        assert np.array_equal(reactions, runs[0][0])
        for f in ("position", "velocity", "affine", "deformation"):
            assert np.array_equal(getattr(pa, f), getattr(runs[0][1], f)), f


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_reactions_support_more_than_sixteen_obstacles(transfers, mode):
    # gpu.rs:2651-2720: 16 far spheres and one the particle hits; reactions are
    # kept for every obstacle, the far ones get none, the hit one matches the
    # CPU.  Obstacles past the 16th of a scene take the direct-atomic path.
    particle = MpmParticle.new([1.02, 0.0, 0.0], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    particle.velocity = np.array([-1.0, 0.0, 0.0])
    w = MpmWorld([particle], MpmParams(gravity=np.zeros(3)))
    obstacles = [RigidObstacle.sphere([10.0 + 3.0 * i, 0.0, 0.0], 1.0) for i in range(16)]
    obstacles.append(RigidObstacle.sphere([0.0, 0.0, 0.0], 1.0))
    w.set_obstacles(obstacles)
    cpu = O.OracleWorld(w)
    expected = cpu.step_with_obstacle_reactions(0.001, 17)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    s.set_reactions(True)
    s.step_adaptive(0.001)
    got = s.reactions()
    assert got.shape == (17, 6)
    assert np.all(np.linalg.norm(got[:16, :3], axis=1) < 1e-6)
    assert got[16, 0] < -0.01, got[16]
    assert np.linalg.norm(got[16, :3] - expected[16, :3]) < 1e-4


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_reactions_keep_ordered_overlap_attribution(transfers, mode):
    # gpu.rs:2723-2770: two overlapping spheres, the second moving; each
    # projection is credited to the obstacle that made it, in obstacle order.
    particle = MpmParticle.new([1.02, 0.0, 0.0], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    particle.velocity = np.array([-1.0, 0.0, 0.0])
    w = MpmWorld([particle], MpmParams(gravity=np.zeros(3)))
    second = RigidObstacle.sphere([0.1, 0.0, 0.0], 1.0)
    second.linear_velocity = np.array([1.0, 0.0, 0.0])
    w.set_obstacles([RigidObstacle.sphere([0.0, 0.0, 0.0], 1.0), second])
    cpu = O.OracleWorld(w)
    expected = cpu.step_with_obstacle_reactions(0.001, 2)
    assert expected[0, 0] < 0.0 and expected[1, 0] < 0.0
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    s.set_reactions(True)
    s.step_adaptive(0.001)
    got = s.reactions()
    for k in range(2):
        assert np.linalg.norm(got[k, :3] - expected[k, :3]) < 1e-4, (k, got[k], expected[k])
        assert np.linalg.norm(got[k, 3:] - expected[k, 3:]) < 1e-4, (k, got[k], expected[k])


def box_on_slab(seed=0):
    rng = np.random.default_rng(seed)
    h = 1.0 / 64
    spacing = h / 2
    lo = np.array([0.35, 0.2, 0.35])
    counts = [38, 8, 38]
    pos = jittered_lattice(rng, counts, lo, lo + np.array(counts) * spacing)
    pa = _particles(pos, np.zeros_like(pos), spacing / 2, 1.0e3, MaterialModel.elastic(5.0e4, 0.3))
    params = MpmParams(cell_width=h, gravity=GRAVITY_Y.copy(), max_substep=1.0e-4,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    floor = RigidObstacle.cuboid([0.5, 0.15, 0.5], [0.3, 0.05, 0.3])
    floor.boundary = ObstacleBoundary.STICK
    top = float(pos[:, 1].max())
    # the box's lower face starts half a cell above the slab: in contact from the first substep
    body = RigidBody.box([0.5, top + 0.03 + 0.5 * h, 0.52], [0.05, 0.03, 0.04], density=600.0)
    body.linear_velocity = np.array([0.0, -0.6, 0.0])
    body.angular_velocity = np.array([0.0, 0.5, 1.5])
    shell = RigidObstacle.cuboid(body.position, [0.05, 0.03, 0.04])
    w.set_obstacles([floor, shell])
    maps = [RigidMapping(None), RigidMapping(0)]
    return w, [body], maps


def oracle_coupled(w, bodies, maps, dt, steps, gravity):
    cpu = O.OracleWorld(w)
    state = [b.as_dict() for b in bodies]
    mp = [m.as_dict() for m in maps]
    obstacles = list(w.obstacles)
    for _ in range(steps):
        posed = [R.sync_obstacle(state[m["body"]], m, o) if m["body"] is not None else o
                 for m, o in zip(mp, obstacles)]
        cpu.set_obstacles(posed)
        r = cpu.step_with_obstacle_reactions(dt, len(posed))
        state = R.apply_reactions(state, mp, posed, [(r[i, :3], r[i, 3:]) for i in range(len(posed))], dt,
                                  gravity, 0.0, True)
        obstacles = posed
    return cpu, state


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_mpm_owned_rigid_box_tracks_the_oracle_coupling(transfers, mode):
    w, bodies, maps = box_on_slab()
    dt = min(5.0e-5, 0.5 * w.stable_timestep())
    g = np.array([0.0, -9.8, 0.0])
    steps = 60
    s = HipMpmResidentSession(transfers, w, dt, mode)
    s.set_rigid_bodies(bodies, maps, gravity=g, integrate=True)
    for _ in range(steps // 20):
        s.submit_steps(20)
    s.synchronize()
    got = s.rigid_bodies()[0]
    cpu, ref = oracle_coupled(w, bodies, maps, dt, steps, g)
    ref = ref[0]
    assert np.abs(got.position - ref["position"]).max() < 2e-6
    assert np.abs(got.linear_velocity - ref["linear_velocity"]).max() < 2e-3 * max(1.0, np.abs(ref["linear_velocity"]).max())
    assert np.abs(got.angular_velocity - ref["angular_velocity"]).max() < 5e-3 * max(1.0, np.abs(ref["angular_velocity"]).max())
    q = got.orientation * np.sign(got.orientation[3])
    qr = ref["orientation"] * np.sign(ref["orientation"][3])
    assert np.abs(q - qr).max() < 1e-5
    # the body moved and was pushed by the slab (not a free fall)
    free_fall_v = bodies[0].linear_velocity[1] + g[1] * dt * steps
    assert abs(ref["linear_velocity"][1] - free_fall_v) > 1e-3
    # the driven obstacle was read back at the pose of the last substep's start:
    # one substep's motion behind the body
    shell = s.obstacles()[1]
    lag = np.abs(shell.center - got.position).max()
    assert 0.0 < lag <= 2.0 * np.abs(got.linear_velocity).max() * dt + 1e-6
    pa = s.read_particles()
    ref_p = cpu.read()
    assert np.max(np.linalg.norm(ref_p["position"] - pa.position, axis=1)) < 1e-5


def test_rigid_coupling_in_a_batch_matches_single_sessions_and_repeats(transfers):
    worlds, bodies, maps = [], [], []
    for seed in (0, 1):
        w, b, m = box_on_slab(seed)
        worlds.append(w)
        bodies.append(b)
        maps.append(m)
    dts = [min(5.0e-5, 0.5 * w.stable_timestep()) for w in worlds]
    g = np.array([0.0, -9.8, 0.0])
    batch = HipMpmBatch(transfers, worlds, dts, "fixed64")
    for k in range(2):
        batch.set_rigid_bodies(k, bodies[k], maps[k], gravity=g)
    batch.step(30)
    for k in range(2):
        alone = HipMpmResidentSession(transfers, worlds[k], dts[k], "fixed64")
        alone.set_rigid_bodies(bodies[k], maps[k], gravity=g)
        alone.step(30)
        a, b = alone.rigid_bodies()[0], batch.rigid_bodies(k)[0]
        for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert np.array_equal(alone.read_particles().position, batch.read_particles(k).position)
