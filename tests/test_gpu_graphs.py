"""Substep graphs and the launch-shape autotuner on the GPU.

With graphs on, each substep's launches are captured into a HIP graph once
per distinct launch plan and replayed afterwards.  A replay must equal the
eager launches bit for bit, on every replay: the same world is run with
graphs on and off, with host work, synchronizations and edits between the
replays (force edits, obstacle edits, added particles, scenes admitted and
retired), over several reorder cycles, and every synchronized state is
compared."""

import numpy as np
import pytest

from tessera_amd.mpm import HipMpmBatch, HipMpmResidentSession, MaterialModel, MpmParticle, ParticleArrays
from tessera_amd.scenes import four_cubes, serving_request, sticky_floor, uniform_cube

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "affine", "deformation", "plastic")


def request(seed, n=6000, boxes=2):
    return serving_request(np.random.default_rng(seed), n, f"g{seed}", 100, 0.0, boxes)


def same(a, b, where):
    for f in FIELDS:
        assert np.array_equal(getattr(a, f), getattr(b, f)), (where, f)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_graph_replays_equal_eager_launches_with_edits_between(transfers, mode):
    req = request(1)
    runs = []
    for graphs in (True, False):
        s = HipMpmResidentSession(transfers, req.world, req.dt, mode)
        s.set_graphs(graphs)
        s.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
        states = []
        for k, steps in enumerate([3, 8, 1, 16, 5, 9, 12]):
            s.submit_steps(steps)
            s.synchronize()
            states.append((s.read_particles(), s.rigid_bodies()))
            if k == 2:
                s.set_particle_force(7, [0.0, 3.0, 0.0])
            if k == 4:
                extra = ParticleArrays.from_particles(
                    [MpmParticle.new([0.5 + 0.01 * i, 0.6, 0.5], 0.004, 1000.0, MaterialModel.snow(2.0e4, 0.3))
                     for i in range(5)])
                s.add_particles(extra)
            if k == 5:
                s.set_obstacles([sticky_floor(0.05)] + list(s.obstacles())[1:])
        runs.append(states)
    for k, ((pa, ba), (pb, bb)) in enumerate(zip(*runs)):
        same(pa, pb, k)
        for x, y in zip(ba, bb):
            assert np.array_equal(x.position, y.position) and np.array_equal(x.angular_velocity, y.angular_velocity)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_graph_replays_in_a_serving_batch_equal_eager_launches(transfers, mode):
    reqs = [request(10 + k, n=3000 + 1500 * k, boxes=1 + k % 2) for k in range(4)]
    results = []
    for graphs in (True, False):
        b = HipMpmBatch(transfers, [reqs[0].world, reqs[1].world], [reqs[0].dt, reqs[1].dt], mode)
        b.set_graphs(graphs)
        for k in range(2):
            b.set_rigid_bodies(k, list(reqs[k].bodies), list(reqs[k].mappings))
        b.step(10)
        slot = b.add_scene(reqs[2].world, reqs[2].dt)
        b.set_rigid_bodies(slot, list(reqs[2].bodies), list(reqs[2].mappings))
        b.step(7)
        first = b.read_particles(0)
        b.remove_scene(0)
        slot = b.add_scene(reqs[3].world, reqs[3].dt)
        b.step(19)
        results.append([first] + [b.read_particles(k) for k in b.scenes()])
    for k, (a, c) in enumerate(zip(*results)):
        same(a, c, k)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_autotuned_launch_shape_computes_the_same_substeps(transfers, mode):
    # the tuner's trials are real substeps under different launch shapes; the
    # result must equal a session that ran the same substeps with the default shape
    req = request(3, n=20000, boxes=1)
    w = req.world
    w.set_obstacles(list(w.obstacles)[:1])  # no rigid bodies: reactions off, every knob is tuned
    tuned = HipMpmResidentSession(transfers, w, req.dt, mode)
    cfg = tuned.autotune(16)
    assert cfg["substeps_run"] >= 16 * 12
    assert cfg["p2g_threads"] in (64, 128) and cfg["g2p_threads"] in (64, 128, 256)
    assert cfg["reorder_interval"] in (4, 8, 16)
    tuned.synchronize()
    plain = HipMpmResidentSession(transfers, w, req.dt, mode)
    left = cfg["substeps_run"]
    while left:
        k = min(left, 64)
        plain.step(k)
        left -= k
    same(tuned.read_particles(), plain.read_particles(), "tuned")
    assert tuned.substeps == plain.substeps


def test_autotuned_launch_shape_with_rigid_reactions_computes_the_same_substeps(transfers):
    # reactions sum rounded per-particle integers, so the tuner may change every
    # knob with rigid bodies coupled and the bodies still move bit for bit alike
    req = request(4, n=8000, boxes=1)
    tuned = HipMpmResidentSession(transfers, req.world, req.dt)
    tuned.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
    cfg = tuned.autotune(8)
    tuned.synchronize()
    assert tuned.substeps == cfg["substeps_run"]
    plain = HipMpmResidentSession(transfers, req.world, req.dt)
    plain.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
    left = cfg["substeps_run"]
    while left:
        k = min(left, 64)
        plain.step(k)
        left -= k
    same(tuned.read_particles(), plain.read_particles(), "tuned")
    for x, y in zip(tuned.rigid_bodies(), plain.rigid_bodies()):
        assert np.array_equal(x.position, y.position) and np.array_equal(x.angular_velocity, y.angular_velocity)


def test_default_launch_shape_follows_the_particle_block_count(transfers):
    # one workgroup per particle block: few blocks take wider workgroups so the
    # launch still fills the chip; reactions do not change the choice
    small = uniform_cube()  # 8,192 particles in 64 blocks
    s = HipMpmResidentSession(transfers, small.world, small.dt)
    cfg = s.launch_config()
    assert (cfg["p2g_threads"], cfg["p2g_copies"], cfg["g2p_threads"]) == (128, 1, 256)
    s.set_reactions(True)
    assert s.launch_config()["g2p_threads"] == 256
    big = four_cubes()  # 1,048,576 particles in ~11,000 blocks
    b = HipMpmResidentSession(transfers, big.world, big.dt)
    cfg = b.launch_config()
    assert (cfg["p2g_threads"], cfg["p2g_copies"], cfg["g2p_threads"]) == (64, 1, 64)
