"""Session entry points beside stepping: the non-blocking synchronize, the
two-half batch admission, the FP64 audit and the autotuner's cache and
capacity handling.

* ``synchronize_async`` (gpu.rs:413-450) is a poll here: while the device
  still runs the submitted substeps it changes nothing and the substeps stay
  pending; once they finish it validates and commits exactly as
  ``synchronize`` (gpu.rs:382-411) does, failures included.
* A staged admission (host half while substeps run, device half at the next
  synchronize point) equals ``add_scene`` of the same world bit for bit.
* The audit's f64 sums are checked against a numpy reduction of the same
  synchronized f32 state, and its per-substep momentum against the gravity
  impulse (partition of unity: the APIC transfers conserve momentum, the grid
  update adds m g dt per node).
"""

import dataclasses
import time

import numpy as np
import pytest

from tessera_amd.mpm import (Capacity, HipMpmBatch, HipMpmResidentSession, MaterialModel, MpmParams, MpmWorld,
                             ResidentUnavailable, StagedScene, tune_cache_clear, tune_cache_size)
from tessera_amd.scenes import _particles, four_cubes, uniform_cube

from .test_gpu_batch import FIELDS, assert_same, dts_for, mixed_worlds

pytestmark = pytest.mark.gpu


def test_synchronize_async_polls_while_pending_then_commits_like_synchronize(transfers):
    sc = four_cubes()
    a = HipMpmResidentSession(transfers, sc.world, sc.dt)
    b = HipMpmResidentSession(transfers, sc.world, sc.dt)
    a.submit_steps(64)  # ~13 ms of device work on the 1M-particle scene
    first = a.synchronize_async()
    assert not first, "64 substeps of a million particles cannot have finished already"
    assert a.pending_substeps() == 64 and a.valid
    polls = 1
    while not a.synchronize_async():
        polls += 1
        time.sleep(0.0005)
    assert a.pending_substeps() == 0 and a.world().substeps == 64
    assert a.synchronize_async() is True  # nothing pending: done at once
    b.submit_steps(64)
    b.synchronize()
    assert_same(a.read_particles(), b.read_particles())
    print(f"synchronize_async: {polls} polls")


def test_synchronize_async_failure_invalidates_and_keeps_the_last_world(transfers):
    sc = uniform_cube(n=2048)
    s = HipMpmResidentSession(transfers, sc.world, sc.dt)
    s.step(3)
    before = s.read_particles()
    m = float(sc.world.particles.mass[0])
    s.set_particle_force(0, [300.0 * m / sc.dt, 0.0, 0.0])  # 300 m/s: beyond the CFL bound
    s.submit_steps(5)
    with pytest.raises(ResidentUnavailable):
        while not s.synchronize_async():
            time.sleep(0.0005)
    assert not s.valid
    assert s.world().substeps == 3
    assert_same(s.read_particles(), before)
    with pytest.raises(ResidentUnavailable):
        s.synchronize_async()


def test_staged_admission_equals_add_scene_bit_for_bit(transfers):
    worlds = mixed_worlds()
    dts = dts_for(worlds)
    a = HipMpmBatch(transfers, worlds[:2], dts[:2])
    b = HipMpmBatch(transfers, worlds[:2], dts[:2])
    a.submit_steps(9)
    staged = StagedScene(worlds[2], dts[2])  # host half while a's substeps run
    a.synchronize()
    b.step(9)
    ka = a.add_staged(staged)
    kb = b.add_scene(worlds[2], dts[2])
    assert ka == kb == 2
    with pytest.raises(Exception):
        a.add_staged(staged)  # consumed
    a.remove_scene(0)
    b.remove_scene(0)
    a.submit_steps(4)
    staged = StagedScene(worlds[3], dts[3])
    a.synchronize()
    b.step(4)
    assert a.add_staged(staged) == b.add_scene(worlds[3], dts[3]) == 0
    a.step(6)
    b.step(6)
    for k in (0, 1, 2):
        assert a.substeps(k) == b.substeps(k)
        assert_same(a.read_particles(k), b.read_particles(k))


def _audit_reference(pa):
    # The audit sums the device's f32 state.  A fixed particle's host position
    # stays the f64 value it was created with (it never moves), so every input
    # is rounded to f32 as the device holds it.
    f32 = lambda a: np.asarray(a, np.float64).astype(np.float32).astype(np.float64)  # noqa: E731
    m = f32(pa.mass) * np.asarray(pa.enabled, bool)
    x, v = f32(pa.position), f32(pa.velocity)
    terms = np.concatenate([m[:, None], m[:, None] * v, (0.5 * m * (v * v).sum(1))[:, None], m[:, None] * x,
                            m[:, None] * np.cross(x, v)], axis=1)
    return terms.sum(0), np.abs(terms).sum(0)


def test_audit_matches_numpy_and_momentum_follows_gravity(transfers):
    sc = uniform_cube(n=8192, seed=1)
    w = sc.world.clone()
    w.set_obstacles([])  # free fall: the cube never reaches a wall in 12 substeps
    s = HipMpmResidentSession(transfers, w, sc.dt)
    s.set_audit(True)
    s.step(12)
    rec = s.audit()
    assert rec.shape == (12, 12)
    got = s.read_particles()
    ref, scale = _audit_reference(got)
    assert rec[-1, 11] == np.count_nonzero(got.enabled)
    err = np.abs(rec[-1, :11] - ref) / np.maximum(scale, 1e-300)
    assert np.all(err < 1e-12), err
    # mass is conserved exactly; momentum changes by M g dt per substep
    mass = rec[:, 0]
    assert np.allclose(mass, mass[0], rtol=1e-14, atol=0.0)
    g = np.asarray(w.params.gravity)
    dp = np.diff(rec[:, 1:4], axis=0)
    expect = mass[0] * g * sc.dt
    assert np.max(np.abs(dp - expect)) < 1e-4 * np.linalg.norm(expect), (dp[:3], expect)
    # a batch audits each scene separately
    worlds = mixed_worlds()[:3]
    dts = dts_for(worlds)
    bt = HipMpmBatch(transfers, worlds, dts)
    bt.set_audit(True)
    bt.step(5)
    for k in range(3):
        r = bt.audit(k)
        assert r.shape == (5, 12)
        ref, scale = _audit_reference(bt.read_particles(k))
        err = np.abs(r[-1, :11] - ref) / np.maximum(scale, 1e-300)
        assert np.all(err < 1e-12), (k, err)


def test_autotune_cache_skips_trials_in_a_tuned_bucket(transfers):
    tune_cache_clear()
    sc = uniform_cube(n=8192)
    a = HipMpmResidentSession(transfers, sc.world, sc.dt)
    cfg_a = a.autotune(8)
    assert cfg_a["substeps_run"] > 0
    assert tune_cache_size() >= 1
    b = HipMpmResidentSession(transfers, sc.world, sc.dt)
    cfg_b = b.autotune(8)
    assert cfg_b["substeps_run"] == 0
    assert {k: v for k, v in cfg_a.items() if k != "substeps_run"} == \
        {k: v for k, v in cfg_b.items() if k != "substeps_run"}
    # tuning changes no result: both equal an untuned session at the same substep count
    a.synchronize()
    ref = HipMpmResidentSession(transfers, sc.world, sc.dt)
    n = cfg_a["substeps_run"]
    while n:
        k = min(64, n)
        ref.submit_steps(k)
        n -= k
    ref.synchronize()
    assert_same(a.read_particles(), ref.read_particles())
    tune_cache_clear()


def _spreading_cloud(n_axis: int = 16) -> tuple[MpmWorld, float]:
    """A soft fluid ball in an unbounded world flying apart at 2 m/s: its
    particle-block count grows several times over within 3,000 substeps."""
    rng = np.random.default_rng(11)
    h = 1.0 / 64
    g = (np.arange(n_axis) + 0.5) / n_axis
    pos = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3) * (n_axis * h / 2) + 0.5
    pos += rng.uniform(-0.1, 0.1, size=pos.shape) * h / 2
    d = pos - pos.mean(0)
    vel = 2.0 * d / np.maximum(np.linalg.norm(d, axis=1, keepdims=True), 1e-9)
    pa = _particles(pos, vel, h / 4, 1.0e3, MaterialModel.fluid(1.0e2, 7.0, 0.0))
    params = MpmParams(cell_width=h, gravity=np.zeros(3), max_substep=1e-4, bounds=None)
    return MpmWorld(pa, params), 1.0e-4


def test_autotune_grows_the_topology_of_a_spreading_scene(transfers):
    tune_cache_clear()
    world, dt = _spreading_cloud()
    probe = HipMpmResidentSession(transfers, world, dt)
    probe.step(1)
    initial = probe.topology_stats()["particle_blocks"]
    probe.close()
    s = HipMpmResidentSession(transfers, world, dt)
    cfg = s.autotune(256)
    steps = cfg["substeps_run"]
    assert steps == 16 + 256 * 12
    s.synchronize()  # no Capacity: the topology grew between the tuning chunks
    assert s.valid
    ref = HipMpmResidentSession(transfers, world, dt)
    left = steps
    while left:
        k = min(64, left)
        ref.step(k)
        left -= k
    st = ref.topology_stats()
    # beyond the 4x headroom the topology was sized with at creation
    assert st["particle_blocks"] > 4 * initial, (initial, st)
    assert_same(s.read_particles(), ref.read_particles())
    tune_cache_clear()


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_fixed_point_range_is_a_capacity_error_in_both_modes(transfers, mode):
    """Both P2G modes accumulate 64-bit fixed point in LDS, so a contribution
    beyond 2^45 units (the heaviest particle at 64 cells per substep is 2^43)
    is a Capacity error in FP32 mode too, where the reference's f32
    compare-exchange loops (gpu.wgsl:739-766) would add it.  Such a particle
    has already broken the CFL bound in the substep before, so the failure
    only changes which error a doomed submission reports."""
    sc = uniform_cube(n=512)
    w = sc.world.clone()
    w.params = dataclasses.replace(w.params, bounds=None)
    w.set_obstacles([])
    s = HipMpmResidentSession(transfers, w, sc.dt, mode)
    s.step(1)
    before = s.read_particles()
    m = float(w.particles.mass[0])
    v = 1024 * w.params.cell_width / sc.dt  # 1024 cells per substep
    s.set_particle_force(0, [v * m / sc.dt, 0.0, 0.0])
    s.submit_steps(2)
    with pytest.raises(Capacity):
        s.synchronize()
    assert not s.valid and s.world().substeps == 1
    assert_same(s.read_particles(), before)


def _exploding_clusters() -> tuple[MpmWorld, float]:
    """Eight clusters of 2,048 particles, each within ~2^3 blocks, expanding homologously (v
    proportional to the distance from the cluster's centre) at up to 50 m/s, 0.32 cells per
    substep, inside the CFL bound.  Four such clusters occupy ~1,030 particle blocks after 65
    substeps (the transfers slow the edge to ~40 m/s), so these eight need ~2,000, past the
    1,024 the session sizes from the initial 64."""
    rng = np.random.default_rng(7)
    h = 1.0 / 64
    r0 = 2 * h
    pos, vel = [], []
    for c in range(8):
        center = np.array([0.5 + 1.5 * (c % 4), 0.5 + 1.5 * (c // 4), 0.5])
        d = rng.normal(size=(2048, 3))
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        p = center + d * r0 * rng.uniform(0.0, 1.0, size=(2048, 1)) ** (1.0 / 3.0)
        pos.append(p)
        vel.append((p - center) / r0 * 50.0)
    pa = _particles(np.concatenate(pos), np.concatenate(vel), h / 4, 1.0e3, MaterialModel.fluid(1.0e2, 7.0, 0.0))
    params = MpmParams(cell_width=h, gravity=np.zeros(3), max_substep=1e-4, bounds=None)
    return MpmWorld(pa, params), 1.0e-4


def test_topology_overflow_within_a_submission_is_a_capacity_error(transfers):
    """The topology grows only between submissions.  Particle blocks beyond its capacity within
    one submission are left out of the substep (their particles are neither scattered nor
    moved) and flagged, and synchronize reports Capacity: the session is invalid and keeps the
    last synchronized world (gpu.rs:377-381, as for any failure after submission)."""
    world, dt = _exploding_clusters()
    s = HipMpmResidentSession(transfers, world, dt)
    s.step(1)
    assert s.topology_stats()["particle_blocks"] <= 128
    before = s.read_particles()
    s.submit_steps(64)
    with pytest.raises(Capacity):
        s.synchronize()
    assert not s.valid and s.world().substeps == 1
    assert_same(s.read_particles(), before)
    # the same substeps in chunks that synchronize in between grow the topology instead
    ok = HipMpmResidentSession(transfers, world, dt)
    for _ in range(16):
        ok.step(4)
    assert ok.valid and ok.world().substeps == 64
    assert ok.topology_stats()["particle_blocks"] > 1500


def _dense_blocks() -> tuple[MpmWorld, float]:
    """Three particle blocks, one for each way the reorder lays a block out: 400 particles over the
    64 cells of one block (round-robin by cell), 300 inside a single cell of another (more than 128
    in one cell: the block keeps its order) and 1,500 over a third (more than 1,024 particles: kept
    as well).  Cell c holds x in [(c + 0.5) h, (c + 1.5) h)."""
    rng = np.random.default_rng(11)
    h = 1.0 / 64

    def box(n, first_cell, cells):
        lo = (np.array(first_cell, dtype=np.float64) + 0.55) * h
        return lo + rng.uniform(0.0, (cells - 0.1) * h, size=(n, 3))

    pos = np.concatenate([box(400, (16, 16, 16), 4), box(300, (25, 17, 17), 1), box(1500, (32, 16, 16), 4)])
    vel = rng.normal(0.0, 0.01, size=pos.shape)
    pa = _particles(pos, vel, h / 8, 1.0e3, MaterialModel.elastic(1.0e3, 0.3))
    params = MpmParams(cell_width=h, gravity=np.zeros(3), max_substep=1e-4, bounds=None)
    return MpmWorld(pa, params), 1.0e-4


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_reorder_layouts_do_not_change_results(transfers, mode, monkeypatch):
    """The reorder writes each block's particles round-robin by cell, or keeps their order in a
    block of more than 1,024 particles or with more than 128 in one cell.  P2G's sums do not depend
    on the order inside a block, so reordering every substep and only once give the same state bit
    for bit."""
    world, dt = _dense_blocks()
    monkeypatch.setenv("TMPM_REORDER_INTERVAL", "1")
    every = HipMpmResidentSession(transfers, world, dt, mode)
    monkeypatch.setenv("TMPM_REORDER_INTERVAL", "1000000")
    once = HipMpmResidentSession(transfers, world, dt, mode)
    assert every.launch_config()["reorder_interval"] == 1 and once.launch_config()["reorder_interval"] == 1000000
    every.step(40)
    once.step(40)
    assert every.valid and once.valid and every.world().substeps == 40
    assert every.topology_stats()["particle_blocks"] >= 3
    assert_same(every.read_particles(), once.read_particles())
