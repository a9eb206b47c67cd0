"""GPU parity of the HIP backend against the FP64 oracle, following the
reference's own GPU<->CPU tests and tolerance tiers
(crates/tessera-mpm/src/gpu.rs:2332-5113).  Norms are Euclidean (vectors)
and Frobenius (matrices), as in the reference's assertions."""

import math

import numpy as np
import pytest

from oracle import oracle as O
from tessera_amd.mpm import (Capacity, HipMpmResidentSession, InvalidInput, MaterialModel, MpmParams, MpmParticle,
                             MpmWorld, ParticleArrays, PendingSteps, ResidentUnavailable, RigidObstacle, UnknownChunk,
                             WorldBounds)

pytestmark = pytest.mark.gpu


def world_of(particles, params, obstacles=()):
    w = MpmWorld(particles, params)
    w.set_obstacles(list(obstacles))
    return w


def max_err(a, b):
    return float(np.max(np.linalg.norm((np.asarray(a) - np.asarray(b)).reshape(len(a), -1), axis=1))) if len(a) else 0.0


def assert_tracks(ref: dict, got: ParticleArrays, pos, vel, aff=None, F=None, mask=None):
    sel = slice(None) if mask is None else mask
    assert max_err(ref["position"][sel], got.position[sel]) < pos
    assert max_err(ref["velocity"][sel], got.velocity[sel]) < vel
    if aff is not None:
        assert max_err(ref["affine"][sel], got.affine[sel]) < aff
    if F is not None:
        assert max_err(ref["deformation"][sel], got.deformation[sel]) < F


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_transfers_track_cpu_reference_for_elastic_particles(transfers, mode):
    # gpu.rs:4043-4098 (step of 0.002 with CFL substeps; 1e-5/1e-4/1e-3/1e-5)
    a = MpmParticle.new([0.35, 0.43, 0.53], 0.04, 1000.0, MaterialModel.elastic(2000.0, 0.2))
    b = MpmParticle.new([0.41, 0.47, 0.56], 0.04, 1000.0, MaterialModel.neo_hookean(2000.0, 0.2))
    a.velocity = np.array([0.2, -0.1, -0.5])
    b.velocity = np.array([-0.1, 0.3, 0.2])
    a.deformation[0, 0] = 0.97
    b.deformation[1, 1] = 1.04
    params = MpmParams(bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = world_of([a, b], params)
    cpu = O.OracleWorld(w)
    cpu.step(0.002)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    s.step_adaptive(0.002)
    assert s.substeps == cpu.substeps
    assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4, 1e-3, 1e-5)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_grid_boundary_and_sparse_span_match_cpu(transfers, mode):
    # gpu.rs:4574-4703: a particle thrown at a wall stays inside it; two
    # particles 10^4 apart keep a sparse grid of exactly their 2 x 27 nodes;
    # adjacent particles at negative coordinates; unrepresentable particles
    # are rejected.
    elastic = MaterialModel.elastic(1000.0, 0.2)
    falling = MpmParticle.new([0.5, 0.5, 0.05], 0.04, 1000.0, elastic)
    falling.velocity = np.array([0.0, 0.0, -10.0])
    w = world_of([falling], MpmParams(gravity=np.zeros(3), bounds=WorldBounds(np.zeros(3), np.ones(3))))
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    s.step_adaptive(0.001)
    got = s.read_particles()
    assert max_err(cpu.read()["velocity"], got.velocity) < 1e-4
    assert got.position[0, 2] >= got.radius[0]

    far = [MpmParticle.new([0.1] * 3, 0.04, 1000.0, elastic), MpmParticle.new([10000.0] * 3, 0.04, 1000.0, elastic)]
    w = world_of(far, MpmParams())
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    assert s.topology_stats()["live_nodes"] == 54
    s.step_adaptive(0.001)
    assert s.substeps == cpu.substeps
    assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4, F=1e-5)

    near = [MpmParticle.new([-0.11, -0.08, -0.05], 0.04, 1000.0, elastic),
            MpmParticle.new([-0.08, -0.07, -0.04], 0.04, 1000.0, elastic)]
    w = world_of(near, MpmParams())
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep(), mode)
    s.step_adaptive(0.001)
    assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4)

    for bad in (MpmParticle.new([0.5] * 3, 1e-20, 1000.0, elastic),
                MpmParticle.new([300_000_000.0] * 3, 0.04, 1000.0, elastic)):
        w = world_of([bad], MpmParams())
        with pytest.raises(InvalidInput):
            HipMpmResidentSession(transfers, w, 1e-4, mode)


def test_transfers_track_cpu_with_materials_fixed_and_disabled_particles(transfers):
    # gpu.rs:4288-4350: all five materials, a force, damping, fixed and disabled particles
    mats = [MaterialModel.elastic(1000.0, 0.2), MaterialModel.neo_hookean(1000.0, 0.2),
            MaterialModel.sand(1000.0, 0.2, math.radians(35.0), 0.0), MaterialModel.fluid(2000.0, 7.0, 0.01),
            MaterialModel.snow(1000.0, 0.2)]
    ps = [MpmParticle.new([0.3 + i * 0.035, 0.5, 0.5], 0.03, 1000.0, m) for i, m in enumerate(mats)]
    ps[0].force = np.array([0.1, -0.2, 0.3])
    ps[1].damping = 0.4
    fixed = MpmParticle.new([0.48, 0.5, 0.5], 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    fixed.fixed = True
    disabled = MpmParticle.new([0.52, 0.5, 0.5], 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    disabled.enabled = False
    ps += [fixed, disabled]
    w = world_of(ps, MpmParams())
    cpu = O.OracleWorld(w)
    cpu.step(0.002)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    s.step_adaptive(0.002)
    got = s.read_particles()
    ref = cpu.read()
    assert_tracks(ref, got, 1e-5, 1e-4, F=1e-5)
    assert np.abs(ref["plastic"] - got.plastic).max() < 1e-5
    assert list(got.enabled) == [True] * 6 + [False]
    assert list(got.fixed) == [False] * 5 + [True, False]


def test_corotated_stress_tracks_cpu_for_elastic_sand_and_snow(transfers):
    # gpu.rs:4423-4516: sheared elastic, sand and snow (plastic_det 0.8) plus two
    # sand particles compressed and stretched past yield; 12 steps of 0.001
    mats = [MaterialModel.elastic(4000.0, 0.2), MaterialModel.sand(4000.0, 0.2, math.radians(30.0), 0.0),
            MaterialModel.snow(4000.0, 0.2)]
    ps = []
    for i, m in enumerate(mats):
        p = MpmParticle.new([0.25 + i * 0.25, 0.4, 0.5], 0.04, 1000.0, m)
        p.deformation = np.array([[1.1, 0.15, 0.02], [0.0, 0.88, 0.08], [0.0, 0.0, 1.04]])
        if i == 2:
            p.plastic.plastic_det = 0.8
        ps.append(p)
    for x, diag in [(0.25, [0.95, 0.8, 0.75]), (0.5, [1.2, 0.9, 0.7])]:
        p = MpmParticle.new([x, 0.7, 0.5], 0.04, 1000.0, MaterialModel.sand(4000.0, 0.2, math.radians(30.0), 0.0))
        p.deformation = np.diag(diag)
        ps.append(p)
    w = world_of(ps, MpmParams(gravity=np.zeros(3)))
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    # the first step's return mapping, to the reference's per-transfer bound
    cpu.step(0.001)
    s.step_adaptive(0.001)
    ref, got = cpu.read(), s.read_particles()
    for i in (1, 2, 3, 4):
        assert np.linalg.norm(ref["deformation"][i] - got.deformation[i]) < 2e-5
        assert np.abs(ref["plastic"][i] - got.plastic[i]).max() < 2e-5
    assert got.plastic[4, 1] > 0.0  # the stretched sand particle hardened
    for _ in range(11):
        cpu.step(0.001)
        s.step_adaptive(0.001)
    assert s.substeps == cpu.substeps
    ref, got = cpu.read(), s.read_particles()
    assert_tracks(ref, got, 2e-5, 3e-4, F=2e-4)
    assert np.abs(ref["plastic"][:, 0] - got.plastic[:, 0]).max() < 2e-4


def test_near_singular_plastic_deformation_takes_the_fp64_fallback(transfers):
    # gpu.rs:4518-4570: F(0, 0) = 1e-6; the reference falls back to the CPU, this
    # backend to its FP64 SVD on the device
    mats = [MaterialModel.elastic(2000.0, 0.2), MaterialModel.sand(2000.0, 0.2, math.radians(30.0), 0.0),
            MaterialModel.snow(2000.0, 0.2)]
    ps = []
    for i, m in enumerate(mats):
        p = MpmParticle.new([0.2 + i * 0.25, 0.4, 0.4], 0.04, 1000.0, m)
        p.deformation[0, 0] = 1e-6
        ps.append(p)
    w = world_of(ps, MpmParams(gravity=np.zeros(3)))
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    s.step_adaptive(0.001)
    ref, got = cpu.read(), s.read_particles()
    assert_tracks(ref, got, 1e-5, 1e-4, F=2e-4)
    assert np.abs(ref["plastic"][:, 0] - got.plastic[:, 0]).max() < 2e-4


def _euler(roll, pitch, yaw):
    """nalgebra's from_euler_angles: Rz(yaw) Ry(pitch) Rx(roll)."""
    cr, sr, cp, sp, cy, sy = (math.cos(roll), math.sin(roll), math.cos(pitch), math.sin(pitch), math.cos(yaw),
                              math.sin(yaw))
    rx = np.array([[1, 0, 0], [0, cr, -sr], [0, sr, cr]])
    ry = np.array([[cp, 0, sp], [0, 1, 0], [-sp, 0, cp]])
    rz = np.array([[cy, -sy, 0], [sy, cy, 0], [0, 0, 1]])
    return rz @ ry @ rx


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_sand_neo_hookean_tracks_cpu(transfers, mode):
    # gpu.rs:2339-2440: identity, rotation, isotropic stretch and compression,
    # rotated anisotropic stretch; hardening 0.2 i; 4 substeps of 1e-4
    r = _euler(0.2, -0.4, 0.7)
    defs = [np.eye(3), r, np.eye(3) * 1.1, np.diag([0.8, 0.79, 0.81]),
            r @ np.diag([math.exp(0.15), math.exp(-0.25), math.exp(-0.1)])]
    ps = []
    for i, F in enumerate(defs):
        p = MpmParticle.new([0.5 + i * 0.4, 0.5, 0.5], 0.04, 1000.0,
                            MaterialModel.sand_neo_hookean(4000.0, 0.2, math.radians(35.0), 0.02))
        p.deformation = F
        p.plastic.hardening += i * 0.2
        ps.append(p)
    w = world_of(ps, MpmParams(gravity=np.zeros(3), bounds=WorldBounds(np.zeros(3), np.full(3, 3.0))))
    cpu = O.OracleWorld(w)
    cpu.substeps_fixed(1e-4, 4)
    s = HipMpmResidentSession(transfers, w, 1e-4, mode)
    s.step(4)
    ref, got = cpu.read(), s.read_particles()
    assert_tracks(ref, got, 2e-5, 3e-4, 2e-3, 2e-4)
    assert np.abs(ref["plastic"] - got.plastic).max() < 2e-4
    assert got.plastic[2, 1] > ps[2].plastic.hardening and got.plastic[4, 1] > ps[4].plastic.hardening


def test_sand_hardening_tracks_multiload_cpu_and_resident_state(transfers):
    # gpu.rs:2442-2524: sand and sand-NH loaded four times (a fresh substep from
    # each load, carrying the device's plastic state), then 5 resident substeps
    mats = [MaterialModel.sand(4000.0, 0.2, math.radians(35.0), 0.0),
            MaterialModel.sand_neo_hookean(4000.0, 0.2, math.radians(35.0), 0.0)]
    ps = [MpmParticle.new([0.3 + i * 0.4, 0.5, 0.5], 0.04, 1000.0, m) for i, m in enumerate(mats)]
    params = MpmParams(gravity=np.zeros(3))
    dt = 1e-5
    for load in [np.array([0.15, -0.25, -0.1])] * 2 + [np.full(3, math.log(1.2)), np.array([0.15, -0.25, -0.1])]:
        for p in ps:
            p.deformation = np.diag(np.exp(load))
        w = world_of(ps, params)
        cpu = O.OracleWorld(w)
        cpu.substeps_fixed(dt, 1)
        s = HipMpmResidentSession(transfers, w, dt)
        s.step(1)
        ref, got = cpu.read(), s.read_particles()
        for i, p in enumerate(ps):
            assert np.linalg.norm(ref["deformation"][i] - got.deformation[i]) < 2e-5
            assert np.abs(ref["plastic"][i] - got.plastic[i]).max() < 2e-5
            assert got.plastic[i, 1] > p.plastic.hardening
            p.plastic.plastic_det, p.plastic.hardening, p.plastic.log_volume_gain = map(float, got.plastic[i])
            p.deformation = got.deformation[i].copy()
            p.velocity = np.zeros(3)
            p.affine = np.zeros((3, 3))
    w = world_of(ps, params)
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, dt)
    s.submit_steps(5)
    cpu.substeps_fixed(dt, 5)
    s.synchronize()
    ref, got = cpu.read(), s.read_particles()
    for i in range(2):
        assert np.linalg.norm(ref["deformation"][i] - got.deformation[i]) < 3e-5
        assert np.abs(ref["plastic"][i] - got.plastic[i]).max() < 3e-5


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_plastic_blocks_in_a_batch_equal_the_same_scene_alone(transfers, mode):
    # plastic state follows its particle through reorders and batch readback
    from tessera_amd.mpm import HipMpmBatch
    from tessera_amd.scenes import jittered_lattice, _particles

    worlds = []
    for k, m in enumerate([MaterialModel.snow(2.0e4, 0.3), MaterialModel.sand(2.0e4, 0.3, math.radians(30.0), 0.0),
                           MaterialModel.sand_neo_hookean(2.0e4, 0.3, math.radians(30.0), 0.01)]):
        rng = np.random.default_rng(k)
        h = 1.0 / 64
        pos = jittered_lattice(rng, [10, 10, 10], [0.4, 0.1, 0.4], [0.4 + 10 * h / 2, 0.1 + 10 * h / 2, 0.4 + 10 * h / 2])
        vel = rng.normal(0.0, 0.3, size=pos.shape)
        pa = _particles(pos, vel, h / 4, 1000.0, m)
        floor = RigidObstacle.ground([0.5, 0.08, 0.5], [0.4, 0.4], quat_x(-math.pi / 2))
        floor.boundary = 1
        w = MpmWorld(pa, MpmParams(cell_width=h, gravity=np.array([0.0, -9.8, 0.0]), max_substep=1e-4,
                                   bounds=WorldBounds(np.zeros(3), np.ones(3))))
        w.set_obstacles([floor])
        worlds.append(w)
    b = HipMpmBatch(transfers, worlds, [1e-4] * 3, mode)
    b.step(40)
    for k, w in enumerate(worlds):
        s = HipMpmResidentSession(transfers, w, 1e-4, mode)
        s.step(40)
        a, g = s.read_particles(), b.read_particles(k)
        for f in ("position", "velocity", "deformation", "plastic"):
            assert np.array_equal(getattr(a, f), getattr(g, f)), (k, f)
        assert np.any(g.plastic != np.array([1.0, 0.0, 0.0]))  # the floor impact yielded


def quat_x(angle):
    from tessera_amd.mpm import quat_from_axis_angle

    return quat_from_axis_angle([1, 0, 0], angle)


def test_neo_hookean_and_viscous_fluid_stress_track_cpu(transfers):
    # gpu.rs:4353-4420 (3 x 0.001; 2e-5 / 3e-4 / 2e-4)
    solid = MpmParticle.new([0.3, 0.4, 0.5], 0.04, 1000.0, MaterialModel.neo_hookean(4000.0, 0.25))
    solid.deformation = np.array([[1.2, 0.1, 0.0], [0.0, 0.85, 0.05], [0.0, 0.0, 1.1]])
    solid.affine = np.array([[0.2, 0.1, 0.0], [0.0, -0.1, 0.0], [0.0, 0.0, 0.1]])
    fluid = MpmParticle.new([0.7, 0.6, 0.5], 0.04, 1000.0, MaterialModel(MaterialModel.FLUID, (3000.0, 6.0, 12.0, 0.3)))
    fluid.deformation = np.diag([0.9, 1.05, 1.0])
    fluid.affine = np.array([[0.4, 0.2, 0.0], [-0.1, -0.2, 0.0], [0.0, 0.0, 0.1]])
    stretched = fluid.clone()
    stretched.position = np.array([0.7, 0.2, 0.5])
    stretched.deformation = np.diag([1.1, 1.03, 1.0])
    w = world_of([solid, fluid, stretched], MpmParams(gravity=np.zeros(3)))
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    s.step_adaptive(0.001)
    first = s.read_particles()
    for F in first.deformation[1:]:
        assert np.linalg.norm(F - np.eye(3) * F[0, 0]) < 1e-6
    cpu.step(0.001)
    for _ in range(2):
        cpu.step(0.001)
        s.step_adaptive(0.001)
    assert s.substeps == cpu.substeps
    assert_tracks(cpu.read(), s.read_particles(), 2e-5, 3e-4, F=2e-4)


def test_resident_substeps_match_cpu_with_force_boundary_and_fixed_particles(transfers):
    # gpu.rs:3634-3774
    ps = [MpmParticle.new([0.42, 0.48, 0.51], 0.04, 1000.0, MaterialModel.neo_hookean(2000.0, 0.2)),
          MpmParticle.new([0.48, 0.48, 0.0401], 0.04, 1000.0, MaterialModel.fluid(2000.0, 7.0, 0.1)),
          MpmParticle.new([0.6, 0.5, 0.5], 0.04, 1000.0, MaterialModel.neo_hookean(2000.0, 0.2))]
    ps[0].force = np.array([0.2, -0.1, 0.3])
    ps[0].damping = 0.4
    ps[1].velocity[2] = -2.0
    ps[2].fixed = True
    ps[2].velocity[0] = 1.0
    params = MpmParams(bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = world_of(ps, params)
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, 0.0001)
    for _ in range(2):
        cpu.substeps_fixed(0.0001, 4)
        s.step(4)
    assert s.substeps == 8
    assert_tracks(cpu.read(), s.read_particles(), 2e-5, 2e-4, 2e-3, 2e-4)
    # obstacle added between batches
    sphere = RigidObstacle.sphere([0.5, 0.5, 0.5], 0.1)
    cpu.set_obstacles([sphere])
    s.set_obstacles([sphere])
    cpu.substeps_fixed(0.0001, 4)
    s.step(4)
    assert_tracks(cpu.read(), s.read_particles(), 2e-5, 2e-4)

    # an obstacle that pushes the particle out of the bounded grid fails the
    # whole batch and leaves the world unchanged
    escape = MpmParticle.new([0.95, 0.5, 0.5], 0.04, 1000.0, MaterialModel.fluid(2000.0, 7.0, 0.1))
    ew = world_of([escape], params, [RigidObstacle.sphere([0.9, 0.5, 0.5], 10.0)])
    es = HipMpmResidentSession(transfers, ew, 0.0001)
    with pytest.raises(ResidentUnavailable):
        es.step(4)
    assert es.substeps == 0
    assert np.array_equal(es.read_particles().position[0], [0.95, 0.5, 0.5])

    # a force that breaks the CFL bound fails; the world keeps its force
    acc = MpmParticle.new([0.5, 0.5, 0.5], 0.04, 1000.0, MaterialModel.neo_hookean(2000.0, 0.2))
    acc.force[0] = 4.0e6
    uw = world_of([acc], params)
    us = HipMpmResidentSession(transfers, uw, 0.0001)
    with pytest.raises(ResidentUnavailable):
        us.step(4)
    assert us.substeps == 0
    with pytest.raises(ResidentUnavailable):
        us.step(1)
    assert us.read_particles().force[0, 0] == 4.0e6
    ok = HipMpmResidentSession(transfers, uw, 0.000001)
    ok.step(4)
    assert ok.substeps == 4
    assert ok.read_particles().force[0, 0] == 0.0


def test_resident_sparse_topology_rebuilds_for_separated_moving_particles(transfers):
    # gpu.rs:3473-3559 (unbounded sparse grid, 32 substeps; 2e-4/2e-3/2e-2)
    ps = [MpmParticle.new([x, 0.5, 0.5], 0.03, 1000.0, MaterialModel.fluid(2000.0, 7.0, 0.1))
          for x in (-20.055, -20.015, 20.055, 20.095, -50.0, 0.0)]
    for i in range(4):
        ps[i].velocity[0] = 1.0 if i % 2 == 0 else 0.2
    ps[4].fixed = True
    ps[5].enabled = False
    w = world_of(ps, MpmParams(gravity=np.zeros(3), cell_width=0.1))
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, 0.001)
    initial_cell = math.floor(-20.055 / 0.1 - 0.5)
    for _ in range(4):
        for _ in range(8):
            cpu.step(0.001)
        s.step(8)
        assert_tracks(cpu.read(), s.read_particles(), 2e-4, 2e-3, 2e-2)
    got = s.read_particles()
    assert math.floor(got.position[0, 0] / 0.1 - 0.5) != initial_cell
    assert s.substeps == 32
    assert np.array_equal(got.position[4], ps[4].position)
    assert np.array_equal(got.position[5], ps[5].position)


def test_live_node_counts_shared_and_disabled_stencils(transfers):
    # gpu.rs:3163-3303: exact live-node count of the sparse grid (63, 81) and
    # no topology at all when nothing is enabled; checked against the oracle's
    # node set key by key.
    for xs, enabled, expected in (([0.05, 0.15, 0.05, 10.05, 20.05], True, 63),
                                  ([0.05, 10.05, 20.05, 20.05, 30.05], True, 81),
                                  ([0.05, 0.15, 0.05, 10.05, 20.05], False, 0)):
        ps = []
        for x in xs:
            p = MpmParticle.new([x, 0.05, 0.05], 0.03, 1000.0, MaterialModel.fluid(2000.0, 7.0, 0.1))
            p.fixed = True
            p.enabled = enabled
            ps.append(p)
        ps[4].enabled = False
        w = world_of(ps, MpmParams(gravity=np.zeros(3), cell_width=0.1,
                                   bounds=WorldBounds(np.full(3, -100.0), np.full(3, 100.0))))
        s = HipMpmResidentSession(transfers, w, 0.001)
        for _ in range(2):
            s.step(1)
            stats = s.topology_stats()
            assert stats["live_nodes"] == expected
            nodes = s.export_nodes()
            ref = O.node_set(w.particles.position, 0.1, w.particles.enabled.astype(np.uint8))
            assert np.array_equal(nodes, ref)
        assert s.substeps == 2


def test_unbounded_resident_rejects_cell_overflow_without_applying_state(transfers):
    # gpu.rs:3411-3470
    p = MpmParticle.new([214748352.0, 0.05, 0.05], 0.01, 1000.0, MaterialModel.fluid(1.0, 7.0, 0.0))
    p.force[0] = 1e15
    w = world_of([p], MpmParams(cell_width=0.1, gravity=np.zeros(3)))
    s = HipMpmResidentSession(transfers, w, 0.001)
    s.submit_steps(1)
    s.submit_steps(2)
    with pytest.raises(ResidentUnavailable):
        s.synchronize()
    assert s.substeps == 0
    got = s.read_particles()
    assert np.array_equal(got.position[0], p.position)
    assert np.array_equal(got.velocity[0], p.velocity)


def test_pending_steps_block_edits_and_free_fall_matches_gravity(transfers):
    p = MpmParticle.new([0.5, 0.5, 0.5], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    w = world_of([p], MpmParams(max_substep=0.001))
    s = HipMpmResidentSession(transfers, w, 0.001)
    s.submit_steps(1)
    assert s.has_pending_steps() and s.pending_substeps() == 1
    with pytest.raises(PendingSteps):
        s.set_particle_force(0, [1.0, 0.0, 0.0])
    with pytest.raises(PendingSteps):
        s.step(1)
    s.synchronize()
    got = s.read_particles()
    # world.rs:660-672 tolerance relaxed to f32 (gpu.rs:5066: 1e-5)
    assert abs(got.velocity[0, 2] + 0.00981) < 1e-5
    assert abs(got.position[0, 2] - (0.5 - 0.00000981)) < 1e-5
    with pytest.raises(InvalidInput):
        s.submit_steps(0)
    with pytest.raises(InvalidInput):
        s.submit_steps(65)


def test_force_edit_is_consumed_once(transfers):
    # gpu.rs:3306-3408: v = F t / m after one substep, then unchanged
    p = MpmParticle.new([0.5, 0.5, 0.5], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    w = world_of([p], MpmParams(gravity=np.zeros(3), max_substep=0.001))
    s = HipMpmResidentSession(transfers, w, 0.001)
    s.set_particle_force(0, [2.0, 0.0, 0.0])
    s.step(1)
    v1 = s.read_particles().velocity[0, 0]
    assert abs(v1 - 2.0 * 0.001 / p.mass) < 1e-5
    s.step(3)
    assert abs(s.read_particles().velocity[0, 0] - v1) < 1e-5
    with pytest.raises(InvalidInput):
        s.set_particle_force(5, [1.0, 0.0, 0.0])
    with pytest.raises(InvalidInput):
        s.set_particle_force(0, [math.nan, 0.0, 0.0])


def test_chunks_add_and_remove_preserve_order_and_state(transfers):
    # gpu.rs:3934-4040
    base = [MpmParticle.new([0.3 + 0.05 * i, 0.5, 0.5], 0.02, 1000.0, MaterialModel.elastic(1000.0, 0.2))
            for i in range(3)]
    w = world_of(base, MpmParams(gravity=np.zeros(3)))
    s = HipMpmResidentSession(transfers, w, 1e-4)
    s.step(2)
    extra = [MpmParticle.new([0.6, 0.6, 0.6], 0.02, 1000.0, MaterialModel.fluid(1000.0, 7.0, 0.0))]
    chunk = s.add_particles(extra)
    assert chunk == 1
    got = s.read_particles()
    assert len(got) == 4 and list(got.chunk_id) == [0, 0, 0, 1]
    s.step(2)
    assert s.remove_chunk(0) == 3
    assert len(s.read_particles()) == 1
    with pytest.raises(UnknownChunk):
        s.remove_chunk(0)
    assert s.remove_chunk(1) == 1
    assert len(s.read_particles()) == 0
    s.step(1)  # empty world: counts substeps only
    assert s.substeps == 5


def test_snapshot_matches_state_after_steps(transfers):
    # gpu.rs:2964-3160: 32-byte records in world order, exact f32 mass/volume
    from tests.hip_util import DeviceBuffer

    ps = [MpmParticle.new([0.3 + 0.04 * i, 0.5, 0.5], 0.02, 1000.0, MaterialModel.fluid(2000.0, 7.0, 0.1))
          for i in range(5)]
    ps[2].enabled = False
    w = world_of(ps, MpmParams())
    s = HipMpmResidentSession(transfers, w, 1e-4)
    s.step(3)
    out = DeviceBuffer(5 * 32)
    n = s.encode_particle_snapshot(out.ptr, out.nbytes)
    s.wait()
    assert n == 5
    snap = out.read(np.float32).reshape(5, 8)
    got = s.read_particles()
    assert np.max(np.abs(snap[:, :3] - got.position)) < 1e-6
    assert np.max(np.abs(snap[:, 4:7] - got.velocity)) < 1e-6
    assert list(snap[:, 3]) == [np.float32(m) if e else 0.0 for m, e in zip(w.particles.mass, w.particles.enabled)]
    assert np.array_equal(snap[:, 7], w.particles.rest_volume.astype(np.float32))
    with pytest.raises(Capacity):
        s.encode_particle_snapshot(out.ptr, 16)


def test_rigid_obstacles_match_cpu_one_way_coupling(transfers):
    # gpu.rs:4706-4926 for the shapes the HIP backend evaluates on device
    from tessera_amd.mpm import ObstacleBoundary, quat_from_axis_angle

    cases = [
        RigidObstacle.sphere([0.5, 0.5, 0.38], 0.08),
        RigidObstacle.cuboid([0.5, 0.5, 0.36], [0.1, 0.1, 0.06], quat_from_axis_angle([0, 0, 1], 0.3)),
        RigidObstacle.capsule([0.5, 0.5, 0.38], 0.05, 0.06, quat_from_axis_angle([1, 0, 0], 1.1)),
        RigidObstacle.cylinder([0.5, 0.5, 0.38], 0.06, 0.08),
        RigidObstacle.cone([0.5, 0.5, 0.38], 0.06, 0.08),
        RigidObstacle.ground([0.5, 0.5, 0.4], [0.3, 0.3]),
    ]
    for boundary in range(5):
        for ob in cases:
            ob.boundary = boundary
            ob.friction = 0.3
            ob.linear_velocity = np.array([0.1, 0.0, 0.05])
            ps = [MpmParticle.new([0.48 + 0.02 * i, 0.5, 0.47], 0.01, 1000.0, MaterialModel.elastic(2000.0, 0.2))
                  for i in range(3)]
            for p in ps:
                p.velocity = np.array([0.0, 0.1, -0.8])
            w = world_of(ps, MpmParams(), [ob])
            cpu = O.OracleWorld(w)
            s = HipMpmResidentSession(transfers, w, 5e-4)
            for _ in range(4):
                cpu.substeps_fixed(5e-4, 1)
            s.step(4)
            assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4)


def _reference_shape_cases():
    """The 14 (obstacle, particle position) cases of gpu.rs:4719-4797."""
    from tessera_amd.mpm import quat_from_axis_angle

    def moving(ob, lin=(0.0, 0.0, 0.0), ang=(0.0, 0.0, 0.0), friction=0.0):
        ob.linear_velocity = np.array(lin, float)
        ob.angular_velocity = np.array(ang, float)
        ob.friction = friction
        return ob

    c = [0.5, 0.5, 0.5]
    tilted = quat_from_axis_angle([0, 1, 0], math.pi / 2)
    tri = [[-0.15, -0.15, 0.0], [0.15, -0.15, 0.0], [0.0, 0.15, 0.0]]
    X, Y, Z = np.eye(3)
    convex = lambda: moving(RigidObstacle.convex([0.45, 0.45, 0.45], [np.zeros(3), X * 0.2, Y * 0.2, Z * 0.2],
                                                 [-X, -Y, -Z, np.ones(3) / math.sqrt(3.0)]), lin=(-0.5, 0, 0))
    return [
        (moving(RigidObstacle.sphere(c, 0.1), lin=(1.0, 0.2, 0.0), friction=0.3), [0.62, 0.5, 0.5]),
        (moving(RigidObstacle.cuboid(c, [0.1] * 3, quat_from_axis_angle([0, 0, 1], math.pi / 4)), lin=(0.5, 0, 0),
                ang=(0, 0, 0.5), friction=0.2), [0.65, 0.5, 0.5]),
        (moving(RigidObstacle.capsule(c, 0.1, 0.1), lin=(0.5, 0, 0)), [0.62, 0.5, 0.5]),
        (moving(RigidObstacle.capsule(c, 0.1, 0.0), lin=(0.5, 0, 0)), [0.54, 0.5, 0.5]),
        (moving(RigidObstacle.cylinder(c, 0.1, 0.1), lin=(0.5, 0, 0)), [0.62, 0.5, 0.5]),
        (moving(RigidObstacle.cone(c, 0.1, 0.1), lin=(0.5, 0, 0)), [0.56, 0.5, 0.5]),
        (moving(RigidObstacle.capsule(c, 0.1, 0.1, tilted), lin=(0, 0, 0.5)), [0.5, 0.5, 0.62]),
        (moving(RigidObstacle.cylinder(c, 0.1, 0.1, tilted), lin=(0, 0, 0.5)), [0.5, 0.5, 0.62]),
        (moving(RigidObstacle.cone(c, 0.1, 0.1, tilted), lin=(0, 0.5, 0)), [0.5, 0.56, 0.5]),
        (moving(RigidObstacle.triangle_prism(c, tri, 1e-4), lin=(0, 0, 0.5)), [0.5, 0.45, 0.54]),
        (moving(RigidObstacle.triangle_prism(c, tri, 1e-4, tilted), lin=(0.5, 0, 0)), [0.54, 0.45, 0.5]),
        (convex(), [0.43, 0.48, 0.48]),
        (convex(), [0.43, 0.42, 0.48]),
        (convex(), [0.69, 0.47, 0.47]),
    ]


@pytest.mark.parametrize("case", range(14))
def test_every_obstacle_shape_matches_cpu_one_way_coupling(transfers, case):
    # gpu.rs:4706-4827: one particle against each shape, step(0.002), 1e-5 / 1e-4
    ob, position = _reference_shape_cases()[case]
    p = MpmParticle.new(position, 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    w = world_of([p], MpmParams(gravity=np.zeros(3)), [ob])
    cpu = O.OracleWorld(w)
    cpu.step(0.002)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    s.step_adaptive(0.002)
    assert s.substeps == cpu.substeps
    assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4)
    got = s.obstacles()[0]
    assert got.shape == ob.shape and np.allclose(got.center, ob.center)


def test_two_convex_obstacles_use_their_own_planes(transfers):
    # gpu.rs:4858-4920: planes of consecutive convex obstacles are packed back to back
    X, Y, Z = np.eye(3)

    def make(center, speed):
        o = RigidObstacle.convex(center, [np.zeros(3), X * 0.15, Y * 0.15, Z * 0.15],
                                 [-X, -Y, -Z, np.ones(3) / math.sqrt(3.0)])
        o.linear_velocity = np.array([speed, 0.0, 0.0])
        return o

    ps = [MpmParticle.new([x, 0.48, 0.48], 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2)) for x in (0.28, 0.68)]
    w = world_of(ps, MpmParams(gravity=np.zeros(3)), [make([0.3, 0.45, 0.45], -0.5), make([0.7, 0.45, 0.45], -1.0)])
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, w.stable_timestep())
    s.step_adaptive(0.001)
    assert_tracks(cpu.read(), s.read_particles(), 1e-5, 1e-4)
    moved = np.linalg.norm(s.read_particles().velocity, axis=1)
    assert np.all(moved > 0.1)  # both particles were pushed by their own obstacle


def test_convex_with_more_planes_than_the_device_holds_is_a_capacity_error(transfers):
    # gpu.rs:4929-4950: more than 128 planes is Capacity, not silently truncated
    k = 130
    angles = np.linspace(0.0, 2.0 * math.pi, k, endpoint=False)
    normals = [np.array([math.cos(a), math.sin(a), 0.0]) for a in angles[: k - 2]] + [np.array([0, 0, 1.0]),
                                                                                        np.array([0, 0, -1.0])]
    verts = [np.array([math.cos(a), math.sin(a), z]) * 0.1 for a in angles for z in (-1.0, 1.0)]
    ob = RigidObstacle.convex([0.5, 0.5, 0.5], verts, normals)
    p = MpmParticle.new([0.2, 0.2, 0.2], 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    w = world_of([p], MpmParams(gravity=np.zeros(3)), [ob])
    with pytest.raises(Capacity):
        HipMpmResidentSession(transfers, w, 1e-3)


def test_transfer_colors_separate_grid_momentum(transfers):
    # gpu.rs:4151-4198: two coincident particles of opposite velocity and colours 0 / 1
    first = MpmParticle.new([0.5, 0.5, 0.5], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    first.velocity = np.array([1.0, 0.0, 0.0])
    second = MpmParticle.new([0.5, 0.5, 0.5], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    second.velocity = np.array([-1.0, 0.0, 0.0])
    second.transfer_color = 1
    w = world_of([first, second], MpmParams(gravity=np.zeros(3), bounds=WorldBounds(np.zeros(3), np.ones(3))))
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    s = HipMpmResidentSession(transfers, w, 0.001)
    s.step(1)
    ref, got = cpu.read(), s.read_particles()
    assert np.all(np.abs(got.velocity[:, 0] - ref["velocity"][:, 0]) < 1e-4)
    assert np.all(np.abs(got.position[:, 0] - ref["position"][:, 0]) < 1e-5)
    assert got.velocity[0, 0] > 0.9 and got.velocity[1, 0] < -0.9


def _cpic_sheet():
    sheet = RigidObstacle.triangle_prism([0.5, 0.5, 0.5], [[0.0, -1.0, -1.0], [0.0, 1.0, -1.0], [0.0, 0.0, 1.0]], 0.001)
    sheet.boundary = 3  # NonReflecting: the sheet only separates the grid
    return sheet


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
@pytest.mark.parametrize("group,left_color,all_groups,enabled",
                         [(0, 0, False, True), (16, 1, False, True), (31, 0, False, True), (31, 0, True, True),
                          (0, 0, False, False), (31, 1, True, False)])
def test_cpic_groups_separate_the_grid(transfers, mode, group, left_color, all_groups, enabled):
    # gpu.rs:4200-4283: a thin CPIC sheet between two particles moving apart
    import copy

    left = MpmParticle.new([0.48, 0.5, 0.5], 0.005, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    left.velocity = np.array([-1.0, 0.0, 0.0])
    left.transfer_color = left_color
    right = MpmParticle.new([0.52, 0.5, 0.5], 0.005, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    right.velocity = np.array([1.0, 0.0, 0.0])
    sheets = []
    for g in (range(32) if all_groups else [group]):
        sh = copy.deepcopy(_cpic_sheet())
        sh.cpic_group = g
        sheets.append(sh)
    w = world_of([left, right], MpmParams(gravity=np.zeros(3)), sheets)
    w.use_cpic = enabled
    cpu = O.OracleWorld(w)
    cpu.step(0.001)
    ref = cpu.read()
    if enabled or left_color != 0:
        assert ref["velocity"][0, 0] < -0.9 and ref["velocity"][1, 0] > 0.9
    else:
        assert abs(ref["velocity"][0, 0]) < 0.5 and abs(ref["velocity"][1, 0]) < 0.5
    s = HipMpmResidentSession(transfers, w, 0.001, mode)
    s.step(1)
    got = s.read_particles()
    assert np.all(np.abs(got.velocity[:, 0] - ref["velocity"][:, 0]) < 1e-3)
    assert np.all(np.abs(got.position[:, 0] - ref["position"][:, 0]) < 1e-5)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_cpic_blocks_of_particles_on_both_sides_track_the_oracle(transfers, mode):
    # Two lattices pressed against a thin sheet from both sides for 20 substeps:
    # the side masks split every straddled block, and both sides stay apart.
    from tessera_amd.scenes import jittered_lattice, _particles

    rng = np.random.default_rng(5)
    h = 1.0 / 64
    sp = h / 2
    a = jittered_lattice(rng, [6, 8, 8], [0.5 - 6 * sp - 0.3 * h, 0.42, 0.42], [0.5 - 0.3 * h, 0.42 + 8 * sp, 0.42 + 8 * sp])
    b = jittered_lattice(rng, [6, 8, 8], [0.5 + 0.3 * h, 0.42, 0.42], [0.5 + 6 * sp + 0.3 * h, 0.42 + 8 * sp, 0.42 + 8 * sp])
    pos = np.concatenate([a, b])
    vel = np.zeros_like(pos)
    vel[: len(a), 0] = 0.4
    vel[len(a):, 0] = -0.4
    pa = _particles(pos, vel, sp / 2, 1000.0, MaterialModel.elastic(3.0e3, 0.3))
    params = MpmParams(cell_width=h, gravity=np.zeros(3), max_substep=1e-4,
                       bounds=WorldBounds(np.zeros(3), np.ones(3)))
    w = MpmWorld(pa, params)
    sheet = _cpic_sheet()
    sheet.cpic_group = 3
    w.set_obstacles([sheet])
    cpu = O.OracleWorld(w)
    s = HipMpmResidentSession(transfers, w, 1e-4, mode)
    cpu.substeps_fixed(1e-4, 20)
    s.step(20)
    ref, got = cpu.read(), s.read_particles()
    assert_tracks(ref, got, 2e-5, 2e-4, 2e-2, 2e-4)
    # no particle crossed the sheet: the two sides never exchanged momentum
    assert np.all(got.position[: len(a), 0] < 0.5) and np.all(got.position[len(a):, 0] > 0.5)
