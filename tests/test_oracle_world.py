"""Pins the FP64 oracle's substep (oracle/mpm_oracle.cpp) against the
closed-form expectations of the reference's own world tests
(crates/tessera-mpm/src/world.rs:646-968) and against invariants of the
method (partition of unity, conservation) that a dropped term, a wrong sign
or a transposed operand breaks."""

import math

import numpy as np
import pytest

from oracle import oracle as O
from tessera_amd.mpm import (MaterialModel, MpmParams, MpmParticle, MpmWorld, ObstacleBoundary, ParticleArrays,
                             RigidObstacle, WorldBounds)


def particle(position):
    # world.rs:650-657
    return MpmParticle.new(position, 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))


def oworld(particles, params=None, obstacles=()):
    w = MpmWorld(particles, params or MpmParams())
    w.set_obstacles(list(obstacles))
    return O.OracleWorld(w)


def test_one_particle_free_fall_matches_gravity():
    # world.rs:660-672
    w = oworld([particle([0.5, 0.5, 0.5])], MpmParams(max_substep=0.001))
    w.step(0.001)
    r = w.read()
    assert abs(r["velocity"][0, 2] + 0.00981) < 1e-10
    assert abs(r["position"][0, 2] - (0.5 - 0.00000981)) < 1e-10


def test_apic_transfer_preserves_uniform_velocity_without_forces():
    # world.rs:675-693
    a = particle([0.45, 0.5, 0.55])
    b = particle([0.55, 0.5, 0.45])
    a.velocity = np.array([1.0, -2.0, 0.5])
    b.velocity = a.velocity.copy()
    w = oworld([a, b], MpmParams(gravity=np.zeros(3), max_substep=0.001))
    w.step(0.001)
    for v in w.read()["velocity"]:
        assert np.linalg.norm(v - [1.0, -2.0, 0.5]) < 1e-10


def test_transfer_colors_keep_opposed_particles_on_separate_grid_nodes():
    # world.rs:696-718
    left = particle([0.5, 0.5, 0.5])
    left.velocity[0] = 1.0
    right = left.clone()
    right.velocity[0] = -1.0
    right.transfer_color = 1
    params = MpmParams(gravity=np.zeros(3), max_substep=0.001)
    sep = oworld([left.clone(), right.clone()], params)
    sep.step(0.001)
    v = sep.read()["velocity"]
    assert abs(v[0, 0] - 1.0) < 1e-10 and abs(v[1, 0] + 1.0) < 1e-10
    right.transfer_color = 0
    mixed = oworld([left, right], params)
    mixed.step(0.001)
    v = mixed.read()["velocity"]
    assert abs(v[0, 0]) < 1e-10 and abs(v[1, 0]) < 1e-10


def test_cpic_triangle_assigns_opposite_sides_to_separate_grid_regions():
    # world.rs:721-781
    left = MpmParticle.new([0.48, 0.5, 0.5], 0.005, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    left.velocity[0] = -1.0
    right = left.clone()
    right.position[0] = 0.52
    right.velocity[0] = 1.0
    sheet = RigidObstacle.triangle_prism([0.5, 0.5, 0.5], [[0.0, -1.0, -1.0], [0.0, 1.0, -1.0], [0.0, 0.0, 1.0]],
                                         0.001)
    sheet.cpic_group = 0
    sheet.boundary = ObstacleBoundary.NON_REFLECTING
    params = MpmParams(gravity=np.zeros(3))

    disabled = oworld([left.clone(), right.clone()], params, [sheet])
    disabled.set_use_cpic(False)
    disabled.step(0.001)
    v = disabled.read()["velocity"]
    assert abs(v[0, 0]) < 0.5 and abs(v[1, 0]) < 0.5

    r2 = right.clone()
    r2.transfer_color = 1
    explicit = oworld([left.clone(), r2], params, [sheet])
    explicit.set_use_cpic(False)
    explicit.step(0.001)
    v = explicit.read()["velocity"]
    assert v[0, 0] < -0.9 and v[1, 0] > 0.9

    separated = oworld([left.clone(), right.clone()], params, [sheet])
    separated.step(0.001)
    v = separated.read()["velocity"]
    assert v[0, 0] < -0.9 and v[1, 0] > 0.9

    sheet.cpic_group = None
    mixed = oworld([left, right], params, [sheet])
    mixed.step(0.001)
    v = mixed.read()["velocity"]
    assert abs(v[0, 0]) < 0.5 and abs(v[1, 0]) < 0.5


def test_obstacle_reaction_opposes_particle_impact():
    # world.rs:784-807
    imp = particle([1.02, 0.0, 0.0])
    imp.velocity[0] = -1.0
    p0 = imp.velocity * imp.mass
    w = oworld([imp], MpmParams(gravity=np.zeros(3)), [RigidObstacle.sphere([0, 0, 0], 1.0)])
    react = w.step_with_obstacle_reactions(0.001, 1)
    lin, ang = react[0, :3], react[0, 3:]
    assert lin[0] < 0.0
    assert abs(lin[1]) < 1e-10
    assert np.linalg.norm(ang) < 1e-10
    v = w.read()["velocity"][0]
    assert v[0] > -1.0
    assert np.linalg.norm(p0 - v * imp.mass - lin) < 1e-10


def test_fixed_particle_does_not_move():
    # world.rs:810-818
    p = particle([0.5, 0.5, 0.5])
    p.fixed = True
    p.velocity = np.array([1.0, 2.0, 3.0])
    w = oworld([p])
    w.step(0.01)
    r = w.read()
    assert np.array_equal(r["position"][0], [0.5, 0.5, 0.5])
    assert np.array_equal(r["velocity"][0], [0.0, 0.0, 0.0])


def test_world_boundary_prevents_particle_escape():
    # world.rs:821-840
    p = particle([0.5, 0.5, 0.05])
    p.velocity[2] = -10.0
    w = oworld([p], MpmParams(gravity=np.zeros(3), max_substep=0.001, bounds=WorldBounds(np.zeros(3), np.ones(3))))
    w.step(0.01)
    r = w.read()
    assert r["position"][0, 2] >= 0.04
    assert r["velocity"][0, 2] >= 0.0


def test_all_materials_complete_finite_steps():
    # world.rs:843-866
    mats = [MaterialModel.elastic(1000.0, 0.2), MaterialModel.neo_hookean(1000.0, 0.2),
            MaterialModel.sand(1000.0, 0.2, math.radians(35.0), 0.0), MaterialModel.fluid(2000.0, 7.0, 0.01),
            MaterialModel.snow(1000.0, 0.2)]
    ps = [MpmParticle.new([0.3 + i * 0.1, 0.5, 0.5], 0.03, 1000.0, m) for i, m in enumerate(mats)]
    w = oworld(ps)
    w.step(0.002)
    r = w.read()
    for k in ("position", "velocity", "deformation", "affine"):
        assert np.all(np.isfinite(r[k]))


def test_cfl_bound_tightens_for_stiffer_material():
    # world.rs:869-893
    soft = MpmParticle.new([0, 0, 0], 0.04, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    stiff = MpmParticle.new([0, 0, 0], 0.04, 1000.0, MaterialModel.elastic(100000.0, 0.2))
    params = MpmParams(max_substep=1.0)
    assert oworld([stiff], params).stable_timestep() < oworld([soft], params).stable_timestep()
    # closed form: 0.4 h / (c + |v|), c = sqrt((lambda + 2 mu) / rho)
    lam, mu = 1000.0 * 0.2 / (1.2 * 0.6), 1000.0 / 2.4
    assert abs(oworld([soft], params).stable_timestep() - 0.4 * 0.1 / math.sqrt((lam + 2 * mu) / 1000.0)) < 1e-15


def test_rejects_invalid_particle_and_world_parameters():
    # world.rs:896-914
    bad = particle([0, 0, 0])
    bad.mass = 0.0
    with pytest.raises(O.OracleError) as e:
        oworld([bad])
    assert e.value.name == "InvalidInput"
    with pytest.raises(O.OracleError):
        oworld([], MpmParams(cell_width=0.0))


def test_moving_sphere_pushes_particle_without_changing_obstacle():
    # world.rs:917-932
    p = particle([0.62, 0.5, 0.5])
    sphere = RigidObstacle.sphere([0.5, 0.5, 0.5], 0.1)
    sphere.linear_velocity = np.array([1.0, 0.0, 0.0])
    w = oworld([p], MpmParams(gravity=np.zeros(3)), [sphere])
    w.step(0.001)
    r = w.read()
    assert r["velocity"][0, 0] > 0.0
    d, _ = O.obstacle_surface(sphere, r["position"][0])
    assert d >= 0.04 - 1e-12


def test_rejects_invalid_obstacle():
    # world.rs:935-945
    w = oworld([])
    w.set_obstacles([RigidObstacle.sphere([0, 0, 0], 1.0)])
    with pytest.raises(O.OracleError):
        w.set_obstacles([RigidObstacle.sphere([0, 0, 0], -1.0)])


def test_invalid_batch_does_not_append_or_consume_chunk_id():
    # world.rs:948-968
    w = oworld([particle([0.0, 0.0, 0.5])])
    bad = particle([0.2, 0.0, 0.5])
    bad.mass = 0.0
    with pytest.raises(O.OracleError):
        w.add_particles(ParticleArrays.from_particles([particle([0.1, 0.0, 0.5]), bad]))
    assert len(w) == 1
    chunk = w.add_particles(ParticleArrays.from_particles([particle([0.3, 0.0, 0.5])]))
    assert chunk == 1
    assert w.remove_chunk(0) == 1
    assert list(w.read()["chunk_id"]) == [1]
    with pytest.raises(O.OracleError) as e:
        w.remove_chunk(0)
    assert e.value.name == "UnknownChunk"


# --- invariants ---------------------------------------------------------------------------------
def random_world(n, seed, material=None, h=0.05, gravity=(0.0, 0.0, 0.0)):
    rng = np.random.default_rng(seed)
    ps = []
    for _ in range(n):
        p = MpmParticle.new(rng.uniform(0.3, 0.7, 3), 0.01, 1000.0, material or MaterialModel.elastic(2000.0, 0.25))
        p.velocity = rng.normal(0, 0.3, 3)
        p.affine = rng.normal(0, 0.5, (3, 3))
        p.deformation = np.eye(3) + rng.normal(0, 0.05, (3, 3))
        ps.append(p)
    return MpmWorld(ps, MpmParams(cell_width=h, gravity=np.array(gravity), max_substep=1e-3))


def test_p2g_partition_of_unity_conserves_mass_and_momentum():
    # sum of B-spline weights is 1 and sum of w*dpos is 0: node mass equals
    # particle mass and node momentum equals particle momentum (the affine and
    # stress terms integrate to zero).
    w = random_world(40, 1)
    ow = O.OracleWorld(w)
    nodes, m, p = ow.p2g_totals(1e-4)
    pa = w.particles
    assert abs(m - pa.mass.sum()) < 1e-14 * pa.mass.sum() * 100
    pm = (pa.mass[:, None] * pa.velocity).sum(axis=0)
    assert np.linalg.norm(p - pm) < 1e-12 * max(1.0, np.linalg.norm(pm)) * 100
    assert nodes >= 27


def test_substep_conserves_linear_momentum_without_external_forces():
    w = random_world(60, 2)
    ow = O.OracleWorld(w)
    pa = w.particles
    p0 = (pa.mass[:, None] * pa.velocity).sum(axis=0)
    ow.substeps_fixed(1e-4, 5)
    p1 = (pa.mass[:, None] * ow.read()["velocity"]).sum(axis=0)
    assert np.linalg.norm(p1 - p0) < 1e-12


def test_gravity_only_changes_momentum_by_impulse():
    w = random_world(30, 3, gravity=(0.0, -9.8, 0.0))
    ow = O.OracleWorld(w)
    pa = w.particles
    p0 = (pa.mass[:, None] * pa.velocity).sum(axis=0)
    ow.substeps_fixed(1e-4, 3)
    p1 = (pa.mass[:, None] * ow.read()["velocity"]).sum(axis=0)
    expected = p0 + pa.mass.sum() * np.array([0.0, -9.8, 0.0]) * 3e-4
    assert np.linalg.norm(p1 - expected) < 1e-12


def test_affine_velocity_field_is_reproduced_by_the_gather():
    # particles sampling v(x) = v0 + L x densely: P2G/G2P return v and C = L
    # for an interior particle (APIC exactness for affine fields).
    L = np.array([[0.1, -0.2, 0.05], [0.3, 0.0, -0.1], [0.02, 0.04, -0.07]])
    v0 = np.array([0.5, -0.3, 0.2])
    ps = []
    h = 0.1
    for x in np.arange(0.2, 0.81, h / 2):
        for y in np.arange(0.2, 0.81, h / 2):
            for z in np.arange(0.2, 0.81, h / 2):
                p = MpmParticle.new([x, y, z], 0.01, 1000.0, MaterialModel.fluid(1e-6, 1.0, 0.0, 0.0))
                p.velocity = v0 + L @ np.array([x, y, z])
                p.affine = L.copy()
                ps.append(p)
    w = MpmWorld(ps, MpmParams(cell_width=h, gravity=np.zeros(3), max_substep=1.0))
    ow = O.OracleWorld(w)
    ow.substeps_fixed(1e-9, 1)
    r = ow.read()
    centre = np.argmin(np.linalg.norm(w.particles.position - 0.5, axis=1))
    x = w.particles.position[centre]
    assert np.linalg.norm(r["velocity"][centre] - (v0 + L @ x)) < 1e-6
    assert np.linalg.norm(r["affine"][centre] - L) < 1e-5


def test_threaded_substep_matches_the_single_thread_one_and_repeats():
    # the all-cores CPU baseline runs the same substep over host threads; only
    # the association of the per-node sums differs from the single-thread pass
    from tessera_amd.scenes import uniform_cube

    sc = uniform_cube(4096, seed=2)
    single = O.OracleWorld(sc.world)
    single.substeps_fixed(sc.dt, 3)
    runs = []
    for _ in range(2):
        w = O.OracleWorld(sc.world)
        w.substeps_fixed_threads(sc.dt, 3, 4)
        runs.append(w.read())
    for f in ("position", "velocity", "affine", "deformation"):
        assert np.array_equal(runs[0][f], runs[1][f])
        scale = max(1.0, float(np.abs(single.read()[f]).max()))
        assert np.abs(runs[0][f] - single.read()[f]).max() <= 1e-12 * scale
