"""The Python ``MpmWorld`` facade (crates/tessera-python3d/src/mpm.rs:630-1283)
on the HIP backend, following the binding's own tests (mpm.rs:1428-1805);
every step runs on the device."""

import math

import numpy as np
import pytest

from tessera_amd.mpm import MaterialModel, MpmParticle, ObstacleBoundary, RigidObstacle
from tessera_amd.tessera3d import MpmWorld, TesseraError

pytestmark = pytest.mark.gpu


def particle(x):
    return MpmParticle.new([x, 0.5, 0.5], 0.03, 1000.0, MaterialModel.elastic(1000.0, 0.2))


def test_configuration_preserves_state_and_rejects_invalid_params(transfers):
    # mpm.rs:1428-1505
    w = MpmWorld([particle(0.5)], [0.0, 0.0, -1.0], 0.1, 0.001, transfers=transfers)
    chunk = w.add_particles([particle(0.7)])
    w.step(0.0001)
    w.set_particle_force(0, [1.0, 2.0, 3.0])
    before = w.particles()
    substeps = w.substeps()
    w.set_gravity([1.0, 0.0, 0.0])
    w.set_params([0.0, 2.0, 0.0], 0.2, 0.0005, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0])
    assert w.substeps() == substeps
    for a, b in zip(w.particles(), before):
        for f in ("position", "velocity", "deformation", "affine", "force"):
            assert np.array_equal(getattr(a, f), getattr(b, f)), f
        assert a.chunk_id == b.chunk_id
    assert w.gravity()[1] == 2.0
    for cw, ms in [(0.0, 0.001), (-0.1, 0.001), (math.nan, 0.001), (0.1, 0.0), (0.1, math.inf)]:
        with pytest.raises(TesseraError):
            w.set_params([3.0, 0.0, 0.0], cw, ms)
    with pytest.raises(TesseraError):
        w.set_gravity([math.nan, 0.0, 0.0])
    with pytest.raises(TesseraError):
        w.set_params([3.0, 0.0, 0.0], 0.1, 0.001, [0.0, 0.0, 0.0], None)
    with pytest.raises(TesseraError):
        w.set_params([3.0, 0.0, 0.0], 0.1, 0.001, [1.0, 0.0, 0.0], [0.0, 1.0, 1.0])
    assert w.gravity()[1] == 2.0
    assert w.remove_chunk(chunk) == 1
    assert w.add_particles([particle(0.7)]) > chunk


def test_frame_substeps_match_explicit_division(transfers):
    # mpm.rs:1507-1540
    a = MpmWorld([particle(0.5)], [0.0, 0.0, -1.0], 0.1, 0.001, transfers=transfers)
    b = MpmWorld([particle(0.5)], [0.0, 0.0, -1.0], 0.1, 0.001, transfers=transfers)
    assert a.frame_substeps() == 1
    a.set_substeps(0)
    assert a.frame_substeps() == 1
    a.set_substeps(4)
    a.step(0.001)
    for _ in range(4):
        b.step(0.00025)
    assert a.substeps() == 4
    pa, pb = a.particles()[0], b.particles()[0]
    assert np.array_equal(pa.position, pb.position) and np.array_equal(pa.velocity, pb.velocity)
    for bad in (math.nan, 0.0):
        with pytest.raises(TesseraError):
            a.step(bad)
    assert a.substeps() == 4


def test_gpu_controls_preserve_pending_and_resident_state(transfers):
    # mpm.rs:1542-1590
    w = MpmWorld([particle(0.5)], [0.0, 0.0, -1.0], 0.1, 0.001, transfers=transfers)
    w.step_gpu_fixed(0.0001, 1)
    before = w.particles()[0].position.copy()
    with pytest.raises(TesseraError):
        w.set_gravity([np.finfo(np.float64).max, 0.0, 0.0])  # not representable on the device
    assert w.gravity()[2] == -1.0
    w.submit_gpu_fixed(0.0001, 1)
    for edit in (lambda: w.set_gravity([0.0, 0.0, 1.0]), lambda: w.set_substeps(4), lambda: w.set_use_cpic(False)):
        with pytest.raises(TesseraError):
            edit()
    assert w.frame_substeps() == 1 and w.use_cpic() and w.pending_gpu_substeps() == 1
    w.synchronize_gpu()
    synced = w.particles()[0].position.copy()
    assert synced[2] < before[2]
    w.set_gravity([0.0, 0.0, 1.0])
    assert w.particles()[0].position[2] == synced[2]
    w.set_use_cpic(False)
    w.step_gpu_fixed(0.0001, 1)
    assert not w.use_cpic() and w.gravity()[2] == 1.0


def test_cpic_groups_validate_shape_and_range(transfers):
    # mpm.rs:1714-1760
    w = MpmWorld([particle(0.48)], [0.0, 0.0, 0.0], 0.1, 0.001, transfers=transfers)
    w.set_obstacles([RigidObstacle.triangle_prism([0.5, 0.5, 0.5], [[0, -1, -1], [0, 1, -1], [0, 0, 1]], 0.001)])
    w.set_obstacle_cpic_groups([0])
    assert w.obstacle_cpic_groups() == [0]
    w.set_obstacle_cpic_groups([31])
    assert w.use_cpic()
    w.set_use_cpic(False)
    assert not w.use_cpic() and w.obstacle_cpic_groups() == [31]
    w.set_obstacle_cpic_groups([7])
    w.set_particle_transfer_color(0, 2)
    w.set_use_cpic(True)
    assert w.obstacle_cpic_groups() == [7]
    assert w.particles()[0].transfer_color == 2
    w.set_obstacle_cpic_groups([31])
    with pytest.raises(TesseraError):
        w.set_obstacle_cpic_groups([32])
    with pytest.raises(TesseraError):
        w.set_obstacle_cpic_groups([])
    assert w.obstacle_cpic_groups() == [31]


def test_exposes_particles_chunks_and_obstacles(transfers):
    # mpm.rs:1762-1803
    w = MpmWorld([particle(0.5)], [0.0, 0.0, 0.0], 0.1, 0.001, [0.0, 0.0, 0.0], [1.0, 1.0, 1.0],
                 transfers=transfers)
    sphere = RigidObstacle.sphere([0.5, 0.5, 0.5], 0.1)
    sphere.linear_velocity = np.array([1.0, 0.0, 0.0])
    sphere.friction = 0.5
    w.set_obstacles([sphere])
    w.step(0.0001)
    assert w.particles()[0].velocity[0] > 0.0
    assert w.obstacle_count() == 1
    chunk = w.add_particles([particle(0.7)])
    assert w.particles()[1].chunk_id == chunk
    assert w.remove_chunk(chunk) == 1
    assert len(w.particles()) == 1
    with pytest.raises(TesseraError):
        w.remove_chunk(chunk)
    bad = RigidObstacle.sphere([0.5, 0.5, 0.5], 0.1)
    bad.orientation = np.array([0.0, 0.0, 0.0, 0.5])
    with pytest.raises(TesseraError):
        w.set_obstacles([bad])
    assert w.obstacle_count() == 1
    w.set_obstacle_boundaries([ObstacleBoundary.STICK])
    assert w.obstacle_boundaries() == [ObstacleBoundary.STICK]
    mesh = MpmWorld([particle(0.5)], [0.0, 0.0, 0.0], 0.1, 0.001, transfers=transfers)
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
    t = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                  [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]])
    c = mesh.add_closed_mesh_particles(v + 2.0, t, 0.25, 1000.0, MaterialModel.elastic(1000.0, 0.2), 64, 64)
    assert sum(1 for p in mesh.particles() if p.chunk_id == c) == 64
