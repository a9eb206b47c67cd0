"""Pins the oracle's obstacle SDFs and boundary responses against the
analytic values of crates/tessera-mpm/src/obstacle.rs:794-1061."""

import math

import numpy as np

from oracle import oracle as O
from tessera_amd.mpm import ObstacleBoundary, RigidObstacle, quat_from_axis_angle

X, Y, Z = np.eye(3)


def test_rotated_box_reports_world_normal_and_signed_distance():
    box = RigidObstacle.cuboid([1.0, 2.0, 3.0], [0.2, 0.4, 0.6], quat_from_axis_angle(Z, math.pi / 2))
    d, n = O.obstacle_surface(box, [1.5, 2.0, 3.0])
    assert abs(d - 0.1) < 1e-12 and np.linalg.norm(n - X) < 1e-12
    d, _ = O.obstacle_surface(box, [1.0, 2.0, 3.0])
    assert abs(d + 0.2) < 1e-12


def test_boundary_modes_apply_distinct_contact_velocities():
    s = RigidObstacle.sphere([0, 0, 0], 1.0)
    s.linear_velocity = np.array([1.0, 0.0, 0.0])
    s.friction = 0.5
    cases = [(ObstacleBoundary.SLIP, [1.0, 1.5, 0.0]), (ObstacleBoundary.STICK, [1.0, 0.0, 0.0]),
             (ObstacleBoundary.SEPARATE, [1.0, 2.0, 0.0]), (ObstacleBoundary.NON_REFLECTING, [0.0, 2.0, 0.0]),
             (ObstacleBoundary.FREE_SLIP, [1.0, 2.0, 0.0])]
    for b, expected in cases:
        s.boundary = b
        assert np.linalg.norm(O.contact_velocity(s, X, [0.0, 2.0, 0.0], X) - expected) < 1e-12, b


def test_free_slip_removes_both_signs_of_relative_normal_velocity():
    s = RigidObstacle.sphere([0, 0, 0], 1.0)
    s.linear_velocity = np.array([1.0, 2.0, 3.0])
    s.angular_velocity = np.array([0.0, 0.0, 2.0])
    s.friction = 10.0
    surface = s.linear_velocity + np.cross(s.angular_velocity, X)
    for ns in (-2.0, 2.0):
        v = surface + [ns, 4.0, -3.0]
        s.boundary = ObstacleBoundary.FREE_SLIP
        assert np.linalg.norm(O.contact_velocity(s, X, v, X) - (surface + [0.0, 4.0, -3.0])) < 1e-12
        for b in (ObstacleBoundary.SLIP, ObstacleBoundary.SEPARATE):
            s.boundary = b
            r = O.contact_velocity(s, X, v, X)
            if ns > 0:
                assert np.array_equal(r, v)
            else:
                assert abs(r[0] - surface[0]) < 1e-12
                if b == ObstacleBoundary.SEPARATE:
                    assert abs(r[1] - v[1]) < 1e-12 and abs(r[2] - v[2]) < 1e-12


def test_capsule_cylinder_and_cone_report_signed_surfaces():
    cap = RigidObstacle.capsule([0, 0, 0], 0.5, 0.2)
    d, n = O.obstacle_surface(cap, [0.3, 0, 0])
    assert abs(d - 0.1) < 1e-12 and np.linalg.norm(n - X) < 1e-12
    d, n = O.obstacle_surface(cap, [0, 0, 0.8])
    assert abs(d - 0.1) < 1e-12 and np.linalg.norm(n - Z) < 1e-12
    assert abs(O.obstacle_surface(cap, [0, 0, 0])[0] + 0.2) < 1e-12
    line = RigidObstacle.capsule([0, 0, 0], 0.5, 0.0)
    assert O.obstacle_valid(line)
    assert O.may_contact(line, [0, 0, 0.54], 0.04)
    assert not O.may_contact(line, [0, 0, 0.55], 0.04)
    cyl = RigidObstacle.cylinder([0, 0, 0], 0.5, 0.2)
    d, n = O.obstacle_surface(cyl, [0.3, 0, 0.7])
    assert abs(d - math.hypot(0.1, 0.2)) < 1e-12
    assert np.linalg.norm(n - np.array([0.1, 0, 0.2]) / math.hypot(0.1, 0.2)) < 1e-12
    assert abs(O.obstacle_surface(cyl, [0, 0, 0])[0] + 0.2) < 1e-12
    cone = RigidObstacle.cone([0, 0, 0], 0.5, 0.4)
    d, n = O.obstacle_surface(cone, [0, 0, 0.7])
    assert abs(d - 0.2) < 1e-12 and np.linalg.norm(n - Z) < 1e-12
    d, n = O.obstacle_surface(cone, [0, 0, -0.7])
    assert abs(d - 0.2) < 1e-12 and np.linalg.norm(n + Z) < 1e-12
    d, n = O.obstacle_surface(cone, [0, 0, 0])
    assert d < -0.1 and n[0] > 0 and n[2] > 0
    assert O.obstacle_valid(cone)


def test_rotated_capsule_and_invalid_primitives():
    cap = RigidObstacle.capsule([0, 0, 0], 0.5, 0.2, quat_from_axis_angle(Y, math.pi / 2))
    d, n = O.obstacle_surface(cap, [0.8, 0, 0])
    assert abs(d - 0.1) < 1e-12 and np.linalg.norm(n - X) < 1e-12
    assert not O.obstacle_valid(RigidObstacle.cylinder([0, 0, 0], 0.0, 0.2))
    assert not O.obstacle_valid(RigidObstacle.cone([0, 0, 0], 0.5, -0.2))


def test_finite_ground_has_one_sided_interior_and_finite_edges():
    g = RigidObstacle.ground([0, 0, 0], [1.0, 1.0])
    d, n = O.obstacle_surface(g, [0, 0, 0.2])
    assert d == 0.2 and np.array_equal(n, Z)
    d, n = O.obstacle_surface(g, [0, 0, -0.2])
    assert d == -0.2 and np.array_equal(n, Z)
    d, n = O.obstacle_surface(g, [1.3, 0, 0])
    assert abs(d - 0.3) < 1e-12 and np.linalg.norm(n - X) < 1e-12
    d, n = O.obstacle_surface(g, [1.3, 0, -0.4])
    assert abs(d - 0.5) < 1e-12 and np.linalg.norm(n - [0.6, 0, -0.8]) < 1e-12
    assert O.obstacle_valid(g)


def test_triangle_prism_exposes_caps_sides_and_interior():
    p = RigidObstacle.triangle_prism([0, 0, 0], [[-0.2, -0.2, 0], [0.2, -0.2, 0], [0.0, 0.2, 0]], 0.01)
    assert O.obstacle_valid(p)
    d, n = O.obstacle_surface(p, [0, -0.05, 0.05])
    assert abs(d - 0.04) < 1e-12 and np.linalg.norm(n - Z) < 1e-12
    d, n = O.obstacle_surface(p, [0, -0.05, -0.05])
    assert abs(d - 0.04) < 1e-12 and np.linalg.norm(n + Z) < 1e-12
    d, _ = O.obstacle_surface(p, [0, -0.05, 0])
    assert abs(d + 0.01) < 1e-12
    d, n = O.obstacle_surface(p, [0, -0.3, 0])
    assert abs(d - 0.1) < 1e-12 and np.linalg.norm(n + Y) < 1e-12


def test_convex_planes_report_face_edge_vertex_and_interior():
    c = RigidObstacle.convex([0, 0, 0], [[0, 0, 0], X, Y, Z], [-X, -Y, -Z, np.ones(3) / math.sqrt(3)])
    assert O.obstacle_valid(c)
    d, n = O.obstacle_surface(c, [0.1, 0.1, 0.1])
    assert abs(d + 0.1) < 1e-12 and np.linalg.norm(n + X) < 1e-12
    d, n = O.obstacle_surface(c, [-0.2, 0.1, 0.1])
    assert abs(d - 0.2) < 1e-9 and np.linalg.norm(n + X) < 1e-9
    d, n = O.obstacle_surface(c, [-0.2, -0.3, 0.1])
    assert abs(d - math.hypot(0.2, 0.3)) < 1e-8
    assert np.linalg.norm(n - np.array([-0.2, -0.3, 0]) / math.hypot(0.2, 0.3)) < 1e-8
    d, _ = O.obstacle_surface(c, [1.2, 0.1, 0.1])
    assert abs(d - math.sqrt(0.06)) < 1e-7
    bad = RigidObstacle.convex([0, 0, 0], [[math.nan, 0, 0]], [-X] * 4)
    assert not O.obstacle_valid(bad)


def test_convex_plane_distance_matches_analytic_box():
    half = np.array([0.2, 0.3, 0.4])
    verts = [np.array([x, y, z]) * half for x in (-1, 1) for y in (-1, 1) for z in (-1, 1)]
    c = RigidObstacle.convex([0, 0, 0], verts, [X, -X, Y, -Y, Z, -Z])
    box = RigidObstacle.cuboid([0, 0, 0], half)
    for x in (-0.6, -0.2, -0.1, 0.0, 0.1, 0.2, 0.6):
        for y in (-0.7, -0.3, 0.0, 0.3, 0.7):
            for z in (-0.8, -0.4, 0.0, 0.4, 0.8):
                assert abs(O.obstacle_surface(c, [x, y, z])[0] - O.obstacle_surface(box, [x, y, z])[0]) < 1e-7
