"""Pins oracle/rigid.py (the FP64 restatement of the rigid coupling,
gpu_rigid_sync.wgsl:43-62 and gpu_rigid_reaction.wgsl:72-150) with closed
forms that a dropped term, a wrong sign or a transposed rotation would fail:
each check computes its expectation by a different route (rotation
matrices, angle sums, geometry) than the restatement."""

import math

import numpy as np
import pytest

from oracle import rigid as R
from tessera_amd.mpm import RigidObstacle


def body(position=(0.0, 0.0, 0.0), inverse_mass=0.5, orientation=(0.0, 0.0, 0.0, 1.0), v=(0.0, 0.0, 0.0),
         w=(0.0, 0.0, 0.0), inverse_inertia=(2.0, 3.0, 4.0)):
    return {"position": np.array(position, float), "inverse_mass": inverse_mass,
            "orientation": np.array(orientation, float), "linear_velocity": np.array(v, float),
            "angular_velocity": np.array(w, float), "inverse_inertia": np.array(inverse_inertia, float)}


def matrix(q):
    """Rotation matrix of an XYZW unit quaternion (textbook form)."""
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def axis_angle(axis, angle):
    axis = np.asarray(axis, float) / np.linalg.norm(axis)
    return np.concatenate([axis * math.sin(angle / 2), [math.cos(angle / 2)]])


def box_at(center):
    return RigidObstacle.cuboid(np.array(center, float), np.array([0.1, 0.1, 0.1]))


def test_rotate_and_product_agree_with_rotation_matrices():
    rng = np.random.default_rng(3)
    for _ in range(20):
        a = axis_angle(rng.normal(size=3), rng.uniform(-3, 3))
        b = axis_angle(rng.normal(size=3), rng.uniform(-3, 3))
        v = rng.normal(size=3)
        assert np.allclose(R.rotate(a, v), matrix(a) @ v, atol=1e-12)
        assert np.allclose(matrix(R.quat_mul(a, b)), matrix(a) @ matrix(b), atol=1e-12)


def test_impulse_through_the_center_changes_only_linear_velocity():
    b = body(v=(1.0, 0.0, 0.0))
    ob = [box_at(b["position"])]
    out = R.apply_reactions([b], [{"body": 0}], ob, [(np.array([0.2, -0.4, 0.6]), np.zeros(3))], 0.01,
                            np.zeros(3), 0.0, False)[0]
    assert np.allclose(out["linear_velocity"], [1.1, -0.2, 0.3])
    assert np.allclose(out["angular_velocity"], 0.0)
    assert np.allclose(out["position"], b["position"])  # no pose update without integration


def test_principal_torque_on_a_rotated_body_uses_r_diag_rt():
    q = axis_angle([0.3, -0.5, 0.8], 1.1)
    b = body(orientation=q)
    L = np.array([0.5, -0.25, 1.0])
    out = R.apply_reactions([b], [{"body": 0}], [box_at(b["position"])], [(np.zeros(3), L)], 0.01, np.zeros(3),
                            0.0, False)[0]
    Rm = matrix(q)
    expected = Rm @ np.diag(b["inverse_inertia"]) @ Rm.T @ L
    assert np.allclose(out["angular_velocity"], expected, atol=1e-12)


def test_offset_obstacle_impulse_adds_the_moment_about_the_body():
    b = body(position=(1.0, 2.0, 3.0), inverse_inertia=(1.0, 1.0, 1.0))
    ob = box_at([1.5, 2.0, 3.0])  # obstacle center 0.5 along +x of the body
    J = np.array([0.0, 1.0, 0.0])
    out = R.apply_reactions([b], [{"body": 0}], [ob], [(J, np.zeros(3))], 0.01, np.zeros(3), 0.0, False)[0]
    # r x J = (0.5, 0, 0) x (0, 1, 0) = (0, 0, 0.5)
    assert np.allclose(out["angular_velocity"], [0.0, 0.0, 0.5])


def test_full_turn_of_constant_spin_returns_to_the_start_and_gravity_integrates():
    n = 1000
    dt = 1.0 / n
    w = 2.0 * math.pi
    b = body(w=(0.0, 0.0, w), inverse_inertia=(0.0, 0.0, 0.0))
    g = np.array([0.0, -9.8, 0.0])
    for _ in range(n):
        b = R.apply_reactions([b], [{"body": 0}], [box_at(b["position"])], [(np.zeros(3), np.zeros(3))], dt, g,
                              0.0, True)[0]
    q = b["orientation"]
    assert abs(abs(q[3]) - 1.0) < 1e-9  # one full turn about z
    # semi-implicit Euler: v_k = g k dt, x_n = g dt^2 n(n+1)/2
    assert np.allclose(b["linear_velocity"], g, atol=1e-12)
    assert np.allclose(b["position"], g * dt * dt * n * (n + 1) / 2, atol=1e-12)


def test_speed_cap_and_static_bodies():
    b = body(v=(3.0, 4.0, 0.0))
    out = R.apply_reactions([b], [{"body": 0}], [box_at(b["position"])], [(np.array([0.0, 0.0, 2.0]), np.zeros(3))],
                            0.01, np.zeros(3), 2.5, False)[0]
    assert math.isclose(float(np.linalg.norm(out["linear_velocity"])), 2.5)
    assert np.allclose(out["linear_velocity"] / 2.5, np.array([3.0, 4.0, 1.0]) / math.sqrt(26.0))
    static = body(inverse_mass=0.0)
    same = R.apply_reactions([static], [{"body": 0}], [box_at([0, 0, 0])], [(np.ones(3), np.ones(3))], 0.01,
                             np.array([0.0, -9.8, 0.0]), 0.0, True)[0]
    assert np.array_equal(same["linear_velocity"], static["linear_velocity"])
    unmapped = R.apply_reactions([b], [{"body": None}], [box_at([0, 0, 0])], [(np.ones(3), np.ones(3))], 0.01,
                                 np.zeros(3), 0.0, True)[0]
    assert np.array_equal(unmapped["position"], b["position"])  # a body without obstacles is not integrated


def test_sync_poses_an_offset_obstacle_with_the_surface_velocity():
    q = axis_angle([0.0, 0.0, 1.0], math.pi / 2)
    b = body(position=(1.0, 0.0, 0.0), orientation=q, v=(0.1, 0.0, 0.0), w=(0.0, 0.0, 2.0))
    m = {"body": 0, "local_center": np.array([0.5, 0.0, 0.0]), "local_orientation": axis_angle([1, 0, 0], 0.3)}
    o = R.sync_obstacle(b, m, box_at([0, 0, 0]))
    # local +x maps to world +y after a quarter turn about z
    assert np.allclose(o.center, [1.0, 0.5, 0.0], atol=1e-12)
    # v + w x r = (0.1, 0, 0) + (0, 0, 2) x (0, 0.5, 0) = (0.1 - 1.0, 0, 0)
    assert np.allclose(o.linear_velocity, [-0.9, 0.0, 0.0], atol=1e-12)
    assert np.allclose(matrix(o.orientation), matrix(q) @ matrix(m["local_orientation"]), atol=1e-12)
    assert np.allclose(o.angular_velocity, [0.0, 0.0, 2.0])
    assert R.sync_obstacle(b, {"body": None}, box_at([3, 3, 3])).center[0] == pytest.approx(3.0)
