"""TEST INFRASTRUCTURE ONLY — FP64 restatement of the reference's device-side
rigid coupling, used as the checker of the HIP backend's rigid bodies.

* :func:`sync_obstacle` restates the pose update of an obstacle driven by a
  rigid body (crates/tessera-mpm/src/gpu_rigid_sync.wgsl:43-62).
* :func:`apply_reactions` restates the application of one substep's obstacle
  reactions to the bodies, with the MPM-owned integration
  (crates/tessera-mpm/src/gpu_rigid_reaction.wgsl:72-150).

The reference computes both in f32 on the device; this restatement is f64.
Bodies are dicts with numpy arrays: ``position`` (3), ``inverse_mass``,
``orientation`` (XYZW), ``linear_velocity`` (3), ``angular_velocity`` (3),
``inverse_inertia`` (body-frame principal, 3).  Mappings are dicts with
``body`` (index or None), ``local_center`` (3), ``local_orientation`` (4).
Pinned by tests/test_oracle_rigid.py against closed forms (impulse on the
center, principal-axis torque, a rotated body checked through explicit
rotation matrices, a full turn of constant spin, the speed cap, and the
surface velocity of a spinning offset obstacle).
"""

from __future__ import annotations

import copy

import numpy as np


def rotate(q, v):
    """v rotated by the unit quaternion q (XYZW), as in both shaders' `rotate`."""
    q = np.asarray(q, dtype=np.float64)
    v = np.asarray(v, dtype=np.float64)
    t = 2.0 * np.cross(q[:3], v)
    return v + q[3] * t + np.cross(q[:3], t)


def quat_mul(a, b):
    """Hamilton product a b (XYZW), gpu_rigid_sync.wgsl `quaternion_multiply`."""
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    v = a[3] * b[:3] + b[3] * a[:3] + np.cross(a[:3], b[:3])
    return np.array([v[0], v[1], v[2], a[3] * b[3] - np.dot(a[:3], b[:3])])


def quat_from_rotation(rotation):
    """gpu_rigid_reaction.wgsl `quat_from_rotation`: exp of a rotation vector."""
    rotation = np.asarray(rotation, dtype=np.float64)
    angle = float(np.linalg.norm(rotation))
    if angle < 1e-5:
        q = np.array([0.5 * rotation[0], 0.5 * rotation[1], 0.5 * rotation[2], 1.0])
        return q / np.linalg.norm(q)
    half = 0.5 * angle
    return np.concatenate([rotation * (np.sin(half) / angle), [np.cos(half)]])


def sync_obstacle(body: dict, mapping: dict, obstacle):
    """Returns a copy of `obstacle` (a tessera_amd.mpm.RigidObstacle) posed by
    its body (gpu_rigid_sync.wgsl:43-62)."""
    o = copy.deepcopy(obstacle)
    if mapping.get("body") is None:
        return o
    offset = rotate(body["orientation"], mapping["local_center"])
    o.center = np.asarray(body["position"], dtype=np.float64) + offset
    o.orientation = quat_mul(body["orientation"], mapping["local_orientation"])
    o.linear_velocity = np.asarray(body["linear_velocity"], dtype=np.float64) + np.cross(body["angular_velocity"],
                                                                                         offset)
    o.angular_velocity = np.asarray(body["angular_velocity"], dtype=np.float64).copy()
    return o


def apply_reactions(bodies, mappings, obstacles, reactions, dt, gravity, max_speed, integrate):
    """One substep of reactions applied to every body (gpu_rigid_reaction.wgsl:72-150).

    `reactions` holds one (linear, angular) pair of 3-vectors per obstacle,
    the angular impulse about that obstacle's center; `obstacles` are the
    obstacles as posed during the substep.  Returns new body dicts."""
    out = []
    for index, state in enumerate(bodies):
        st = copy.deepcopy(state)
        inv_mass = float(st["inverse_mass"])
        if inv_mass <= 0.0:
            out.append(st)
            continue
        lin = np.zeros(3)
        ang = np.zeros(3)
        mapped = False
        for o, (m, r) in enumerate(zip(mappings, reactions)):
            if m.get("body") != index:
                continue
            mapped = True
            impulse = np.asarray(r[0], dtype=np.float64)
            lin += impulse
            ang += np.asarray(r[1], dtype=np.float64) + np.cross(
                np.asarray(obstacles[o].center, dtype=np.float64) - st["position"], impulse)
        if not (np.all(np.abs(lin) < 1e30) and np.all(np.abs(ang) < 1e30)) or not mapped:
            out.append(st)
            continue
        if not lin.any() and not ang.any() and not integrate:
            out.append(st)
            continue
        q = np.asarray(st["orientation"], dtype=np.float64)
        conj = np.array([-q[0], -q[1], -q[2], q[3]])
        torque_body = rotate(conj, ang)
        st["linear_velocity"] = np.asarray(st["linear_velocity"], dtype=np.float64) + lin * inv_mass
        st["angular_velocity"] = np.asarray(st["angular_velocity"], dtype=np.float64) + rotate(
            q, torque_body * np.asarray(st["inverse_inertia"], dtype=np.float64))
        if integrate:
            st["linear_velocity"] = st["linear_velocity"] + np.asarray(gravity, dtype=np.float64) * dt
        if max_speed > 0.0:
            speed = float(np.linalg.norm(st["linear_velocity"]))
            if speed > max_speed:
                st["linear_velocity"] = st["linear_velocity"] * (max_speed / speed)
        if integrate:
            st["position"] = np.asarray(st["position"], dtype=np.float64) + st["linear_velocity"] * dt
            nq = quat_mul(quat_from_rotation(st["angular_velocity"] * dt), q)
            st["orientation"] = nq / np.linalg.norm(nq)
        out.append(st)
    return out
