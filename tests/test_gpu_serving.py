"""Continuous-batching serving on the GPU.

Requests arrive over time, are admitted under an HBM budget, advance together
in one batch with falling rigid boxes driven by the MPM reactions, and retire
at their horizon.  Every served scene must equal the same world stepped alone
in its own session for the same number of substeps, bit for bit (particles and
rigid bodies): batching, admission order, staging while the device runs and
the other scenes must not change any result.  The plain per-scene server (one
resident session per request) must give the same results.

The stream at the serving workload's own sizes (4,096-524,288 particles, 1-4
boxes, horizons 50-2000 substeps) is served once; sampled requests are
checked alone bit for bit, and one of them against the FP64 oracle's coupled
MPM-rigid substeps (gpu_rigid_sync.rs:187-213) over a window."""

import numpy as np
import pytest

from oracle import oracle as O  # noqa: F401  (the coupling oracle below loads it)
from tessera_amd.mpm import HipMpmResidentSession, InvalidInput, ResidentUnavailable
from tessera_amd.scenes import serving_request, serving_requests
from tessera_amd.serving import PlainServer, Server

from .test_gpu_coupling import oracle_coupled

pytestmark = pytest.mark.gpu

FIELDS = ("position", "velocity", "affine", "deformation")
BODY_FIELDS = ("position", "orientation", "linear_velocity", "angular_velocity")


def alone(transfers, req, mode, steps=None):
    s = HipMpmResidentSession(transfers, req.world, req.dt, mode)
    if req.bodies:
        s.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
    left = req.horizon if steps is None else steps
    while left:
        k = min(left, 64)
        s.step(k)
        left -= k
    out = s.read_particles(), (s.rigid_bodies() if req.bodies else [])
    s.close()
    return out


def assert_served_equals_alone(transfers, done, req, mode):
    assert done.error is None, (done.name, done.error)
    assert done.horizon == req.horizon
    assert done.world.substeps == req.horizon
    pa, bodies = alone(transfers, req, mode)
    for f in FIELDS:
        assert np.array_equal(getattr(done.world.particles, f), getattr(pa, f)), (done.name, f)
    assert len(done.bodies) == len(bodies)
    for a, b in zip(done.bodies, bodies):
        for f in BODY_FIELDS:
            assert np.array_equal(getattr(a, f), getattr(b, f)), (done.name, f)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_served_scenes_equal_the_same_worlds_stepped_alone(transfers, mode):
    reqs = serving_requests(7, rate=200.0, seed=3, n_min=2048, n_max=12000, horizons=(9, 41), max_boxes=2)
    budget = int(2.5 * np.mean([r.device_bytes() for r in reqs]))  # forces some requests to queue
    server = Server(transfers, hbm_budget=budget, steps_per_tick=8, mode=mode)
    report = server.run(reqs)
    server.close()
    assert report["completed"] == len(reqs) and report["failed"] == 0
    assert report["queue_wait_s_p99"] >= 0.0 and report["step_latency_ms_p50"] > 0.0
    assert report["particle_updates"] == sum(r.particles * r.horizon for r in reqs)
    by_name = {r.name: r for r in reqs}
    for done in server.completed:
        assert_served_equals_alone(transfers, done, by_name[done.name], mode)


def test_plain_per_scene_server_gives_the_same_results(transfers):
    reqs = serving_requests(5, rate=300.0, seed=5, n_min=2048, n_max=9000, horizons=(7, 30), max_boxes=2)
    budget = int(2.5 * np.mean([r.device_bytes() for r in reqs]))
    batched = Server(transfers, hbm_budget=budget, steps_per_tick=8)
    rb = batched.run(reqs)
    batched.close()
    plain = PlainServer(transfers, hbm_budget=budget, steps_per_tick=8)
    rp = plain.run(reqs)
    plain.close()
    assert rb["completed"] == rp["completed"] == len(reqs)
    assert rb["particle_updates"] == rp["particle_updates"]
    a = {c.name: c for c in batched.completed}
    for c in plain.completed:
        assert c.error is None
        for f in FIELDS:
            assert np.array_equal(getattr(c.world.particles, f), getattr(a[c.name].world.particles, f)), (c.name, f)
        for x, y in zip(c.bodies, a[c.name].bodies):
            for f in BODY_FIELDS:
                assert np.array_equal(getattr(x, f), getattr(y, f)), (c.name, f)


def test_failed_and_invalid_requests_complete_with_their_error(transfers):
    """A scene thrown past the CFL bound stops alone: it completes with ResidentUnavailable, its
    substeps count no updates, and the other requests still equal their worlds stepped alone.  A
    request with a horizon below 1 is rejected with InvalidInput without touching the batch."""
    rng = np.random.default_rng(9)
    good = [serving_request(rng, 3000, f"good{k}", 24, 0.0, max_boxes=1) for k in range(2)]
    bad = serving_request(rng, 3000, "bad", 24, 0.0, max_boxes=1)
    m = np.asarray(bad.world.particles.mass)
    # one-shot forces worth 300 m/s in one substep: beyond the CFL bound of dt = 5e-5 at h = 1/64
    bad.forces = [(i, (300.0 * m[i] / bad.dt, 0.0, 0.0)) for i in range(bad.particles)]
    empty = serving_request(rng, 3000, "zero", 0, 0.0, max_boxes=1)
    server = Server(transfers, hbm_budget=1 << 34, steps_per_tick=8)
    report = server.run(good + [bad, empty])
    server.close()
    by = {c.name: c for c in server.completed}
    assert report["completed"] == 4 and report["failed"] == 2
    assert isinstance(by["bad"].error, ResidentUnavailable)
    assert isinstance(by["zero"].error, InvalidInput)
    assert report["particle_updates"] == sum(r.particles * r.horizon for r in good)
    for r in good:
        assert_served_equals_alone(transfers, by[r.name], r, "fp32")


def test_serving_stream_at_full_sizes_equals_requests_alone(transfers):
    """Twelve requests drawn from the serving workload's distributions (4,096-524,288 particles,
    1-4 boxes of edge 0.05-0.15 and density 500-5000, horizons 50-2000 substeps), arriving
    together so that they share ticks.  The largest, the smallest and the longest request are
    checked bit for bit against their worlds stepped alone; the largest also against the FP64
    oracle's coupled substeps over its first 16 substeps."""
    reqs = serving_requests(12, rate=1.0e4, seed=21)
    sizes = [r.particles for r in reqs]
    assert min(sizes) >= 3000 and max(sizes) > 200_000
    assert all(50 <= r.horizon <= 2000 for r in reqs)
    server = Server(transfers, hbm_budget=1 << 34, steps_per_tick=32)
    report = server.run(reqs)
    server.close()
    assert report["completed"] == len(reqs) and report["failed"] == 0, report
    print(f"full-size stream: {report['value']:.3e} particle-updates/s, p50 {report['step_latency_ms_p50']:.3f} ms, "
          f"p99 {report['step_latency_ms_p99']:.3f} ms over {report['ticks']} ticks")
    by = {c.name: c for c in server.completed}
    largest = max(reqs, key=lambda r: r.particles)
    sample = {largest.name: largest, min(reqs, key=lambda r: r.particles).name: min(reqs, key=lambda r: r.particles),
              max(reqs, key=lambda r: r.horizon).name: max(reqs, key=lambda r: r.horizon)}
    for name, req in sample.items():
        assert_served_equals_alone(transfers, by[name], req, "fp32")
    # the largest request's first 16 coupled substeps against the FP64 oracle
    steps = 16
    pa, bodies = alone(transfers, largest, "fp32", steps)
    cpu, ref = oracle_coupled(largest.world, list(largest.bodies), list(largest.mappings), largest.dt, steps,
                              np.asarray(largest.gravity))
    ref_p = cpu.read()
    assert np.max(np.linalg.norm(ref_p["position"] - pa.position, axis=1)) < 1e-5
    assert np.max(np.linalg.norm(ref_p["velocity"] - pa.velocity, axis=1)) < 1e-4
    for got, want in zip(bodies, ref):
        assert np.abs(got.position - want["position"]).max() < 2e-6
        assert np.abs(got.linear_velocity - want["linear_velocity"]).max() < 2e-3 * max(
            1.0, np.abs(want["linear_velocity"]).max())
