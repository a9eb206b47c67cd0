"""Continuous-batching serving on the GPU.

Requests arrive over time, are admitted under an HBM budget, advance together
in one batch with falling rigid boxes driven by the MPM reactions, and retire
at their horizon.  Every served scene must equal the same world stepped alone
in its own session for the same number of substeps, bit for bit (particles and
rigid bodies): batching, admission order and the other scenes must not change
any result."""

import numpy as np
import pytest

from tessera_amd.mpm import HipMpmResidentSession
from tessera_amd.scenes import serving_requests
from tessera_amd.serving import Server

pytestmark = pytest.mark.gpu


def alone(transfers, req, mode):
    s = HipMpmResidentSession(transfers, req.world, req.dt, mode)
    if req.bodies:
        s.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
    left = req.horizon
    while left:
        k = min(left, 64)
        s.step(k)
        left -= k
    return s.read_particles(), (s.rigid_bodies() if req.bodies else [])


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_served_scenes_equal_the_same_worlds_stepped_alone(transfers, mode):
    reqs = serving_requests(7, rate=200.0, seed=3, n_min=2048, n_max=12000, horizons=(9, 41), max_boxes=2)
    budget = int(2.5 * np.mean([r.device_bytes() for r in reqs]))  # forces some requests to queue
    server = Server(transfers, hbm_budget=budget, steps_per_tick=8, mode=mode)
    report = server.run(reqs)
    server.close()
    assert report["completed"] == len(reqs) and report["failed"] == 0
    assert report["queue_wait_s_p99"] >= 0.0 and report["step_latency_ms_p50"] > 0.0
    by_name = {r.name: r for r in reqs}
    for done in server.completed:
        req = by_name[done.name]
        assert done.horizon == req.horizon
        assert done.world.substeps == req.horizon
        pa, bodies = alone(transfers, req, mode)
        for f in ("position", "velocity", "affine", "deformation"):
            assert np.array_equal(getattr(done.world.particles, f), getattr(pa, f)), (done.name, f)
        for a, b in zip(done.bodies, bodies):
            for f in ("position", "orientation", "linear_velocity", "angular_velocity"):
                assert np.array_equal(getattr(a, f), getattr(b, f)), (done.name, f)
