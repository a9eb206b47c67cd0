"""The benchmark configurations at their full sizes and horizons against the
FP64 oracle (tests/accuracy.py describes the checkpoint method).

* C2, 1,048,576 elastic particles, 1000 substeps: checkpoints 0, 250, 500
  and 1000 in both P2G modes, plus the whole first 10 substeps.
* C3, 4,194,304 fluid particles, 2000 substeps of a dam break with block
  churn: checkpoints 0 and 2000 in both modes (tests/accuracy_report.py
  adds 500 and 1000).
* C1, 8,192 particles: the whole 200-substep run, device and FP32
  restatement against FP64.

At every checkpoint each field's device error must stay within
max(reference tier, 4 x FP32 floor) (tests/accuracy.py), the block partition
and the live node set must equal the oracle's brute force key for key, and
the device's FP64 momentum audit must follow the FP64 oracle's momentum
change to 1e-3 of the substep's gravity impulse (walls act in C3)."""

import numpy as np
import pytest

from tessera_amd.scenes import dam_break, four_cubes, uniform_cube

from . import accuracy as A

pytestmark = pytest.mark.gpu


def check(results):
    for r in results:
        for f in A.FIELDS:
            assert r["gpu"][f] <= r["tolerance"][f], (r["config"], r["mode"], r["checkpoint"], f, r["gpu"][f],
                                                      r["tolerance"][f])
        assert r["blocks_exact"], r
        assert r["nodes_exact"], r
        assert r["momentum_audit_rel"] < 1e-3, r


@pytest.fixture(scope="module")
def c2():
    return four_cubes()


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_four_cubes_checkpoints_over_1000_substeps(transfers, c2, mode):
    check(A.run_checkpoints(transfers, c2, mode, (0, 250, 500, 1000)))


def test_four_cubes_first_10_substeps_track_fp64(transfers, c2):
    for r in A.trajectories(transfers, c2, ("fp32", "fixed64"), 10, threads=A.oracle_threads()):
        print(f"four cubes {r['mode']}, 10 substeps: gpu {A.fmt(r['gpu'])} floor {A.fmt(r['floor'])}")
        for f in A.FIELDS:
            assert r["gpu"][f] <= r["tolerance"][f], (r["mode"], f, r)


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_dam_break_checkpoints_over_2000_substeps(transfers, mode):
    check(A.run_checkpoints(transfers, dam_break(), mode, (0, 2000)))


@pytest.mark.parametrize("mode", ["fp32", "fixed64"])
def test_uniform_cube_200_substeps_within_tier_and_floor(transfers, mode):
    scene = uniform_cube()
    # The sticky floor at y = 0.05 is never reached in these 200 substeps (the
    # cube falls ~2 mm from y = 0.4), so the FP32 restatement, which has no
    # obstacles, runs the same world without it.
    r = A.trajectory(transfers, scene, mode, scene.steps, fp32=False)
    bare = scene.world.clone()
    bare.set_obstacles([])
    from oracle import oracle as O

    cpu = O.OracleWorld(bare)
    f32 = cpu.fp32_substeps(scene.dt, scene.steps)
    cpu.substeps_fixed(scene.dt, scene.steps)
    ref = cpu.read()
    live = np.asarray(bare.particles.enabled, bool) & ~np.asarray(bare.particles.fixed, bool)
    assert np.min(ref["position"][:, 1]) > 0.05 + 2 * scene.world.params.cell_width
    floor = {f: A.max_err(f32[f], ref[f], live) for f in A.FIELDS}
    tol = A.tolerance(floor)
    print(f"uniform cube {mode}, 200 substeps: gpu {A.fmt(r['gpu'])} floor {A.fmt(floor)}")
    for f in A.FIELDS:
        assert r["gpu"][f] <= tol[f], (f, r["gpu"][f], tol[f])
