"""Accuracy against the FP64 oracle at the benchmark configurations' full
sizes and horizons (shared by tests/test_gpu_accuracy.py and
tests/accuracy_report.py).

A one-million or four-million particle scene cannot be run for its whole
horizon in FP64 in test time, so each configuration is checked at
checkpoints along its horizon: the device session is stepped to the
checkpoint and its state read back (f32 values, exact in f64); from that
identical state one substep runs on the device, in the FP64 oracle and in
the oracle's FP32 restatement of the reference's GPU substep
(gpu.wgsl:428-1019).  Per field the device error and the FP32 floor are both
measured against FP64.  At each checkpoint the topology the device builds is
compared key for key with the oracle's brute-force node set and block
partition of the same positions, and the device's FP64 momentum audit over
the substep with the FP64 oracle's momentum change.

Tolerance rule per field: max(reference tier, K_FLOOR x FP32 floor), the
tier being gpu.rs:4043-4098's (position 1e-5, velocity 1e-4, affine C 1e-3,
deformation F 1e-5).
"""

from __future__ import annotations

import os
import time

import numpy as np

from oracle import oracle as O
from tessera_amd.mpm import HipMpmResidentSession

FIELDS = ("position", "velocity", "affine", "deformation")
# gpu.rs:4043-4098 (gpu_transfers_track_cpu_reference_for_elastic_particles)
TIER = dict(position=1e-5, velocity=1e-4, affine=1e-3, deformation=1e-5)
# multiple of the FP32 restatement's error that a field may reach
K_FLOOR = 4.0


def oracle_threads() -> int:
    return max(1, min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4))


def max_err(a: np.ndarray, b: np.ndarray, mask: np.ndarray) -> float:
    d = (np.asarray(a) - np.asarray(b)).reshape(len(a), -1)[mask]
    return float(np.max(np.linalg.norm(d, axis=1))) if len(d) else 0.0


def tolerance(floor: dict) -> dict:
    return {f: max(TIER[f], K_FLOOR * floor[f]) for f in FIELDS}


def advance(session: HipMpmResidentSession, steps: int) -> None:
    left = steps
    while left:
        k = min(64, left)
        session.submit_steps(k)
        session.synchronize()
        left -= k


def momentum(pa) -> np.ndarray:
    m = np.asarray(pa.mass) * np.asarray(pa.enabled, bool)
    return (m[:, None] * np.asarray(pa.velocity)).sum(0)


def checkpoint(transfers, session: HipMpmResidentSession, dt: float, log=print) -> dict:
    """One substep from the session's current state on the device, in FP64 and
    in FP32, plus the topology and momentum checks; returns the measurements."""
    t0 = time.perf_counter()
    world = session.world()
    pa = world.particles
    live = np.asarray(pa.enabled, bool) & ~np.asarray(pa.fixed, bool)
    h = world.params.cell_width
    # topology of this state: the device's binning against brute force
    keys, part = session.export_blocks()
    ref_keys, ref_part = O.block_partition(pa.position, h, pa.enabled.astype(np.uint8), pa.transfer_color)
    blocks_exact = bool(np.array_equal(keys, ref_keys) and np.array_equal(part, ref_part))
    nodes = session.export_nodes()
    ref_nodes = O.node_set(pa.position, h, pa.enabled.astype(np.uint8), pa.transfer_color)
    nodes_exact = bool(len(nodes) == len(ref_nodes) and np.array_equal(nodes, ref_nodes))
    n_nodes, n_blocks = len(ref_nodes), len(ref_keys)
    del keys, part, nodes, ref_keys, ref_part, ref_nodes
    # one substep three ways from the identical state
    cpu = O.OracleWorld(world)
    fp32 = cpu.fp32_substeps(dt, 1)
    cpu.substeps_fixed_threads(dt, 1, oracle_threads())
    ref = cpu.read()
    del cpu
    session.set_audit(True)
    session.step(1)
    session.set_audit(False)
    audit = session.audit()
    got = session.read_particles()
    gpu_err = {f: max_err(getattr(got, f), ref[f], live) for f in FIELDS}
    floor = {f: max_err(fp32[f], ref[f], live) for f in FIELDS}
    # momentum over the substep: device FP64 audit against the FP64 oracle
    p_before = momentum(pa)
    m = np.asarray(pa.mass) * np.asarray(pa.enabled, bool)
    p_ref = (m[:, None] * ref["velocity"]).sum(0)
    dp_gpu = audit[-1, 1:4] - p_before
    dp_ref = p_ref - p_before
    gdt = float(np.linalg.norm(np.asarray(world.params.gravity)) * m.sum() * dt)
    audit_err = float(np.linalg.norm(dp_gpu - dp_ref) / gdt)
    out = dict(substep=int(world.substeps), gpu=gpu_err, floor=floor, tolerance=tolerance(floor),
               blocks_exact=blocks_exact, nodes_exact=nodes_exact, live_nodes=n_nodes, particle_blocks=n_blocks,
               momentum_audit_rel=audit_err, seconds=time.perf_counter() - t0)
    log(f"checkpoint {out['substep']}: gpu {fmt(gpu_err)} floor {fmt(floor)} nodes {n_nodes} exact={nodes_exact} "
        f"blocks {n_blocks} exact={blocks_exact} dp/(Mg dt) err {audit_err:.2e} ({out['seconds']:.1f}s)")
    return out


def fmt(d: dict) -> str:
    return " ".join(f"{k[:3]}={v:.2e}" for k, v in d.items())


def run_checkpoints(transfers, scene, mode: str, checkpoints, log=print) -> list[dict]:
    """Steps `scene` through its horizon, checking each checkpoint substep."""
    s = HipMpmResidentSession(transfers, scene.world, scene.dt, mode)
    results = []
    done = 0
    for c in checkpoints:
        advance(s, c - done)
        r = checkpoint(transfers, s, scene.dt, log)
        r.update(config=scene.name, mode=mode, checkpoint=c)
        results.append(r)
        done = c + 1
    s.close()
    return results


def trajectory(transfers, scene, mode: str, steps: int, threads: int | None = None, fp32: bool = True) -> dict:
    """Whole-trajectory error after `steps` substeps from the initial state:
    device and FP32 restatement against the FP64 oracle."""
    return trajectories(transfers, scene, (mode,), steps, threads, fp32)[0]


def trajectories(transfers, scene, modes, steps: int, threads: int | None = None, fp32: bool = True) -> list[dict]:
    """`trajectory` for several P2G modes against one FP64 (and FP32) oracle run."""
    got = {}
    for mode in modes:
        s = HipMpmResidentSession(transfers, scene.world, scene.dt, mode)
        advance(s, steps)
        got[mode] = s.read_particles()
        s.close()
    cpu = O.OracleWorld(scene.world)
    f32 = cpu.fp32_substeps(scene.dt, steps) if fp32 else None
    if threads:
        cpu.substeps_fixed_threads(scene.dt, steps, threads)
    else:
        cpu.substeps_fixed(scene.dt, steps)
    ref = cpu.read()
    pa = scene.world.particles
    live = np.asarray(pa.enabled, bool) & ~np.asarray(pa.fixed, bool)
    floor = {f: max_err(f32[f], ref[f], live) for f in FIELDS} if fp32 else None
    out = []
    for mode in modes:
        gpu_err = {f: max_err(getattr(got[mode], f), ref[f], live) for f in FIELDS}
        out.append(dict(config=scene.name, mode=mode, steps=steps, gpu=gpu_err, floor=floor,
                        tolerance=tolerance(floor) if floor else dict(TIER)))
    return out
