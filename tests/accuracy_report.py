"""Writes the per-field accuracy table of the benchmark configurations
against the FP64 oracle (the measurements of tests/accuracy.py, with the
FP32 restatement's floor beside each device error).

    python tests/accuracy_report.py --out accuracy_vs_fp64.md

Needs the GPU.  C1 runs its whole 200 substeps; C2 (1000 substeps) and C3
(2000) are checked at checkpoints, in both P2G modes, and C2 also over a
whole-trajectory window from its initial state (--c2-window substeps, FP64
on all host cores).  C4 runs the 256-scene batch for its whole 100 substeps
and compares a sample of its scenes; C5 compares the largest request of a
full-size serving stream over a coupled MPM-rigid window.
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))

import numpy as np  # noqa: E402

from oracle import oracle as O  # noqa: E402
from tessera_amd.mpm import HipMpmBatch, HipMpmResidentSession, HipMpmTransfers  # noqa: E402
from tessera_amd.scenes import batch_scenes, dam_break, four_cubes, serving_requests, uniform_cube  # noqa: E402
from tests import accuracy as A  # noqa: E402
from tests.test_gpu_coupling import oracle_coupled  # noqa: E402

# MaterialModel.kind (variant order of material.rs)
KIND = {0: "corotated", 1: "neo-Hookean", 2: "snow", 3: "fluid"}


def log(msg: str) -> None:
    print(msg, flush=True)


def c1_rows(transfers) -> list[dict]:
    scene = uniform_cube()
    bare = scene.world.clone()
    bare.set_obstacles([])
    cpu = O.OracleWorld(bare)
    f32 = cpu.fp32_substeps(scene.dt, scene.steps)
    cpu.substeps_fixed(scene.dt, scene.steps)
    ref = cpu.read()
    live = np.asarray(bare.particles.enabled, bool)
    floor = {f: A.max_err(f32[f], ref[f], live) for f in A.FIELDS}
    rows = []
    for mode in ("fp32", "fixed64"):
        r = A.trajectory(transfers, scene, mode, scene.steps, fp32=False)
        r.update(floor=floor, tolerance=A.tolerance(floor), checkpoint="0-200 (whole run)")
        log(f"C1 {mode}: gpu {A.fmt(r['gpu'])} floor {A.fmt(floor)}")
        rows.append(r)
    return rows


def c4_rows(transfers, sample: int = 8) -> list[dict]:
    """C4: the 256-scene batch advanced together for its whole 100-substep horizon in both P2G
    modes; `sample` scenes spread over the size range are compared with the FP64 oracle and the
    FP32 restatement, each run alone from the scene's initial state."""
    scenes = batch_scenes()
    steps = scenes[0].steps
    order = sorted(range(len(scenes)), key=lambda k: len(scenes[k].world.particles))
    picks = [order[int(round(q * (len(order) - 1) / (sample - 1)))] for q in range(sample)]
    got = {}
    for mode in ("fp32", "fixed64"):
        b = HipMpmBatch(transfers, [s.world for s in scenes], [s.dt for s in scenes], mode)
        A.advance(b, steps)
        for k in picks:
            got[(mode, k)] = b.read_particles(k)
        b.close()
    rows = []
    for k in picks:
        sc = scenes[k]
        pa = sc.world.particles
        cpu = O.OracleWorld(sc.world)
        f32 = cpu.fp32_substeps(sc.dt, steps)
        cpu.substeps_fixed_threads(sc.dt, steps, A.oracle_threads())
        ref = cpu.read()
        live = np.asarray(pa.enabled, bool) & ~np.asarray(pa.fixed, bool)
        floor = {f: A.max_err(f32[f], ref[f], live) for f in A.FIELDS}
        kind = KIND.get(pa.materials[0].kind, "?") if pa.materials else "?"
        for mode in ("fp32", "fixed64"):
            g = {f: A.max_err(getattr(got[(mode, k)], f), ref[f], live) for f in A.FIELDS}
            rows.append(dict(config=f"batch scene {k}: {len(pa):,} {kind}, h=1/{round(1 / sc.world.params.cell_width)}",
                             mode=mode, checkpoint=f"0-{steps} (whole run)", gpu=g, floor=floor,
                             tolerance=A.tolerance(floor)))
            log(f"C4 scene {k} {mode}: gpu {A.fmt(g)} floor {A.fmt(floor)}")
    return rows


def c5_rows(transfers, window: int = 16) -> tuple[list[dict], list[dict]]:
    """C5: the largest request of a full-size serving stream (the stream the serving test serves;
    a served scene equals the same world stepped alone bit for bit, which that test checks), its
    first `window` coupled MPM-rigid substeps against the FP64 oracle's coupling.  Obstacles are
    outside the FP32 restatement, so the floor is measured on the same particles with the
    obstacles removed, over the same window (the cube starts a cell above its floor and moves
    ~1e-4 m in 16 substeps, so its contacts barely enter the window)."""
    reqs = serving_requests(12, rate=1.0e4, seed=21)
    req = max(reqs, key=lambda r: r.particles)
    g = np.asarray(req.gravity)
    cpu, ref_bodies = oracle_coupled(req.world, list(req.bodies), list(req.mappings), req.dt, window, g)
    ref = cpu.read()
    pa0 = req.world.particles
    live = np.asarray(pa0.enabled, bool) & ~np.asarray(pa0.fixed, bool)
    bare = req.world.clone()
    bare.set_obstacles([])
    cpu_b = O.OracleWorld(bare)
    f32 = cpu_b.fp32_substeps(req.dt, window)
    cpu_b.substeps_fixed_threads(req.dt, window, A.oracle_threads())
    ref_b = cpu_b.read()
    floor = {f: A.max_err(f32[f], ref_b[f], live) for f in A.FIELDS}
    del cpu_b, ref_b, f32
    rows, body_rows = [], []
    for mode in ("fp32", "fixed64"):
        s = HipMpmResidentSession(transfers, req.world, req.dt, mode)
        s.set_rigid_bodies(list(req.bodies), list(req.mappings), gravity=req.gravity)
        A.advance(s, window)
        pa, bodies = s.read_particles(), s.rigid_bodies()
        s.close()
        err = {f: A.max_err(getattr(pa, f), ref[f], live) for f in A.FIELDS}
        rows.append(dict(config=f"serving request {req.name}: {len(pa0):,} particles, {len(req.bodies)} boxes",
                         mode=mode, checkpoint=f"0-{window} (coupled window)", gpu=err, floor=floor,
                         tolerance=A.tolerance(floor)))
        for j, (b, w) in enumerate(zip(bodies, ref_bodies)):
            body_rows.append(dict(mode=mode, body=j, position=float(np.abs(b.position - w["position"]).max()),
                                  linear_velocity=float(np.abs(b.linear_velocity - w["linear_velocity"]).max()),
                                  angular_velocity=float(np.abs(b.angular_velocity - w["angular_velocity"]).max()),
                                  speed=float(np.abs(w["linear_velocity"]).max())))
        log(f"C5 {mode}: gpu {A.fmt(err)} floor (no obstacles) {A.fmt(floor)}")
    return rows, body_rows


def body_table(rows: list[dict]) -> list[str]:
    out = ["### C5: rigid boxes of that request against the FP64 coupling", "",
           "| mode | box | position error (m) | linear velocity error (m/s) | angular velocity error (1/s) | "
           "oracle speed (m/s) |", "|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| {r['mode']} | {r['body']} | {r['position']:.3e} | {r['linear_velocity']:.3e} | "
                   f"{r['angular_velocity']:.3e} | {r['speed']:.3f} |")
    out.append("")
    return out


def table(rows: list[dict], title: str) -> list[str]:
    out = [f"### {title}", "",
           "| config | mode | substeps | field | GPU error | FP32 floor | tolerance | within |",
           "|---|---|---|---|---|---|---|---|"]
    for r in rows:
        for f in A.FIELDS:
            g, t = r["gpu"][f], r["tolerance"][f]
            fl = f"{r['floor'][f]:.3e}" if r["floor"] else "n/a"
            out.append(f"| {r['config']} | {r['mode']} | {r.get('checkpoint', r.get('steps'))} | {f} | {g:.3e} | "
                       f"{fl} | {t:.1e} | {'yes' if g <= t else '**no**'} |")
    out.append("")
    return out


def topo_table(rows: list[dict]) -> list[str]:
    out = ["### Topology and momentum at the checkpoints", "",
           "| config | mode | substep | live nodes | node set = oracle | particle blocks | partition = oracle | "
           "momentum audit error / (M g dt) |", "|---|---|---|---|---|---|---|---|"]
    for r in rows:
        out.append(f"| {r['config']} | {r['mode']} | {r['checkpoint']} | {r['live_nodes']} | "
                   f"{'yes' if r['nodes_exact'] else '**no**'} | {r['particle_blocks']} | "
                   f"{'yes' if r['blocks_exact'] else '**no**'} | {r['momentum_audit_rel']:.2e} |")
    out.append("")
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default="accuracy_vs_fp64.md")
    ap.add_argument("--c2-window", type=int, default=20)
    ap.add_argument("--skip-c3", action="store_true")
    ap.add_argument("--c4-sample", type=int, default=8, help="batch scenes compared (0: skip C4)")
    ap.add_argument("--c5-window", type=int, default=16, help="coupled substeps of the C5 request (0: skip C5)")
    args = ap.parse_args()
    t0 = time.perf_counter()
    transfers = HipMpmTransfers(0)
    report = {"c1": c1_rows(transfers)}
    c2 = four_cubes()
    log(f"C2: whole-trajectory window of {args.c2_window} substeps")
    report["c2_window"] = A.trajectories(transfers, c2, ("fp32", "fixed64"), args.c2_window,
                                         threads=A.oracle_threads())
    for r in report["c2_window"]:
        r["checkpoint"] = f"0-{args.c2_window} (whole window)"
        log(f"C2 {r['mode']} window: gpu {A.fmt(r['gpu'])} floor {A.fmt(r['floor'])}")
    report["c2"] = []
    for mode in ("fp32", "fixed64"):
        log(f"C2 {mode}: checkpoints")
        report["c2"] += A.run_checkpoints(transfers, c2, mode, (0, 250, 500, 1000), log)
    del c2
    report["c3"] = []
    if not args.skip_c3:
        c3 = dam_break()
        for mode in ("fp32", "fixed64"):
            log(f"C3 {mode}: checkpoints")
            report["c3"] += A.run_checkpoints(transfers, c3, mode, (0, 500, 1000, 2000), log)
    report["c4"] = c4_rows(transfers, args.c4_sample) if args.c4_sample > 1 else []
    body_rows = []
    report["c5"] = []
    if args.c5_window > 0:
        report["c5"], body_rows = c5_rows(transfers, args.c5_window)
        report["c5_bodies"] = body_rows
    lines = ["# Accuracy against the FP64 oracle", "",
             f"Command: `python tests/accuracy_report.py --c2-window {args.c2_window} --c4-sample {args.c4_sample} "
             f"--c5-window {args.c5_window}` on one MI355X; "
             f"host threads for the FP64 oracle: {A.oracle_threads()}; {time.perf_counter() - t0:.0f} s.", "",
             "Errors are the largest per-particle Euclidean norm of the field difference (position m, velocity m/s, "
             "affine C 1/s, deformation F) over enabled, non-fixed particles.  The FP32 floor is the oracle's FP32 "
             "restatement of the reference's GPU substep (gpu.wgsl:428-1019) against FP64 from the same state.  "
             f"Tolerance = max(reference tier of gpu.rs:4043-4098, {A.K_FLOOR:g} x FP32 floor).", "",
             "Checkpoint rows: the device state at that substep, then one substep on the device, in FP64 and in "
             "FP32 from that identical state.", ""]
    lines += table(report["c1"], "C1: 8,192-particle cube, whole 200-substep run")
    lines += table(report["c2_window"], "C2: 1,048,576-particle four cubes, whole-trajectory window")
    lines += table(report["c2"], "C2: four cubes, one substep at checkpoints of the 1000-substep horizon")
    if report["c3"]:
        lines += table(report["c3"], "C3: 4,194,304-particle dam break, one substep at checkpoints of 2000")
    if report["c4"]:
        lines += table(report["c4"], f"C4: 256-scene batch advanced together, {len(report['c4']) // 2} scenes "
                                     "across the size range against each scene run alone in FP64")
    if report["c5"]:
        lines += table(report["c5"], "C5: largest request of a full-size serving stream, coupled MPM-rigid window "
                                     "(FP32 floor: the same particles without the obstacles, which the FP32 "
                                     "restatement leaves out)")
        lines += body_table(body_rows)
    lines += topo_table(report["c2"] + report["c3"])
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text("\n".join(lines) + "\n")
    Path(args.out).with_suffix(".json").write_text(json.dumps(report, indent=1, default=float))
    log(f"wrote {args.out}")
    bad = [r for key, rows in report.items() if key != "c5_bodies" for r in rows
           if any(r["gpu"][f] > r["tolerance"][f] for f in A.FIELDS)
           or not r.get("nodes_exact", True) or not r.get("blocks_exact", True)]
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
