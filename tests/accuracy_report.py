"""Writes the per-field accuracy table of the benchmark configurations
against the FP64 oracle (the measurements of tests/accuracy.py, with the
FP32 restatement's floor beside each device error).

    python tests/accuracy_report.py --out gpurun_out/accuracy_vs_fp64.md

Needs the GPU.  C1 runs its whole 200 substeps; C2 (1000 substeps) and C3
(2000) are checked at checkpoints, in both P2G modes, and C2 also over a
whole-trajectory window from its initial state (--c2-window substeps, FP64
on all host cores).
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
from tessera_amd.mpm import HipMpmTransfers  # noqa: E402
from tessera_amd.scenes import dam_break, four_cubes, uniform_cube  # noqa: E402
from tests import accuracy as A  # noqa: E402


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


def table(rows: list[dict], title: str) -> list[str]:
    out = [f"### {title}", "",
           "| config | mode | substeps | field | GPU error | FP32 floor | tolerance | within |",
           "|---|---|---|---|---|---|---|---|"]
    for r in rows:
        for f in A.FIELDS:
            g, fl, t = r["gpu"][f], (r["floor"] or {}).get(f, float("nan")), r["tolerance"][f]
            out.append(f"| {r['config']} | {r['mode']} | {r.get('checkpoint', r.get('steps'))} | {f} | {g:.3e} | "
                       f"{fl:.3e} | {t:.1e} | {'yes' if g <= t else '**no**'} |")
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
    ap.add_argument("--out", default="gpurun_out/accuracy_vs_fp64.md")
    ap.add_argument("--c2-window", type=int, default=20)
    ap.add_argument("--skip-c3", action="store_true")
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
    lines = ["# Accuracy against the FP64 oracle", "",
             f"Command: `python tests/accuracy_report.py --c2-window {args.c2_window}` on one MI355X; "
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
    lines += topo_table(report["c2"] + report["c3"])
    Path(args.out).parent.mkdir(parents=True, exist_ok=True)
    Path(args.out).write_text("\n".join(lines) + "\n")
    Path(args.out).with_suffix(".json").write_text(json.dumps(report, indent=1, default=float))
    log(f"wrote {args.out}")
    bad = [r for rows in report.values() for r in rows
           if any(r["gpu"][f] > r["tolerance"][f] for f in A.FIELDS)
           or not r.get("nodes_exact", True) or not r.get("blocks_exact", True)]
    return 1 if bad else 0


if __name__ == "__main__":
    sys.exit(main())
