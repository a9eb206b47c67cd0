#!/usr/bin/env python3
"""Throughput benchmark: particle-updates/s of the resident HIP MLS-MPM step.

Workload: the single large elastic scene, 1,048,576 particles in four
jittered cubes of edge 0.2 with N(0, 0.5) m/s velocities, corotated elastic
(E=1e4, nu=0.3, rho=1e3), 4x4x4-cell blocks over a 256^3-cell domain
(h=1/256), dt=5e-5.  A step is one substep of the whole scene (binning,
P2G, grid update, G2P).  The timed region covers K device-resident substeps
with the particle state already in HBM: it starts after a device wait and
ends with a device wait; readback and validation follow outside it.

Printed JSON fields beyond the common contract:
  roofline      dominant kernel's algorithmic bytes per launch over its
                average duration (HIP events on the session stream), against
                MI355X's 8 TB/s HBM peak; plus the whole-substep figure.
                `traffic` is that kernel's measured HBM bytes per launch from
                the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
  cpu_baseline  the FP64 C++ restatement of the reference CPU substep
                (oracle/, one thread) timed on a bounded sample of the same
                scene on this host
"""

from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes (FP32 state): P2G reads x, v, C, F, m, V0 (104 B per
# particle) and writes one 16-B (m, p) accumulator per live node; G2P writes
# x, v, C, F (96 B per particle) and reads one 16-B velocity per live node.
# Together: 200 B per particle-update + 32 B per live node per substep.
P2G_BYTES_PARTICLE, P2G_BYTES_NODE = 104, 16
G2P_BYTES_PARTICLE, G2P_BYTES_NODE = 96, 16


TRAFFIC_TABLE = ROOT / "profiles" / "substep_four_cubes_1M_hbm_traffic.md"


def measured_traffic(kernel: str, mode: str):
    """HBM bytes per launch of `kernel` from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE summary
    (already scaled by the copy-kernel calibration), or None if it has no row for this kernel."""
    if not TRAFFIC_TABLE.exists():
        return None
    prefix = {"k_p2g": "tmpm::k_p2g<0", "k_p2g_fixed64": "tmpm::k_p2g<1"}.get(kernel, "tmpm::" + kernel)
    for line in TRAFFIC_TABLE.read_text().splitlines():
        cells = [c.strip().strip("`") for c in line.strip().strip("|").split("|")]
        if len(cells) == 4 and cells[0].startswith(prefix):
            try:
                return float(cells[3]) * 1e6
            except ValueError:
                return None
    return None


def submit(session, steps: int) -> None:
    left = steps
    while left > 0:
        k = min(64, left)
        session.submit_steps(k)
        left -= k


def batch_leg(transfers, mode: str, steps: int) -> dict:
    """256 independent scenes (N_i log-uniform in [2048, 65536]) advanced as one batch, against
    the same scenes advanced one resident session at a time (each timed from a device wait to a
    device wait; sessions created beforehand)."""
    from tessera_amd.mpm import HipMpmBatch, HipMpmResidentSession
    from tessera_amd.scenes import batch_scenes

    scenes = batch_scenes(256, seed=0)
    total = sum(len(sc.world.particles) for sc in scenes)
    b = HipMpmBatch(transfers, [sc.world for sc in scenes], [sc.dt for sc in scenes], mode)
    submit(b, 8)
    b.wait()
    t0 = time.perf_counter()
    submit(b, steps)
    b.wait()
    t_default = time.perf_counter() - t0
    # the launch-shape autotuner (its trials are real substeps with identical results)
    cfg = b.autotune(16)
    t0 = time.perf_counter()
    submit(b, steps)
    b.wait()
    t_batch = time.perf_counter() - t0
    # the same launches captured once per distinct substep plan and replayed as HIP graphs
    b.set_graphs(True)
    submit(b, 8)
    b.wait()
    t0 = time.perf_counter()
    submit(b, steps)
    b.wait()
    t_graphs = time.perf_counter() - t0
    b.set_graphs(bool(cfg.get("graphs")))
    b.synchronize()
    failed = sum(1 for k in range(len(scenes)) if b.scene_error(k) is not None)
    b.close()
    sessions = [HipMpmResidentSession(transfers, sc.world, sc.dt, mode) for sc in scenes]
    t_alone = 0.0
    for s in sessions:
        submit(s, 8)
        s.wait()
        t0 = time.perf_counter()
        submit(s, steps)
        s.wait()
        t_alone += time.perf_counter() - t0
    for s in sessions:
        s.synchronize()
        s.close()
    value = total * steps / t_batch
    alone = total * steps / t_alone
    print(f"bench: batch of {len(scenes)} scenes ({total} particles): {value:.4e} ({total * steps / t_graphs:.4e} "
          f"replayed as graphs) vs one session at a time {alone:.4e} particle-updates/s", file=sys.stderr, flush=True)
    return {"workload": "batch_256_scenes", "scenes": len(scenes), "particles": total, "steps": steps,
            "value": value, "unit": "particle-updates/s", "ms_per_step": t_batch / steps * 1e3,
            "value_default_shape": total * steps / t_default, "launch_config": cfg,
            "value_graphs": total * steps / t_graphs, "ms_per_step_graphs": t_graphs / steps * 1e3,
            "one_session_at_a_time": alone, "speedup": value / alone, "failed_scenes": failed}


def scenes_leg(transfers, mode: str, steps: int) -> dict:
    """Throughput of the other single-scene workloads: the 8,192-particle cube (C1) and the
    4,194,304-particle fluid dam break (C3), with the default and the autotuned launch shape."""
    from tessera_amd.mpm import HipMpmResidentSession
    from tessera_amd.scenes import dam_break, uniform_cube

    out = {}
    for name, make in (("c1_uniform_cube_8k", uniform_cube), ("c3_dam_break_fluid_4M", dam_break)):
        sc = make()
        n = len(sc.world.particles)
        s = HipMpmResidentSession(transfers, sc.world, sc.dt, mode)
        submit(s, 16)
        s.wait()
        t0 = time.perf_counter()
        submit(s, steps)
        s.wait()
        t_default = time.perf_counter() - t0
        cfg = s.autotune(16)
        t0 = time.perf_counter()
        submit(s, steps)
        s.wait()
        t_tuned = time.perf_counter() - t0
        s.synchronize()
        out[name] = {"particles": n, "steps": steps, "dt": sc.dt, "value": n * steps / t_tuned,
                     "value_default_shape": n * steps / t_default, "unit": "particle-updates/s",
                     "launch_config": cfg}
        print(f"bench: {name}: {n * steps / t_tuned:.4e} particle-updates/s (default shape "
              f"{n * steps / t_default:.4e})", file=sys.stderr, flush=True)
        s.close()
    return out


def serving_leg(transfers, mode: str, seconds: float, single_rate: float, plain_seconds: float = 15.0) -> dict:
    """Continuous batching: Poisson arrivals offering 80% of the measured single-scene rate,
    N log-uniform in [4096, 524288], 1-4 rigid boxes per scene, horizons U{50..2000} substeps,
    admission under an HBM budget; runs until every request has completed.  The same stream is
    then served by the plain per-scene path (one resident session per request on the same GPU)
    for at most `plain_seconds` (0: skip)."""
    from tessera_amd.scenes import serving_requests
    from tessera_amd.serving import PlainServer, Server

    mean_n = (524288 - 4096) / math.log(524288 / 4096)
    work = mean_n * 1025.0  # particle-updates per request
    rate = 0.8 * single_rate / work
    count = max(1, int(round(rate * seconds)))
    reqs = serving_requests(count, rate, seed=7)
    budget = 16 << 30
    server = Server(transfers, hbm_budget=budget, steps_per_tick=32, mode=mode, keep_worlds=False)
    report = server.run(reqs)
    server.close()
    report.update({"workload": "serving_poisson", "arrival_rate_per_s": rate, "offered_fraction": 0.8,
                   "steps_per_tick": 32, "hbm_budget_bytes": budget})
    print(f"bench: serving {report['completed']}/{count} requests in {report['wall_s']:.1f}s: "
          f"{report['value']:.4e} particle-updates/s, step latency p50 {report['step_latency_ms_p50']:.3f} ms "
          f"p99 {report['step_latency_ms_p99']:.3f} ms", file=sys.stderr, flush=True)
    if plain_seconds > 0:
        # the plain path is far slower on the same stream: it serves the stream for a bounded
        # window, its throughput being the updates completed within it over its length
        ps = PlainServer(transfers, hbm_budget=budget, steps_per_tick=32, mode=mode, keep_worlds=False)
        base = ps.run(reqs, time_limit=plain_seconds)
        ps.close()
        base["window_s"] = plain_seconds
        report["plain"] = base
        report["speedup"] = report["value"] / base["value"] if base["value"] > 0 else None
        print(f"bench: plain per-scene path {base['completed']}/{count} requests in {base['wall_s']:.1f}s: "
              f"{base['value']:.4e} particle-updates/s, step latency p50 {base['step_latency_ms_p50']:.3f} ms "
              f"p99 {base['step_latency_ms_p99']:.3f} ms; batched/plain {report['speedup']:.2f}x",
              file=sys.stderr, flush=True)
    return report


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--mode", choices=["fp32", "fixed64"], default="fp32")
    ap.add_argument("--timing-steps", type=int, default=128, help="substeps of the per-kernel event pass")
    ap.add_argument("--cpu-substeps", type=int, default=2, help="substeps of the CPU-baseline sample")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-extra-modes", action="store_true")
    ap.add_argument("--batch-steps", type=int, default=64, help="substeps of the batched-scenes leg (0: skip)")
    ap.add_argument("--serve-seconds", type=float, default=3.0, help="arrival window of the serving leg (0: skip)")
    ap.add_argument("--plain-seconds", type=float, default=15.0,
                    help="window of the plain per-scene path on the serving stream (0: skip)")
    ap.add_argument("--scenes-steps", type=int, default=128, help="substeps of the C1/C3 scene leg (0: skip)")
    args = ap.parse_args()

    world_size = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world_size != 1:
        print("bench.py measures one GPU (the substep shards by scene or by spatial slab across processes, "
              "which it does not launch), so --gpus must be 1.", file=sys.stderr)
        return 2

    from tessera_amd.mpm import HipMpmResidentSession, HipMpmTransfers
    from tessera_amd.scenes import four_cubes, record_hash

    scene = four_cubes()
    n = len(scene.world.particles)
    transfers = HipMpmTransfers(0)
    session = HipMpmResidentSession(transfers, scene.world, scene.dt, args.mode)
    print(f"bench: {scene.name}, {n} particles, mode {args.mode}, warmup {args.warmup}, steps {args.steps}",
          file=sys.stderr, flush=True)

    submit(session, args.warmup)
    session.wait()
    t0 = time.perf_counter()
    submit(session, args.steps)
    session.wait()
    t1 = time.perf_counter()
    session.synchronize()  # validates device flags and reads the state back (untimed)
    elapsed = t1 - t0
    value = n * args.steps / elapsed
    print(f"bench: timed {args.steps} substeps in {elapsed:.4f}s -> {value:.4e} particle-updates/s",
          file=sys.stderr, flush=True)

    # Per-kernel durations from HIP events on the session stream.
    session.set_timing(True)
    submit(session, args.timing_steps)
    session.synchronize()
    times = session.kernel_times()
    session.set_timing(False)
    stats = session.topology_stats()
    live = stats["live_nodes"]

    substep_ms = sum(ms for ms, _ in times.values())
    kernels = {k: {"avg_ms": ms, "calls": c} for k, (ms, c) in sorted(times.items(), key=lambda kv: -kv[1][0])}
    p2g_name = "k_p2g_fixed64" if args.mode == "fixed64" else "k_p2g"
    alg = {
        p2g_name: P2G_BYTES_PARTICLE * n + P2G_BYTES_NODE * live,
        "k_g2p": G2P_BYTES_PARTICLE * n + G2P_BYTES_NODE * live,
    }
    dominant = max(alg, key=lambda k: times.get(k, (0.0, 0))[0])
    dom_ms = times[dominant][0]
    achieved = alg[dominant] / (dom_ms * 1e-3) / 1e9
    substep_bytes = (P2G_BYTES_PARTICLE + G2P_BYTES_PARTICLE) * n + (P2G_BYTES_NODE + G2P_BYTES_NODE) * live
    wall_substep_gbs = substep_bytes / (elapsed / args.steps) / 1e9

    extra_modes = {}
    if not args.no_extra_modes:
        other = "fixed64" if args.mode == "fp32" else "fp32"
        s2 = HipMpmResidentSession(transfers, scene.world, scene.dt, other)
        submit(s2, args.warmup)
        s2.wait()
        k2 = min(args.steps, 256)
        ta = time.perf_counter()
        submit(s2, k2)
        s2.wait()
        tb = time.perf_counter()
        s2.synchronize()
        extra_modes[other] = {"value": n * k2 / (tb - ta), "steps": k2}
        s2.close()

    batch = None
    if args.batch_steps > 0:
        batch = batch_leg(transfers, args.mode, args.batch_steps)

    scenes = None
    if args.scenes_steps > 0:
        scenes = scenes_leg(transfers, args.mode, args.scenes_steps)

    serving = None
    if args.serve_seconds > 0:
        serving = serving_leg(transfers, args.mode, args.serve_seconds, value, args.plain_seconds)

    cpu = cpu_all = None
    if not args.no_cpu_baseline:
        from oracle import oracle as O

        ow = O.OracleWorld(scene.world)
        ca = time.perf_counter()
        ow.substeps_fixed(scene.dt, args.cpu_substeps)
        cb = time.perf_counter()
        cpu = {
            "value": n * args.cpu_substeps / (cb - ca),
            "unit": "particle-updates/s",
            "cores": 1,
            "kind": "port",
            "sample": f"{args.cpu_substeps} substeps of the full {n}-particle scene, FP64 C++ restatement of "
                      f"MpmWorld::substep (single thread, {cb - ca:.1f}s)",
        }
        # the same restatement over every host core this process may use (per-chunk P2G grids
        # merged in thread order); reported beside the single-thread port
        threads = max(1, min(16, len(os.sched_getaffinity(0))))
        ow = O.OracleWorld(scene.world)
        ca = time.perf_counter()
        ow.substeps_fixed_threads(scene.dt, args.cpu_substeps, threads)
        cb = time.perf_counter()
        cpu_all = {
            "value": n * args.cpu_substeps / (cb - ca),
            "unit": "particle-updates/s",
            "cores": threads,
            "kind": "port",
            "sample": f"{args.cpu_substeps} substeps of the full {n}-particle scene, FP64 C++ restatement on "
                      f"{threads} threads ({cb - ca:.1f}s)",
        }

    out = {
        "metric": "particle-updates/s",
        "value": value,
        "unit": "particle-updates/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "four_cubes_elastic_1M",
            "particles": n,
            "cell_width": scene.world.params.cell_width,
            "dt": scene.dt,
            "block_cells": 4,
            "p2g_mode": args.mode,
            "scene_hash": record_hash(scene.world.particles)[:16],
        },
        "roofline": {
            "bound": "hbm",
            "kernel": dominant,
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(dominant, args.mode),
            "traffic_source": str(TRAFFIC_TABLE.relative_to(ROOT)),
            "algorithmic_bytes": alg[dominant],
            "avg_ms": dom_ms,
            "live_nodes": live,
            "substep_algorithmic_bytes": substep_bytes,
            "substep_wall_gbs": wall_substep_gbs,
            "substep_wall_frac": wall_substep_gbs / HBM_PEAK_GBS,
        },
        "kernels": kernels,
        "event_substep_ms": substep_ms,
        "other_modes": extra_modes,
        "cpu_baseline": cpu,
        "cpu_baseline_all_cores": cpu_all,
        "batch": batch,
        "scenes": scenes,
        "serving": serving,
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
