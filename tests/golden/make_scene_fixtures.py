"""Writes the scene fixtures that pin the synthetic workload generators
(tessera_amd/scenes.py): the 8,192 C1 positions as float64, and SHA-256
record hashes (first 4096 position/velocity/mass/volume records) of C1-C4
and of a fixed 16-request serving stream.

    python tests/golden/make_scene_fixtures.py

Rerun only when a generator changes on purpose; tests/test_scene_fixtures.py
fails on any other change, since the benchmark numbers rest on these inputs.
"""

from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent.parent))

import numpy as np  # noqa: E402

from tessera_amd.scenes import batch_scenes, dam_break, four_cubes, record_hash, serving_requests, uniform_cube  # noqa: E402


def hashes() -> dict:
    out = {"c1_uniform_cube_8k": record_hash(uniform_cube().world.particles),
           "c2_four_cubes_1M": record_hash(four_cubes().world.particles),
           "c3_dam_break_4M": record_hash(dam_break().world.particles)}
    batch = batch_scenes(256, seed=0)
    out["c4_batch_256_scenes"] = hashlib.sha256(
        "".join(record_hash(s.world.particles) for s in batch).encode()).hexdigest()
    out["c4_batch_particles"] = int(sum(len(s.world.particles) for s in batch))
    reqs = serving_requests(16, 1.0, seed=7)
    out["c5_serving_16_requests"] = hashlib.sha256(
        "".join(record_hash(r.world.particles) + f"{r.horizon}:{len(r.bodies)}" for r in reqs).encode()).hexdigest()
    out["c5_serving_16_particles"] = int(sum(r.particles for r in reqs))
    return out


def main() -> None:
    np.save(HERE / "c1_positions.npy", uniform_cube().world.particles.position.astype(np.float64))
    (HERE / "scene_hashes.json").write_text(json.dumps(hashes(), indent=1) + "\n")
    print("wrote", HERE / "c1_positions.npy", HERE / "scene_hashes.json")


if __name__ == "__main__":
    main()
