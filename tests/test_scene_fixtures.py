"""The synthetic benchmark inputs are pinned: regenerating every scene gives
the committed record hashes (tests/golden/scene_hashes.json) and, for the
8,192-particle cube, exactly the committed positions.  The C2 hash is the
`scene_hash` bench.py prints beside its result."""

import json
from pathlib import Path

import numpy as np

from tessera_amd.scenes import uniform_cube

GOLDEN = Path(__file__).resolve().parent / "golden"


def test_c1_positions_equal_the_committed_fixture():
    ref = np.load(GOLDEN / "c1_positions.npy", allow_pickle=False)
    assert np.array_equal(uniform_cube().world.particles.position, ref)


def test_every_scene_generator_reproduces_its_committed_hash():
    from tests.golden.make_scene_fixtures import hashes

    assert hashes() == json.loads((GOLDEN / "scene_hashes.json").read_text())
