"""Pins the binning oracle (f32 cell rule, live-node set, block partition)
against the reference's exact topology expectations
(crates/tessera-mpm/src/gpu.rs:3183-3186, 3777-3811) and brute force."""

import itertools

import numpy as np

from oracle import oracle as O


def test_live_node_counts_of_shared_duplicate_and_disabled_stencils():
    # gpu.rs:3183-3226: five particles on a row at y = z = 0.05, h = 0.1, the
    # last one disabled; shared and duplicate stencil nodes count once.
    for xs, expected in (([0.05, 0.15, 0.05, 10.05, 20.05], 63), ([0.05, 10.05, 20.05, 20.05, 30.05], 81)):
        pos = [[x, 0.05, 0.05] for x in xs]
        active = [1, 1, 1, 1, 0]
        assert len(O.node_set(pos, 0.1, active)) == expected
    pos = [[x, 0.05, 0.05] for x in (0.05, 0.15, 0.05, 10.05, 20.05)]
    assert len(O.node_set(pos, 0.1, [0] * 5)) == 0


def test_dense_and_far_apart_stencils():
    # gpu.rs:3784-3799: cells [-3,2,4] and [-2,2,4] span 4x3x3 = 36 nodes;
    # far-apart stencils (one at the lowest representable cell) do not merge.
    h = 1.0
    pos = [[-3 + 1.0, 2 + 1.0, 4 + 1.0], [-2 + 1.0, 2 + 1.0, 4 + 1.0]]
    assert [tuple(O.cell_f32(p, h)) for p in pos] == [(-3, 2, 4), (-2, 2, 4)]
    assert len(O.node_set(pos, h)) == 36
    low = [-2147483647.0, 1.0, 1.0]
    assert O.cell_f32(low, h)[0] == -2147483648
    assert len(O.node_set([low, [1.0e9, 1.0, 1.0]], h)) == 54


def test_cell_rule_rejects_unrepresentable_and_non_finite_positions():
    assert O.cell_f32([np.nan, 0, 0], 0.1) is None
    assert O.cell_f32([np.inf, 0, 0], 0.1) is None
    assert O.cell_f32([3.0e9, 0, 0], 1.0) is None
    # i32::MAX - 2 is the highest base cell the stencil allows
    assert O.cell_f32([-3.0e9, 0, 0], 1.0) is None


def test_cell_rule_rounds_each_operation_to_f32_without_fma():
    # x * inv_h - 0.5 evaluated with two f32 roundings; brute-force compare
    # against numpy float32 arithmetic on values that straddle cell faces.
    rng = np.random.default_rng(11)
    h = np.float32(1.0 / 64.0)
    inv_h = np.float32(1.0) / h
    for _ in range(2000):
        c = rng.integers(-200, 200)
        x = np.float32((c + 0.5) / 64.0) + np.float32(rng.integers(-3, 4)) * np.float32(1e-7)
        expected = int(np.floor(np.float32(np.float32(x * inv_h) - np.float32(0.5))))
        assert O.cell_f32([x, 0.5, 0.5], float(h))[0] == expected


def brute_nodes(cells, regions):
    s = set()
    for c, r in zip(cells, regions):
        for d in itertools.product(range(3), repeat=3):
            s.add((c[0] + d[0], c[1] + d[1], c[2] + d[2], r))
    return sorted(s)


def test_node_set_and_partition_match_brute_force():
    rng = np.random.default_rng(3)
    h = 0.05
    pos = rng.uniform(-0.5, 0.5, size=(300, 3))
    region = rng.integers(0, 3, 300).astype(np.uint32)
    active = rng.random(300) > 0.1
    cells = [O.cell_f32(p, h) for p in pos]
    ref = brute_nodes([c for c, a in zip(cells, active) if a], [r for r, a in zip(region, active) if a])
    got = [tuple(k) for k in O.node_set(pos, h, active.astype(np.uint8), region)]
    assert got == ref
    blocks, part = O.block_partition(pos, h, active.astype(np.uint8), region)
    for i in range(300):
        if not active[i]:
            assert part[i] == -1
            continue
        b = blocks[part[i]]
        c = cells[i]
        assert tuple(b) == (c[0] // 4, c[1] // 4, c[2] // 4, region[i])
    assert len(blocks) == len({(c[0] // 4, c[1] // 4, c[2] // 4, r) for c, r, a in zip(cells, region, active) if a})
