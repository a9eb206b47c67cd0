"""The allocation-free mesh sampler (tmpm_sample_closed_mesh_volume_noalloc)
against the allocating one (tmpm_sample_closed_mesh_volume), through the C ABI.

Both must give the same samples bit for bit and the same errors on two
calls in a row, the second with a smaller mesh and lattice (the scratch keeps
the first call's contents, so the second call must not depend on them).
Beyond its fixed scratch the allocation-free one fails with Capacity."""

import ctypes as C

import numpy as np
import pytest

from tessera_amd.mpm import Capacity, InvalidInput, _check, _dp, load_library

pytestmark = pytest.mark.gpu

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
CUBE_T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                   [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)


def icosphere(levels: int):
    """A closed, outward-wound icosphere of radius 0.3 subdivided `levels` times."""
    t = (1.0 + 5 ** 0.5) / 2
    v = [np.array(p, float) for p in [[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                                      [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]]]
    f = [[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
         [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5], [2, 4, 11],
         [6, 2, 10], [8, 6, 7], [9, 8, 1]]
    v = [p / np.linalg.norm(p) for p in v]
    for _ in range(levels):
        mid = {}

        def midpoint(a, b):
            key = (min(a, b), max(a, b))
            if key not in mid:
                p = v[a] + v[b]
                v.append(p / np.linalg.norm(p))
                mid[key] = len(v) - 1
            return mid[key]

        nf = []
        for a, b, c in f:
            ab, bc, ca = midpoint(a, b), midpoint(b, c), midpoint(c, a)
            nf += [[a, ab, ca], [b, bc, ab], [c, ca, bc], [ab, bc, ca]]
        f = nf
    return np.array(v) * 0.3 + 0.5, np.array(f, np.uint32)


def sample(noalloc: bool, ctx, vertices, triangles, spacing, max_candidates, max_samples):
    lib = load_library()
    fn = lib.tmpm_sample_closed_mesh_volume_noalloc if noalloc else lib.tmpm_sample_closed_mesh_volume
    v = np.ascontiguousarray(np.asarray(vertices, np.float64).reshape(-1, 3))
    t = np.ascontiguousarray(np.asarray(triangles, np.uint32).reshape(-1, 3))
    out = np.full((int(max_samples), 3), np.nan)
    count = C.c_uint64()
    _check(fn(ctx.handle, v.ctypes.data_as(_dp), len(v), t.ctypes.data_as(C.POINTER(C.c_uint32)), len(t),
              float(spacing), int(max_candidates), int(max_samples), out.ctypes.data_as(_dp), C.byref(count)), lib)
    assert np.all(np.isnan(out[count.value:]))  # nothing written past the count
    return out[: count.value].copy()


def test_noalloc_sampler_equals_the_allocating_one_on_two_calls(transfers):
    big_v, big_t = icosphere(3)  # 1,280 triangles
    cases = [(big_v, big_t, 0.011, 10 ** 7, 10 ** 6),  # ~200k candidates
             (CUBE_V, CUBE_T, 0.25, 64, 64)]           # then a smaller mesh and lattice
    for v, t, spacing, max_c, max_s in cases:
        want = sample(False, transfers, v, t, spacing, max_c, max_s)
        got = sample(True, transfers, v, t, spacing, max_c, max_s)
        assert len(want) > 0
        assert np.array_equal(got, want)
    # the same twice in a row on the scratch the previous call left
    again = sample(True, transfers, big_v, big_t, 0.011, 10 ** 7, 10 ** 6)
    assert np.array_equal(again, sample(False, transfers, big_v, big_t, 0.011, 10 ** 7, 10 ** 6))


def test_noalloc_sampler_reports_the_same_errors(transfers):
    flipped = CUBE_T.copy()
    flipped[0] = flipped[0][[0, 2, 1]]
    bad = [((CUBE_V, CUBE_T[:11], 0.25, 64, 64), InvalidInput),   # open surface
           ((CUBE_V, flipped, 0.25, 64, 64), InvalidInput),       # one face wound the other way
           ((CUBE_V, CUBE_T, 0.25, 63, 64), Capacity),            # max_candidates
           ((CUBE_V, CUBE_T, 0.25, 64, 63), Capacity),            # max_samples
           ((CUBE_V, CUBE_T, float("nan"), 64, 64), InvalidInput)]
    for args, err in bad:
        for noalloc in (False, True):
            with pytest.raises(err):
                sample(noalloc, transfers, *args)
    # a good call after the failures still matches
    assert np.array_equal(sample(True, transfers, CUBE_V, CUBE_T, 0.125, 512, 512),
                          sample(False, transfers, CUBE_V, CUBE_T, 0.125, 512, 512))


def test_noalloc_sampler_rejects_what_exceeds_its_scratch(transfers):
    # a lattice of 400^3 = 64M candidate flags does not fit the 32 MiB scratch
    with pytest.raises(Capacity):
        sample(True, transfers, CUBE_V, CUBE_T, 1.0 / 400, 10 ** 9, 10)
    # more than 65,536 triangles: disjoint tetrahedra, each closed
    tet_v = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float) * 0.01
    tet_t = np.array([[0, 2, 1], [0, 1, 3], [0, 3, 2], [1, 2, 3]], np.uint32)
    k = 16_385
    shifts = np.arange(k, dtype=float)[:, None] * np.array([0.02, 0.0, 0.0])
    v = (tet_v[None] + shifts[:, None]).reshape(-1, 3)
    t = (tet_t[None] + 4 * np.arange(k, dtype=np.uint32)[:, None, None]).reshape(-1, 3)
    with pytest.raises(Capacity):
        sample(True, transfers, v, t, 0.005, 10 ** 9, 10)
