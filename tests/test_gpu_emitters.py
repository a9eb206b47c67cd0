"""Mesh volume sampling on the device and emitters into resident sessions,
following the reference's own tests (sampling.rs:319-370, emitter.rs:236-347)."""

import math

import numpy as np
import pytest

from tessera_amd.emitters import BoxEmitter, MeshEmitter, sample_closed_mesh_volume
from tessera_amd.mpm import (Capacity, HipMpmResidentSession, InvalidInput, MaterialModel, MpmParams, MpmParticle,
                             MpmWorld, ParticleArrays, UnknownChunk)

pytestmark = pytest.mark.gpu

CUBE_V = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float)
CUBE_T = np.array([[0, 2, 1], [0, 3, 2], [4, 5, 6], [4, 6, 7], [0, 1, 5], [0, 5, 4], [3, 7, 6], [3, 6, 2],
                   [0, 4, 7], [0, 7, 3], [1, 2, 6], [1, 6, 5]], np.uint32)


def test_closed_mesh_volume_samples_lattice_and_reversed_winding(transfers):
    s = sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T, 0.25, 64, 64)
    assert len(s) == 64
    assert tuple(s[0]) == (0.125, 0.125, 0.125) and tuple(s[63]) == (0.875, 0.875, 0.875)
    assert np.array_equal(s, sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T[:, [0, 2, 1]], 0.25, 64, 64))
    joined_v = np.concatenate([CUBE_V, CUBE_V + [2.0, 0.0, 0.0]])
    joined_t = np.concatenate([CUBE_T, CUBE_T + 8])
    j = sample_closed_mesh_volume(transfers, joined_v, joined_t, 0.25, 192, 128)
    assert len(j) == 128
    assert np.all((j[:, 0] < 1.0) | (j[:, 0] > 2.0))


def test_closed_mesh_volume_rejects_open_or_badly_oriented_surfaces_and_limits(transfers):
    with pytest.raises(InvalidInput):
        sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T[:11], 0.25, 64, 64)
    flipped = CUBE_T.copy()
    flipped[0] = flipped[0][[0, 2, 1]]
    with pytest.raises(InvalidInput):
        sample_closed_mesh_volume(transfers, CUBE_V, flipped, 0.25, 64, 64)
    with pytest.raises(Capacity):
        sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T, 0.25, 63, 64)
    with pytest.raises(Capacity):
        sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T, 0.25, 64, 63)
    with pytest.raises(InvalidInput):
        sample_closed_mesh_volume(transfers, CUBE_V, CUBE_T, math.nan, 64, 64)


def test_sphere_mesh_volume_matches_the_analytic_inside_test(transfers):
    # an icosphere: every accepted centre lies inside the polyhedron and
    # every rejected one outside (checked against its face planes)
    t = (1.0 + 5 ** 0.5) / 2
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t], [0, -1, -t], [0, 1, -t],
                  [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], float) * 0.3
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4], [11, 10, 2],
                  [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8], [3, 8, 9], [4, 9, 5],
                  [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], np.uint32)
    spacing = 0.037
    s = sample_closed_mesh_volume(transfers, v, f, spacing, 10 ** 6, 10 ** 5)
    normals = np.cross(v[f[:, 1]] - v[f[:, 0]], v[f[:, 2]] - v[f[:, 0]])
    offsets = np.einsum("ij,ij->i", normals, v[f[:, 0]])
    lo = v.min(0)
    counts = np.ceil((v.max(0) - lo) / spacing).astype(int)
    grid = np.stack(np.meshgrid(*[np.arange(c) for c in counts], indexing="ij"), -1).reshape(-1, 3)
    pts = lo + (grid + 0.5) * spacing
    inside = np.all(pts @ normals.T - offsets < 0, axis=1) & np.all(pts < v.max(0), axis=1)
    assert np.array_equal(s, pts[inside])


def test_mesh_emitter_creates_mass_and_a_removable_session_chunk(transfers):
    # emitter.rs:236-287
    seed = MpmParticle.new([3.0, 3.0, 3.0], 0.05, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    s = HipMpmResidentSession(transfers, MpmWorld(ParticleArrays.from_particles([seed]), MpmParams()), 1e-3)
    e = MeshEmitter(CUBE_V, CUBE_T, 0.25, 1000.0, MaterialModel.elastic(1000.0, 0.2), 64, 64,
                    velocity=np.array([0.0, 0.0, 1.0]), group_id=7)
    chunk = e.emit(s)
    pa = s.read_particles()
    new = pa.chunk_id == chunk
    assert new.sum() == 64
    assert np.all(pa.rest_volume[new] == 0.25 ** 3) and np.all(pa.mass[new] == 1000.0 * 0.25 ** 3)
    assert np.all(pa.radius[new] == 0.125) and np.all(pa.group_id[new] == 7)
    assert np.all(pa.velocity[new] == [0.0, 0.0, 1.0])
    s.step(1)
    assert s.remove_chunk(chunk) == 64
    assert len(s.read_particles()) == 1
    e.max_particles = 63
    with pytest.raises(Capacity):
        e.emit(s)
    assert len(s.read_particles()) == 1


def test_emitted_chunks_can_be_removed_oldest_first(transfers):
    # emitter.rs:316-347
    seed = MpmParticle.new([3.0, 3.0, 3.0], 0.05, 1000.0, MaterialModel.elastic(1000.0, 0.2))
    s = HipMpmResidentSession(transfers, MpmWorld(ParticleArrays.from_particles([seed]), MpmParams()), 1e-3)
    e = BoxEmitter(np.array([0.0, 0.0, 0.5]), (2, 1, 1), np.full(3, 0.1), 0.04, 1000.0,
                   MaterialModel.elastic(1000.0, 0.2))
    first = e.emit(s)
    e.center = np.array([1.0, 0.0, 0.5])
    second = e.emit(s)
    assert first != second and len(s.read_particles()) == 5
    assert s.remove_chunk(first) == 2
    rest = s.read_particles()
    assert len(rest) == 3 and np.sum(rest.chunk_id == second) == 2
    s.step(1)
    assert s.remove_chunk(second) == 2
    with pytest.raises(UnknownChunk):
        s.remove_chunk(first)
