"""CPU checks of the C-ABI boundary: the library loads, exports every symbol
include/tessera_mpm.h declares, and the ctypes structures match the C
layouts (sizes and offsets computed by the C compiler)."""

import ctypes as C
import subprocess
import textwrap
from pathlib import Path

import pytest

from tessera_amd import mpm

ROOT = Path(__file__).resolve().parent.parent


def test_library_exports_every_declared_symbol():
    lib = mpm.load_library()
    declared = mpm.exported_symbols()
    assert len(declared) >= 30
    missing = [name for name in declared if not hasattr(lib, name)]
    assert not missing, missing


def test_version_and_error_strings_without_a_device():
    lib = mpm.load_library()
    assert b"gfx950" in lib.tmpm_version()
    assert lib.tmpm_last_error() is not None


def test_oracle_library_loads():
    from oracle import oracle as O

    assert O.lib() is not None


@pytest.fixture(scope="module")
def c_layout(tmp_path_factory):
    src = textwrap.dedent(
        """
        #include <stdio.h>
        #include <stddef.h>
        #include "tessera_mpm.h"
        #define S(t) printf(#t " %zu\\n", sizeof(t));
        #define O(t, f) printf(#t "." #f " %zu\\n", offsetof(t, f));
        int main(void) {
          S(tmpm_material) S(tmpm_params) S(tmpm_obstacle) S(tmpm_particles) S(tmpm_world_desc)
          S(tmpm_topology_stats)
          O(tmpm_obstacle, planes) O(tmpm_obstacle, dims) O(tmpm_params, bound_max)
          O(tmpm_world_desc, obstacles) O(tmpm_world_desc, substeps) O(tmpm_particles, chunk_id)
          return 0;
        }
        """
    )
    d = tmp_path_factory.mktemp("layout")
    (d / "l.c").write_text(src)
    exe = d / "l"
    subprocess.run(["gcc", "-I", str(ROOT / "include"), str(d / "l.c"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, stdout=subprocess.PIPE, text=True).stdout
    return dict(line.split() for line in out.strip().splitlines())


def test_ctypes_structures_match_c_layout(c_layout):
    pairs = {
        "tmpm_material": mpm.CMaterial, "tmpm_params": mpm.CParams, "tmpm_obstacle": mpm.CObstacle,
        "tmpm_particles": mpm.CParticles, "tmpm_world_desc": mpm.CWorld, "tmpm_topology_stats": mpm.CTopologyStats,
    }
    for cname, py in pairs.items():
        assert int(c_layout[cname]) == C.sizeof(py), cname
    assert int(c_layout["tmpm_obstacle.planes"]) == mpm.CObstacle.planes.offset
    assert int(c_layout["tmpm_obstacle.dims"]) == mpm.CObstacle.dims.offset
    assert int(c_layout["tmpm_params.bound_max"]) == mpm.CParams.bound_max.offset
    assert int(c_layout["tmpm_world_desc.obstacles"]) == mpm.CWorld.obstacles.offset
    assert int(c_layout["tmpm_world_desc.substeps"]) == mpm.CWorld.substeps.offset
    assert int(c_layout["tmpm_particles.chunk_id"]) == mpm.CParticles.chunk_id.offset


def test_context_without_gfx950_device_reports_no_device():
    # Without a visible gfx950 device the context reports NoDevice instead of
    # falling back to anything; with one, the context opens.
    try:
        ctx = mpm.HipMpmTransfers(0)
    except mpm.NoDevice:
        return
    ctx.close()
