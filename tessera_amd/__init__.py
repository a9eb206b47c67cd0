"""MI355X (gfx950) HIP backend for tessera-mpm's sparse-grid MLS-MPM substep.

The compute path is the hand-written HIP library ``libtessera_mpm_hip.so``
(C ABI in include/tessera_mpm.h); :mod:`tessera_amd.mpm` binds it.
"""

from .mpm import (  # noqa: F401
    GpuMpmError,
    HipMpmResidentSession,
    HipMpmTransfers,
    MaterialModel,
    MpmParams,
    MpmParticle,
    MpmWorld,
    ObstacleBoundary,
    ParticleArrays,
    RigidObstacle,
    WorldBounds,
)
