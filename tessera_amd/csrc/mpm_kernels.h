// Launch wrappers of the substep kernels (mpm_kernels.hip).
#pragma once

#include <hip/hip_runtime.h>

#include "mpm_device.h"

namespace tmpm {

void launch_topo_clear(const Topo& t, hipStream_t s);
// cpic != 0: block keys carry the particles' CPIC side masks (t.pmask, t.bmask)
void launch_bin_insert(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int cpic,
                       hipStream_t s);
void launch_bin_blocks(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int cpic,
                       hipStream_t s);
void launch_cpic_mask(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, const DevObstacle* o,
                      hipStream_t s);
void launch_bin_offsets(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, hipStream_t s);
void launch_bin_scatter(const Parts& p, const Topo& t, hipStream_t s);
void launch_p2g(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, int fixed64, uint32_t blocks,
                uint32_t threads, int copies, int plastic, hipStream_t s);
// shapes != 0: some obstacle is a triangle prism or convex polyhedron
void launch_grid_update(const Topo& t, const DevScene* sc, const DevObstacle* o, int fixed64, int react, int shapes,
                        hipStream_t s);
// q: the reorder target (the whole state moved into block order), or empty
void launch_g2p(const Parts& p, const Parts& q, const Topo& t, const DevClass* c, const DevScene* sc,
                const DevObstacle* o, uint32_t blocks, uint32_t threads, int react, int shapes, int plastic,
                hipStream_t s);
void launch_react_finalize(const Topo& t, const DevObstacle* o, hipStream_t s);
void launch_rigid_sync(const DevRigid* bodies, const DevMapping* maps, DevObstacle* o, uint32_t count, hipStream_t s);
void launch_rigid_apply(const Topo& t, const DevScene* sc, DevRigid* bodies, const DevMapping* maps,
                        const DevObstacle* o, const uint32_t* body_scene, uint32_t count, hipStream_t s);
void launch_snapshot(const Parts& p, float4* out, hipStream_t s);
void launch_snapshot_disabled(const float4* rec, uint32_t count, float4* out, hipStream_t s);
void launch_retire(const Parts& p, uint32_t first, uint32_t count, hipStream_t s);
void launch_renumber(uint32_t* ins_id, uint32_t n, uint32_t first, uint32_t count, hipStream_t s);
void launch_gather_scene(const Parts& p, uint32_t first, uint32_t count, float* out, hipStream_t s);
void launch_copy(const float4* src, float4* dst, size_t n, hipStream_t s);
// flags[c] = 1 for lattice candidate c (x-major, then y, then z) inside the closed mesh
void launch_mesh_inside(const double* verts, const uint32_t* tris, uint64_t ntri, const double lower[3],
                        const double upper[3], double spacing, const uint64_t counts[3], uint8_t* flags,
                        hipStream_t s);
// *bad = 1 unless every directed edge of the triangles occurs exactly once
// and its reverse exactly once (the caller zeroes *bad)
void launch_mesh_edges(const uint32_t* tris, uint64_t ntri, uint32_t* bad, hipStream_t s);
// The first n particles' state (positions, velocities, classes, insertion
// indices, C, F and the plastic state) from src to dst, each with its stride;
// both P arrays must exist
void launch_copy_state(const Parts& dst, const Parts& src, uint32_t n, hipStream_t s);
// FP64 per-scene sums of the particle state (kAuditFields doubles per scene,
// added into out, which the caller zeroes)
void launch_audit(const Parts& p, const DevClass* c, double* out, hipStream_t s);
void launch_live_nodes(const Parts& p, const Topo& t, const DevClass* c, const DevScene* sc, hipStream_t s);

}  // namespace tmpm
