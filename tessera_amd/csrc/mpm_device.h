// Device data layout shared by the HIP kernels (mpm_kernels.hip) and the
// host session (mpm_session.hip).
//
// HBM layout (structure of arrays, particle capacity `stride`):
//   pos_mass[i]  float4  x, y, z, mass (mass 0 marks a disabled particle)
//   vel_vol[i]   float4  vx, vy, vz, rest volume
//   C[k*stride+i] float  APIC affine matrix, k = column*3 + row
//   F[k*stride+i] float  deformation gradient, same indexing
//   P[k*stride+i] float  plastic state {plastic_det, hardening, log_volume_gain}
//   cls[i]       u32     particle class (material, radius, damping, flags, region)
//   ins_id[i]    u32     insertion index of the particle stored in slot i
// Grid: 4x4x4-cell blocks; node (slot s, local l) at s*64 + l with
// l = (x&3)*16 + (y&3)*4 + (z&3).  P2G writes one 6^3-node tile per particle
// block, each of its 8 sub-boxes into the area of the grid block it covers
// (ptile); the grid update sums, in a fixed order, the up to 8 parts on each
// node (grid_rev says which exist), so no global atomics are needed.
#pragma once

#include <cstdint>

#include <hip/hip_runtime.h>

namespace tmpm {

constexpr uint32_t kInvalid = 0xffffffffu;
constexpr int kBlockCells = 4;        // cells per block edge
constexpr int kBlockNodes = 64;       // nodes per block
constexpr int kTileEdge = 6;          // P2G/G2P tile: block + 2-cell halo
constexpr int kTileNodes = 216;
constexpr int kMinBuckets = 1024;     // smallest block-hash table
constexpr int kMaxObstacles = 256;
constexpr int32_t kNoKey = INT32_MIN;  // no block key has x = INT32_MIN (cells are >> 2)

// counters[] slots
enum Counter : uint32_t {
  CNT_GRID_BLOCKS = 0,      // grid blocks of the current substep (clamped to capacity)
  CNT_PARTICLE_BLOCKS = 1,  // particle blocks of the current substep (clamped)
  CNT_BINNED = 2,
  CNT_ERROR = 3,
  CNT_LIVE_NODES = 5,
  CNT_REACT_SATURATED = 6,  // particle-side reaction impulses clamped to the bins' range
  CNT_COUNT = 8
};
// error bits in counters[CNT_ERROR]
enum ErrorBit : uint32_t {
  ERR_RESIDENT = 1u,        // CFL, out of grid, non-finite: ResidentUnavailable
  ERR_TOPOLOGY = 2u,        // hash or grid capacity exhausted: Capacity
  ERR_FIXED_RANGE = 4u,     // fixed-point contribution out of range: Capacity
  ERR_INTERNAL = 8u         // inconsistent binning (must never happen)
};

// Device material codes.  Snow and sand use the corotated stress, sand-NH the
// neo-Hookean one; all three carry a plastic state and a return mapping.
enum DevMaterial : int32_t {
  DM_COROTATED = 0,
  DM_NEO_HOOKEAN = 1,
  DM_FLUID = 2,
  DM_SNOW = 3,
  DM_SAND = 4,
  DM_SAND_NH = 5
};
__host__ __device__ constexpr bool dm_plastic(int32_t k) { return k >= DM_SNOW; }

struct DevClass {
  int32_t kind;
  uint32_t region;   // (scene << 16) | transfer colour; the block key's 4th field
  float inv_h;       // the scene's 1 / cell width, so binning reads one table per particle
  uint32_t fixed;
  float radius;
  float damping;
  float lambda, mu;                         // elastic Lame parameters
  float bulk, gamma, viscosity, tensile;    // fluid
  float wave_numerator;                     // lambda + 2 mu, or bulk * gamma
  float hardening;                          // snow: hardening coefficient
  float crit_compression, crit_stretch;     // snow: stretch clamp [1 - c, 1 + s]
  float friction_angle, cohesion;           // sand, sand-NH (radians; log-volume units)
};

struct DevScene {
  float h, inv_h, inv_d, dt;
  float gravity[3];
  int32_t has_bounds;
  float bmin[3];
  float bmax[3];
  int32_t first_base[3];
  int32_t last_base[3];
  int32_t obstacle_begin;
  int32_t obstacle_count;
  int32_t apply_force;     // read the force array in this substep's G2P
  int32_t mass_shift;      // fixed-point scale exponents
  int32_t mom_shift;
  int32_t body_begin;      // rigid bodies of this scene: [body_begin, body_begin + body_count)
  int32_t body_count;
  int32_t rigid_integrate; // MPM-owned rigid step: gravity, speed cap and pose update
  float rigid_gravity[3];
  float rigid_max_speed;   // 0: no cap
};

struct DevObstacle {
  int32_t shape;
  int32_t boundary;
  float friction;
  float pad0;
  float center[3];
  float pad1;
  float q[4];          // i, j, k, w
  float lin[3];
  float pad2;
  float ang[3];
  float pad3;
  float dims[4];
  // reactions: contributions are added as int64 floor(value * 2^shift)
  float rscale_lin, rscale_ang;  // 2^(shift - 32), the fixed64_of input scale
  int32_t rshift_lin, rshift_ang;
  float tri[9];          // triangle prism: local vertices a, b, c
  float bound;           // prism, convex: radius of the contact sphere about center
  int32_t plane_begin;   // convex: planes [plane_begin, plane_begin + plane_count) of Topo::planes
  int32_t plane_count;
  int32_t cpic_group;    // -1: none
  int32_t pad4[3];
};
constexpr int kMaxConvexPlanes = 128;
constexpr int kGatherFloats = 29;  // per particle in a scene gather (k_gather_scene)  // planes per convex obstacle (gpu.rs pack_obstacle)

// Rigid body state (GpuRigidBodyState, phys/src/gpu_rigid_state.rs:17-28):
// position and inverse mass, XYZW orientation, world velocities, body-frame
// inverse principal inertia.  Zero inverse mass: static or kinematic.
struct DevRigid {
  float pos[3];
  float inv_mass;
  float q[4];
  float lin[3];
  float pad0;
  float ang[3];
  float pad1;
  float inv_inertia[3];
  float pad2;
};

// Obstacle -> rigid body mapping (gpu_rigid_sync.rs:15-30): body index in the
// session's body array (-1: obstacle not driven), pose in the body frame.
struct DevMapping {
  int32_t body;
  float local_center[3];
  float local_q[4];
};

// FP64 audit record per scene and substep: mass; momentum (3); kinetic
// energy; first moment of mass (3); angular momentum about the origin (3);
// particles counted.
constexpr int kAuditFields = 12;

constexpr int kReactBins = 16;       // int64 accumulators per obstacle component
constexpr int kMaxReactObstacles = 16;  // per scene, G2P's LDS reaction sums (later obstacles: direct atomics)

struct Parts {
  float4* pos_mass;
  float4* vel_vol;
  float* C;
  float* F;
  float* P;          // plastic state, 3 per particle (component-major): plastic_det,
                     // hardening, log_volume_gain; null when no scene has a plastic material
  uint32_t* cls;
  float4* force;
  uint32_t* ins_id;
  uint32_t n;        // particles in storage
  uint32_t stride;   // SoA capacity
};

struct Topo {
  uint32_t* owner;          // bucket -> owner id (0 empty, 1..n particle, > n neighbour insert)
  int4* bkey;               // bucket -> key, written by the claimer; x = kNoKey until then
  uint32_t* bucket_count;   // particles binned into the bucket's block
  uint32_t* bucket_start;   // bucket -> first block-sorted position of its particle block
  uint32_t* bucket_slot;    // bucket -> grid slot
  uint32_t* bucket_pblock;  // bucket -> particle-block id
  uint32_t bucket_mask;     // bucket count - 1 (power of two)
  int4* grid_key;           // slot -> (bx, by, bz, region)
  uint32_t grid_cap;
  uint32_t pblock_cap;
  uint32_t* pblock_nbr;     // 8 per particle block: buckets of block + {0,1}^3
  uint32_t* pblock_slot8;   // 8 grid slots per particle block
  uint32_t* pblock_start;
  uint32_t* pblock_len;
  uint32_t* pslot;          // particle -> bucket
  uint32_t* prank;          // particle -> rank within its block
  uint32_t* sorted;         // binned order -> particle storage index
  uint32_t* counters;
  uint32_t* scene_err;      // per scene: error bits of its particles and blocks
  uint32_t isolate;         // batches: binning skips the particles of scenes whose error bits are set
  // binning lists: insert workgroup g bins particles [g * bin_span, (g + 1) * bin_span)
  uint32_t bin_span;        // particles per binning workgroup (multiple of 256)
  uint32_t bin_groups;      // binning workgroups
  uint32_t* bin_stage;      // bin_span per group: buckets the group claimed (particle blocks)
  uint32_t* bin_nstage;     // 7 * bin_span per group: neighbour-only blocks it claimed
  uint32_t* bin_claims;     // per group: particle blocks claimed
  uint32_t* bin_nclaims;    // per group: neighbour-only blocks claimed
  uint32_t* bin_csum;       // per group: particles in its particle blocks
  // reactions (when enabled): [obstacle][bin][6] int64, then per obstacle the
  // substep's reaction (6 floats) and the sum since the last readback (6 doubles)
  long long* react_bins;
  float* react_sub;
  double* react_total;
  uint32_t obstacle_total;
  void* ptile;              // per grid slot, 216 nodes: the tile parts covering it (FP32: float4; FIXED64: 4 x int64)
  uint32_t* grid_rev;       // per grid slot, 8 entries: particle block whose tile covers it at offset d
  float4* grid_vel;
  uint32_t* node_mask;      // 2 words per grid slot (diagnostic live-node set)
  const float4* planes;     // convex obstacle planes {nx, ny, nz, offset}, local frame
  // CPIC (sessions with CPIC obstacle groups): per particle slot, the side
  // mask of this substep; per bucket, the mask of its block key
  uint32_t* pmask;
  uint32_t* bmask;
};

}  // namespace tmpm
