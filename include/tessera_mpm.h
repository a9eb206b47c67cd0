/*
 * tessera_mpm.h — C ABI of the MI355X (gfx950) HIP backend for tessera-mpm's
 * sparse-grid MLS-MPM substep.
 *
 * The backend is a drop-in for the reference's resident WebGPU session
 * (crates/tessera-mpm/src/gpu.rs:206-662, `GpuMpmResidentSession`) and its
 * world-level entry points (gpu.rs:2177-2293).  Every function below names the
 * reference interface it replaces.  Plain pointers and sizes only; no
 * exceptions cross the boundary; every call returns a tmpm_status whose
 * values map one-to-one onto the reference's error enums
 * (`GpuMpmError`, gpu.rs:135-154, and `MpmError`, world.rs:165-181).
 *
 * Threading: a session is externally serialized (the reference takes
 * `&mut self` for every mutating call).  Each session owns one HIP stream on
 * its context's device.  Different sessions may be used from different
 * threads concurrently.
 *
 * Units and layouts follow the reference world exactly: f64 SI values,
 * vectors as 3 consecutive doubles, 3x3 matrices as 9 doubles in nalgebra's
 * column-major order (element (row r, column c) at index c*3 + r), particles
 * in insertion order.
 */
#ifndef TESSERA_MPM_H
#define TESSERA_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes.  gpu.rs:135-154 (GpuMpmError) and world.rs:165-181 (MpmError). */
typedef enum tmpm_status {
  TMPM_OK = 0,
  TMPM_INVALID_INPUT = 1,           /* GpuMpmError::InvalidInput / MpmError::InvalidInput */
  TMPM_CAPACITY = 2,                /* GpuMpmError::Capacity (grid, hash or fixed-point range exhausted) */
  TMPM_PENDING_STEPS = 3,           /* GpuMpmError::PendingSteps */
  TMPM_READBACK = 4,                /* GpuMpmError::Readback (HIP runtime failure; message in tmpm_last_error) */
  TMPM_RESIDENT_UNAVAILABLE = 5,    /* GpuMpmError::ResidentUnavailable */
  TMPM_EXCESSIVE_SUBSTEPS = 6,      /* MpmError::ExcessiveSubsteps */
  TMPM_NON_FINITE_STATE = 7,        /* MpmError::NonFiniteState */
  TMPM_UNKNOWN_CHUNK = 8,           /* MpmError::UnknownChunk */
  TMPM_PARTICLE_CAPACITY = 9,       /* MpmError::ParticleCapacity */
  TMPM_NO_DEVICE = 10               /* no usable gfx950 device for the requested ordinal */
} tmpm_status;

/* Material kinds in `MaterialModel` variant order (material.rs:28-91).
 * Parameters p[0..4] in the variant's field order:
 *   LINEAR_ELASTIC:   young_modulus, poisson_ratio              (corotated)
 *   NEO_HOOKEAN:      young_modulus, poisson_ratio
 *   SAND:             young_modulus, poisson_ratio, friction_angle, cohesion
 *   FLUID:            bulk_modulus, gamma, viscosity, tensile_stiffness
 *   SNOW:             young_modulus, poisson_ratio, critical_compression, critical_stretch, hardening
 *   SAND_NEO_HOOKEAN: young_modulus, poisson_ratio, friction_angle, cohesion */
enum {
  TMPM_MATERIAL_LINEAR_ELASTIC = 0,
  TMPM_MATERIAL_NEO_HOOKEAN = 1,
  TMPM_MATERIAL_SAND = 2,
  TMPM_MATERIAL_FLUID = 3,
  TMPM_MATERIAL_SNOW = 4,
  TMPM_MATERIAL_SAND_NEO_HOOKEAN = 5
};

typedef struct tmpm_material {
  int32_t kind;
  int32_t reserved;
  double p[5];
} tmpm_material;

/* `MpmParams` (world.rs:27-49) with `use_cpic` (world.rs:217-219). */
typedef struct tmpm_params {
  double cell_width;
  double gravity[3];
  double max_substep;
  int32_t has_bounds;
  int32_t use_cpic;
  double bound_min[3];
  double bound_max[3];
} tmpm_params;

/* Obstacle shapes in `ObstacleShape` order (obstacle.rs:10-61) and boundary
 * responses in `ObstacleBoundary` order (obstacle.rs:65-77). */
enum {
  TMPM_SHAPE_SPHERE = 0,
  TMPM_SHAPE_BOX = 1,
  TMPM_SHAPE_CAPSULE = 2,
  TMPM_SHAPE_CYLINDER = 3,
  TMPM_SHAPE_CONE = 4,
  TMPM_SHAPE_GROUND = 5,
  TMPM_SHAPE_TRIANGLE_PRISM = 6,
  TMPM_SHAPE_CONVEX = 7
};
enum {
  TMPM_BOUNDARY_SLIP = 0,
  TMPM_BOUNDARY_STICK = 1,
  TMPM_BOUNDARY_SEPARATE = 2,
  TMPM_BOUNDARY_NON_REFLECTING = 3,
  TMPM_BOUNDARY_FREE_SLIP = 4
};

/* `RigidObstacle` (obstacle.rs:79-98).  dims: sphere {radius}; box {half
 * extents xyz}; capsule/cylinder/cone {half_height, radius}; ground {half
 * extents xy}; triangle prism {half_thickness}; convex {bound_radius}.
 * orientation is the unit quaternion (i, j, k, w). */
typedef struct tmpm_obstacle {
  int32_t shape;
  int32_t boundary;
  int32_t cpic_group; /* -1: none */
  uint32_t plane_count;
  double center[3];
  double orientation[4];
  double linear_velocity[3];
  double angular_velocity[3];
  double friction;
  double dims[4];
  double vertices[9];
  const double* planes; /* convex: plane_count x {nx, ny, nz, offset} */
} tmpm_obstacle;

/* Structure-of-arrays view of `MpmParticle` records (world.rs:67-102), all
 * arrays of `count` entries in insertion order.  material_index selects an
 * entry of the accompanying material table.  plastic holds
 * {plastic_det, hardening, log_volume_gain} (material.rs:5-14). */
typedef struct tmpm_particles {
  uint64_t count;
  double* position;      /* 3 per particle */
  double* velocity;      /* 3 */
  double* deformation;   /* 9, column-major */
  double* affine;        /* 9, column-major */
  double* force;         /* 3 */
  double* mass;          /* 1 */
  double* rest_volume;   /* 1 */
  double* radius;        /* 1 */
  double* damping;       /* 1 */
  double* plastic;       /* 3 */
  uint32_t* material_index;
  uint32_t* group_id;
  uint16_t* transfer_color;
  uint8_t* enabled;
  uint8_t* fixed;
  uint64_t* chunk_id;    /* session creation: the particles' chunks (NULL: chunk 0);
                            add_particles ignores it (the batch becomes a new chunk) */
} tmpm_particles;

/* A whole `MpmWorld` (world.rs:206-221) handed over at session creation. */
typedef struct tmpm_world_desc {
  tmpm_params params;
  tmpm_particles particles;
  const tmpm_material* materials;
  uint32_t material_count;
  uint32_t obstacle_count;
  const tmpm_obstacle* obstacles;
  uint64_t substeps;
  uint64_t next_chunk_id; /* id of the next added chunk; 0: one above the largest particle chunk */
} tmpm_world_desc;

/* P2G precision (the reference accumulates with f32 CAS loops,
 * gpu.wgsl:739-766).  Both modes add 64-bit fixed-point integers in LDS and
 * are bit-reproducible run to run.  FP32 rounds each particle block's tile
 * to f32 and sums up to 8 tiles per node in f32 in a fixed order; FIXED64
 * sums the integers exactly and rounds each node once. */
typedef enum tmpm_p2g_mode { TMPM_P2G_FP32 = 0, TMPM_P2G_FIXED64 = 1 } tmpm_p2g_mode;

typedef struct tmpm_context tmpm_context;
typedef struct tmpm_session tmpm_session;

/* Message of the last failing call on this thread ("" when none). */
const char* tmpm_last_error(void);

/* Library version string. */
const char* tmpm_version(void);

/* Device context; replaces `GpuMpmTransfers::new(&Device)` (gpu.rs:927-992),
 * which compiles the pipelines once per device.  Fails with TMPM_NO_DEVICE
 * when the ordinal does not name a gfx950 device. */
tmpm_status tmpm_context_create(int32_t device_ordinal, tmpm_context** out);
void tmpm_context_destroy(tmpm_context* ctx);

/* `GpuMpmResidentSession::new(gpu, device, queue, world, substep_dt)`
 * (gpu.rs:253-280, validation gpu.rs:664-677 and world.rs:225-241).  The
 * world is copied; the caller's arrays are not retained. */
tmpm_status tmpm_session_create(tmpm_context* ctx, const tmpm_world_desc* world, double substep_dt,
                                tmpm_p2g_mode mode, tmpm_session** out);
void tmpm_session_destroy(tmpm_session* s);

/* `submit_steps(steps)` (gpu.rs:303-322): 1..=64 substeps, asynchronous. */
tmpm_status tmpm_session_submit_steps(tmpm_session* s, uint32_t steps);
/* `synchronize()` (gpu.rs:382-411): wait, validate device flags, update the
 * host world.  Any failure invalidates the session and keeps the last world. */
tmpm_status tmpm_session_synchronize(tmpm_session* s);
/* Wait for the submitted substeps to finish on the device without reading
 * the particle state back or validating it (no reference counterpart; the
 * throughput benchmark uses it to time device-resident stepping alone). */
tmpm_status tmpm_session_wait(tmpm_session* s);
/* `synchronize_async()` (gpu.rs:413-450) as a poll: while the device still
 * runs the submitted substeps, returns TMPM_OK with *done_out = 0 and changes
 * nothing (the substeps stay pending; a caller may stop polling, as the
 * reference's cancellation keeps them pending too).  Once they have finished
 * it validates and commits exactly as tmpm_session_synchronize and sets
 * *done_out = 1, also when that fails (the failure invalidates the session
 * and keeps the last world).  No pending substeps: *done_out = 1. */
tmpm_status tmpm_session_synchronize_async(tmpm_session* s, int32_t* done_out);
/* `step(steps)` (gpu.rs:286-292) = submit_steps + synchronize. */
tmpm_status tmpm_session_step(tmpm_session* s, uint32_t steps);
/* `MpmWorld::step_with_gpu_transfers(dt)` (gpu.rs:2208-2240): CFL-limited
 * substeps of min(remaining, stable_timestep), synchronized after each. */
tmpm_status tmpm_session_step_adaptive(tmpm_session* s, double dt);
/* `has_pending_steps()` / `pending_substeps()` (gpu.rs:367-375). */
uint64_t tmpm_session_pending_substeps(const tmpm_session* s);
/* Device memory the session holds (no reference counterpart: wgpu sizes
 * its buffers per prepared transfer, gpu.rs:1391-1405): the bytes of every
 * device allocation, and the particle capacity of the state buffers.  A batch
 * keeps its capacity when scenes retire and reuses it for later admissions;
 * the serving loop budgets admissions on these figures. */
tmpm_status tmpm_session_memory(const tmpm_session* s, uint64_t* device_bytes, uint64_t* particle_capacity);
/* Whether the session is still usable (no failure since creation). */
int32_t tmpm_session_valid(const tmpm_session* s);
/* `world().substeps` (world.rs:215). */
uint64_t tmpm_session_substeps(const tmpm_session* s);
/* `world().particles.len()`. */
uint64_t tmpm_session_particle_count(const tmpm_session* s);

/* Copy the last synchronized world's particle state out, insertion order.
 * Any array pointer may be NULL to skip that field.  `world()` (gpu.rs:653-656). */
tmpm_status tmpm_session_read_particles(const tmpm_session* s, tmpm_particles* out);

/* `set_particle_force(index, force)` (gpu.rs:457-485): one-shot force
 * consumed by the next substep. */
tmpm_status tmpm_session_set_particle_force(tmpm_session* s, uint64_t index, const double force[3]);
/* `set_obstacles(obstacles)` (gpu.rs:491-529). */
tmpm_status tmpm_session_set_obstacles(tmpm_session* s, const tmpm_obstacle* obstacles, uint32_t count);
/* `add_particles(particles) -> ParticleChunkId` (gpu.rs:535-549). */
tmpm_status tmpm_session_add_particles(tmpm_session* s, const tmpm_particles* particles,
                                       const tmpm_material* materials, uint32_t material_count,
                                       uint64_t* chunk_out);
/* `remove_chunk(chunk)` (gpu.rs:553-564). */
tmpm_status tmpm_session_remove_chunk(tmpm_session* s, uint64_t chunk, uint64_t* removed_out);

/* `encode_particle_snapshot(encoder, output)` (gpu.rs:595-651): writes one
 * 32-byte record {x, y, z, mass; vx, vy, vz, rest_volume} per particle in
 * insertion order (mass 0 for disabled particles) into caller-owned device
 * memory, enqueued on `stream` (NULL: the session's stream).  Returns the
 * record count through count_out (0 and nothing written when no particle is
 * enabled). */
tmpm_status tmpm_session_encode_snapshot(tmpm_session* s, void* device_output, uint64_t output_bytes,
                                         void* stream, uint32_t* count_out);

/* `MpmWorld::stable_timestep()` (world.rs:289-303) on the session's last
 * synchronized world. */
double tmpm_session_stable_timestep(const tmpm_session* s);

/* Topology of the most recent substep, for inspection (the reference's tests
 * read the indirect node count, gpu.rs:3227-3297).  live_nodes counts the
 * exact union of particle stencils (computed on demand); grid_blocks and
 * particle_blocks count 4x4x4-cell blocks. */
typedef struct tmpm_topology_stats {
  uint64_t live_nodes;
  uint64_t grid_blocks;
  uint64_t particle_blocks;
  uint64_t binned_particles;
} tmpm_topology_stats;
tmpm_status tmpm_session_topology_stats(tmpm_session* s, tmpm_topology_stats* out);

/* Export the live node set of the most recent substep as sorted (x, y, z,
 * region) int32 quadruples (region = transfer colour).  Writes up to `cap`
 * keys and the total through count_out. */
tmpm_status tmpm_session_export_nodes(tmpm_session* s, int32_t* keys, uint64_t cap, uint64_t* count_out);
/* Export the particle-block partition of the most recent binning: sorted
 * unique block keys (bx, by, bz, region), and for each particle (insertion
 * order) the index of its block in that list or -1 when not binned. */
tmpm_status tmpm_session_export_blocks(tmpm_session* s, int32_t* block_keys, uint64_t cap, int64_t* particle_block,
                                       uint64_t* count_out);

/* ---------------------------------------------------------------------------
 * Obstacle reactions and rigid coupling
 * ------------------------------------------------------------------------- */
/* Reactions (ObstacleReaction, world.rs:190-204): for every obstacle, the
 * linear impulse and the angular impulse about its center that the MPM
 * transferred to it, from grid-node projection and particle collision
 * projection (step_with_obstacle_reactions, world.rs:309-330; the resident
 * GPU path's reaction passes, gpu.wgsl:1021-1167).  When enabled, each
 * synchronize replaces the readback with the sums over the substeps it
 * synchronized.  Any number of obstacles is supported (the reference's
 * reaction capacity grows with the obstacle count, gpu.rs:2651-2720). */
tmpm_status tmpm_session_set_reactions(tmpm_session* s, int32_t enabled);
/* Particle-side reaction impulses clamped since the session was created (no
 * reference counterpart: the reference adds f32 impulses, which do not
 * saturate).  Each particle's impulse is rounded to the obstacle's int64
 * units, the heaviest particle at 64 cells per substep being 2^43 units; an
 * impulse component beyond 2^51 units (a surface moving over 16,000 cells
 * per substep against the particle) is clamped there and counted here, read
 * at each synchronize.  0 means every reaction summed exactly. */
uint64_t tmpm_session_reaction_saturations(const tmpm_session* s);
/* 6 doubles per obstacle in obstacle order: linear xyz, angular xyz. */
tmpm_status tmpm_session_read_reactions(const tmpm_session* s, double* out, uint32_t obstacle_cap,
                                        uint32_t* count_out);

/* Rigid body state (GpuRigidBodyState, phys/src/gpu_rigid_state.rs:17-28):
 * zero inverse mass is static or kinematic; orientation is XYZW; inertia is
 * the body-frame inverse principal inertia. */
typedef struct tmpm_rigid_body {
  double position[3];
  double inverse_mass;
  double orientation[4];
  double linear_velocity[3];
  double angular_velocity[3];
  double inverse_inertia[3];
} tmpm_rigid_body;

/* One per obstacle, in obstacle order (the coupler's Mapping,
 * gpu_rigid_sync.rs:15-30): the body that drives it (UINT32_MAX: none) and
 * its pose in that body's frame. */
typedef struct tmpm_rigid_mapping {
  uint32_t body;
  uint32_t reserved;
  double local_center[3];
  double local_orientation[4];
} tmpm_rigid_mapping;

/* Couples rigid bodies to the session's obstacles, device-resident, as
 * GpuRigidMpmCoupler's MPM-owned step does (gpu_rigid_sync.rs:187-213): before
 * every substep each mapped obstacle takes its body's pose and velocity
 * (gpu_rigid_sync.wgsl:43-62); after it each body takes the substep's
 * reactions of its obstacles and, when `integrate` is set, gravity, the speed
 * cap (0: none) and its pose update (gpu_rigid_reaction.wgsl:72-150).
 * Reactions are computed while bodies are set.  count 0 removes the
 * coupling.  Bodies and obstacle poses are read back at synchronize. */
tmpm_status tmpm_session_set_rigid_bodies(tmpm_session* s, const tmpm_rigid_body* bodies, uint32_t body_count,
                                          const tmpm_rigid_mapping* mappings, uint32_t mapping_count,
                                          const double gravity[3], double max_linear_speed, int32_t integrate);
tmpm_status tmpm_session_read_rigid_bodies(const tmpm_session* s, tmpm_rigid_body* out, uint32_t cap,
                                           uint32_t* count_out);
/* Obstacles as last synchronized (poses of mapped obstacles follow their
 * bodies); planes of convex obstacles are not returned (NULL). */
tmpm_status tmpm_session_read_obstacles(const tmpm_session* s, tmpm_obstacle* out, uint32_t cap, uint32_t* count_out);

/* ---------------------------------------------------------------------------
 * Batches.  The reference has one world per GpuMpmResidentSession and no
 * batch type; these entry points are the batch/serving boundary the
 * replacement adds (many independent scenes advanced together, one kernel
 * launch per stage per substep for all of them).  A batch is a session
 * handle: tmpm_session_submit_steps / _wait / _synchronize / _step /
 * _pending_substeps / _valid / _set_timing / _kernel_times / _destroy apply to
 * all its scenes.  Per scene it keeps the resident-session semantics of
 * GpuMpmResidentSession (gpu.rs:206-662): its own parameters, substep dt,
 * obstacles and insertion order; transactional edits; and a failure after
 * submission (CFL, particle out of the grid, non-finite state, fixed-point
 * range) stops that scene only, keeping its last synchronized world
 * readable; tmpm_batch_scene_status reports it.  Grid or hash capacity and
 * device failures invalidate the whole batch.
 * ------------------------------------------------------------------------- */
/* New batch from `scene_count` worlds, each with its substep dt (each dt
 * validated like GpuMpmResidentSession::new, gpu.rs:253-280). */
tmpm_status tmpm_batch_create(tmpm_context* ctx, const tmpm_world_desc* worlds, const double* substep_dt,
                              uint32_t scene_count, tmpm_p2g_mode mode, tmpm_session** out);
/* Scene slots (including removed ones, whose status is INVALID_INPUT). */
uint32_t tmpm_batch_scene_count(const tmpm_session* s);
/* TMPM_OK while the scene steps; otherwise the error that stopped it. */
tmpm_status tmpm_batch_scene_status(const tmpm_session* s, uint32_t scene);
uint64_t tmpm_batch_scene_particle_count(const tmpm_session* s, uint32_t scene);
uint64_t tmpm_batch_scene_substeps(const tmpm_session* s, uint32_t scene);
/* `world()` of one scene (gpu.rs:653-656): its last synchronized state. */
tmpm_status tmpm_batch_read_particles(const tmpm_session* s, uint32_t scene, tmpm_particles* out);
/* `set_particle_force` / `set_obstacles` of one scene (gpu.rs:457-529). */
tmpm_status tmpm_batch_set_particle_force(tmpm_session* s, uint32_t scene, uint64_t index, const double force[3]);
tmpm_status tmpm_batch_set_obstacles(tmpm_session* s, uint32_t scene, const tmpm_obstacle* obstacles,
                                     uint32_t count);
/* Per-scene forms of the reaction and rigid-coupling entry points above. */
tmpm_status tmpm_batch_read_reactions(const tmpm_session* s, uint32_t scene, double* out, uint32_t obstacle_cap,
                                      uint32_t* count_out);
tmpm_status tmpm_batch_set_rigid_bodies(tmpm_session* s, uint32_t scene, const tmpm_rigid_body* bodies,
                                        uint32_t body_count, const tmpm_rigid_mapping* mappings,
                                        uint32_t mapping_count, const double gravity[3], double max_linear_speed,
                                        int32_t integrate);
tmpm_status tmpm_batch_read_rigid_bodies(const tmpm_session* s, uint32_t scene, tmpm_rigid_body* out, uint32_t cap,
                                         uint32_t* count_out);
tmpm_status tmpm_batch_read_obstacles(const tmpm_session* s, uint32_t scene, tmpm_obstacle* out, uint32_t cap,
                                      uint32_t* count_out);
/* Admits a new scene (serving): it takes the first free slot. */
tmpm_status tmpm_batch_add_scene(tmpm_session* s, const tmpm_world_desc* world, double substep_dt,
                                 uint32_t* scene_out);
/* Admission in two halves, so that the host work of an admission can run
 * while a batch has substeps in flight (no reference counterpart: the
 * reference has no batch).  tmpm_batch_stage_scene does the host half of
 * tmpm_batch_add_scene without any session or device: validation, the f64
 * copy of the world, the f32 packing and the block count; it may run on any
 * thread, concurrently with calls on sessions.  tmpm_batch_add_staged admits
 * the staged scene into a batch with no pending substeps (the device half:
 * uploads and capacity); it consumes the staged scene, also on failure.
 * Results equal tmpm_batch_add_scene with the same world. */
typedef struct tmpm_staged_scene tmpm_staged_scene;
tmpm_status tmpm_batch_stage_scene(const tmpm_world_desc* world, double substep_dt, tmpm_staged_scene** out);
tmpm_status tmpm_batch_add_staged(tmpm_session* s, tmpm_staged_scene* staged, uint32_t* scene_out);
void tmpm_staged_scene_destroy(tmpm_staged_scene* staged);
/* Retires a scene; its slot becomes free. */
tmpm_status tmpm_batch_remove_scene(tmpm_session* s, uint32_t scene);

/* Per-kernel average durations (milliseconds) recorded with HIP events on the
 * session stream over the substeps since the last reset; names are the
 * kernel names.  Timing is off unless enabled (it adds one event pair per
 * kernel launch). */
tmpm_status tmpm_session_set_timing(tmpm_session* s, int32_t enabled);
/* No counterpart in the reference (its encoder records one command buffer per
 * submit).  With graphs on (off by default; TMPM_GRAPHS=1 turns them on) each
 * substep's launches are captured once per distinct launch plan into a HIP
 * graph and replayed, which lowers the host cost per substep.  Results are
 * identical either way.  Per-kernel timing launches eagerly. */
tmpm_status tmpm_session_set_graphs(tmpm_session* s, int32_t enabled);

/* `sample_closed_mesh_volume` (sampling.rs:47-146), the sampler behind
 * `MeshEmitter` (emitter.rs:70-113): lattice centres lower + (i + 1/2) *
 * spacing inside a closed, consistently oriented triangle mesh (solid angle
 * above 2 pi, either winding), in lattice order (x, then y, then z).  The
 * solid-angle test runs on the device.  vertices: 3 doubles each; triangles:
 * 3 indices each; samples_out holds 3 * max_samples doubles.  InvalidInput for
 * bad geometry or an open/badly oriented surface, Capacity when the lattice
 * exceeds max_candidates or the accepted points exceed max_samples. */
tmpm_status tmpm_sample_closed_mesh_volume(tmpm_context* ctx, const double* vertices, uint64_t vertex_count,
                                           const uint32_t* triangles, uint64_t triangle_count, double spacing,
                                           uint64_t max_candidates, uint64_t max_samples, double* samples_out,
                                           uint64_t* count_out);
/* No counterpart in the reference: tmpm_sample_closed_mesh_volume for
 * callers that cannot allocate (code under graph capture, tight emission
 * loops, harnesses that forbid device allocation).  Same arguments, results
 * and errors, but it allocates nothing, on the device or the host, and
 * creates no stream: the mesh, the candidate flags and a status word live in
 * a fixed 32 MiB device array that each device holds for the life of the
 * process (set up by tmpm_context_create), and the closed-surface test runs
 * on the device.  Meshes of more than 65,536 triangles, or whose vertices,
 * triangles and candidate lattice do not fit the 32 MiB, fail with
 * TMPM_CAPACITY.  Every call on a device shares that device's array and
 * leaves its contents behind; calls of this process on one device are
 * serialized by a lock, but calls from several processes on one device race. */
tmpm_status tmpm_sample_closed_mesh_volume_noalloc(tmpm_context* ctx, const double* vertices, uint64_t vertex_count,
                                                   const uint32_t* triangles, uint64_t triangle_count,
                                                   double spacing, uint64_t max_candidates, uint64_t max_samples,
                                                   double* samples_out, uint64_t* count_out);
/* No counterpart in the reference.  Chooses the launch shape (P2G workgroup
 * size and LDS tile copies, G2P workgroup size, reorder interval, graphs) by
 * timing each candidate over `trial` real substeps (8..=256), one knob at a
 * time, and keeps the fastest.  The trials advance the session like
 * tmpm_session_submit_steps (they become pending substeps; *substeps_run
 * says how many) and compute exactly what any shape computes, obstacle
 * reactions included.  They are queued 64 substeps at a time; between
 * chunks the topology grows as synchronize grows it, and an error stops the
 * tuning (the next synchronize reports it, as for untuned stepping).  A
 * bucket tuned before runs no trials (see tmpm_tune_cache_clear). */
tmpm_status tmpm_session_autotune(tmpm_session* s, uint32_t trial, uint32_t* substeps_run);
/* The current launch shape: {P2G threads, P2G tile copies, G2P threads,
 * reorder interval, graphs on}.  Before tuning, the shape follows the
 * particle-block count of the last synchronized substep and its mean
 * particles per block: from 8,192 blocks of fewer than 192 particles P2G
 * runs 64 threads and G2P 64 threads.  Fewer blocks or denser ones take
 * wider workgroups, so that a launch still fills the chip: P2G 128 threads;
 * G2P 128 threads below 8,192 blocks or from 192 particles per block, 256
 * below 4,096 blocks or from 288 particles per block.  P2G uses one tile
 * copy unless tuned to two. */
tmpm_status tmpm_session_launch_config(tmpm_session* s, int32_t out[5]);
tmpm_status tmpm_session_kernel_times(tmpm_session* s, char* names_csv, uint64_t names_cap, double* avg_ms,
                                      uint64_t* calls, uint32_t cap, uint32_t* count_out);

/* FP64 audit of every substep (the reference's momentum balance check,
 * world.rs:784-807, computes such sums on the host).  While enabled (set
 * with no pending substeps; off by default), each substep ends with a device
 * reduction in f64 over each scene's particles, and synchronize keeps the
 * records of the substeps it synchronized.  A record is 12 doubles: mass;
 * momentum xyz; kinetic energy; first moment of mass xyz (sum of m x, so
 * the gravitational potential energy is -g . that); angular momentum about
 * the origin xyz (sum of m x cross v); particles counted.  Sums are f64 but
 * their order follows the scheduler.  read_audit writes up to cap_records
 * records of `scene` (0 for a single session) in substep order and the
 * number available through count_out. */
tmpm_status tmpm_session_set_audit(tmpm_session* s, int32_t enabled);
tmpm_status tmpm_session_read_audit(const tmpm_session* s, uint32_t scene, double* out, uint64_t cap_records,
                                    uint64_t* count_out);

/* Autotuner cache (no reference counterpart): tmpm_session_autotune keeps
 * the shape it picks per workload bucket (device; particle count and
 * particles per block, each rounded down to a power of two; reactions,
 * plastic materials, obstacle shapes, P2G mode, batch, CPIC).  A later
 * autotune in a tuned bucket runs no trials, and a session whose shape is not
 * fixed launches with the cached shape of its current bucket.  Process-wide;
 * with TMPM_TUNE_CACHE naming a file, also loaded from and appended to it. */
void tmpm_tune_cache_clear(void);
uint64_t tmpm_tune_cache_size(void);

/* Bandwidth probe (no reference counterpart): copies `bytes` bytes between
 * two device buffers `iterations` times with a plain float4 copy kernel and
 * reports the achieved bandwidth (read + write bytes per second, in GB/s).
 * Used to confirm the device's achievable HBM bandwidth and to calibrate
 * hardware byte counters against a kernel of known traffic. */
tmpm_status tmpm_bandwidth_probe(int32_t device_ordinal, uint64_t bytes, uint32_t iterations, double* gbs_out);

/* Host-side validation helpers shared with the binding layer. */
double tmpm_timestep_bound(const tmpm_material* material, double density, const double velocity[3],
                           double cell_width, double deformation_det);
int32_t tmpm_material_valid(const tmpm_material* material);
int32_t tmpm_obstacle_valid(const tmpm_obstacle* obstacle);

#ifdef __cplusplus
}
#endif

#endif /* TESSERA_MPM_H */
