/*
 * nexus_mpm.h — C ABI of the MI355X (gfx950) sparse-grid MLS-MPM substep.
 *
 * Drop-in for the nexus_mpm pipeline API of dimforge/nexus (non-CPIC path):
 * particle / grid / model / collider setup on an MpmState-like scene object and a
 * step() that enqueues one substep (binning sort -> P2G -> grid update with gravity,
 * velocity clamp and one-way cuboid colliders -> G2P fused with the constitutive
 * update).  Every entry point below names the reference item it replaces
 * (file:line in the dimforge/nexus tree).
 *
 * Conventions
 *  - Plain C types only; host arrays are caller-owned and copied in/out; no pointer
 *    passed in is retained after a call returns.
 *  - Every function returns an nxm_status (0 = success).  nxm_last_error() gives a
 *    thread-local message for the last failure on the calling thread.
 *  - Matrices are 3x3 column-major (9 floats: column 0, 1, 2), the layout of the
 *    reference's glam Mat3.
 *  - nxm_pipeline_step / nxm_batch_step enqueue work on the scene's (batch's) HIP
 *    stream and return; errors detected on the device (grid capacity, domain range)
 *    are latched and reported by the next synchronizing call (nxm_*_synchronize,
 *    nxm_scene_read_*).  The reference discards step errors and leaves hash-map
 *    overflow undefined (src/pipeline.rs:83-85, src_mpm_shaders/grid/grid.rs:455-457).
 *  - One host thread at a time per scene or batch; a context may be shared by
 *    threads working on different scenes.
 */
#ifndef NEXUS_MPM_H
#define NEXUS_MPM_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int nxm_status;
#define NXM_OK 0
#define NXM_ERR_INVALID_ARGUMENT 1
#define NXM_ERR_OUT_OF_MEMORY 2
#define NXM_ERR_GRID_CAPACITY 3     /* more active blocks than the grid capacity */
#define NXM_ERR_DOMAIN 4            /* a block coordinate outside the packable range */
#define NXM_ERR_TOO_MANY_COLLIDERS 5
#define NXM_ERR_UNSUPPORTED_MODEL 6 /* model tag above NXM_MODEL_SNOW */
#define NXM_ERR_DEVICE 7
#define NXM_ERR_BUSY 8              /* scene is attached to a batch */

/* Model tags of GpuParticleModel (src_mpm_shaders/models/default.rs:20-31). */
#define NXM_MODEL_ELASTIC_LINEAR 0
#define NXM_MODEL_ELASTIC_NEO_HOOKEAN 1
#define NXM_MODEL_SAND_LINEAR 2
#define NXM_MODEL_SAND_NEO_HOOKEAN 3
#define NXM_MODEL_FLUID 4
#define NXM_MODEL_SNOW 5            /* models/snow.rs; tags above 5 are rejected */
#define NXM_MODEL_WORDS 24

/* BoundaryCondition::ty (src_mpm_shaders/solver/boundary_condition.rs:4-7). */
#define NXM_BC_STICK 0
#define NXM_BC_SLIP 1
#define NXM_BC_SEPARATE 2
#define NXM_BC_NON_REFLECTING 3

#define NXM_MAX_COLLIDERS 16        /* MAX_COLLISION_BODIES, boundary_condition.rs:44 */

typedef struct nxm_ctx nxm_ctx;
typedef struct nxm_scene nxm_scene;
typedef struct nxm_batch nxm_batch;

/* MpmCapacities (src_mpm/pipeline.rs:28-43). */
typedef struct {
    uint32_t grid_size;          /* block capacity, rounded up to a power of two */
    uint32_t particles_capacity; /* particles reserved up front */
} nxm_capacities;

/* SimulationParams (src_mpm_shaders/solver/params.rs:9-17). */
typedef struct {
    float gravity[3];
    float dt;
} nxm_sim_params;

/* Particle (src_mpm/solver/particle.rs:26-33) = position + ParticleDynamics
 * (:65-98) + GpuParticleModel (src_mpm_shaders/models/default.rs:41-44). */
typedef struct {
    float position[3];
    float velocity[3];
    float def_grad[9];
    float affine[9];
    float init_volume;
    float init_radius;
    float mass;
    float damping;
    float phase;
    float boundary_friction;
    uint32_t group_id;
    uint32_t enabled;
    uint32_t fixed;
    uint32_t model_tag;
    uint32_t model_data[NXM_MODEL_WORDS];
} nxm_particle;

/* Structure-of-arrays particle upload/readback (same content as nxm_particle,
 * without per-particle padding).  Any pointer may be NULL on readback to skip it;
 * on upload NULL means the ParticleDynamics::new default (velocity 0, F = I,
 * affine 0, damping 0, phase 1, boundary_friction 1, group 0, enabled, not fixed). */
typedef struct {
    float *position;     /* n*3 */
    float *velocity;     /* n*3 */
    float *def_grad;     /* n*9 */
    float *affine;       /* n*9 */
    float *init_volume;  /* n */
    float *init_radius;  /* n */
    float *mass;         /* n */
    float *damping;      /* n */
    float *phase;        /* n */
    float *boundary_friction; /* n */
    uint32_t *group_id;  /* n */
    uint32_t *enabled;   /* n */
    uint32_t *fixed;     /* n */
    uint32_t *model_tag; /* n */
    uint32_t *model_data;/* n*24 */
} nxm_particles_soa;

/* A cuboid collider and the body it is attached to (rapier collider + body +
 * BoundaryCondition, src/state.rs:418-450; projection by parry3d Cuboid through
 * src_rbd_shaders/shapes/shape.rs:378-384).  The collider's pose is the body's pose.
 *
 * inv_mass and local_com are the body's mass properties as MPM holds them
 * (GpuBodySet::from_rapier, src_rbd/dynamics/body.rs:153-170): zero for a body coupled
 * one way and for fixed or kinematic bodies; a dynamic body coupled two ways carries
 * its world inverse mass per axis (zero along locked axes) and its centre of mass in
 * the body frame.  MPM integrates the body each substep (rigid_impulses.rs:66-140):
 * the pose advances about the world centre of mass pose * local_com, then gravity is
 * added to the linear velocity along the axes where inv_mass is nonzero.  The
 * particles' impulses on the body are compiled out in the reference
 * (TWO_WAYS_COUPLING_ENABLED = false, src_mpm_shaders/lib.rs:38), so they are zero here
 * too and the body's velocity is never capped. */
typedef struct {
    float half_extents[3];
    float rotation[4];   /* unit quaternion x, y, z, w */
    float translation[3];
    float linvel[3];
    float angvel[3];
    uint32_t bc_type;
    float bc_friction;
    float inv_mass[3];
    float local_com[3];
} nxm_cuboid;

/* Analytic collider shapes: the reference's Shape tags (src_rbd_shaders/shapes/shape.rs:
 * 29-33), converted from rapier's TypedShape by src_rbd/shapes/shape.rs:35-41, 252-254. */
#define NXM_SHAPE_BALL 0u
#define NXM_SHAPE_CUBOID 1u
#define NXM_SHAPE_CAPSULE 2u
#define NXM_SHAPE_CONE 3u      /* apex at +half_height on the body's y axis, base at -half_height */
#define NXM_SHAPE_CYLINDER 4u  /* along the body's y axis */

/* An analytic collider on its body.  The body record is nxm_cuboid's (pose, velocities,
 * boundary condition, mass properties); its half_extents are the shape's for
 * NXM_SHAPE_CUBOID and ignored otherwise.  dims, in the body frame:
 *   NXM_SHAPE_BALL      radius
 *   NXM_SHAPE_CUBOID    unused
 *   NXM_SHAPE_CAPSULE   segment end a (3), end b (3), radius
 *   NXM_SHAPE_CONE      half height, base radius
 *   NXM_SHAPE_CYLINDER  half height, radius
 * Each grid node projects onto the shape's boundary as the reference's
 * Shape::project_point_on_boundary (shape.rs:378-399): parry3d 0.30's Ball and Cuboid,
 * Capsule (capsule.rs:62-91, segment.rs:43-60), Cone (cone.rs:98-155), Cylinder
 * (cylinder.rs:90-155). */
typedef struct {
    nxm_cuboid body;
    uint32_t shape;   /* NXM_SHAPE_* */
    float dims[7];
} nxm_collider;

/* A triangle-mesh collider on its body (a rapier TriMesh or HeightField collider coupled
 * to MPM; GpuRigidParticles::from_rapier, src_mpm/solver/particle.rs:295-345): vertices
 * and triangles in the body frame, and the body as in nxm_cuboid (pose, velocities,
 * boundary condition, mass properties; half_extents ignored).  A heightfield is given
 * as its triangles (HeightField::to_trimesh). */
typedef struct {
    const float *vertices;      /* num_vertices*3 */
    const uint32_t *triangles;  /* num_triangles*3 vertex indices */
    uint32_t num_vertices;
    uint32_t num_triangles;
    nxm_cuboid body;
} nxm_trimesh;

/* Per-stage device times of the last step, milliseconds (the reference's
 * GpuTimestamps passes, src_mpm/pipeline.rs:394-511). */
typedef struct {
    float sort_ms;         /* "[MPM] Grid sort" */
    float p2g_ms;          /* "[MPM] P2G" */
    float grid_update_ms;  /* "[MPM] Grid update" */
    float g2p_update_ms;   /* "[MPM] G2P" + "[MPM] Particle update" (fused) */
    float total_ms;
    /* The P2G and the G2P + update kernels alone, from the kernels' own start and end
     * timestamps (the stage times above also hold the launch gap before the kernel). */
    float p2g_kernel_ms;
    float g2p_update_kernel_ms;
} nxm_stage_times;

/* Launch shapes of a substep (no reference counterpart: wgpu picks none of these).
 * Results do not depend on them (fixed-point P2G, per-particle and per-node stages), so
 * they can be tuned while a scene runs. */
typedef struct {
    uint32_t p2g_groups;      /* workgroups (one wave each) of the P2G kernel */
    uint32_t p2g_block_cost;  /* per-block cost, in particles, of the P2G work split */
    uint32_t grid_groups;     /* workgroup cap of the grid-stride kernels (256 threads each) */
} nxm_launch_config;

/* Canonical form of the block sort (the part of the reference's sort that does not
 * depend on atomic ordering): per active block its virtual id, the number of
 * primary particles, of particles including spilled extras, the ten bucket counts
 * (grid/sort.rs:41-62), and per particle its primary block and local cell. */
typedef struct {
    int32_t *block_vid;        /* nblocks*3 */
    uint32_t *block_np;        /* nblocks: num_particles */
    uint32_t *block_npe;       /* nblocks: num_particles_with_extras */
    uint32_t *block_buckets;   /* nblocks*10 */
    uint32_t *particle_block;  /* n: index into block_vid of the primary block */
    uint32_t *particle_cell;   /* n: local cell x + 4y + 16z inside the primary block */
} nxm_sort_canonical;

/* ---- context (MpmPipeline::new, src_mpm/pipeline.rs:366) ---- */
nxm_status nxm_ctx_create(int device, nxm_ctx **out);
void nxm_ctx_destroy(nxm_ctx *ctx);
const char *nxm_last_error(void);
const char *nxm_version(void);

/* ---- scene (MpmState, src_mpm/pipeline.rs:65-362) ---- */
/* MpmState::empty (src_mpm/pipeline.rs:107): gravity 0, dt 1/60, cell width 1. */
nxm_status nxm_scene_create(nxm_ctx *ctx, const nxm_capacities *caps, nxm_scene **out);
void nxm_scene_destroy(nxm_scene *scene);
/* MpmState::set_simulation_params (src_mpm/pipeline.rs:169). */
nxm_status nxm_scene_set_simulation_params(nxm_scene *scene, const nxm_sim_params *params);
/* MpmState::write_substep_params (src_mpm/pipeline.rs:183): dt = base dt / max(n, 1). */
nxm_status nxm_scene_write_substep_params(nxm_scene *scene, uint32_t num_substeps);
/* MpmState::set_cell_width (src_mpm/pipeline.rs:200). */
nxm_status nxm_scene_set_cell_width(nxm_scene *scene, float cell_width, uint32_t grid_capacity);
/* GpuParticles::append (src_mpm/solver/particle.rs:533) / from_particles (:391). */
nxm_status nxm_scene_add_particles(nxm_scene *scene, const nxm_particle *particles, size_t n);
nxm_status nxm_scene_add_particles_soa(nxm_scene *scene, const nxm_particles_soa *p, size_t n);
/* MpmState::set_coupling (src_mpm/pipeline.rs:220) for cuboid colliders (with the mesh
 * colliders at most 16, the reference's GpuMaterials limit).  Sets the poses; from then on the device
 * integrates them every substep (nxm_scene_read_colliders). */
nxm_status nxm_scene_set_colliders(nxm_scene *scene, const nxm_cuboid *colliders, size_t n);
/* MpmState::set_coupling (src_mpm/pipeline.rs:220) for analytic colliders of any
 * NXM_SHAPE_* kind, in the caller's order (ids 0..n-1: the collider pass keeps the first
 * of equally near colliders, grid_update_collide.rs:89-93).  Replaces the colliders set by
 * nxm_scene_set_colliders; mesh colliders keep their ids after these.  Rejects unknown
 * kinds and negative or non-finite dimensions (NXM_ERR_INVALID_ARGUMENT).
 * nxm_scene_read_colliders lists them with half extents zero except for cuboids. */
nxm_status nxm_scene_set_shape_colliders(nxm_scene *scene, const nxm_collider *colliders, size_t n);
/* The mesh colliders of a scene (MpmState::with_select_coupling with TriMesh or
 * HeightField colliders, src_mpm/pipeline.rs:297-316).  Their ids follow the analytic
 * colliders' (all together <= 16); setting either kind keeps the other with its current
 * poses.  With CPIC the meshes act through the contact distance field: each substep
 * their rigid particles (nxm_sample_trimesh at the cell width) reach the grid nodes around
 * them (gpu_p2g_cdf, src_mpm_shaders/solver/p2g_cdf.rs:60-259).  Without CPIC the grid
 * collider pass projects each node onto the nearest triangle within 1.5 cell widths
 * (grid_update_collide.rs:53-67, TriMesh::project_local_point, src_rbd_shaders/shapes/
 * trimesh.rs:155-207), inside by the pseudo-normal of the nearest feature; nodes deeper
 * than that inside a closed mesh see no triangle and no collider, as in the reference.
 * CPIC may be switched either way while meshes are set.  Rejects out-of-range vertex
 * indices and non-finite vertices.  The bodies are integrated like the other colliders
 * (nxm_scene_read_colliders lists them after them, half extents zero). */
nxm_status nxm_scene_set_trimesh_colliders(nxm_scene *scene, const nxm_trimesh *meshes, size_t n);

/* ReadbackData (src_mpm_shaders/solver/prep_readback.rs:84-92): a particle's render
 * colour, its deformation (column-major 4x4, the 3x3 in the upper left) and position. */
typedef struct {
    float color[4];
    float deformation[16];
    float position[3];
    float pad;
} nxm_readback_data;

/* RenderConfig::mode values (prep_readback.rs:16-21); any other value is the default
 * (the particle group's colour). */
#define NXM_RENDER_DEFAULT 0u
#define NXM_RENDER_VOLUME 1u
#define NXM_RENDER_VELOCITY 2u
#define NXM_RENDER_PHASE 3u
#define NXM_RENDER_CDF_NORMALS 4u
#define NXM_RENDER_CDF_DISTANCES 5u
#define NXM_RENDER_CDF_SIGNS 6u

/* WgPrepReadback::launch (src_mpm/solver/prep_readback.rs:153-210, gpu_prep_readback):
 * per particle, in id order, its render data computed on the device: the deformation
 * diag(2 init_radius) F with F's entries clamped to [-4, 4], and the colour by mode from
 * the group palette (group_colors: num_groups RGBA colours, group id modulo their count;
 * NULL / 0: the reference's six-colour default palette), the velocity, the singular values
 * of F (largest first), the phase, or the contact field (zero for scenes without CPIC);
 * disabled particles in red.  out sized n >= the scene's particle count; synchronizes. */
nxm_status nxm_scene_prep_readback(nxm_scene *scene, uint32_t mode, const float *group_colors, uint32_t num_groups,
                                   nxm_readback_data *out, size_t n);
/* gpu_prep_readback_rigid (prep_readback.rs:429-469): the mesh colliders' rigid particles
 * at their bodies' current poses, rigid particle i coloured by palette entry i modulo its
 * size, deformation diag(0.4 h).  Writes up to n; *count receives how many there are. */
nxm_status nxm_scene_prep_readback_rigid(nxm_scene *scene, const float *group_colors, uint32_t num_groups,
                                         nxm_readback_data *out, size_t n, size_t *count);
/* sample_mesh (src_mpm/sampling/sample_trimesh.rs:60-231): the rigid particles of a
 * triangle mesh, points on every triangle and along every edge once, at most
 * spacing / sqrt 2 apart.  Writes up to capacity points (3 floats each) and the triangle
 * of each; *count receives how many there are.  Host-only: no device call. */
nxm_status nxm_sample_trimesh(const float *vertices, uint32_t num_vertices, const uint32_t *triangles,
                              uint32_t num_triangles, float spacing, float *points, uint32_t *triangle_of,
                              size_t capacity, size_t *count);
/* Current poses and velocities of the scene's colliders.  MPM integrates its coupled
 * bodies every substep (src_mpm/pipeline.rs:510-520, gpu_rigid_impulses_update in
 * src_mpm_shaders/solver/rigid_impulses.rs:66-140): each pose advances with its linear
 * and angular velocity about the body's centre of mass, and a dynamic body coupled two
 * ways gains g dt of velocity along the axes of nonzero inverse mass.  Counterpart of
 * MpmPipeline::writeback_body_poses (src_mpm/pipeline.rs:536).  Synchronizes when the
 * scene is resident; out sized n >= the number of colliders. */
nxm_status nxm_scene_read_colliders(nxm_scene *scene, nxm_cuboid *out, size_t n);
/* MpmState::use_cpic (src_mpm/pipeline.rs:69, default off here): compatible
 * particle-in-cell transfers (pipeline.rs:433-460 with the use_cpic branches of p2g.rs,
 * grid_update.rs, g2p.rs and particle_update.rs) instead of the grid collider pass.
 * The contact distance field comes from the analytic colliders (grid_update_cdf.rs)
 * and from the mesh colliders' rigid particles (p2g_cdf.rs); with it off, mesh colliders
 * act through the grid collider pass's projection.  Switching it resets every
 * particle's contact field to zero (as for new particles); synchronizes when the scene
 * is resident. */
nxm_status nxm_scene_set_cpic(nxm_scene *scene, int enabled);
/* Per-particle contact field of the last step (Kinematics::cdf, solver/particle.rs:
 * 35-50 without rigid_vel, which lives only inside the step): affinity word (bit i:
 * collider i in reach, bit 16 + i: its side), normal (n*3) and signed distance (n), in
 * original particle order; any pointer may be NULL.  Synchronizes. */
nxm_status nxm_scene_read_cdf(nxm_scene *scene, uint32_t *affinity, float *normal, float *signed_distance,
                              size_t n);
size_t nxm_scene_len(const nxm_scene *scene);
/* GpuParticles::swap_remove (src_mpm/solver/particle.rs:575-631): removes the particles
 * with the given ids (duplicates and ids >= len ignored); processing ids in descending
 * order, the current last particle moves into each freed id.  The moves are reported as
 * (moved_from[k], moved_to[k]), k < *n_moved (arrays sized n, may be NULL).  Synchronizes. */
nxm_status nxm_scene_swap_remove(nxm_scene *scene, const uint32_t *ids, size_t n, uint32_t *moved_from,
                                 uint32_t *moved_to, size_t *n_moved);
/* CFL-based upper bound on the substep length of the scene's current state:
 * WgTimestepBounds::compute_bounds (src_mpm/solver/timestep_bound.rs:25-54), kernel
 * src_mpm_shaders/solver/timestep_bound.rs:61-111.  Same fixed-point result (1e-12 s
 * units, minimum over enabled particles, 4.294967 ms when none is enabled).  Works for
 * standalone and batched scenes; synchronizes. */
nxm_status nxm_scene_timestep_bound(nxm_scene *scene, float *max_dt);
/* Standalone steps of this scene replay one captured HIP graph (default off); as
 * nxm_batch_set_graph. */
nxm_status nxm_scene_set_graph(nxm_scene *scene, int enabled);
/* Device memory the scene occupies while resident (particle arenas, sort arrays and its
 * grid capacity's block tables); used for admission against an HBM budget.  No
 * reference counterpart (the reference sizes one scene's buffers at creation,
 * src_mpm/pipeline.rs:107-160). */
size_t nxm_scene_device_bytes(const nxm_scene *scene);
/* Packs the particles of a scene that is not resident yet for its next upload (host work
 * only, into pinned memory shared by the context; safe on another thread than the one
 * stepping the worlds, as long as nothing else uses this scene meanwhile): adding it to a
 * batch then only queues the copies.  Used by a server's ingest thread.  No reference
 * counterpart (GpuParticles::append copies synchronously, src_mpm/solver/particle.rs:533).
 * NXM_ERR_BUSY if the scene is resident. */
nxm_status nxm_scene_stage_upload(nxm_scene *scene);
/* GpuParticles::read_positions (src_mpm/solver/particle.rs:634-647); synchronizes. */
nxm_status nxm_scene_read_positions(nxm_scene *scene, float *xyz, size_t n);
/* Full particle state in original particle order; synchronizes. */
nxm_status nxm_scene_read_particles_soa(nxm_scene *scene, const nxm_particles_soa *out, size_t n);
/* Block count of the last step (0 before the first step); synchronizes. */
nxm_status nxm_scene_sort_size(nxm_scene *scene, size_t *nblocks);
/* Canonical sort of the last step; arrays sized by nxm_scene_sort_size and len. */
nxm_status nxm_scene_read_sort_canonical(nxm_scene *scene, const nxm_sort_canonical *out);
/* Grid of the last step: per block (same order as read_sort_canonical) 64 nodes
 * (x + 4y + 16z) of {mass, momentum.xyz} after P2G and velocity.xyz after the grid
 * update and collider boundary.  mass: nblocks*64, momentum/velocity: nblocks*64*3. */
nxm_status nxm_scene_read_grid(nxm_scene *scene, float *mass, float *momentum, float *velocity);
/* CPIC grid of the last step, in read_grid's order: per node the contact distance field
 * (NodeCdf, grid/grid.rs:367-374: distance, affinity word, closest collider or
 * 0xFFFFFFFF) and the incompatible mass after P2G and velocity after the grid update
 * (Node::mass_incompatible / momentum_velocity_incompatible).  Any pointer may be NULL. */
nxm_status nxm_scene_read_grid_cpic(nxm_scene *scene, float *distance, uint32_t *affinity, uint32_t *closest,
                                    float *mass_incompatible, float *velocity_incompatible);
nxm_status nxm_scene_synchronize(nxm_scene *scene);
/* Keep the per-node momentum after P2G for nxm_scene_read_grid (costs one extra
 * write per node per step).  No counterpart in the reference: a test hook. */
nxm_status nxm_scene_set_debug_grid(nxm_scene *scene, int enabled);
/* Test hook, no counterpart in the reference: the next `count` particle uploads into a
 * world (a scene's first step, joining a batch, particles appended to a resident scene)
 * fail with NXM_ERR_DEVICE before any copy is queued, as a failed device copy would.  The
 * scene is then left on the host, detached, and every object can still be stepped and
 * destroyed in any order.  0 ends it. */
nxm_status nxm_debug_fail_uploads(uint32_t count);

/* ---- pipeline step (MpmPipeline::step, src_mpm/pipeline.rs:385) ---- */
/* Enqueues one substep of a standalone scene.  times != NULL synchronizes and
 * fills per-stage device times. */
nxm_status nxm_pipeline_step(nxm_ctx *ctx, nxm_scene *scene, nxm_stage_times *times);
/* Enqueues n substeps (the NexusPipeline::simulate substep loop, src/pipeline.rs:84-86). */
nxm_status nxm_pipeline_run(nxm_ctx *ctx, nxm_scene *scene, uint32_t substeps);

/* Adaptive substeps (no reference counterpart; the reference computes the bound,
 * WgTimestepBounds::compute_bounds, src_mpm/solver/timestep_bound.rs:25-54, and leaves the
 * choice of dt to its caller).  When enabled, each substep's length is
 * min(dt, the CFL bound of the scene's state at the start of the substep), with dt the
 * scene's substep length (nxm_scene_set_simulation_params / write_substep_params) as the
 * maximum and the bound as nxm_scene_timestep_bound computes it.  The bound is taken on
 * the device (the G2P kernel min-reduces it for the next substep), so a batch of adaptive
 * scenes steps without a host round trip, each scene at its own length. */
nxm_status nxm_scene_set_adaptive_dt(nxm_scene *scene, int enabled);
/* The scene's clock: simulated time (the sum of its substeps' lengths since creation),
 * the length of its last substep and the number of substeps.  Synchronizes.  Any pointer
 * may be NULL. */
nxm_status nxm_scene_read_clock(nxm_scene *scene, double *time, float *last_dt, uint64_t *substeps);

/* ---- one substep without allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device memory:
 * code under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc.
 * nxm_pipeline_step_noalloc runs one substep (the kernels of nxm_pipeline_step, non-CPIC
 * path) of a scene whose particles live in caller-owned device arrays, updated in place
 * in particle-id order.  Everything else a substep needs (particle arenas, sort scratch,
 * hash, block tables, staging tiles, counters) lives in one static device array of
 * NXM_NOALLOC_SCRATCH_BYTES per GPU, which the library holds from the first
 * nxm_ctx_create on that GPU until the process ends.  A scene needs about
 * 308 B per particle + 4.6 KB per hash slot (2 x next_pow2(grid_capacity) + 1 slots)
 * + 60 KB; a larger one is rejected with NXM_ERR_INVALID_ARGUMENT before anything is
 * enqueued.  The call allocates nothing (no hipMalloc, no pinned memory, no library
 * allocation), enqueues on `stream` and returns without synchronizing.
 *
 * Every call on one GPU uses that GPU's single scratch array, which keeps the last
 * call's contents (each call sets up what it reads): calls that overlap on one GPU
 * (different streams, or different threads without ordering) race and must not.
 * Colliders are the bodies' poses and velocities at the start of the substep; the
 * substep does not write back their integrated poses (a caller with moving bodies
 * advances them itself between calls).  CPIC and mesh colliders are not supported here. */
#define NXM_NOALLOC_SCRATCH_BYTES (64u << 20)
#define NXM_NOALLOC_MAX_MATERIALS 16

/* A material: model tag and words (GpuParticleModel) with the particle damping and
 * boundary friction factor it is used with. */
typedef struct {
    uint32_t model_tag;
    uint32_t model_data[NXM_MODEL_WORDS];
    float damping;
    float boundary_friction;
} nxm_material;

/* Device arrays (n particles, id order), read and written by the substep.  plastic may
 * be NULL: the sand tags' (det Fp, q, log volume gain) and snow's det Fp then start from
 * the material's leading words and are not written back.  flags: bit 0 enabled, bit 1
 * fixed, bits 2..15 the material index (into nxm_noalloc_scene.materials); the substep
 * clears bit 0 (and the mass) of a particle that fails (NaN position or det F <= 0). */
typedef struct {
    float *position;     /* n*3 */
    float *velocity;     /* n*3 */
    float *def_grad;     /* n*9 */
    float *affine;       /* n*9 */
    float *mass;         /* n */
    float *init_volume;  /* n */
    float *plastic;      /* n*3, or NULL */
    uint32_t *flags;     /* n */
} nxm_device_particles;

/* The scene of one no-allocation substep (host memory, read during the call). */
typedef struct {
    float cell_width;
    uint32_t grid_capacity;          /* blocks, rounded up to a power of two */
    nxm_sim_params params;           /* gravity and the substep length */
    const nxm_material *materials;   /* 1..NXM_NOALLOC_MAX_MATERIALS */
    uint32_t num_materials;
    const nxm_cuboid *colliders;     /* 0..NXM_MAX_COLLIDERS */
    uint32_t num_colliders;
} nxm_noalloc_scene;

/* errors: optional device word that receives the substep's error bits (1 grid capacity
 * exceeded, 2 a block outside the packable domain, 4 a material index out of range; 0
 * when none).  stream: a hipStream_t (NULL: the null stream). */
nxm_status nxm_pipeline_step_noalloc(nxm_ctx *ctx, const nxm_noalloc_scene *scene,
                                     const nxm_device_particles *particles, uint32_t n,
                                     uint32_t *errors, void *stream);
/* Per-stage device timestamps of every step from begin to end, recorded on the
 * scene's stream without synchronizing and resolved at end (the reference's
 * GpuTimestamps passed to step and resolved afterwards, src_mpm/pipeline.rs:389,
 * 521-523; src_viewer/viewer.rs:675-697).  end synchronizes and returns the
 * per-step averages. */
nxm_status nxm_scene_profile_begin(nxm_scene *scene);
/* As nxm_scene_profile_begin, recording the stage events on one substep in `every` only
 * (an event record between kernels costs GPU time of its own; sampling keeps the timed
 * region's throughput intact).  No reference counterpart (GpuTimestamps times every pass). */
nxm_status nxm_scene_profile_sampled(nxm_scene *scene, uint32_t every);
nxm_status nxm_scene_profile_end(nxm_scene *scene, nxm_stage_times *avg, uint32_t *steps);

/* ---- batching: many independent scenes stepped by one launch per stage ----
 * No counterpart in the reference (its MPM state is single-scene,
 * src/state.rs:444-446).  A scene added to a batch is stepped only through the
 * batch until it is removed. */
nxm_status nxm_batch_create(nxm_ctx *ctx, nxm_batch **out);
void nxm_batch_destroy(nxm_batch *batch);
nxm_status nxm_batch_add_scene(nxm_batch *batch, nxm_scene *scene);
nxm_status nxm_batch_remove_scene(nxm_batch *batch, nxm_scene *scene);
nxm_status nxm_batch_step(nxm_batch *batch, nxm_stage_times *times);
nxm_status nxm_batch_run(nxm_batch *batch, uint32_t substeps);
nxm_status nxm_batch_synchronize(nxm_batch *batch);
size_t nxm_batch_len(const nxm_batch *batch);
/* Replay each substep of this batch as one captured HIP graph (default off).  The graph
 * is recaptured when scenes join or leave or buffers grow; stage timing steps run
 * uncaptured.  No reference counterpart (the reference records one compute pass per
 * step, src_mpm/pipeline.rs:385-526). */
nxm_status nxm_batch_set_graph(nxm_batch *batch, int enabled);
/* Online autotuner: the next substeps of the batch try candidate launch shapes
 * (coordinate descent over p2g_groups, p2g_block_cost and grid_groups, each candidate
 * timed with HIP events over a few substeps) and keep the fastest.  The substeps run
 * normally meanwhile; retunes when the particle count changes by more than 2x.  No
 * reference counterpart. */
nxm_status nxm_batch_set_autotune(nxm_batch *batch, int enabled);
nxm_status nxm_scene_set_autotune(nxm_scene *scene, int enabled);
/* Current launch shapes (the tuner's choice once it has finished) and a manual setting
 * (which stops the tuner).  No reference counterpart. */
nxm_status nxm_batch_get_launch_config(nxm_batch *batch, nxm_launch_config *out);
nxm_status nxm_batch_set_launch_config(nxm_batch *batch, const nxm_launch_config *cfg);
nxm_status nxm_scene_get_launch_config(nxm_scene *scene, nxm_launch_config *out);
/* Tuner state: *tuning = 1 while the tuner is still trying shapes (0 once it has kept
 * one, or when it is off); *substeps = the substeps its current or last run has used.
 * Either pointer may be null.  No reference counterpart. */
nxm_status nxm_batch_get_autotune(nxm_batch *batch, int *tuning, uint32_t *substeps);
nxm_status nxm_scene_get_autotune(nxm_scene *scene, int *tuning, uint32_t *substeps);
nxm_status nxm_batch_profile_begin(nxm_batch *batch);
nxm_status nxm_batch_profile_end(nxm_batch *batch, nxm_stage_times *avg, uint32_t *steps);

#ifdef __cplusplus
}
#endif
#endif
