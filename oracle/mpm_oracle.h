/*
 * mpm_oracle.h — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of dimforge/nexus's sparse-grid MLS-MPM substep
 * (binning/sort -> P2G -> grid update + cuboid colliders -> G2P -> particle
 * update with the constitutive models).  It is the checker for the HIP
 * library in nexus_amd/: only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it.  The product path never links or calls it.
 *
 * The same source is compiled twice:
 *   liboracle32.so  REAL=float   op-for-op FP32 mirror of the reference arithmetic
 *   liboracle64.so  REAL=double  FP64 oracle (same algorithm, double precision)
 *
 * Every array argument is caller-owned; nothing is retained after a call except
 * the inspection data kept in an orc_ws workspace.
 *
 * Matrices are 3x3, column-major (9 values: column 0, column 1, column 2), the
 * memory layout of the reference's glam Mat3.
 */
#ifndef NEXUS_MPM_ORACLE_H
#define NEXUS_MPM_ORACLE_H

#include <stdint.h>

#ifndef REAL
#define REAL float
#endif

#ifdef __cplusplus
extern "C" {
#endif

/* Shape kinds: the reference's Shape tags (src_rbd_shaders/shapes/shape.rs:29-35). */
#define ORC_SHAPE_BALL 0u
#define ORC_SHAPE_CUBOID 1u
#define ORC_SHAPE_CAPSULE 2u
#define ORC_SHAPE_CONE 3u
#define ORC_SHAPE_CYLINDER 4u
#define ORC_SHAPE_TRIMESH 6u

/* A collider on its body, with a one-way boundary condition.  Same layout as
 * nxm_collider in include/nexus_mpm.h (the nxm_cuboid body, then the shape).
 * inv_mass / local_com: the body's mass properties as MPM holds them (zero unless the
 * body is dynamic and coupled two ways): the world centre of mass is pose * local_com,
 * and gravity acts on the body along the axes where inv_mass is nonzero.  dims: ball
 * radius; capsule segment ends a, b and radius; cone and cylinder half height and
 * radius (y axis); a cuboid's half extents are half_extents. */
typedef struct {
    float half_extents[3];
    float rotation[4];      /* unit quaternion x, y, z, w */
    float translation[3];
    float linvel[3];
    float angvel[3];
    uint32_t bc_type;       /* 0 stick, 1 slip, 2 separate, 3 non-reflecting */
    float bc_friction;
    float inv_mass[3];
    float local_com[3];
    uint32_t shape;         /* ORC_SHAPE_* */
    float dims[7];
} orc_cuboid;

typedef struct orc_ws orc_ws;

orc_ws *orc_ws_create(void);
void orc_ws_destroy(orc_ws *ws);

/*
 * One substep, in place.  sim = {gx, gy, gz, dt, cell_width}.
 * x, v: n*3; C (the fused APIC "affine"), F: n*9; pstate: n*3 Drucker-Prager
 * plastic state {det Fp, hardening q, log volume gain} (only read/written for
 * the sand tags); params: n*24 model words (f32 bit patterns, words 0..2 of the
 * sand tags are ignored in favour of pstate).  keep_debug != 0 keeps the sort and
 * grid of this substep in ws for inspection.  The colliders' poses are integrated
 * with their velocities at the end of the substep (in place).  Returns 0 on success.
 */
int orc_step(orc_ws *ws, int64_t n, REAL *x, REAL *v, REAL *C, REAL *F, REAL *mass,
             const REAL *vol0, const REAL *damping, REAL *pstate, uint32_t *enabled,
             const uint32_t *fixed, const uint32_t *tag, const float *params,
             const float *sim, orc_cuboid *cols, int ncol, int keep_debug);

/*
 * CPIC state (use_cpic = true in the reference, src_mpm/pipeline.rs:412-460): per
 * particle the affinity word (bit i: collider i within reach, bit 16 + i: its side;
 * read as the previous substep's and overwritten), and the outputs of the substep's
 * contact distance field: the normal (n*3), signed distance (n) and the bodies'
 * velocity at the particle (rigid_vel, n*3).  boundary_friction (n) scales the
 * colliders' friction per particle.  At most 16 colliders.
 */
typedef struct orc_trimesh orc_trimesh;
typedef struct {
    uint32_t *affinity;
    REAL *normal;
    REAL *signed_distance;
    REAL *rigid_vel;
    const REAL *boundary_friction;
    const orc_trimesh *meshes;  /* the last nmesh colliders are these meshes' bodies */
    int nmesh;
} orc_cpic;

/* A triangle-mesh collider (src_mpm/solver/particle.rs:295-345): vertices and triangles
 * in the body frame and its rigid particles, the surface samples of orc_sample_trimesh
 * with the triangle each belongs to.  Its body (pose, velocities, boundary condition,
 * mass properties) is an orc_cuboid entry of the collider array, half extents unused. */
struct orc_trimesh {
    const float *vertices;          /* nv*3 */
    const uint32_t *triangles;      /* nt*3 */
    const float *samples;           /* ns*3 */
    const uint32_t *sample_triangle;/* ns */
    int64_t nv, nt, ns;
};

/* sample_mesh (src_mpm/sampling/sample_trimesh.rs:60-231): points on every triangle
 * (rows spaced spacing / sqrt 2 along the longest edge and its height) and along every
 * edge once, in f32.  Writes up to cap samples (points*3, triangle ids) and returns how
 * many there are. */
int64_t orc_sample_trimesh(const float *vertices, int64_t nv, const uint32_t *triangles, int64_t nt,
                           float spacing, float *points, uint32_t *triangle_of, int64_t cap);

/* orc_step with the CPIC transfers when cp != NULL: node contact distance field (the
 * cuboids' by projection, the meshes' from their rigid particles),
 * particle CDF, compatible / incompatible P2G, both grid fields (no collider pass),
 * CPIC G2P and the particle update's reprojection and penalty.  Returns -3 for more
 * than 16 colliders. */
int orc_step_ex(orc_ws *ws, int64_t n, REAL *x, REAL *v, REAL *C, REAL *F, REAL *mass,
                const REAL *vol0, const REAL *damping, REAL *pstate, uint32_t *enabled,
                const uint32_t *fixed, const uint32_t *tag, const float *params,
                const float *sim, orc_cuboid *cols, int ncol, int keep_debug, const orc_cpic *cp);

/* Mesh colliders in the grid collider pass of substeps without CPIC
 * (grid_update_collide.rs:53-67, TriMesh::project_local_point, trimesh.rs:155-207): the
 * meshes of the last nmesh colliders, kept by pointer (with their pseudo-normals
 * computed here) until the next call; nmesh = 0 forgets them. */
int orc_ws_set_meshes(orc_ws *ws, const orc_trimesh *meshes, int nmesh);
/* One point's projection onto a posed mesh collider within max_dist: 1 with the world
 * point and the pseudo-normal side, 0 when no triangle is that near.  Only the
 * vertices and triangles of m are read. */
int orc_project_trimesh(const orc_trimesh *m, const orc_cuboid *body, REAL max_dist, const REAL *pt, REAL *proj,
                        int *inside);

/* Inspection of the last substep run with keep_debug != 0. */
int64_t orc_ws_nblocks(const orc_ws *ws);
int64_t orc_ws_nprimary_blocks(const orc_ws *ws);
int64_t orc_ws_nsorted(const orc_ws *ws);
/* vid: nb*3; np, npe, first: nb; buckets: nb*10 (per-bucket counts). */
int orc_ws_blocks(const orc_ws *ws, int32_t *vid, uint32_t *np, uint32_t *npe,
                  uint32_t *buckets, uint32_t *first);
int orc_ws_sorted_ids(const orc_ws *ws, uint32_t *ids);
/* Grid after P2G (mv: nb*64*3, m: nb*64) and after the grid update with the
 * collider boundary (vel: nb*64*3).  Node order inside a block: x + 4y + 16z. */
int orc_ws_grid(const orc_ws *ws, REAL *mv, REAL *m, REAL *vel);

/* CPIC grid data of the last substep (orc_step_ex with cp): node CDF distance, affinity
 * word and closest collider (NONE = 0xFFFFFFFF) per node (nb*64), incompatible momentum
 * and mass after P2G, incompatible velocity after the grid update.  -1 without CPIC. */
int orc_ws_grid_cpic(const orc_ws *ws, REAL *dist, uint32_t *aff, uint32_t *closest,
                     REAL *mv_i, REAL *m_i, REAL *vel_i);

/* Building blocks exposed for unit tests. */
int orc_svd3(const REAL *F, REAL *U, REAL *s, REAL *V);
int orc_model_update(uint32_t tag, const float *params, REAL *pstate, REAL *F,
                     const REAL *gradv, REAL *tau);
/* Shape::project_point_on_boundary of any analytic collider (the name predates the
 * other shapes). */
int orc_project_cuboid(const orc_cuboid *c, const REAL *pt, REAL *proj, int *inside);
int orc_grid_node_bc(const orc_cuboid *cols, int ncol, const float *sim,
                     const REAL *node_pos, REAL *vel);
int orc_block_keys(int64_t n, const REAL *x, float cell_width, int32_t *vid,
                   uint32_t *local);
/* gpu_particle_update for one particle (solver/particle_update.rs:60-201) from its G2P
 * outputs: velocity gv (3), velocity gradient gradv (9) and divergence div, and its contact
 * field (signed distance psd, normal and the bodies' velocity rigid_vel, 3 each; psd = 0
 * without CPIC).  Updates x, v, C, F, pstate (3, sand and snow tags), mass and enabled in
 * place.  Returns nonzero for an unknown model tag. */
int orc_particle_update(uint32_t tag, const float *params, float cell_width, float dt, REAL damping, uint32_t fixed,
                        REAL vol0, const REAL *gv, const REAL *gradv, REAL div, REAL psd, const REAL *normal,
                        const REAL *rigid_vel, REAL *x, REAL *v, REAL *C, REAL *F, REAL *pstate, REAL *mass,
                        uint32_t *enabled);
/* Node CDF at pt (one collider pass of grid_update_cdf); particle CDF from the 27 stencil
 * nodes in NBH_SHIFTS order (refp = base node - particle); 4x4 solve m x = b (m
 * row-major), returning det m. */
int orc_node_cdf(const orc_cuboid *cols, int ncol, float cell_width, const REAL *pt, REAL *dist,
                 uint32_t *aff, uint32_t *closest);
int orc_particle_cdf(const REAL *refp, float cell_width, const REAL *dist27, const uint32_t *aff27,
                     uint32_t prev, uint32_t *aff, REAL *normal, REAL *sd);
REAL orc_m4_solve(const REAL *m, const REAL *b, REAL *out);
/* CFL timestep bound over the enabled particles, in units of 1e-12 s (the reference's
 * GpuTimestepBounds::computed_max_dt_as_uint); u32::MAX when none is enabled. */
uint32_t orc_timestep_bound(int64_t n, const REAL *v, const REAL *C, const REAL *F, const REAL *mass,
                            const REAL *vol0, const uint32_t *enabled, const uint32_t *tag, const float *params,
                            float cell_width);

#ifdef __cplusplus
}
#endif
#endif
