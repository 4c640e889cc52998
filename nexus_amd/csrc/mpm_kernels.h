// mpm_kernels.h — device view of a world (all scenes stepped together) and the
// launch entry points of mpm_kernels.hip.
#pragma once

#include "mpm_device.h"
#include <cstddef>

namespace nxm {

struct WorldDev {
    Arena in, out;                 // particle arenas: read (previous order), written (sorted order)
    uint32_t *pkey;                // per input slot: packed block key
    uint32_t *pdense;              // per input slot: dense block id
    uint32_t *prank;               // per input slot: rank inside its block
    uint32_t *perm;                // per sorted position: input slot
    uint32_t *sblock;              // per sorted position: dense block id
    uint32_t *bscene;              // per block: scene slot (NONE = unused id)
    int4 *bvid;                    // per block: virtual block coordinates
    uint32_t *bnp;                 // per block: primary particles
    uint32_t *bfirst;              // per block: first sorted position
    uint32_t *bup;                 // per block: 8 ids of vid + {0,1}^3 (reference order)
    uint32_t *bdn;                 // per block: 8 ids of vid - {0,1}^3 (NONE unless primary)
    float *stage;                  // per block: P2G tile, 4 x 216 floats
    float4 *gridv;                 // per node: velocity after the grid update, mass in w
    uint32_t *hkeys, *hvals;       // hash regions; hvals: dense block id of each occupied slot
    uint32_t *hmark;               // per hash slot: 1 if a particle's primary block
    uint32_t *bprim;               // per block: 1 if primary (holds particles)
    uint32_t *bmask;               // per block: colliders that can reach its nodes
    const uint32_t *region_start;  // per active scene, ascending: first hash slot of its region
    const uint32_t *region_scene;  // per active scene: scene slot
    uint32_t n_regions;
    const SceneDev *scenes;
    SceneCtr *ctr;
    const ClassDev *classes;
    ColliderDev *cols;             // integrated in place at the end of each substep
    Globals *g;
    uint32_t *p2g_start;           // P2G workgroup g: blocks [p2g_start[g], p2g_start[g+1])
    uint32_t p2g_groups;           // workgroups of the P2G launch
    uint32_t p2g_block_cost;       // per-block cost of the P2G split, in particles
    unsigned long long *scan_state;   // single-pass scan tile status, 2 x scan_stride words
    uint32_t scan_stride;
    float *dbg_mom;                // optional per-node momentum after P2G (3 floats)
    // CPIC (only when some scene uses it): per node the contact distance field; the
    // incompatible P2G tiles and velocities.
    uint4 *ncdf;                   // {distance bits, affinity, closest collider, its bc type}
    float4 *nbody;                 // per node with a closest collider: its velocity there, its friction
    uint32_t *binc;                // per block: its P2G tile has incompatible contributions
    uint32_t *bany;                // per block: some node is in reach of a collider
    uint32_t *breach;              // per primary block: some node of its particles' stencils is (null: no CPIC)
    uint32_t *clist;               // sorted positions of the contact particles (n_contact of them)
    float *stage_i;
    float4 *gridv_i;
    uint32_t any_cpic;
    // Mesh colliders of CPIC scenes: geometry in the body frame and the rigid particles
    // (surface samples); per node the meshes' part of the field, merged into ncdf and
    // reset to empty by k_grid_cdf.
    const float4 *mesh_v;          // vertices
    const uint4 *mesh_t;           // triangles: three vertex ids, (world scene slot << 4) | collider
    const float4 *mesh_rp;         // rigid particles: point, triangle id (bits)
    unsigned long long *mkey;      // per node: least (distance bits << 32) | collider, ~0 when empty
    uint32_t *maff;                // per node: the meshes' affinity bits
    uint32_t n_rp;
    // Mesh colliders of scenes without CPIC (the grid collider pass projects onto them):
    // vertex and edge pseudo-normals, cell grids and lists, each world collider's mesh.
    const float4 *mesh_vn;
    const float4 *mesh_en;
    const MeshGridDev *mesh_grid;
    const uint32_t *mesh_cell;
    const uint32_t *mesh_ctri;
    const uint32_t *col_mesh;
    uint32_t mesh_pass;            // some scene without CPIC has mesh colliders
    uint32_t p_cap, b_total, h_total, s_count;   // b_total, h_total: allocated block ids, hash slots
    // Per scene slot: simulated time and the CFL bound for adaptive substeps (null: no
    // clocks kept, as in the no-allocation substep); any_adaptive: some scene is adaptive.
    SceneClock *clock;
    uint32_t any_adaptive;
};

struct LaunchConfig {
    uint64_t p_hint;               // upper bound of particles (grid sizing only)
    int persistent_blocks;         // cap on workgroups of grid-stride kernels
    int p2g_blocks;                // workgroups of the per-block kernels
    // CPIC worlds: a second stream (default priority) on which the contact particles' G2P
    // runs beside the plain particles' one, forked and joined with these two events.
    hipStream_t side = nullptr;
    hipEvent_t fork = nullptr, join = nullptr;
};

// keys: run k_keys_touch first (the hash and counters emptied by launch_clear); without
// it the substep uses the keys and block insertions of the previous substep's G2P.
// ev (null, or SUBSTEP_EVENTS events): ev[0..4] are recorded on the stream between the
// stages (sort | P2G | grid update | G2P + update); ev[5..6] and ev[7..8] take the start
// and end of the P2G and the G2P + update kernels themselves (hipExtLaunchKernelGGL's
// kernel events, without the launch gaps the stage events include).
constexpr int SUBSTEP_EVENTS = 9;
void launch_substep(const WorldDev &w, const LaunchConfig &lc, hipStream_t st, hipEvent_t *ev, bool keys);
void launch_clear(const WorldDev &w, const LaunchConfig &lc, hipStream_t st);
// Render data (24 floats per particle, at its id) of arena positions [start, start + count).
void launch_prep_readback(const Arena &a, uint32_t start, uint32_t count, const uint32_t *group, const float *phase,
                          const float *radius, const float4 *palette, uint32_t npal, uint32_t mode, float h, float dt,
                          bool cdf, float *out, int cap, hipStream_t st);
void launch_set_particles(Globals *g, uint32_t p_in, hipStream_t st);
// Minimum over each scene's enabled particles of the CFL bound (1e-12 s units), into
// out[s * stride] by atomicMin (the caller fills it with NONE first).
void launch_timestep_bound(const WorldDev &w, uint32_t n, uint32_t *out, uint32_t stride, int max_blocks, hipStream_t st);
void launch_canonical(const WorldDev &w, const Arena &prev, uint32_t *pcell, uint32_t *buckets, uint32_t p_hint,
                      hipStream_t st);

// ---- substep without allocation (nxm_pipeline_step_noalloc) ----
// The scene's small tables travel as one kernel argument (under 4 KB), so that a captured
// launch carries them by value and no host buffer has to outlive the call.
constexpr int NOALLOC_MAX_CLASSES = 16;
constexpr int NOALLOC_MAX_COLLIDERS = 16;
// A cuboid of that path: ColliderDev's leading body fields (the nxm_cuboid words) and its
// bounding radius; k_noalloc_tables completes the record.
constexpr int COLLIDER_BODY_WORDS = 24;
static_assert(offsetof(ColliderDev, shape) == COLLIDER_BODY_WORDS * sizeof(float), "body words lead ColliderDev");
struct NoallocCuboid {
    float body[COLLIDER_BODY_WORDS];
    float reach;
};
struct NoallocTables {
    SceneDev scene;
    ClassDev classes[NOALLOC_MAX_CLASSES];
    NoallocCuboid cols[NOALLOC_MAX_COLLIDERS];
    uint32_t n_classes, n_cols, n;
};
static_assert(sizeof(NoallocTables) <= 4000, "kernel arguments are limited to 4 KB");
// The caller's particles: device structure-of-arrays in particle-id order, read before
// the substep and written back after it.
struct NoallocParticles {
    float *x, *v, *F, *C, *mass, *vol0, *pstate;   // n*3, n*3, n*9, n*9, n, n, n*3 (pstate may be null)
    uint32_t *flags;                               // n: bit 0 enabled, bit 1 fixed, bits 2..15 material
    uint32_t *errors;                              // 1 word (may be null): the substep's error bits
};
// Writes the tables, the counters and the input arena (from the caller's arrays).
void launch_noalloc_setup(const WorldDev &w, const NoallocTables &t, const NoallocParticles &p, int cap, hipStream_t st);
// Copies the output arena back to the caller's arrays by particle id.
void launch_noalloc_unpack(const WorldDev &w, const NoallocParticles &p, uint32_t n, int cap, hipStream_t st);

}  // namespace nxm
