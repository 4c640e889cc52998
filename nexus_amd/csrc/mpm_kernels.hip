// mpm_kernels.hip — gfx950 kernels of the MLS-MPM substep.
//
// One substep over every active scene of a world (device data in WorldDev):
//   k_clear         clear hash regions, block tables and counters
//   k_keys_touch    per particle: block key; wave-deduplicated insertion of primary blocks
//   k_touch_nbrs    per primary block: insert its 7 +1 neighbours
//   k_block_info    per block: dense ids of the 7 +1 and 7 -1 neighbours
//   k_count         per particle: dense block id, wave-aggregated block counts and ranks
//   k_scan_*        exclusive scan of block counts -> first particle of each block
//   k_scatter       per particle: sorted position (block-major permutation)
//   k_p2g           per block: LDS-staged scatter of its particles into a 6^3 node tile
//   k_grid_update   per node: sum of the tiles covering it, momentum -> velocity, gravity,
//                   clamp, cuboid colliders
//   k_g2p_update    per sorted particle: 27-node gather + deformation / constitutive update,
//                   written in sorted order to the output arena
// The reference runs the same stages as ~21 dispatches (src_mpm/pipeline.rs:385-526).
#include "mpm_kernels.h"

namespace nxm {

// ------------------------------------------------------------------------------------
// Wave helpers (wave64).

__device__ inline uint32_t lane_id() { return __lane_id(); }

__device__ inline uint64_t lanes_le(uint32_t l) { return l == 63 ? ~0ull : ((2ull << l) - 1ull); }

// Index of the run head at or below this lane given the 64-bit mask of run heads.
__device__ inline uint32_t run_head(uint64_t heads, uint32_t l)
{
    uint64_t m = heads & lanes_le(l);
    return 63u - (uint32_t)__clzll(m);
}

// Index one past the end of the run starting at head h.
__device__ inline uint32_t run_end(uint64_t heads, uint32_t h)
{
    uint64_t above = h == 63 ? 0ull : (heads & ~lanes_le(h));
    return above ? (uint32_t)__ffsll((unsigned long long)above) - 1u : 64u;
}

// XCD-contiguous work ranges.  Workgroups are dealt to the 8 XCDs round-robin
// (workgroup i runs on XCD i mod 8), and each XCD has its own L2.  Renumbering the
// workgroups so that those of one XCD are consecutive, and giving each a contiguous
// range of the (Morton-ordered) work, keeps neighbouring blocks and their particles in
// the same L2.  The grid must be a multiple of 8 workgroups.
constexpr uint32_t NUM_XCD = 8;

__device__ inline void xcd_range(uint32_t n, uint32_t &begin, uint32_t &end)
{
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t logical = (g % NUM_XCD == 0) ? (b % NUM_XCD) * (g / NUM_XCD) + b / NUM_XCD : b;
    const uint32_t per = (n + g - 1) / g;
    begin = min(n, logical * per);
    end = min(n, begin + per);
}

// ------------------------------------------------------------------------------------
// Hash regions: open addressing, linear probing, 32-bit packed block keys
// (src_mpm_shaders/grid/grid.rs:447-536).  Values are dense block ids.

__device__ inline bool hash_insert(uint32_t *keys, uint32_t mask, uint32_t key, uint32_t &slot)
{
    uint32_t s = hash_key(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        uint32_t prev = atomicCAS(keys + s, EMPTY_KEY, key);
        if (prev == EMPTY_KEY) { slot = s; return true; }
        if (prev == key) { slot = s; return false; }
        s = (s + 1) & mask;
    }
    slot = NONE;
    return false;
}

__device__ inline uint32_t hash_find(const uint32_t *keys, const uint32_t *vals, uint32_t mask, uint32_t key)
{
    uint32_t s = hash_key(key) & mask;
    for (uint32_t probe = 0; probe <= mask; ++probe) {
        uint32_t k = keys[s];
        if (k == key) return vals[s];
        if (k == EMPTY_KEY) return NONE;
        s = (s + 1) & mask;
    }
    return NONE;
}

__device__ inline uint32_t find_block(const WorldDev &w, const SceneDev &sc, int32_t x, int32_t y, int32_t z)
{
    return hash_find(w.hkeys + sc.h_base, w.hvals + sc.h_base, sc.h_cap - 1u, pack_vid(x, y, z));
}

// Key of a particle that cannot get a block: the x field 0x7FF never occurs for
// in-range blocks or their +1 neighbours.
constexpr uint32_t OVERFLOW_KEY = 0x000007FFu;

// ------------------------------------------------------------------------------------
// Block building without counters: keys are inserted into the scenes' hash regions,
// then the occupied slots are compacted by a prefix scan into dense block ids.  Each
// region ends with one virtual slot, always occupied, that becomes the scene's
// overflow block, so the dense ids of a scene are contiguous and end with it.

__device__ inline uint32_t scene_of_slot(const WorldDev &w, uint32_t j)
{
    // last active region whose start is <= j
    uint32_t lo = 0, hi = w.n_regions;
    while (hi - lo > 1) {
        uint32_t mid = (lo + hi) >> 1;
        if (w.region_start[mid] <= j) lo = mid; else hi = mid;
    }
    return w.region_scene[lo];
}

__device__ inline int4 unpack_vid(uint32_t key)
{
    return make_int4((int)(key & 0x7FFu) - 0x3FF, (int)((key >> 11) & 0x3FFu) - 0x1FF, (int)(key >> 21) - 0x3FF, 0);
}

__global__ void k_clear(WorldDev w)
{
    uint32_t n = max(w.h_total, w.b_total);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < w.h_total) { w.hkeys[i] = EMPTY_KEY; w.hmark[i] = 0u; }
        if (i < w.b_total) w.bnp[i] = 0u;
        if (i == 0) { w.g->p_out = 0u; w.g->n_blocks = 0u; }
    }
}

__global__ void k_keys_touch(WorldDev w)
{
    // The virtual overflow slot closing each active scene's region.
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.s_count; s += gridDim.x * blockDim.x) {
        const SceneDev sc = w.scenes[s];
        if (sc.active) w.hkeys[sc.h_base + sc.h_cap] = OVERFLOW_KEY;
    }
    const uint32_t n = w.g->p_in;
    const uint32_t nr = (n + WAVE - 1) / WAVE * WAVE;
    const uint32_t l = lane_id();
    const float *X = w.in.plane(PX), *Y = w.in.plane(PY), *Z = w.in.plane(PZ);
    const uint32_t *FL = w.in.uplane(PFLAGS);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        bool valid = i < n;
        uint32_t s = 0, key = EMPTY_KEY;
        SceneDev sc;
        if (valid) {
            s = flags_scene(FL[i]);
            sc = w.scenes[s];
            valid = sc.active != 0u;
        }
        if (valid) {
            int32_t bx, by, bz;
            uint32_t lx, ly, lz;
            cell_of(X[i], sc.h, bx, lx);
            cell_of(Y[i], sc.h, by, ly);
            cell_of(Z[i], sc.h, bz, lz);
            if (vid_in_range(bx, by, bz)) {
                key = pack_vid(bx, by, bz);
            } else {
                key = OVERFLOW_KEY;
                atomicOr(&w.ctr[s].err, ERR_DOMAIN);
            }
            w.pkey[i] = key;
        }
        // Run heads: adjacent lanes with the same (scene, key) insert once.
        uint32_t pk = __shfl_up(key, 1, WAVE);
        uint32_t ps = __shfl_up(s, 1, WAVE);
        bool pv = __shfl_up((int)valid, 1, WAVE) != 0;
        bool head = valid && key != OVERFLOW_KEY && (l == 0 || !pv || pk != key || ps != s);
        if (head) {
            uint32_t slot;
            hash_insert(w.hkeys + sc.h_base, sc.h_cap - 1u, key, slot);
            if (slot != NONE) w.hmark[sc.h_base + slot] = 1u;
            else atomicOr(&w.ctr[s].err, ERR_GRID_CAPACITY);
        }
    }
}

// Inserts the 7 +1 neighbours of every primary block (hash slots marked by the
// touch): 8 threads per slot, thread 1..7 inserting one neighbour each.
__global__ void k_touch_nbrs(WorldDev w)
{
    const uint64_t n = (uint64_t)w.h_total * 8;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t j = (uint32_t)(t >> 3), i = (uint32_t)(t & 7);
        if (i == 0 || w.hmark[j] != 1u) continue;
        const uint32_t key = w.hkeys[j];
        const uint32_t s = scene_of_slot(w, j);
        const SceneDev sc = w.scenes[s];
        const int4 v = unpack_vid(key);
        uint32_t slot;
        hash_insert(w.hkeys + sc.h_base, sc.h_cap - 1u, pack_vid(v.x + ASSOC[i][0], v.y + ASSOC[i][1], v.z + ASSOC[i][2]), slot);
        if (slot == NONE) atomicOr(&w.ctr[s].err, ERR_GRID_CAPACITY);
    }
}

// Exclusive scans, 1024 elements per tile.  MODE_SLOTS counts occupied hash slots
// (dense block ids); MODE_COUNTS sums the particles of each block (first sorted
// position of each block).
constexpr int SCAN_TILE = 1024;
enum ScanMode { MODE_SLOTS = 0, MODE_COUNTS = 1 };

template <int MODE>
__device__ inline uint32_t scan_len(const WorldDev &w) { return MODE == MODE_SLOTS ? w.h_total : w.b_total; }

template <int MODE>
__device__ inline uint32_t scan_in(const WorldDev &w, uint32_t i)
{
    if (MODE == MODE_SLOTS) return w.hkeys[i] != EMPTY_KEY ? 1u : 0u;
    return w.bnp[i];
}

__device__ inline uint32_t block_reduce_256(uint32_t v, uint32_t *sh)
{
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, WAVE);
    if (lane_id() == 0) sh[threadIdx.x / WAVE] = v;
    __syncthreads();
    uint32_t t = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return t;
}

// Inclusive wave scan.
__device__ inline uint32_t wave_incl_scan(uint32_t v)
{
    uint32_t l = lane_id();
    for (int o = 1; o < WAVE; o <<= 1) {
        uint32_t u = __shfl_up(v, o, WAVE);
        if (l >= (uint32_t)o) v += u;
    }
    return v;
}

// Exclusive scan over a 256-thread block, one element per thread.
__device__ inline uint32_t block_excl_scan_256(uint32_t v, uint32_t *sh, uint32_t &total)
{
    uint32_t incl = wave_incl_scan(v);
    uint32_t wv = threadIdx.x / WAVE;
    if (lane_id() == WAVE - 1) sh[wv] = incl;
    __syncthreads();
    uint32_t off = 0;
    for (uint32_t k = 0; k < wv; ++k) off += sh[k];
    total = sh[0] + sh[1] + sh[2] + sh[3];
    __syncthreads();
    return off + incl - v;
}

template <int MODE>
__global__ void __launch_bounds__(256) k_scan_reduce(WorldDev w)
{
    __shared__ uint32_t sh[4];
    const uint32_t len = scan_len<MODE>(w);
    const uint32_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t v = 0;
        for (int k = 0; k < 4; ++k) {
            uint32_t i = t * SCAN_TILE + k * 256 + threadIdx.x;
            if (i < len) v += scan_in<MODE>(w, i);
        }
        uint32_t total = block_reduce_256(v, sh);
        if (threadIdx.x == 0) w.scan_tmp[t] = total;
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_scan_tiles(WorldDev w)
{
    __shared__ uint32_t sh[4];
    const uint32_t ntiles = (scan_len<MODE>(w) + SCAN_TILE - 1) / SCAN_TILE;
    uint32_t carry = 0;
    for (uint32_t base = 0; base < ntiles; base += 256) {
        uint32_t i = base + threadIdx.x;
        uint32_t v = i < ntiles ? w.scan_tmp[i] : 0u;
        uint32_t total;
        uint32_t ex = block_excl_scan_256(v, sh, total);
        if (i < ntiles) w.scan_tmp[i] = carry + ex;
        carry += total;
    }
    if (threadIdx.x == 0) {
        if (MODE == MODE_SLOTS) w.g->n_blocks = min(carry, w.b_total);
        else w.g->p_out = carry;
    }
}

template <int MODE>
__global__ void __launch_bounds__(256) k_scan_down(WorldDev w)
{
    __shared__ uint32_t sh[4];
    const uint32_t len = scan_len<MODE>(w);
    const uint32_t ntiles = (len + SCAN_TILE - 1) / SCAN_TILE;
    for (uint32_t t = blockIdx.x; t < ntiles; t += gridDim.x) {
        uint32_t carry = w.scan_tmp[t];
        for (int k = 0; k < 4; ++k) {
            const uint32_t i = t * SCAN_TILE + k * 256 + threadIdx.x;
            const uint32_t v = i < len ? scan_in<MODE>(w, i) : 0u;
            uint32_t total;
            const uint32_t ex = block_excl_scan_256(v, sh, total);
            if (i < len) {
                const uint32_t r = carry + ex;
                if (MODE == MODE_COUNTS) {
                    w.bfirst[i] = r;
                } else {
                    // Dense block id of slot i: its rank among occupied slots.
                    w.hvals[i] = r;
                    if (v) {
                        const uint32_t key = w.hkeys[i];
                        const uint32_t s = scene_of_slot(w, i);
                        if (r < w.b_total) {
                            w.bscene[r] = s;
                            w.bvid[r] = key == OVERFLOW_KEY ? make_int4(OVERFLOW_VID, 0, 0, 0) : unpack_vid(key);
                            w.bprim[r] = w.hmark[i];
                        } else {
                            w.hvals[i] = NONE;
                            atomicOr(&w.ctr[s].err, ERR_GRID_CAPACITY);
                        }
                    }
                }
            }
            carry += total;
        }
    }
}

// Per block, 16 threads: threads 0..7 look up the 8 blocks vid - {0,1}^3 (grid update
// sources), threads 8..15 the 8 blocks vid + {0,1}^3 (G2P tile, primaries only); thread j
// also tests collider j against the block (bounding spheres) for the collider mask.
__global__ void k_block_info(WorldDev w)
{
    // Per scene: its dense block range; the last block of the range is the overflow block.
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.s_count; s += gridDim.x * blockDim.x) {
        const SceneDev sc = w.scenes[s];
        if (!sc.active) continue;
        const uint32_t first = w.hvals[sc.h_base];
        const uint32_t last = w.hvals[sc.h_base + sc.h_cap];
        w.ctr[s].b_start = first;
        w.ctr[s].b_count = last == NONE ? 0u : last + 1u - first;
        if (last != NONE && last - first > sc.b_cap) atomicOr(&w.ctr[s].err, ERR_GRID_CAPACITY);
    }
    const uint64_t n = (uint64_t)w.g->n_blocks * 16;
    for (uint64_t t = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += (uint64_t)gridDim.x * blockDim.x) {
        const uint32_t d = (uint32_t)(t >> 4), j = (uint32_t)(t & 15);
        const int4 v = w.bvid[d];
        if (v.x == OVERFLOW_VID) continue;
        const SceneDev sc = w.scenes[w.bscene[d]];
        if (j < 8) {
            w.bdn[8 * d + j] = j == 0 ? d : find_block(w, sc, v.x - ASSOC[j][0], v.y - ASSOC[j][1], v.z - ASSOC[j][2]);
        } else if (w.bprim[d]) {
            const uint32_t i = j - 8;
            w.bup[8 * d + i] = i == 0 ? d : find_block(w, sc, v.x + ASSOC[i][0], v.y + ASSOC[i][1], v.z + ASSOC[i][2]);
        }
        // Collider mask: collider j can touch a node of this block only if its bounding
        // sphere comes within 1.5 h (the projection reach) of the block's node sphere.
        bool hit = false;
        if (j < sc.col_count) {
            const ColliderDev &c = w.cols[sc.col_start + j];
            const float h = sc.h;
            const f3 center = mk3(((float)(4 * v.x) + 1.5f) * h, ((float)(4 * v.y) + 1.5f) * h, ((float)(4 * v.z) + 1.5f) * h);
            const float rb = 1.5f * h * 1.7320509f;                  // node sphere of the block
            const float rc = sqrtf(c.he[0] * c.he[0] + c.he[1] * c.he[1] + c.he[2] * c.he[2]);
            const f3 dc = center - mk3(c.t[0], c.t[1], c.t[2]);
            const float reach = rb + rc + 1.5f * h * 1.7320509f + 2.0f * h;
            hit = dot(dc, dc) <= reach * reach;
        }
        // 16 consecutive lanes of a wave hold one block: combine their bits.
        const uint64_t bits = __ballot(hit);
        if (j == 0) {
            const uint32_t lane0 = __lane_id() & ~15u;
            w.bmask[d] = (uint32_t)((bits >> lane0) & 0xFFFFull);
        }
    }
}

__global__ void k_count(WorldDev w)
{
    const uint32_t n = w.g->p_in;
    const uint32_t nr = (n + WAVE - 1) / WAVE * WAVE;
    const uint32_t l = lane_id();
    const uint32_t *FL = w.in.uplane(PFLAGS);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        bool valid = i < n;
        uint32_t d = NONE;
        if (valid) {
            uint32_t s = flags_scene(FL[i]);
            const SceneDev sc = w.scenes[s];
            valid = sc.active != 0u;
            if (valid) {
                uint32_t key = w.pkey[i];
                d = key == OVERFLOW_KEY ? NONE
                                        : hash_find(w.hkeys + sc.h_base, w.hvals + sc.h_base, sc.h_cap - 1u, key);
                if (d == NONE) d = w.hvals[sc.h_base + sc.h_cap];
                w.pdense[i] = d;
            }
        }
        uint32_t pd = __shfl_up(d, 1, WAVE);
        bool pv = __shfl_up((int)valid, 1, WAVE) != 0;
        bool head = !valid || l == 0 || !pv || pd != d;
        uint64_t heads = __ballot(head);
        uint32_t h = run_head(heads, l);
        uint32_t base = 0;
        if (head && valid) base = atomicAdd(&w.bnp[d], run_end(heads, l) - l);
        base = __shfl(base, (int)h, WAVE);
        if (valid) w.prank[i] = base + (l - h);
    }
}

__global__ void k_scatter(WorldDev w)
{
    const uint32_t n = w.g->p_in;
    const uint32_t *FL = w.in.uplane(PFLAGS);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        uint32_t s = flags_scene(FL[i]);
        if (!w.scenes[s].active) continue;
        uint32_t d = w.pdense[i];
        uint32_t pos = w.bfirst[d] + w.prank[i];
        w.perm[pos] = i;
        w.sblock[pos] = d;
    }
}

// ------------------------------------------------------------------------------------
// P2G: one wave per block.  Each particle adds its 27-node quadratic-B-spline stencil
// (momentum m v + C dpt and mass, weighted) into a 6^3 LDS tile covering the block and
// its +1 halo; the tile is then written to the block's staging slot.  Same per-pair
// arithmetic as gpu_p2g_generic (p2g.rs:240-311): dpt = node - x, contribution
// (C dpt + m v) w and m w.
//
// The tile accumulates in 64-bit fixed point with integer LDS atomics (ds_add_u64):
// float LDS atomics serialize on gfx950 (~200 cycles per wave-op against ~8 for the
// 64-bit integer add).  A first pass bounds the block's contributions by 2^e; a tile
// node receives at most np of them (one per particle of the block), so the scale
// 2^(62 - e - ceil(log2 np)) keeps every node sum below 2^62 while resolving
// contributions down to 2^-62 np of the largest.  Integer sums do not depend on the
// order of the adds, so the tile is bit-reproducible whatever the particle order.

__device__ inline float fixed_scale(int e, uint32_t np)
{
    int lg = 0;
    while ((1u << lg) < np && lg < 31) ++lg;
    int k = 62 - e - lg;
    k = k > 126 ? 126 : (k < -126 ? -126 : k);
    return ldexpf(1.0f, k);
}

// Staging tile regions, in the order of the +1 offsets e = 4 ex + 2 ey + ez (the
// reference's blocks_associated_to_block order): sizes 64, 32, 32, 16, 32, 16, 16, 8.
__device__ __constant__ const int REGION_OFFSET[8] = {0, 64, 96, 128, 144, 176, 192, 208};

__device__ inline int bound_exponent(float b)
{
    if (!(b > 0.0f) || !isfinite(b)) return 0;
    int e;
    frexpf(b, &e);   // b = m 2^e, m in [0.5, 1): b < 2^e
    return e;
}

__device__ inline unsigned long long to_fixed(float v, float scale)
{
    float x = v * scale;
    x = fminf(fmaxf(x, -4.6e18f), 4.6e18f);
    return (unsigned long long)__float2ll_rn(x);
}

__device__ inline float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
    return v;
}

__global__ void __launch_bounds__(64) k_p2g(WorldDev w)
{
    __shared__ unsigned long long tile[4 * TILE_NODES];
    __shared__ uint32_t bad[(4 * TILE_NODES + 31) / 32];   // non-finite contributions
    const uint32_t l = threadIdx.x;
    const float *X = w.in.plane(PX), *Y = w.in.plane(PY), *Z = w.in.plane(PZ);
    const float *VX = w.in.plane(PVX), *VY = w.in.plane(PVY), *VZ = w.in.plane(PVZ);
    const float *M = w.in.plane(PMASS);
    const uint32_t *FL = w.in.uplane(PFLAGS);
    uint32_t dbegin, dend;
    xcd_range(w.g->n_blocks, dbegin, dend);
    for (uint32_t d = dbegin; d < dend; ++d) {
        const uint32_t s = w.bscene[d];
        const uint32_t np = w.bnp[d];
        const int4 v = w.bvid[d];
        if (np == 0u || v.x == OVERFLOW_VID) continue;
        const SceneDev sc = w.scenes[s];
        const float h = sc.h;
        const uint32_t first = w.bfirst[d];
        // Pass 1: bound of |m v + C dpt| (|dpt_j| <= 1.5 h) and of m over the block.
        float bmom = 0.0f, bm = 0.0f;
        for (uint32_t c = 0; c < np; c += WAVE) {
            const uint32_t pos = first + c + l;
            if (pos < first + np) {
                const uint32_t slot = w.perm[pos];
                if (FL[slot] & FLAG_ENABLED) {
                    const float m = M[slot];
                    float cmax = 0.0f;
                    #pragma unroll
                    for (int k = 0; k < 9; ++k) cmax = fmaxf(cmax, fabsf(w.in.plane(PC0 + k)[slot]));
                    const float vmax = fmaxf(fabsf(VX[slot]), fmaxf(fabsf(VY[slot]), fabsf(VZ[slot])));
                    bmom = fmaxf(bmom, fabsf(m) * vmax + 4.5f * h * cmax);
                    bm = fmaxf(bm, fabsf(m));
                }
            }
        }
        bmom = wave_max(bmom);
        bm = wave_max(bm);
        const float smom = fixed_scale(bound_exponent(bmom), np);
        const float sm = fixed_scale(bound_exponent(bm), np);
        for (int k = l; k < 4 * TILE_NODES; k += WAVE) tile[k] = 0ull;
        if (l < (4 * TILE_NODES + 31) / 32) bad[l] = 0u;
        __syncthreads();
        // Pass 2: scatter.
        for (uint32_t c = 0; c < np; c += WAVE) {
            const uint32_t pos = first + c + l;
            if (pos < first + np) {
                const uint32_t slot = w.perm[pos];
                const uint32_t fl = FL[slot];
                if (fl & FLAG_ENABLED) {
                    const float px = X[slot], py = Y[slot], pz = Z[slot];
                    const float m = M[slot];
                    const f3 mom = mk3(VX[slot], VY[slot], VZ[slot]) * m;
                    m3 C;
                    C.c0 = mk3(w.in.plane(PC0)[slot], w.in.plane(PC1)[slot], w.in.plane(PC2)[slot]);
                    C.c1 = mk3(w.in.plane(PC3)[slot], w.in.plane(PC4)[slot], w.in.plane(PC5)[slot]);
                    C.c2 = mk3(w.in.plane(PC6)[slot], w.in.plane(PC7)[slot], w.in.plane(PC8)[slot]);
                    // Base node of the stencil and its offset from the particle.
                    const float cx = roundf(px / h) - 1.0f, cy = roundf(py / h) - 1.0f, cz = roundf(pz / h) - 1.0f;
                    const int bx = (int)cx - 4 * v.x, by = (int)cy - 4 * v.y, bz = (int)cz - 4 * v.z;
                    // A particle's primary block holds its base cell: bx, by, bz are in
                    // [0, 3] for every finite position; anything else contributes nothing.
                    const bool inside = (unsigned)bx < 4u && (unsigned)by < 4u && (unsigned)bz < 4u;
                    const f3 d0 = mk3(cx * h - px, cy * h - py, cz * h - pz);
                    float wx[3], wy[3], wz[3];
                    {
                        const float fx = -d0.x / h, fy = -d0.y / h, fz = -d0.z / h;
                        wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx); wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f); wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
                        wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy); wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f); wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
                        wz[0] = 0.5f * (1.5f - fz) * (1.5f - fz); wz[1] = 0.75f - (fz - 1.0f) * (fz - 1.0f); wz[2] = 0.5f * (fz - 0.5f) * (fz - 0.5f);
                    }
                    #pragma unroll
                    for (int sz = 0; sz < 3 && inside; ++sz) {
                        #pragma unroll
                        for (int sy = 0; sy < 3; ++sy) {
                            #pragma unroll
                            for (int sx = 0; sx < 3; ++sx) {
                                const float wgt = wx[sx] * wy[sy] * wz[sz];
                                const f3 dpt = mk3(d0.x + (float)sx * h, d0.y + (float)sy * h, d0.z + (float)sz * h);
                                const f3 vc = (mul(C, dpt) + mom) * wgt;
                                const float mw = m * wgt;
                                const int t = (bx + sx) + TILE * (by + sy) + TILE * TILE * (bz + sz);
                                atomicAdd(&tile[t], to_fixed(vc.x, smom));
                                atomicAdd(&tile[TILE_NODES + t], to_fixed(vc.y, smom));
                                atomicAdd(&tile[2 * TILE_NODES + t], to_fixed(vc.z, smom));
                                atomicAdd(&tile[3 * TILE_NODES + t], to_fixed(mw, sm));
                                if (!isfinite(vc.x) || !isfinite(vc.y) || !isfinite(vc.z) || !isfinite(mw)) {
                                    const int q[4] = {t, TILE_NODES + t, 2 * TILE_NODES + t, 3 * TILE_NODES + t};
                                    #pragma unroll
                                    for (int k = 0; k < 4; ++k) atomicOr(&bad[q[k] >> 5], 1u << (q[k] & 31));
                                }
                            }
                        }
                    }
                }
            }
        }
        __syncthreads();
        // Flush: node t of the 6^3 tile goes to the region of the block it belongs to
        // (offset e = (tx >= 4, ty >= 4, tz >= 4)), stored contiguously in that block's
        // node order, so the consumer reads each region as one run of float4.
        float4 *st = reinterpret_cast<float4 *>(w.stage) + (size_t)d * TILE_NODES;
        const double inv_smom = 1.0 / (double)smom, inv_sm = 1.0 / (double)sm;
        for (int t = l; t < TILE_NODES; t += WAVE) {
            const int tx = t % TILE, ty = (t / TILE) % TILE, tz = t / (TILE * TILE);
            const int ex = tx >> 2, ey = ty >> 2, ez = tz >> 2;
            const int e = 4 * ex + 2 * ey + ez;
            const int nx = ex ? 2 : 4, ny = ey ? 2 : 4;
            const int idx = (tx - 4 * ex) + nx * ((ty - 4 * ey) + ny * (tz - 4 * ez));
            float v4[4];
            #pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int k = c * TILE_NODES + t;
                float val = (float)((double)(long long)tile[k] * (c < 3 ? inv_smom : inv_sm));
                if (bad[k >> 5] & (1u << (k & 31))) val = __builtin_nanf("");
                v4[c] = val;
            }
            st[REGION_OFFSET[e] + idx] = make_float4(v4[0], v4[1], v4[2], v4[3]);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------------------------------
// Grid update: gpu_grid_update + gpu_grid_update_collide (grid_update.rs:21-88,
// grid_update_collide.rs:105-183) on the summed tiles.
__global__ void __launch_bounds__(256) k_grid_update(WorldDev w)
{
    uint32_t nbegin, nend;
    xcd_range(w.g->n_blocks * NODES_PER_BLOCK, nbegin, nend);
    for (uint32_t n = nbegin + threadIdx.x; n < nend; n += blockDim.x) {
        const uint32_t d = n >> 6;
        const uint32_t s = w.bscene[d];
        const int4 v = w.bvid[d];
        if (v.x == OVERFLOW_VID) continue;
        const uint32_t li = n & 63u;
        const int i = li & 3, j = (li >> 2) & 3, k = li >> 4;
        float mx = 0.0f, my = 0.0f, mz = 0.0f, m = 0.0f;
        #pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ex = ASSOC[e][0], ey = ASSOC[e][1], ez = ASSOC[e][2];
            if ((ex && i > 1) || (ey && j > 1) || (ez && k > 1)) continue;
            const uint32_t src = w.bdn[8 * d + e];
            if (src == NONE || w.bnp[src] == 0u) continue;
            const int nx = ex ? 2 : 4, ny = ey ? 2 : 4;
            const float4 q = reinterpret_cast<const float4 *>(w.stage)[(size_t)src * TILE_NODES + REGION_OFFSET[e] + i + nx * (j + ny * k)];
            mx += q.x;
            my += q.y;
            mz += q.z;
            m += q.w;
        }
        if (w.dbg_mom) {
            w.dbg_mom[3 * (size_t)n] = mx;
            w.dbg_mom[3 * (size_t)n + 1] = my;
            w.dbg_mom[3 * (size_t)n + 2] = mz;
        }
        const SceneDev sc = w.scenes[s];
        const float inv_mass = m > 0.0f ? 1.0f / m : 0.0f;
        const float mdt = m * sc.dt;
        f3 vel = mk3((mx + sc.gx * mdt) * inv_mass, (my + sc.gy * mdt) * inv_mass, (mz + sc.gz * mdt) * inv_mass);
        const float lim = sc.vlim;
        vel.x = fminf(fmaxf(vel.x, -lim), lim);
        vel.y = fminf(fmaxf(vel.y, -lim), lim);
        vel.z = fminf(fmaxf(vel.z, -lim), lim);
        const uint32_t mask = w.bmask[d];
        if (mask) {
            const f3 pt = mk3((float)(v.x * 4 + i), (float)(v.y * 4 + j), (float)(v.z * 4 + k)) * sc.h;
            vel = collide_node(w.cols + sc.col_start, sc.col_count, mask, sc.h, sc.dt, pt, vel);
        }
        w.gridv[n] = make_float4(vel.x, vel.y, vel.z, m);
    }
}

// ------------------------------------------------------------------------------------
// G2P + particle update, one thread per sorted particle (g2p.rs:150-270,
// particle_update.rs:60-201, non-CPIC so the contact-distance terms vanish).

// NBH_SHIFTS (grid/kernel.rs:42-70), the reference's G2P summation order.
__device__ __constant__ const int8_t NBH[27][3] = {
    {2, 2, 2}, {2, 0, 2}, {2, 1, 2}, {0, 2, 2}, {0, 0, 2}, {0, 1, 2}, {1, 2, 2}, {1, 0, 2}, {1, 1, 2},
    {2, 2, 0}, {2, 0, 0}, {2, 1, 0}, {0, 2, 0}, {0, 0, 0}, {0, 1, 0}, {1, 2, 0}, {1, 0, 0}, {1, 1, 0},
    {2, 2, 1}, {2, 0, 1}, {2, 1, 1}, {0, 2, 1}, {0, 0, 1}, {0, 1, 1}, {1, 2, 1}, {1, 0, 1}, {1, 1, 1}};

__global__ void __launch_bounds__(256) k_g2p_update(WorldDev w)
{
    uint32_t pbegin, pend;
    xcd_range(w.g->p_out, pbegin, pend);
    for (uint32_t pos = pbegin + threadIdx.x; pos < pend; pos += blockDim.x) {
        const uint32_t slot = w.perm[pos];
        const uint32_t d = w.sblock[pos];
        const uint32_t s = w.bscene[d];
        const SceneDev sc = w.scenes[s];
        const int4 bv = w.bvid[d];
        uint32_t fl = w.in.uplane(PFLAGS)[slot];
        const float px = w.in.plane(PX)[slot], py = w.in.plane(PY)[slot], pz = w.in.plane(PZ)[slot];
        float mass = w.in.plane(PMASS)[slot];
        const float vol0 = w.in.plane(PVOL0)[slot];
        const uint32_t id = w.in.uplane(PID)[slot];
        m3 F;
        F.c0 = mk3(w.in.plane(PF0)[slot], w.in.plane(PF1)[slot], w.in.plane(PF2)[slot]);
        F.c1 = mk3(w.in.plane(PF3)[slot], w.in.plane(PF4)[slot], w.in.plane(PF5)[slot]);
        F.c2 = mk3(w.in.plane(PF6)[slot], w.in.plane(PF7)[slot], w.in.plane(PF8)[slot]);
        f3 vel;
        m3 C;
        f3 npos = mk3(px, py, pz);
        const ClassDev &cls = w.classes[sc.class_base + flags_class(fl)];
        const uint32_t tag = cls.tag;
        const bool sand = tag == 2u || tag == 3u;
        float ps0 = 0.0f, ps1 = 0.0f, ps2 = 0.0f;
        if (sand) { ps0 = w.in.plane(PS0)[slot]; ps1 = w.in.plane(PS1)[slot]; ps2 = w.in.plane(PS2)[slot]; }
        if (bv.x == OVERFLOW_VID) {
            // No block: keep the particle unchanged (the scene reports an error).
            vel = mk3(w.in.plane(PVX)[slot], w.in.plane(PVY)[slot], w.in.plane(PVZ)[slot]);
            C.c0 = mk3(w.in.plane(PC0)[slot], w.in.plane(PC1)[slot], w.in.plane(PC2)[slot]);
            C.c1 = mk3(w.in.plane(PC3)[slot], w.in.plane(PC4)[slot], w.in.plane(PC5)[slot]);
            C.c2 = mk3(w.in.plane(PC6)[slot], w.in.plane(PC7)[slot], w.in.plane(PC8)[slot]);
        } else {
            const float h = sc.h, dt = sc.dt, inv_d = sc.inv_d;
            vel = mk3(0, 0, 0);
            m3 gradv = mzero();
            float div = 0.0f;
            const float cx = roundf(px / h) - 1.0f, cy = roundf(py / h) - 1.0f, cz = roundf(pz / h) - 1.0f;
            const int bx = (int)cx - 4 * bv.x, by = (int)cy - 4 * bv.y, bz = (int)cz - 4 * bv.z;
            const bool inside = (unsigned)bx < 4u && (unsigned)by < 4u && (unsigned)bz < 4u;
            if ((fl & FLAG_ENABLED) && inside) {
                const f3 ref = mk3(cx * h - px, cy * h - py, cz * h - pz);
                float wx[3], wy[3], wz[3];
                {
                    const float fx = -ref.x / h, fy = -ref.y / h, fz = -ref.z / h;
                    wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx); wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f); wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
                    wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy); wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f); wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
                    wz[0] = 0.5f * (1.5f - fz) * (1.5f - fz); wz[1] = 0.75f - (fz - 1.0f) * (fz - 1.0f); wz[2] = 0.5f * (fz - 0.5f) * (fz - 0.5f);
                }
                uint32_t nb[8];
                #pragma unroll
                for (int o = 0; o < 8; ++o) nb[o] = w.bup[8 * d + o];
                #pragma unroll
                for (int q = 0; q < 27; ++q) {
                    const int sx = NBH[q][0], sy = NBH[q][1], sz = NBH[q][2];
                    const int tx = bx + sx, ty = by + sy, tz = bz + sz;
                    const int ox = tx >> 2, oy = ty >> 2, oz = tz >> 2;
                    const uint32_t blk = nb[ox * 4 + oy * 2 + oz];
                    f3 cv = mk3(0, 0, 0);
                    if (blk != NONE) {
                        const float4 g = w.gridv[(size_t)blk * NODES_PER_BLOCK + (tx & 3) + 4 * (ty & 3) + 16 * (tz & 3)];
                        cv = mk3(g.x, g.y, g.z);
                    }
                    const f3 dpt = mk3(ref.x + (float)sx * h, ref.y + (float)sy * h, ref.z + (float)sz * h);
                    const float wgt = wx[sx] * wy[sy] * wz[sz];
                    const float wi = wgt * inv_d;
                    vel = vel + cv * wgt;
                    gradv.c0 = gradv.c0 + cv * (dpt.x * wi);
                    gradv.c1 = gradv.c1 + cv * (dpt.y * wi);
                    gradv.c2 = gradv.c2 + cv * (dpt.z * wi);
                    div = div + wi * dot(cv, dpt);
                }
            }
            // Particle update.
            const float vl = length(vel);
            if (vl > h / dt) vel = vel / vl * h / dt;
            vel = vel / (1.0f + cls.damping * dt);
            if (fl & FLAG_FIXED) vel = mk3(0, 0, 0);
            npos = mk3(px, py, pz) + vel * dt;
            m3 tau;
            if (tag != 4u) {
                F = F + mul(gradv * dt, F);
            } else {
                const float f0 = F.c0.x;
                float rate = div * dt / 3.0f;
                rate = fminf(fmaxf(rate, -0.1f), 0.1f);
                float nd = f0 + rate * f0;
                nd = nd < 0.5f ? 0.5f : (nd > 2.0f ? 2.0f : nd);
                F = m3{{nd, 0, 0}, {0, nd, 0}, {0, 0, nd}};
            }
            switch (tag) {
            case 0u: {
                m3 U, V;
                f3 sv;
                svd3(F, U, sv, V);
                tau = linear_stress_svd(cls.w[0], cls.w[1], U, sv);
                break;
            }
            case 1u: tau = nh_stress(cls.w[0], cls.w[1], F); break;
            case 2u:
            case 3u: {
                m3 U, V;
                f3 sv;
                dp_project(cls.w + 3, ps0, ps1, ps2, F, U, sv, V);
                tau = tag == 2u ? linear_stress_svd(cls.w[10], cls.w[11], U, sv) : nh_stress(cls.w[10], cls.w[11], F);
                break;
            }
            case 4u: tau = fluid_stress(cls.w, F, gradv); break;
            default: tau = mzero(); break;
            }
            C = gradv * mass - tau * (vol0 * inv_d * dt);
            const bool nan_pos = npos.x != npos.x || npos.y != npos.y || npos.z != npos.z;
            if (nan_pos || !(det(F) > 0.0f)) {
                fl &= ~FLAG_ENABLED;
                vel = mk3(0, 0, 0);
                F = mident();
                C = mzero();
                mass = 0.0f;
                npos = mk3(px, py, pz);
            }
        }
        float *o = nullptr;
        (void)o;
        w.out.plane(PX)[pos] = npos.x; w.out.plane(PY)[pos] = npos.y; w.out.plane(PZ)[pos] = npos.z;
        w.out.plane(PVX)[pos] = vel.x; w.out.plane(PVY)[pos] = vel.y; w.out.plane(PVZ)[pos] = vel.z;
        w.out.plane(PC0)[pos] = C.c0.x; w.out.plane(PC1)[pos] = C.c0.y; w.out.plane(PC2)[pos] = C.c0.z;
        w.out.plane(PC3)[pos] = C.c1.x; w.out.plane(PC4)[pos] = C.c1.y; w.out.plane(PC5)[pos] = C.c1.z;
        w.out.plane(PC6)[pos] = C.c2.x; w.out.plane(PC7)[pos] = C.c2.y; w.out.plane(PC8)[pos] = C.c2.z;
        w.out.plane(PF0)[pos] = F.c0.x; w.out.plane(PF1)[pos] = F.c0.y; w.out.plane(PF2)[pos] = F.c0.z;
        w.out.plane(PF3)[pos] = F.c1.x; w.out.plane(PF4)[pos] = F.c1.y; w.out.plane(PF5)[pos] = F.c1.z;
        w.out.plane(PF6)[pos] = F.c2.x; w.out.plane(PF7)[pos] = F.c2.y; w.out.plane(PF8)[pos] = F.c2.z;
        w.out.plane(PMASS)[pos] = mass;
        w.out.plane(PVOL0)[pos] = vol0;
        if (sand) { w.out.plane(PS0)[pos] = ps0; w.out.plane(PS1)[pos] = ps1; w.out.plane(PS2)[pos] = ps2; }
        w.out.uplane(PFLAGS)[pos] = fl;
        w.out.uplane(PID)[pos] = id;
    }
}

// ------------------------------------------------------------------------------------
// Canonical sort readback (not on the hot path): per sorted particle its primary
// block and local cell, and the ten bucket counts of every block including the
// spilled extras (grid/sort.rs:287-349).
__global__ void k_canonical(WorldDev w, Arena prev, uint32_t *pcell_out, uint32_t *buckets)
{
    const uint32_t n = w.g->p_out;
    for (uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += gridDim.x * blockDim.x) {
        const uint32_t slot = w.perm[pos];
        const uint32_t d = w.sblock[pos];
        const int4 bv = w.bvid[d];
        const SceneDev sc = w.scenes[w.bscene[d]];
        if (bv.x == OVERFLOW_VID) { pcell_out[pos] = NONE; continue; }
        int32_t b[3];
        uint32_t loc[3];
        cell_of(prev.plane(PX)[slot], sc.h, b[0], loc[0]);
        cell_of(prev.plane(PY)[slot], sc.h, b[1], loc[1]);
        cell_of(prev.plane(PZ)[slot], sc.h, b[2], loc[2]);
        pcell_out[pos] = loc[0] + 4 * loc[1] + 16 * loc[2];
        atomicAdd(&buckets[10 * d + loc[2]], 1u);
        for (int i = 1; i < 8; ++i) {
            const int sx = ASSOC[i][0], sy = ASSOC[i][1], sz = ASSOC[i][2];
            if ((sx && loc[0] < 2) || (sy && loc[1] < 2) || (sz && loc[2] < 2)) continue;
            const uint32_t dn = w.bup[8 * d + i];
            if (dn == NONE) continue;
            int lz = (int)loc[2] - sz * 4;
            if (lz < -2) lz = -2;
            atomicAdd(&buckets[10 * dn + 4 + (lz + 2)], 1u);
        }
    }
}

// ------------------------------------------------------------------------------------
// Launch sequence.

static inline int grid_for(uint64_t n, int threads, int cap)
{
    uint64_t g = (n + threads - 1) / threads;
    if (g < 1) g = 1;
    if (g > (uint64_t)cap) g = cap;
    g = (g + NUM_XCD - 1) / NUM_XCD * NUM_XCD;   // whole XCD rounds (xcd_range)
    return (int)g;
}

void launch_substep(const WorldDev &w, const LaunchConfig &lc, hipStream_t st, hipEvent_t *ev)
{
    const int T = 256;
    const int cap = lc.persistent_blocks;
    if (ev) (void)hipEventRecord(ev[0], st);
    k_clear<<<grid_for(std::max(w.h_total, w.b_total), T, cap), T, 0, st>>>(w);
    k_keys_touch<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    k_touch_nbrs<<<grid_for((uint64_t)w.h_total * 8, T, cap), T, 0, st>>>(w);
    const int hscan = std::min((int)((w.h_total + SCAN_TILE - 1) / SCAN_TILE), cap);
    k_scan_reduce<MODE_SLOTS><<<hscan, 256, 0, st>>>(w);
    k_scan_tiles<MODE_SLOTS><<<1, 256, 0, st>>>(w);
    k_scan_down<MODE_SLOTS><<<hscan, 256, 0, st>>>(w);
    k_block_info<<<grid_for((uint64_t)w.b_total * 16, T, cap), T, 0, st>>>(w);
    k_count<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    const int bscan = std::min((int)((w.b_total + SCAN_TILE - 1) / SCAN_TILE), cap);
    k_scan_reduce<MODE_COUNTS><<<bscan, 256, 0, st>>>(w);
    k_scan_tiles<MODE_COUNTS><<<1, 256, 0, st>>>(w);
    k_scan_down<MODE_COUNTS><<<bscan, 256, 0, st>>>(w);
    k_scatter<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    if (ev) (void)hipEventRecord(ev[1], st);
    k_p2g<<<lc.p2g_blocks, 64, 0, st>>>(w);
    if (ev) (void)hipEventRecord(ev[2], st);
    k_grid_update<<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);
    if (ev) (void)hipEventRecord(ev[3], st);
    k_g2p_update<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    if (ev) (void)hipEventRecord(ev[4], st);
}

void launch_canonical(const WorldDev &w, const Arena &prev, uint32_t *pcell, uint32_t *buckets, uint32_t p_hint,
                      hipStream_t st)
{
    k_canonical<<<grid_for(p_hint, 256, 4096), 256, 0, st>>>(w, prev, pcell, buckets);
}

}  // namespace nxm
