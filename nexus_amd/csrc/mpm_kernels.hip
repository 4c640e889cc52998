// mpm_kernels.hip — gfx950 kernels of the MLS-MPM substep.
//
// One substep over every active scene of a world (device data in WorldDev):
//   k_clear         clear hash regions, block tables and counters, then
//   k_keys_touch    per particle: block key; per run of equal keys in a wave, insertion
//                   of the block and its 7 +1 neighbours (both only for the first substep
//                   and after the host changed the arena or the tables: otherwise the
//                   previous substep's k_g2p_update made the keys and insertions)
//   k_scan<SLOTS>   single-pass scan of occupied hash slots -> dense block ids
//   k_block_count   per block: dense ids of the 7 +1 and 7 -1 neighbours, collider mask;
//                   per particle: dense block id, wave-aggregated block counts and ranks
//   k_scan<COUNTS>  single-pass scan of block counts -> first particle of each block
//   k_scatter       per particle: sorted position (block-major permutation)
//   k_p2g           per block: LDS-staged scatter of its particles into a 6^3 node tile
//   k_grid_update   per node: sum of the tiles covering it, momentum -> velocity, gravity,
//                   clamp, cuboid colliders; empties the hash
//   k_g2p_update    per sorted particle: 27-node gather + deformation / constitutive update,
//                   written in sorted order to the output arena; then the next substep's
//                   block keys and insertions
// CPIC worlds add the contact fields (k_mesh_cdf, k_grid_cdf, k_particle_cdf) before P2G,
// the CPIC variants of the transfers, k_g2p_contact and k_integrate_colliders.
// The reference runs the same stages as ~21 dispatches (src_mpm/pipeline.rs:385-526).
#include <type_traits>

#include <hip/hip_ext.h>

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

// Workgroup size of k_g2p_update (one particle per thread): two waves, measured fastest for
// the plain kernel (256: sand column G2P 1 us, clumped fluid 5-7 % slower; 192: the fluid 6 %
// slower; 64 and 512 slower too); CPIC worlds keep 256, which shares the GPU better with the contact G2P running
// beside it (128: CPIC sand column 4 % slower).
template <bool CPIC> constexpr int g2p_threads() { return CPIC ? 256 : 128; }

__device__ inline void xcd_range(uint32_t n, uint32_t &begin, uint32_t &end)
{
    const uint32_t g = gridDim.x, b = blockIdx.x;
    const uint32_t logical = (g % NUM_XCD == 0) ? (b % NUM_XCD) * (g / NUM_XCD) + b / NUM_XCD : b;
    // Balanced split: every workgroup gets floor(n / g) or ceil(n / g) items.
    begin = (uint32_t)(((uint64_t)logical * n) / g);
    end = (uint32_t)(((uint64_t)(logical + 1) * n) / g);
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
    uint32_t n = max(max(max(w.h_total, w.b_total), 2u * w.scan_stride), w.s_count);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        if (i < w.h_total) { w.hkeys[i] = EMPTY_KEY; w.hmark[i] = 0u; }
        if (i < w.b_total) w.bnp[i] = 0u;
        if (i < 2u * w.scan_stride) w.scan_state[i] = 0ull;
        if (i == 0) { w.g->p_out = 0u; w.g->n_blocks = 0u; w.g->ticket[0] = 0u; w.g->ticket[1] = 0u; }
        if (w.clock && i < w.s_count) w.clock[i].bound = NONE;   // recomputed by the key pass's bound
    }
}

// Block key of particle slot i at position (x, y, z) of active scene s (written to
// pkey[i]), and the wave's block insertions: adjacent lanes with the same (scene, key)
// form a run; each run inserts its block (marked primary) and the block's 7 +1
// neighbours (the active set is P + {0,1}^3, src_mpm/grid/grid.rs:193-200).  Called by
// every lane of a wave together (invalid lanes with valid = false), so that the lanes'
// neighbours in the wave are the particles' neighbours in the arena.
__device__ inline void touch_keys(const WorldDev &w, uint32_t i, bool valid, uint32_t s, float x, float y, float z)
{
    const uint32_t l = lane_id();
    uint32_t key = EMPTY_KEY;
    if (valid) {
        const SceneDev sc = w.scenes[s];
        int32_t bx, by, bz;
        uint32_t lx, ly, lz;
        cell_of(x, sc.h, sc.inv_h, bx, lx);
        cell_of(y, sc.h, sc.inv_h, by, ly);
        cell_of(z, sc.h, sc.inv_h, bz, lz);
        if (vid_in_range(bx, by, bz)) {
            key = pack_vid(bx, by, bz);
        } else {
            key = OVERFLOW_KEY;
            atomicOr(&w.ctr[s].err, ERR_DOMAIN);
        }
        w.pkey[i] = key;
    }
    const uint32_t pk = __shfl_up(key, 1, WAVE);
    const uint32_t ps = __shfl_up(s, 1, WAVE);
    const bool pv = __shfl_up((int)valid, 1, WAVE) != 0;
    const bool head = valid && key != OVERFLOW_KEY && (l == 0 || !pv || pk != key || ps != s);
    // Up to 8 runs per round: lane l inserts neighbour (l mod 8) of run (l / 8), so a
    // wave's insertions overlap instead of waiting for each other's atomics.
    uint64_t heads = __ballot(head);
    while (heads) {
        uint32_t hsrc[8];
        #pragma unroll
        for (int k = 0; k < 8; ++k) {
            hsrc[k] = heads ? (uint32_t)__ffsll((unsigned long long)heads) - 1u : 64u;
            heads = heads ? heads & (heads - 1ull) : 0ull;
        }
        const uint32_t r = l >> 3, j = l & 7u;
        uint32_t src = hsrc[0];
        #pragma unroll
        for (int k = 1; k < 8; ++k) src = r == (uint32_t)k ? hsrc[k] : src;
        const uint32_t hk = __shfl(key, (int)(src & 63u), WAVE);
        const uint32_t hs = __shfl(s, (int)(src & 63u), WAVE);
        if (src < 64u) {
            const SceneDev hsc = w.scenes[hs];
            const int4 v = unpack_vid(hk);
            uint32_t slot;
            hash_insert(w.hkeys + hsc.h_base, hsc.h_cap - 1u,
                        pack_vid(v.x + ASSOC[j][0], v.y + ASSOC[j][1], v.z + ASSOC[j][2]), slot);
            if (slot == NONE) atomicOr(&w.ctr[hs].err, ERR_GRID_CAPACITY);
            else if (j == 0) w.hmark[hsc.h_base + slot] = 1u;
        }
    }
}

// The virtual overflow slot closing each active scene's hash region.
__device__ inline void mark_overflow_slots(const WorldDev &w)
{
    for (uint32_t s = blockIdx.x * blockDim.x + threadIdx.x; s < w.s_count; s += gridDim.x * blockDim.x) {
        const SceneDev sc = w.scenes[s];
        if (sc.active) w.hkeys[sc.h_base + sc.h_cap] = OVERFLOW_KEY;
    }
}

// Keys and block insertions of every particle of the input arena.  Used for a substep
// whose keys the previous substep's G2P did not insert (the first one, and after the
// host changed the arena or the tables): k_clear has emptied the hash and the counters
// before it.
__global__ void k_keys_touch(WorldDev w)
{
    mark_overflow_slots(w);
    const uint32_t n = w.g->p_in;
    const uint32_t nr = (n + WAVE - 1) / WAVE * WAVE;
    const float *X = w.in.plane(PX), *Y = w.in.plane(PY), *Z = w.in.plane(PZ);
    const uint32_t *FL = w.in.uplane(PFLAGS);
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        bool valid = i < n;
        uint32_t s = 0;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (valid) {
            s = flags_scene(FL[i]);
            valid = w.scenes[s].active != 0u;
            x = X[i];
            y = Y[i];
            z = Z[i];
        }
        touch_keys(w, i, valid, s, x, y, z);
    }
}

// Exclusive scans, SCAN_TILE elements per tile.  MODE_SLOTS counts occupied hash slots
// (dense block ids); MODE_COUNTS sums the particles of each block (first sorted
// position of each block).
enum ScanMode { MODE_SLOTS = 0, MODE_COUNTS = 1 };

template <int MODE>
__device__ inline uint32_t scan_len(const WorldDev &w) { return MODE == MODE_SLOTS ? w.h_total : w.b_total; }

// Runs of SCAN_TILE elements per scan tile for a scan of len elements (at most
// SCAN_MAX_TILES tiles).
constexpr uint32_t SCAN_MAX_TILES = 4096;
__host__ __device__ inline uint32_t scan_chunks(uint32_t len)
{
    const uint32_t per = (uint32_t)SCAN_TILE * SCAN_MAX_TILES;
    return len <= per ? 1u : (uint32_t)(((uint64_t)len + per - 1) / per);
}

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

// Output of one scanned element: its exclusive prefix r and its value v.
template <int MODE>
__device__ inline void scan_out(const WorldDev &w, uint32_t i, uint32_t r, uint32_t v)
{
    if (MODE == MODE_COUNTS) {
        w.bfirst[i] = r;
        return;
    }
    // Dense block id of slot i: its rank among occupied slots.
    w.hvals[i] = r;
    if (v) {
        const uint32_t key = w.hkeys[i];
        const uint32_t s = scene_of_slot(w, i);
        if (r < w.b_total) {
            w.bscene[r] = s;
            w.bvid[r] = key == OVERFLOW_KEY ? make_int4(OVERFLOW_VID, 0, 0, 0) : unpack_vid(key);
            w.bprim[r] = w.hmark[i];
            // This substep's count starts at zero (k_block_count adds to it; the previous
            // substep's stay readable until here; only the dense ids handed out are read).
            w.bnp[r] = 0u;
        } else {
            w.hvals[i] = NONE;
            atomicOr(&w.ctr[s].err, ERR_GRID_CAPACITY);
        }
    }
}

// Single-pass exclusive scan with decoupled look-back (Merrill and Garland, "Single-pass
// Parallel Prefix Scan with Decoupled Look-back", 2016).  Workgroups take tiles in ticket
// order, publish the tile's aggregate, then add the published values of the preceding
// tiles until one carries an inclusive prefix.  Status and value share one 64-bit word
// (bits 32-33: 1 aggregate, 2 inclusive prefix), so a reader sees both or neither.  A
// tile waits only for tiles with smaller tickets; those belong to workgroups already
// running, which publish their aggregate before waiting themselves, so every wait ends.
constexpr unsigned long long ST_AGG = 1ull << 32, ST_PREFIX = 2ull << 32;

__device__ inline void advance_clocks(const WorldDev &w);

template <int MODE>
__global__ void __launch_bounds__(256) k_scan(WorldDev w)
{
    constexpr int PER = SCAN_TILE / 256;
    if (MODE == MODE_SLOTS && blockIdx.x == 0) advance_clocks(w);   // the substep's first kernel
    __shared__ uint32_t sh[4];
    __shared__ uint32_t s_tile, s_prefix;
    const uint32_t len = MODE == MODE_SLOTS ? w.h_total : min(w.g->n_blocks, w.b_total);
    // A tile is `chunks` runs of SCAN_TILE elements, so that there are at most
    // SCAN_MAX_TILES tiles: every tile takes a ticket, one same-address atomic, and on a
    // hash of millions of slots those tickets, not the data, would set the time.  The state
    // array has scan_stride words per scan (the host sizes it for SCAN_TILE tiles); the bound
    // keeps every tile's word inside it whatever the host passed.
    const uint32_t chunks = scan_chunks(len);
    const uint32_t tile_len = SCAN_TILE * chunks;
    const uint32_t ntiles = min((len + tile_len - 1) / tile_len, w.scan_stride - 1u);
    unsigned long long *state = w.scan_state + (MODE == MODE_SLOTS ? 0u : w.scan_stride);
    for (;;) {
        if (threadIdx.x == 0) s_tile = atomicAdd(&w.g->ticket[MODE], 1u);
        __syncthreads();
        const uint32_t t = s_tile;
        if (t >= ntiles) break;
        // One run (a small scan): thread elements i1 .. i1 + PER - 1, one block scan.  Several
        // runs (a large scan): element i0 + 256 k of each run (k < PER) belongs to this
        // thread, row k of the run being 256 consecutive elements, so that loads and stores
        // are coalesced; a row's exclusive scan plus the rows before it give the rank.
        uint32_t v[PER], sum = 0;
        const uint32_t i0 = t * tile_len + threadIdx.x, i1 = t * tile_len + PER * threadIdx.x;
        uint32_t total = 0, ex = 0;
        if (chunks == 1u) {
            #pragma unroll
            for (int k = 0; k < PER; ++k) {
                v[k] = i1 + k < len ? scan_in<MODE>(w, i1 + k) : 0u;
                sum += v[k];
            }
            ex = block_excl_scan_256(sum, sh, total);
        } else {
            // First pass: the tile's total; the second (below) reads the runs again.
            for (uint32_t c = 0; c < chunks; ++c) {
                const uint32_t ic = i0 + c * SCAN_TILE;
                #pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const uint32_t i = ic + 256u * k;
                    sum += i < len ? scan_in<MODE>(w, i) : 0u;
                }
            }
            total = block_reduce_256(sum, sh);
        }
        if (threadIdx.x < WAVE) {
            // Look-back by the first wave, 64 predecessors at a time: lane l reads tile
            // hi - l; the nearest tile with an inclusive prefix ends the walk once every
            // tile up to it has published.
            const uint32_t l = threadIdx.x;
            uint32_t prefix = 0;
            if (t == 0) {
                if (l == 0) __hip_atomic_store(&state[0], ST_PREFIX | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            } else {
                if (l == 0) __hip_atomic_store(&state[t], ST_AGG | total, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                int64_t hi = (int64_t)t - 1;
                for (;;) {
                    const int64_t j = hi - (int64_t)l;
                    // Below tile 0 (never needed: tile 0 carries a prefix) read as prefix 0.
                    const unsigned long long st =
                        j >= 0 ? __hip_atomic_load(&state[j], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) : ST_PREFIX;
                    const unsigned long long flag = st & (3ull << 32);
                    const uint64_t pmask = __ballot(flag == ST_PREFIX);
                    const uint32_t stop = pmask ? (uint32_t)__ffsll((unsigned long long)pmask) - 1u : 64u;
                    const uint64_t need = stop >= 63u ? ~0ull : ((2ull << stop) - 1ull);
                    if (__ballot(flag == 0ull) & need) continue;   // a tile in range has not published yet
                    uint32_t val = l <= stop ? (uint32_t)st : 0u;
                    for (int o = 32; o > 0; o >>= 1) val += __shfl_xor(val, o, WAVE);
                    prefix += val;
                    if (stop < 64u) break;
                    hi -= 64;                                         // 64 aggregates: walk further back
                }
                if (l == 0)
                    __hip_atomic_store(&state[t], ST_PREFIX | (unsigned long long)(prefix + total), __ATOMIC_RELAXED,
                                       __HIP_MEMORY_SCOPE_AGENT);
            }
            if (l == 0) {
                s_prefix = prefix;
                if (t == ntiles - 1) {
                    if (MODE == MODE_SLOTS) w.g->n_blocks = min(prefix + total, w.b_total);
                    else w.g->p_out = prefix + total;
                }
            }
        }
        __syncthreads();
        if (chunks == 1u) {
            uint32_t r = s_prefix + ex;
            for (int k = 0; k < PER; ++k) {
                if (i1 + k < len) scan_out<MODE>(w, i1 + k, r, v[k]);
                r += v[k];
            }
        } else {
            uint32_t run = s_prefix;
            for (uint32_t c = 0; c < chunks; ++c) {
                const uint32_t ic = i0 + c * SCAN_TILE;
                #pragma unroll
                for (int k = 0; k < PER; ++k) {
                    const uint32_t i = ic + 256u * k;
                    v[k] = i < len ? scan_in<MODE>(w, i) : 0u;
                }
                #pragma unroll
                for (int k = 0; k < PER; ++k) {
                    uint32_t rt;
                    const uint32_t e = block_excl_scan_256(v[k], sh, rt);
                    const uint32_t i = ic + 256u * k;
                    if (i < len) scan_out<MODE>(w, i, run + e, v[k]);
                    run += rt;
                }
            }
        }
    }
}

// Per block, 16 threads: threads 0..7 look up the 8 blocks vid - {0,1}^3 (grid update
// sources), threads 8..15 the 8 blocks vid + {0,1}^3 (G2P tile, primaries only); thread j
// also tests collider j against the block (bounding spheres) for the collider mask.
__device__ inline void block_info_part(const WorldDev &w)
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
        if (j == 0 && w.breach) w.breach[d] = 0u;   // set by k_grid_cdf (CPIC worlds)
        const int4 v = w.bvid[d];
        if (v.x == OVERFLOW_VID) continue;
        const SceneDev sc = w.scenes[w.bscene[d]];
        if (j < 8) {
            // Sources of the grid update: only primary blocks have a P2G staging tile.
            const uint32_t src = j == 0 ? d : find_block(w, sc, v.x - ASSOC[j][0], v.y - ASSOC[j][1], v.z - ASSOC[j][2]);
            w.bdn[8 * d + j] = (src != NONE && w.bprim[src]) ? src : NONE;
        } else if (w.bprim[d]) {
            const uint32_t i = j - 8;
            w.bup[8 * d + i] = i == 0 ? d : find_block(w, sc, v.x + ASSOC[i][0], v.y + ASSOC[i][1], v.z + ASSOC[i][2]);
        }
        // Collider mask: collider j can touch a node of this block only if its bounding
        // sphere (radius reach about the body origin) comes within 1.5 h (the projection
        // reach) of the block's node sphere.
        bool hit = false;
        if (j < sc.col_count) {
            const ColliderDev &c = w.cols[sc.col_start + j];
            const float h = sc.h;
            const f3 center = mk3(((float)(4 * v.x) + 1.5f) * h, ((float)(4 * v.y) + 1.5f) * h, ((float)(4 * v.z) + 1.5f) * h);
            const float rb = 1.5f * h * 1.7320509f;                  // node sphere of the block
            const float rc = c.reach;
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

__device__ inline void count_part(const WorldDev &w)
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

// Block tables and particle counts, both needing only the dense block ids.
__global__ void k_block_count(WorldDev w)
{
    count_part(w);
    block_info_part(w);
}

// P2G work split: workgroup g of the P2G grid takes blocks [p2g_start[g], p2g_start[g+1]),
// so that every workgroup gets about the same cost, counted as its particles plus
// w.p2g_block_cost per block (the tile clear and flush).  cost(d) = first particle of
// block d + p2g_block_cost d is non-decreasing; p2g_start[g] = min{d : cost(d) G >= g T},
// T the total cost, G the number of workgroups.

__device__ inline void p2g_split(const WorldDev &w)
{
    const uint32_t nb = w.g->n_blocks, G = w.p2g_groups;
    if (nb == 0u || G == 0u) return;
    const uint64_t K = w.p2g_block_cost;
    const uint64_t T = (uint64_t)w.g->p_out + K * nb;
    for (uint32_t d = blockIdx.x * blockDim.x + threadIdx.x; d < nb; d += gridDim.x * blockDim.x) {
        const uint64_t a = ((uint64_t)w.bfirst[d] + K * d) * G;
        const uint64_t g_hi = a / T;                                  // last g with start d
        uint64_t g_lo = 0;                                            // first g with start d
        if (d > 0) g_lo = (((uint64_t)w.bfirst[d - 1] + K * (d - 1)) * G) / T + 1;
        for (uint64_t g = g_lo; g <= g_hi && g < G; ++g) w.p2g_start[g] = d;
        if (d == nb - 1)
            for (uint64_t g = g_hi + 1; g <= G; ++g) w.p2g_start[g] = nb;
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
    p2g_split(w);
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

// 64-bit fixed-point accumulator words of the P2G tile.
struct Fixed {
    using word = unsigned long long;
    static constexpr int BITS = 62;
    // r = v * scale as a two's-complement 64-bit integer hi 2^32 + lo, for |r| < 2^63:
    // hi = floor(r / 2^32) is exact; lo = r - hi 2^32 lies in [0, 2^32) and is rounded to
    // float and truncated to an integer.  It is exact when hi >= 0 and |r| >= 2^24 or when
    // r is a non-negative integer below 2^24; otherwise (a fractional part, or hi = -1 with
    // lo close to 2^32) it is off by less than 2^8 units, against 2^62 / np units for the
    // block's largest contribution.  The result depends only on v and the scale,
    // so integer sums stay independent of the order of the adds.  Non-finite inputs give
    // garbage that the caller's non-finite mask overrides.
    __device__ static word to(float v, float scale)
    {
        const float r = v * scale;
        const float hi = floorf(r * 0x1p-32f);
        const float lo = __builtin_fmaf(hi, -0x1p32f, r);
        const uint32_t uhi = (uint32_t)(int32_t)hi;
        const uint32_t ulo = (uint32_t)lo;
        return ((unsigned long long)uhi << 32) | ulo;
    }
    // scale is a power of two: multiplying by its reciprocal is exact
    __device__ static float back(word x, double inv_scale) { return (float)((double)(long long)x * inv_scale); }
};

__device__ inline float fixed_scale(int bits, int e, uint32_t np)
{
    int lg = 0;
    while ((1u << lg) < np && lg < 31) ++lg;
    int k = bits - e - lg;
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


__device__ inline float wave_max(float v)
{
    for (int o = 32; o > 0; o >>= 1) v = fmaxf(v, __shfl_xor(v, o, WAVE));
    return v;
}

// Scatter of one particle's 27 contributions into the block's fixed-point tile.  Lanes
// start the stencil at different z slabs (lane mod 3): particles of one cell in one wave
// then add to different nodes at the same instruction instead of serializing on one LDS
// address; the integer sums do not depend on the order.  CHECK adds the non-finite mask
// (only for suspect particles, so the common path carries none of the tests).
// Words of the non-finite mask (one bit per tile word); the word after them flags any.
// With CPIC the tile holds a second (incompatible) momentum and mass after the first.
// The tile holds 4 fixed-point fields per node (momentum, mass).  CPIC blocks with
// particles that carry affinity bits fill it twice: the compatible contributions, flushed
// to the staging tiles, then the incompatible ones, flushed to the second staging tiles.
template <bool CPIC> constexpr int tile_fields() { return 4; }
template <bool CPIC> constexpr int bad_any() { return (tile_fields<CPIC>() * TILE_NODES + 31) / 32; }

// Which contributions a scatter pass adds: all (no affinity in the block), the compatible
// ones, or the incompatible ones.
constexpr int PART_ALL = 0, PART_COMPATIBLE = 1, PART_INCOMPATIBLE = 2;

// K lanes share one particle: K = 1 scatters all 27 stencil nodes, starting at z slab
// sel (the lane's rotation, so that particles of one cell in one wave add to different
// nodes at the same instruction); K = 3 the 9 nodes of z slab sel; K = 9 the 3 nodes of
// row (sz, sy) = (sel / 3, sel % 3); K = 27 node (sel / 9, sel / 3 % 3, sel % 3).  Every
// node's contribution is computed by the same operations whatever K is.
template <bool CHECK, bool CPIC, int PART, int K>
__device__ __forceinline__ void p2g_scatter(unsigned long long *tile, uint32_t *bad, int bx, int by, int bz, int sel,
                                            const f3 base, const f3 hx, const f3 hy, const f3 hz, const float (&wx)[3],
                                            const float (&wy)[3], const float (&wz)[3], float smom, float msm,
                                            const uint32_t *naff, uint32_t pa)
{
    constexpr int NZ = K == 1 ? 3 : 1, NY = K <= 3 ? 3 : 1, NX = K <= 9 ? 3 : 1;
    #pragma unroll
    for (int kz = 0; kz < NZ; ++kz) {
        int sz;
        if (K == 1) sz = kz + sel >= 3 ? kz + sel - 3 : kz + sel;
        else if (K == 3) sz = sel;
        else if (K == 9) sz = sel / 3;
        else sz = sel / 9;
        // base + sz hz, exactly as base + hz or base + 2 hz (sz hz is exact).
        const float fsz = (float)sz;
        const f3 az = mk3(__builtin_fmaf(hz.x, fsz, base.x), __builtin_fmaf(hz.y, fsz, base.y),
                          __builtin_fmaf(hz.z, fsz, base.z));
        float wzs = wz[0];
        wzs = sz == 1 ? wz[1] : wzs;
        wzs = sz == 2 ? wz[2] : wzs;
        #pragma unroll
        for (int ky = 0; ky < NY; ++ky) {
            const int sy = NY == 3 ? ky : (K == 9 ? sel % 3 : (sel / 3) % 3);
            f3 ay;
            float wys;
            if (NY == 3) {
                ay = sy == 0 ? az : (sy == 1 ? az + hy : az + hy * 2.0f);
                wys = wy[sy];
            } else {
                const f3 a1 = az + hy, a2 = az + hy * 2.0f;
                ay = sy == 0 ? az : (sy == 1 ? a1 : a2);
                wys = sy == 0 ? wy[0] : (sy == 1 ? wy[1] : wy[2]);
            }
            #pragma unroll
            for (int kx = 0; kx < NX; ++kx) {
                const int sx = NX == 3 ? kx : sel % 3;
                f3 a;
                float wxs;
                if (NX == 3) {
                    a = sx == 0 ? ay : (sx == 1 ? ay + hx : ay + hx * 2.0f);
                    wxs = wx[sx];
                } else {
                    const f3 a1 = ay + hx, a2 = ay + hx * 2.0f;
                    a = sx == 0 ? ay : (sx == 1 ? a1 : a2);
                    wxs = sx == 0 ? wx[0] : (sx == 1 ? wx[1] : wx[2]);
                }
                const float wgt = (wxs * wys) * wzs;
                const f3 vc = a * (wgt * smom);
                const float mw = msm * wgt;
                const int tn = (bx + sx) + TILE * (by + sy) + TILE * TILE * (bz + sz);
                // CPIC (p2g.rs:268-302): a particle not compatible with the node adds to
                // the node's incompatible fields (the second pass of the block).
                if (PART != PART_ALL) {
                    const bool incomp = (pa & 0xFFFFu) != 0u && !aff_compatible(pa, naff[tn]);
                    if (incomp != (PART == PART_INCOMPATIBLE)) continue;
                }
                const int t = tn;
                atomicAdd(&tile[t], Fixed::to(vc.x, 1.0f));
                atomicAdd(&tile[TILE_NODES + t], Fixed::to(vc.y, 1.0f));
                atomicAdd(&tile[2 * TILE_NODES + t], Fixed::to(vc.z, 1.0f));
                atomicAdd(&tile[3 * TILE_NODES + t], Fixed::to(mw, 1.0f));
                if (CHECK && (!isfinite(vc.x) || !isfinite(vc.y) || !isfinite(vc.z) || !isfinite(mw))) {
                    atomicOr(&bad[bad_any<CPIC>()], 1u);
                    atomicOr(&bad[t >> 5], 1u << (t & 31));
                    atomicOr(&bad[(TILE_NODES + t) >> 5], 1u << ((TILE_NODES + t) & 31));
                    atomicOr(&bad[(2 * TILE_NODES + t) >> 5], 1u << ((2 * TILE_NODES + t) & 31));
                    atomicOr(&bad[(3 * TILE_NODES + t) >> 5], 1u << ((3 * TILE_NODES + t) & 31));
                }
            }
        }
    }
}

// Flush: after the block's scatter pass, node t of the 6^3 tile goes to the region of the
// block it belongs to (offset e = (tx >= 4, ty >= 4, tz >= 4)), stored contiguously in that
// block's node order in the staging array st, so the consumer reads each region as one run
// of float4.  fdst, the kernel's LDS copy of FLUSH_CODE, gives each tile node's region e
// (bits 8..10) and its offset in the block's 216-node staging slot (bits 0..7).  Ends with
// the tile free for the next pass.
struct alignas(8) FlushCodes {
    uint16_t c[TILE_NODES];
};
constexpr FlushCodes flush_codes()
{
    FlushCodes f{};
    const int off[8] = {0, 64, 96, 128, 144, 176, 192, 208};   // REGION_OFFSET
    for (int t = 0; t < TILE_NODES; ++t) {
        const int tx = t % TILE, ty = (t / TILE) % TILE, tz = t / (TILE * TILE);
        const int ex = tx >> 2, ey = ty >> 2, ez = tz >> 2;
        const int e = 4 * ex + 2 * ey + ez;
        const int nx = ex ? 2 : 4, ny = ey ? 2 : 4;
        const int idx = (tx - 4 * ex) + nx * ((ty - 4 * ey) + ny * (tz - 4 * ez));
        f.c[t] = (uint16_t)((e << 8) | (off[e] + idx));
    }
    return f;
}
__device__ __constant__ const FlushCodes FLUSH_CODE = flush_codes();

// Copies FLUSH_CODE into the kernel's LDS table (54 lanes, 8 bytes each).
__device__ inline void load_flush_table(uint16_t *fdst, uint32_t l)
{
    static_assert(TILE_NODES * sizeof(uint16_t) == 54 * sizeof(uint64_t), "216 codes in 54 words");
    if (l < 54u) reinterpret_cast<uint64_t *>(fdst)[l] = reinterpret_cast<const uint64_t *>(FLUSH_CODE.c)[l];
}

// With zero (a block whose particles do not reach every region), the flush also zeroes the
// words it reads, which are all the words the pass added to: the tile is then clean for the
// next pass without a full clear.
__device__ __forceinline__ void p2g_flush(unsigned long long *tile, const uint32_t *bad, float4 *stage,
                                          uint32_t d, float smom, float sm, uint32_t l, uint32_t touched,
                                          const uint16_t *fdst, bool zero)
{
    __syncthreads();
    const double inv_smom = 1.0 / (double)smom, inv_sm = 1.0 / (double)sm;   // exact: powers of two
    const bool any_bad = bad[bad_any<false>()] != 0u;
    float4 *st = stage + (size_t)d * TILE_NODES;
    for (int t = (int)l; t < TILE_NODES; t += WAVE) {
        const uint32_t code = fdst[t];
        if (!((touched >> (code >> 8)) & 1u)) continue;   // no particle's stencil reaches region e
        float v4[4];
        #pragma unroll
        for (int c = 0; c < 4; ++c) v4[c] = Fixed::back(tile[c * TILE_NODES + t], c < 3 ? inv_smom : inv_sm);
        if (zero) {
            #pragma unroll
            for (int c = 0; c < 4; ++c) tile[c * TILE_NODES + t] = 0ull;
        }
        if (any_bad) {
            #pragma unroll
            for (int c = 0; c < 4; ++c) {
                const int q = c * TILE_NODES + t;
                if (bad[q >> 5] & (1u << (q & 31))) v4[c] = __builtin_nanf("");
            }
        }
        st[code & 0xFFu] = make_float4(v4[0], v4[1], v4[2], v4[3]);
    }
    __syncthreads();
}

// Empties the tile (unless the last flush left it clean) and its non-finite marks for the
// next scatter pass.  Blocks that reach all 8 regions (dense ones) clear here: one tight
// loop costs them less than zeroing inside the flush.
__device__ __forceinline__ void p2g_clear(unsigned long long *tile, uint32_t *bad, uint32_t l, bool dirty)
{
    if (dirty)
        for (int k = l; k < 4 * TILE_NODES; k += WAVE) tile[k] = 0;
    if (l <= (uint32_t)bad_any<false>()) bad[l] = 0u;
}

// One particle's P2G inputs.
struct P2GIn {
    float px, py, pz, m;
    f3 v;
    m3 C;
};

__device__ inline P2GIn p2g_load(const WorldDev &w, uint32_t slot)
{
    P2GIn p;
    p.px = w.in.plane(PX)[slot]; p.py = w.in.plane(PY)[slot]; p.pz = w.in.plane(PZ)[slot];
    p.m = w.in.plane(PMASS)[slot];
    p.v = mk3(w.in.plane(PVX)[slot], w.in.plane(PVY)[slot], w.in.plane(PVZ)[slot]);
    p.C.c0 = mk3(w.in.plane(PC0)[slot], w.in.plane(PC1)[slot], w.in.plane(PC2)[slot]);
    p.C.c1 = mk3(w.in.plane(PC3)[slot], w.in.plane(PC4)[slot], w.in.plane(PC5)[slot]);
    p.C.c2 = mk3(w.in.plane(PC6)[slot], w.in.plane(PC7)[slot], w.in.plane(PC8)[slot]);
    return p;
}

// Bounds of |m v + C dpt| (|dpt_j| <= 1.5 h) and of m for one particle.
__device__ inline void p2g_bound(const P2GIn &p, float h, float &bmom, float &bm)
{
    float cmax = fmaxf(fmaxf(fabsf(p.C.c0.x), fmaxf(fabsf(p.C.c0.y), fabsf(p.C.c0.z))),
                       fmaxf(fmaxf(fabsf(p.C.c1.x), fmaxf(fabsf(p.C.c1.y), fabsf(p.C.c1.z))),
                             fmaxf(fabsf(p.C.c2.x), fmaxf(fabsf(p.C.c2.y), fabsf(p.C.c2.z)))));
    const float vmax = fmaxf(fabsf(p.v.x), fmaxf(fabsf(p.v.y), fabsf(p.v.z)));
    bmom = fmaxf(bmom, fabsf(p.m) * vmax + 4.5f * h * cmax);
    bm = fmaxf(bm, fabsf(p.m));
}

// Staging regions of a block's tile that a particle with base cell (bx, by, bz) in the
// block reaches: region e = 4 ex + 2 ey + ez (the part of the +e neighbour in the tile)
// needs base cell coordinates >= 2 on the axes where e is 1 (the reference's spill rule,
// grid/sort.rs:322-328).
__device__ inline uint32_t spill_mask(int bx, int by, int bz)
{
    const bool sx = bx >= 2, sy = by >= 2, sz = bz >= 2;
    uint32_t m = 1u;
    if (sz) m |= 1u << 1;
    if (sy) m |= 1u << 2;
    if (sy && sz) m |= 1u << 3;
    if (sx) m |= 1u << 4;
    if (sx && sz) m |= 1u << 5;
    if (sx && sy) m |= 1u << 6;
    if (sx && sy && sz) m |= 1u << 7;
    return m;
}

// Scatter of one enabled particle of block vid into the tile.
template <bool CPIC, int PART, int K = 1>
__device__ __forceinline__ uint32_t p2g_one(unsigned long long *tile, uint32_t *bad, const P2GIn &p, const int4 v,
                                        const SceneDev &sc, float smom, float sm, float bmom, float bm, uint32_t sel,
                                        const uint32_t *naff, uint32_t pa)
{
    const float h = sc.h;
    const float px = p.px, py = p.py, pz = p.pz, m = p.m;
    const f3 mom = p.v * m;
    const m3 &C = p.C;
    // With finite inputs and a finite block bound far below FLT_MAX every contribution is
    // finite (|contribution| <= bound), so only suspect particles check each contribution
    // for the non-finite mask.
    float amax = fmaxf(fabsf(px), fmaxf(fabsf(py), fabsf(pz)));
    amax = fmaxf(amax, fmaxf(fabsf(mom.x), fmaxf(fabsf(mom.y), fabsf(mom.z))));
    amax = fmaxf(amax, fabsf(m));
    #pragma unroll
    for (int k = 0; k < 3; ++k) {
        const f3 col = k == 0 ? C.c0 : (k == 1 ? C.c1 : C.c2);
        amax = fmaxf(amax, fmaxf(fabsf(col.x), fmaxf(fabsf(col.y), fabsf(col.z))));
    }
    // fmaxf drops NaN operands: test them separately.
    const bool any_nan = px != px || py != py || pz != pz || mom.x != mom.x || mom.y != mom.y ||
                         mom.z != mom.z || m != m || C.c0.x != C.c0.x || C.c0.y != C.c0.y ||
                         C.c0.z != C.c0.z || C.c1.x != C.c1.x || C.c1.y != C.c1.y ||
                         C.c1.z != C.c1.z || C.c2.x != C.c2.x || C.c2.y != C.c2.y || C.c2.z != C.c2.z;
    const bool suspect = any_nan || !(amax < 0x1p100f) || !(bmom < 0x1p100f) || !(bm < 0x1p100f);
    // Base node of the stencil and its offset from the particle.
    const float ih = sc.inv_h;
    const float cx = round_div(px, h, ih) - 1.0f, cy = round_div(py, h, ih) - 1.0f, cz = round_div(pz, h, ih) - 1.0f;
    const int bx = (int)cx - 4 * v.x, by = (int)cy - 4 * v.y, bz = (int)cz - 4 * v.z;
    // A particle's primary block holds its base cell: bx, by, bz are in [0, 3] for every
    // finite position; anything else contributes nothing.
    const bool inside = (unsigned)bx < 4u && (unsigned)by < 4u && (unsigned)bz < 4u;
    const f3 d0 = mk3(cx * h - px, cy * h - py, cz * h - pz);
    float wx[3], wy[3], wz[3];
    {
        const float fx = -d0.x * ih, fy = -d0.y * ih, fz = -d0.z * ih;
        wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx); wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f); wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
        wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy); wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f); wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
        wz[0] = 0.5f * (1.5f - fz) * (1.5f - fz); wz[1] = 0.75f - (fz - 1.0f) * (fz - 1.0f); wz[2] = 0.5f * (fz - 0.5f) * (fz - 0.5f);
    }
    // C dpt + m v at node (sx, sy, sz) = (C d0 + m v) + sx (h C.c0) + sy (h C.c1) + sz (h C.c2):
    // three adds per node instead of a matrix product.  The fixed-point scales are powers of
    // two, so folding them into the weight changes no bits.
    const f3 base = mul(C, d0) + mom;
    const f3 hx = C.c0 * h, hy = C.c1 * h, hz = C.c2 * h;
    const float msm = m * sm;
    if (!inside) return 0u;
    if (suspect)
        p2g_scatter<true, CPIC, PART, K>(tile, bad, bx, by, bz, (int)sel, base, hx, hy, hz, wx, wy, wz, smom, msm, naff, pa);
    else
        p2g_scatter<false, CPIC, PART, K>(tile, bad, bx, by, bz, (int)sel, base, hx, hy, hz, wx, wy, wz, smom, msm, naff, pa);
    return spill_mask(bx, by, bz);
}

// Scatter of the particle at sorted position pos (when enabled; the incompatible pass
// takes only particles with affinity bits).
template <bool CPIC, int PART, int K>
__device__ __forceinline__ uint32_t p2g_pos(const WorldDev &w, unsigned long long *tile, uint32_t *bad, uint32_t pos,
                                            const int4 v, const SceneDev &sc, float smom, float sm, float bmom, float bm,
                                            uint32_t sel, const uint32_t *naff)
{
    const uint32_t slot = w.perm[pos];
    if (!(w.in.uplane(PFLAGS)[slot] & FLAG_ENABLED)) return 0u;
    const uint32_t pa = PART == PART_ALL ? 0u : w.out.uplane(PAFF)[pos];
    if (PART == PART_INCOMPATIBLE && (pa & 0xFFFFu) == 0u) return 0u;
    return p2g_one<CPIC, PART, K>(tile, bad, p2g_load(w, slot), v, sc, smom, sm, bmom, bm, sel, naff, pa);
}

// One scatter pass over the np particles of a block starting at sorted position first:
// whole chunks of 64 one particle per lane; the remainder r with K lanes per particle,
// K = 27, 9, 3 or 1 for r <= 2, 7, 21 or more, so that a small remainder costs a fraction
// of a full 27-node pass.
template <bool CPIC, int PART>
__device__ __forceinline__ uint32_t p2g_block_pass(const WorldDev &w, unsigned long long *tile, uint32_t *bad,
                                                   uint32_t first, uint32_t np, const int4 v, const SceneDev &sc,
                                                   float smom, float sm, float bmom, float bm, uint32_t l,
                                                   const uint32_t *naff)
{
    const uint32_t full = np / WAVE * WAVE, rem = np - full;
    uint32_t spill = 0u;
    for (uint32_t c = 0; c < full; c += WAVE)
        spill |= p2g_pos<CPIC, PART, 1>(w, tile, bad, first + c + l, v, sc, smom, sm, bmom, bm, l % 3u, naff);
    if (rem == 0u) return spill;
    const uint32_t at = first + full;
    if (rem <= 2u) {
        if (l < 27u * rem)
            spill |= p2g_pos<CPIC, PART, 27>(w, tile, bad, at + l / 27u, v, sc, smom, sm, bmom, bm, l % 27u, naff);
    } else if (rem <= 7u) {
        if (l < 9u * rem)
            spill |= p2g_pos<CPIC, PART, 9>(w, tile, bad, at + l / 9u, v, sc, smom, sm, bmom, bm, l % 9u, naff);
    } else if (rem <= 21u) {
        if (l < 3u * rem)
            spill |= p2g_pos<CPIC, PART, 3>(w, tile, bad, at + l / 3u, v, sc, smom, sm, bmom, bm, l % 3u, naff);
    } else if (l < rem) {
        spill |= p2g_pos<CPIC, PART, 1>(w, tile, bad, at + l, v, sc, smom, sm, bmom, bm, l % 3u, naff);
    }
    return spill;
}

// Node affinity words of block d's 6^3 tile (CPIC scenes; zero, compatible with every
// particle, elsewhere and for missing blocks).
__device__ inline void load_tile_affinity(const WorldDev &w, uint32_t d, bool cpic, uint32_t *naff, uint32_t l)
{
    for (int t = l; t < TILE_NODES; t += WAVE) {
        const int tx = t % TILE, ty = (t / TILE) % TILE, tz = t / (TILE * TILE);
        const int e = 4 * (tx >> 2) + 2 * (ty >> 2) + (tz >> 2);
        const uint32_t blk = cpic ? w.bup[8 * (size_t)d + e] : NONE;
        naff[t] = blk == NONE ? 0u : w.ncdf[(size_t)blk * NODES_PER_BLOCK + (tx & 3) + 4 * (ty & 3) + 16 * (tz & 3)].y;
    }
}

// The union over the wave of the lanes' region masks.
__device__ inline uint32_t wave_or_mask8(uint32_t m)
{
    uint32_t out = 0u;
    #pragma unroll
    for (int e = 0; e < 8; ++e) out |= __ballot((m >> e) & 1u) ? (1u << e) : 0u;
    return out;
}

// Regions of block d that no particle reached are neither flushed nor read: the block
// receiving region e (bup[d][e]) stops listing d as its source e (bdn), so the grid update
// skips it.  Only block d writes that entry, and the grid update reads it after P2G.
__device__ inline void unlink_untouched(const WorldDev &w, uint32_t d, uint32_t touched, uint32_t l)
{
    if (l < 8u && !((touched >> l) & 1u)) {
        const uint32_t dst = w.bup[8 * (size_t)d + l];
        if (dst != NONE) w.bdn[8 * (size_t)dst + l] = NONE;
    }
}

template <bool CPIC>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(4))) k_p2g(WorldDev w)
{
    using FX = Fixed;
    constexpr int NF = tile_fields<CPIC>();
    __shared__ typename FX::word tile[NF * TILE_NODES];
    __shared__ uint32_t bad[bad_any<CPIC>() + 1];   // non-finite contributions; last word: any at all
    __shared__ uint32_t naff[CPIC ? TILE_NODES : 1];
    const uint32_t l = threadIdx.x;
    const uint32_t *FL = w.in.uplane(PFLAGS);
    __shared__ __align__(8) uint16_t fdst[TILE_NODES];
    load_flush_table(fdst, l);   // read after the first __syncthreads
    bool dirty = true;           // the tile may hold non-zero words (p2g_clear, p2g_flush)
    // Cost-balanced block range, XCD-contiguous (see p2g_split and xcd_range).
    const uint32_t G = gridDim.x, b = blockIdx.x;
    const uint32_t g = (G % NUM_XCD == 0) ? (b % NUM_XCD) * (G / NUM_XCD) + b / NUM_XCD : b;
    // Per-block words by scalar loads in the plain variant (the CPIC one spills more with them).
    const uint32_t dbegin = ld_maybe_uniform<!CPIC>(w.p2g_start + g), dend = ld_maybe_uniform<!CPIC>(w.p2g_start + g + 1);
    for (uint32_t d = dbegin; d < dend; ++d) {
        const uint32_t s = ld_maybe_uniform<!CPIC>(w.bscene + d);
        const uint32_t np = ld_maybe_uniform<!CPIC>(w.bnp + d);
        const int4 v = ld_maybe_uniform<!CPIC>(w.bvid + d);
        if (np == 0u || v.x == OVERFLOW_VID) continue;
        const SceneDev sc = w.scenes[s];
        const float h = sc.h;
        const uint32_t first = ld_maybe_uniform<!CPIC>(w.bfirst + d);
        float bmom = 0.0f, bm = 0.0f;
        if (np <= (uint32_t)WAVE && np > 21u) {
            // One particle per lane: load it once, bound, then scatter from registers.
            P2GIn p{};
            bool on = false;
            uint32_t pa = 0u, spill = 0u;
            if (l < np) {
                const uint32_t slot = w.perm[first + l];
                on = (FL[slot] & FLAG_ENABLED) != 0u;
                if (on) {
                    p = p2g_load(w, slot);
                    p2g_bound(p, h, bmom, bm);
                    if (CPIC && w.breach[d]) pa = w.out.uplane(PAFF)[first + l];
                }
            }
            bmom = wave_max(bmom);
            bm = wave_max(bm);
            const float smom = fixed_scale(FX::BITS, bound_exponent(bmom), np);
            const float sm = fixed_scale(FX::BITS, bound_exponent(bm), np);
            const bool inc = CPIC && __ballot((pa & 0xFFFFu) != 0u) != 0ull;
            p2g_clear(tile, bad, l, dirty);
            if (inc) load_tile_affinity(w, d, true, naff, l);
            __syncthreads();
            if (on) {
                if (inc) spill = p2g_one<CPIC, PART_COMPATIBLE>(tile, bad, p, v, sc, smom, sm, bmom, bm, l % 3u, naff, pa);
                else spill = p2g_one<CPIC, PART_ALL>(tile, bad, p, v, sc, smom, sm, bmom, bm, l % 3u, naff, pa);
            }
            const uint32_t touched = wave_or_mask8(spill);
            p2g_flush(tile, bad, reinterpret_cast<float4 *>(w.stage), d, smom, sm, l, touched, fdst, touched != 0xFFu);
            dirty = touched == 0xFFu;
            if (CPIC && l == 0) w.binc[d] = inc ? 1u : 0u;
            if (CPIC && inc) {
                // Second pass: the incompatible contributions of the particles with affinity.
                p2g_clear(tile, bad, l, dirty);
                __syncthreads();
                if (on && (pa & 0xFFFFu) != 0u)
                    p2g_one<CPIC, PART_INCOMPATIBLE>(tile, bad, p, v, sc, smom, sm, bmom, bm, l % 3u, naff, pa);
                p2g_flush(tile, bad, reinterpret_cast<float4 *>(w.stage_i), d, smom, sm, l, touched, fdst, touched != 0xFFu);
                dirty = touched == 0xFFu;
            }
            unlink_untouched(w, d, touched, l);
            continue;
        }
        // Pass 1: bounds over the block; pass 2: the scatter (p2g_block_pass).
        bool anyc = false;
        for (uint32_t c = 0; c < np; c += WAVE) {
            const uint32_t pos = first + c + l;
            if (pos < first + np) {
                const uint32_t slot = w.perm[pos];
                if (FL[slot] & FLAG_ENABLED) {
                    p2g_bound(p2g_load(w, slot), h, bmom, bm);
                    if (CPIC && w.breach[d]) anyc |= (w.out.uplane(PAFF)[pos] & 0xFFFFu) != 0u;
                }
            }
        }
        bmom = wave_max(bmom);
        bm = wave_max(bm);
        const float smom = fixed_scale(FX::BITS, bound_exponent(bmom), np);
        const float sm = fixed_scale(FX::BITS, bound_exponent(bm), np);
        const bool inc = CPIC && __ballot(anyc) != 0ull;
        p2g_clear(tile, bad, l, dirty);
        if (inc) load_tile_affinity(w, d, true, naff, l);
        __syncthreads();
        const uint32_t spill =
            inc ? p2g_block_pass<CPIC, PART_COMPATIBLE>(w, tile, bad, first, np, v, sc, smom, sm, bmom, bm, l, naff)
                : p2g_block_pass<CPIC, PART_ALL>(w, tile, bad, first, np, v, sc, smom, sm, bmom, bm, l, naff);
        const uint32_t touched = wave_or_mask8(spill);
        p2g_flush(tile, bad, reinterpret_cast<float4 *>(w.stage), d, smom, sm, l, touched, fdst, touched != 0xFFu);
        dirty = touched == 0xFFu;
        if (CPIC && l == 0) w.binc[d] = inc ? 1u : 0u;
        if (CPIC && inc) {
            // Second pass: the incompatible contributions of the particles with affinity.
            p2g_clear(tile, bad, l, dirty);
            __syncthreads();
            p2g_block_pass<CPIC, PART_INCOMPATIBLE>(w, tile, bad, first, np, v, sc, smom, sm, bmom, bm, l, naff);
            p2g_flush(tile, bad, reinterpret_cast<float4 *>(w.stage_i), d, smom, sm, l, touched, fdst, touched != 0xFFu);
            dirty = touched == 0xFFu;
        }
        unlink_untouched(w, d, touched, l);
    }
}

// Empties the hash regions (keys and primary marks) for the next substep's insertions.
__device__ inline void clear_hash(const WorldDev &w)
{
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    for (uint32_t i = gid; i < w.h_total; i += gsz) {
        w.hkeys[i] = EMPTY_KEY;
        w.hmark[i] = 0u;
    }
}

// ------------------------------------------------------------------------------------
// Grid update: gpu_grid_update + gpu_grid_update_collide (grid_update.rs:21-88,
// grid_update_collide.rs:105-183) on the summed tiles.  CPIC scenes update both velocity
// fields and skip the collider pass (solver/grid_update.rs:26-50).  MESH: the world has
// mesh colliders in scenes without CPIC, whose projection the collider pass then runs
// (its registers kept out of the common variant).
template <bool CPIC, bool MESH>
__global__ void __launch_bounds__(256) k_grid_update(WorldDev w)
{
    clear_hash(w);   // not read after the block tables (k_block_count, k_mesh_cdf) are built
    // Whole blocks per workgroup: a wave's 64 consecutive nodes are then one block, and the
    // per-block reads (scene, block id, tile sources, collider mask) are wave-uniform.
    uint32_t bbegin, bend;
    xcd_range(w.g->n_blocks, bbegin, bend);
    const uint32_t nend = bend * NODES_PER_BLOCK;
    for (uint32_t n = bbegin * NODES_PER_BLOCK + threadIdx.x; n < nend; n += blockDim.x) {
        const uint32_t d = __builtin_amdgcn_readfirstlane(n >> 6);
        const uint32_t s = ld_uniform(w.bscene + d);
        const int4 v = ld_uniform(w.bvid + d);
        if (v.x == OVERFLOW_VID) continue;
        const uint32_t li = n & 63u;
        const int i = li & 3, j = (li >> 2) & 3, k = li >> 4;
        const SceneDev sc = ld_uniform(w.scenes + s);
        const bool cp = CPIC && sc.cpic;
        float mx = 0.0f, my = 0.0f, mz = 0.0f, m = 0.0f;
        float ix = 0.0f, iy = 0.0f, iz = 0.0f, im = 0.0f;
        #pragma unroll
        for (int e = 0; e < 8; ++e) {
            const int ex = ASSOC[e][0], ey = ASSOC[e][1], ez = ASSOC[e][2];
            if ((ex && i > 1) || (ey && j > 1) || (ez && k > 1)) continue;
            const uint32_t src = ld_uniform(w.bdn + 8 * d + e);
            if (src == NONE) continue;
            const int nx = ex ? 2 : 4, ny = ey ? 2 : 4;
            const size_t at = (size_t)src * TILE_NODES + REGION_OFFSET[e] + i + nx * (j + ny * k);
            const float4 q = reinterpret_cast<const float4 *>(w.stage)[at];
            mx += q.x;
            my += q.y;
            mz += q.z;
            m += q.w;
            if (cp && ld_uniform(w.binc + src)) {
                const float4 qi = reinterpret_cast<const float4 *>(w.stage_i)[at];
                ix += qi.x;
                iy += qi.y;
                iz += qi.z;
                im += qi.w;
            }
        }
        if (w.dbg_mom) {
            w.dbg_mom[3 * (size_t)n] = mx;
            w.dbg_mom[3 * (size_t)n + 1] = my;
            w.dbg_mom[3 * (size_t)n + 2] = mz;
        }
        const float inv_mass = m > 0.0f ? 1.0f / m : 0.0f;
        const float mdt = m * sc.dt;
        f3 vel = mk3((mx + sc.gx * mdt) * inv_mass, (my + sc.gy * mdt) * inv_mass, (mz + sc.gz * mdt) * inv_mass);
        const float lim = sc.vlim;
        vel.x = fminf(fmaxf(vel.x, -lim), lim);
        vel.y = fminf(fmaxf(vel.y, -lim), lim);
        vel.z = fminf(fmaxf(vel.z, -lim), lim);
        if (cp) {
            const float inv_mi = im > 0.0f ? 1.0f / im : 0.0f;
            const float midt = im * sc.dt;
            f3 vi = mk3((ix + sc.gx * midt) * inv_mi, (iy + sc.gy * midt) * inv_mi, (iz + sc.gz * midt) * inv_mi);
            vi.x = fminf(fmaxf(vi.x, -lim), lim);
            vi.y = fminf(fmaxf(vi.y, -lim), lim);
            vi.z = fminf(fmaxf(vi.z, -lim), lim);
            w.gridv_i[n] = make_float4(vi.x, vi.y, vi.z, im);
        }
        const uint32_t mask = cp ? 0u : ld_uniform(w.bmask + d);
        if (mask) {
            const f3 pt = mk3((float)(v.x * 4 + i), (float)(v.y * 4 + j), (float)(v.z * 4 + k)) * sc.h;
            const MeshSetDev ms{w.mesh_v, w.mesh_t, w.mesh_vn, w.mesh_en, w.mesh_grid, w.mesh_cell, w.mesh_ctri, w.col_mesh};
            vel = collide_node<MESH>(w.cols + sc.col_start, sc.col_count, mask, sc.h, sc.dt, pt, vel, ms, sc.col_start);
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

// Node velocities of the particle's 6^3 stencil tile, read from the grid (the 8 blocks
// vid + {0,1}^3; missing blocks read as zero).
struct GatherGrid {
    const uint32_t *bup;
    const float4 *gridv;
    const float4 *gridv_i;                     // CPIC: incompatible velocities
    const uint4 *ncdf;                         // CPIC: node contact distance fields
    uint32_t n0, n1, n2, n3, n4, n5, n6, n7;   // vid + {0,1}^3 in ASSOC order (4 ox + 2 oy + oz)
    __device__ void prepare(uint32_t d)
    {
        const uint32_t *p = bup + 8 * (size_t)d;
        n0 = p[0]; n1 = p[1]; n2 = p[2]; n3 = p[3]; n4 = p[4]; n5 = p[5]; n6 = p[6]; n7 = p[7];
    }
    // Grid node of tile position (tx, ty, tz), NONE in a missing block.
    __device__ uint32_t index(int tx, int ty, int tz) const
    {
        // block of the node by selects on the three offset bits (no indexed register array)
        const bool ox = tx >= 4, oy = ty >= 4, oz = tz >= 4;
        const uint32_t z0 = oz ? n1 : n0, z1 = oz ? n3 : n2, z2 = oz ? n5 : n4, z3 = oz ? n7 : n6;
        const uint32_t y0 = oy ? z1 : z0, y1 = oy ? z3 : z2;
        const uint32_t blk = ox ? y1 : y0;
        if (blk == NONE) return NONE;
        return blk * NODES_PER_BLOCK + (tx & 3) + 4 * (ty & 3) + 16 * (tz & 3);
    }
    // Without branches: a missing block's node reads node 0 (always allocated) and is
    // replaced by zero, so the 27 reads of a particle go out back to back.
    __device__ f3 at(int tx, int ty, int tz) const
    {
        const bool ox = tx >= 4, oy = ty >= 4, oz = tz >= 4;
        const uint32_t z0 = oz ? n1 : n0, z1 = oz ? n3 : n2, z2 = oz ? n5 : n4, z3 = oz ? n7 : n6;
        const uint32_t y0 = oy ? z1 : z0, y1 = oy ? z3 : z2;
        const uint32_t blk = ox ? y1 : y0;
        const bool none = blk == NONE;
        const uint32_t i = none ? 0u : blk * NODES_PER_BLOCK + (tx & 3) + 4 * (ty & 3) + 16 * (tz & 3);
        const float4 g = gridv[i];
        return none ? mk3(0, 0, 0) : mk3(g.x, g.y, g.z);
    }
};

// Returns the position and flags it wrote (xo, flo), for the caller's key insertion.
template <class Gather, bool CPIC>
__device__ __forceinline__ void g2p_particle(const WorldDev &w, uint32_t pos, uint32_t d, Gather &gather, f3 &xo,
                                             uint32_t &flo)
{
    {
        const uint32_t slot = w.perm[pos];
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
        // Per-particle plastic state: three words for the sand tags, one for snow.
        const bool sand = tag == 2u || tag == 3u, snow = tag == 5u;
        float ps0 = 0.0f, ps1 = 0.0f, ps2 = 0.0f;
        if (sand || snow) ps0 = w.in.plane(PS0)[slot];
        if (sand) { ps1 = w.in.plane(PS1)[slot]; ps2 = w.in.plane(PS2)[slot]; }
        // CPIC: this substep's particle contact field (written at the sorted position).  A
        // particle without affinity bits is compatible with every node and its contact
        // field is zero (a regular fit needs a node sharing an affinity), so it takes the
        // plain path.
        uint32_t pa = 0u;
        if (CPIC && sc.cpic) pa = w.out.uplane(PAFF)[pos];
        const bool cp = CPIC && (pa & 0xFFFFu) != 0u;
        f3 rigid_vel = mk3(0, 0, 0);
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
            const float ih = sc.inv_h;
            const float cx = round_div(px, h, ih) - 1.0f, cy = round_div(py, h, ih) - 1.0f,
                        cz = round_div(pz, h, ih) - 1.0f;
            const int bx = (int)cx - 4 * bv.x, by = (int)cy - 4 * bv.y, bz = (int)cz - 4 * bv.z;
            const bool inside = (unsigned)bx < 4u && (unsigned)by < 4u && (unsigned)bz < 4u;
            if ((fl & FLAG_ENABLED) && inside) {
                const f3 ref = mk3(cx * h - px, cy * h - py, cz * h - pz);
                float wx[3], wy[3], wz[3];
                {
                    const float fx = -ref.x * ih, fy = -ref.y * ih, fz = -ref.z * ih;
                    wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx); wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f); wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
                    wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy); wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f); wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
                    wz[0] = 0.5f * (1.5f - fz) * (1.5f - fz); wz[1] = 0.75f - (fz - 1.0f) * (fz - 1.0f); wz[2] = 0.5f * (fz - 0.5f) * (fz - 0.5f);
                }
                gather.prepare(d);
                // With dpt = ref + s h (s in {0,1,2}^3): sum w v dpt^T = V ref^T + h [Gx Gy Gz],
                // V = sum w v and Gj = sum w v s_j, so each node costs a product and a few
                // adds instead of an outer product; the velocity gradient and its trace
                // (the divergence) follow after the loop.
                f3 gx = mk3(0, 0, 0), gy = mk3(0, 0, 0), gz = mk3(0, 0, 0);
                if (cp) {
                    // g2p.rs:206-233: a node the particle is not compatible with gives its
                    // incompatible velocity, seen through its closest collider's boundary
                    // condition (friction scaled by the particle's factor) relative to the
                    // body's velocity there.  Fully unrolled: the contact particles are few,
                    // so their kernel runs about one wave per SIMD and needs every load in
                    // flight rather than occupancy.
                    const f3 pn = mk3(w.out.plane(PNX)[pos], w.out.plane(PNY)[pos], w.out.plane(PNZ)[pos]);
                    #pragma unroll
                    for (int q = 0; q < 27; ++q) {
                        const int sx = NBH[q][0], sy = NBH[q][1], sz = NBH[q][2];
                        const uint32_t ni = gather.index(bx + sx, by + sy, bz + sz);
                        f3 cv = mk3(0, 0, 0);
                        if (ni != NONE) {
                            const uint4 nc = gather.ncdf[ni];
                            const bool comp = aff_compatible(pa, nc.y);
                            const float4 g = comp ? gather.gridv[ni] : gather.gridv_i[ni];
                            cv = mk3(g.x, g.y, g.z);
                            if (!comp && nc.z != NONE) {
                                // the closest body's velocity at the node and its boundary
                                // condition {type in ncdf.w, friction}, stored by k_grid_cdf
                                const float4 nb = w.nbody[ni];
                                const f3 bpv = mk3(nb.x, nb.y, nb.z);
                                cv = bpv + bc_project(nc.w, nb.w * cls.bfric, cv - bpv, pn);
                            }
                        }
                        const float wsx = sx == 0 ? wx[0] : (sx == 1 ? wx[1] : wx[2]);
                        const float wsy = sy == 0 ? wy[0] : (sy == 1 ? wy[1] : wy[2]);
                        const float wsz = sz == 0 ? wz[0] : (sz == 1 ? wz[1] : wz[2]);
                        const f3 cw = cv * (wsx * wsy * wsz);
                        vel = vel + cw;
                        gx = gx + cw * (float)sx;
                        gy = gy + cw * (float)sy;
                        gz = gz + cw * (float)sz;
                    }
                } else {
                    #pragma unroll
                    for (int q = 0; q < 27; ++q) {
                        const int sx = NBH[q][0], sy = NBH[q][1], sz = NBH[q][2];
                        const f3 cv = gather.at(bx + sx, by + sy, bz + sz);
                        const float wgt = wx[sx] * wy[sy] * wz[sz];
                        const f3 cw = cv * wgt;
                        vel = vel + cw;
                        if (sx == 1) gx = gx + cw; else if (sx == 2) gx = gx + cw * 2.0f;
                        if (sy == 1) gy = gy + cw; else if (sy == 2) gy = gy + cw * 2.0f;
                        if (sz == 1) gz = gz + cw; else if (sz == 2) gz = gz + cw * 2.0f;
                    }
                }
                gradv.c0 = (vel * ref.x + gx * h) * inv_d;
                gradv.c1 = (vel * ref.y + gy * h) * inv_d;
                gradv.c2 = (vel * ref.z + gz * h) * inv_d;
                div = gradv.c0.x + gradv.c1.y + gradv.c2.z;
                if (cp) {
                    // g2p.rs:248-258: the bodies' velocities at the particle over its affinities
                    for (uint32_t m = pa & ((1u << sc.col_count) - 1u); m != 0u; m &= m - 1u)   // col_count <= 16
                        rigid_vel = rigid_vel + body_point_velocity(w.cols[sc.col_start + (uint32_t)__ffs(m) - 1u], mk3(px, py, pz));
                }
            }
            // Particle update.  CPIC: a particle deeper than 0.05 h behind a surface is
            // reprojected (slip relative to the bodies) and pushed out by a penalty
            // impulse after the position update (particle_update.rs:92-136).
            // (The contact field is re-read here so that it is not live across the node loop.)
            f3 pn = mk3(0, 0, 0);
            float psd = 0.0f;
            if (cp) {
                pn = mk3(w.out.plane(PNX)[pos], w.out.plane(PNY)[pos], w.out.plane(PNZ)[pos]);
                psd = w.out.plane(PSD)[pos];
            }
            const bool pen = cp && psd < -0.05f * h;
            if (pen) vel = rigid_vel + bc_project(1u, 0.0f, vel - rigid_vel, pn);
            const float vl = length(vel);
            if (vl > sc.vlim) vel = vel / vl * h / dt;
            vel = vel * rcp_hw(1.0f + cls.damping * dt);
            if (fl & FLAG_FIXED) vel = mk3(0, 0, 0);
            npos = mk3(px, py, pz) + vel * dt;
            if (pen) {
                const float corrected = fmaxf(psd, -0.3f * h);
                vel = vel + pn * (dt * -corrected * 1.0e3f);
            }
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
            case 5u: {
                m3 U, V;
                f3 sv;
                const float hard = snow_project(cls.w + 1, ps0, F, U, sv, V);
                tau = linear_stress_svd(cls.w[4] * hard, cls.w[5] * hard, U, sv);
                break;
            }
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
        if (sand || snow) w.out.plane(PS0)[pos] = ps0;
        if (sand) { w.out.plane(PS1)[pos] = ps1; w.out.plane(PS2)[pos] = ps2; }
        w.out.uplane(PFLAGS)[pos] = fl;
        w.out.uplane(PID)[pos] = id;
        xo = npos;
        flo = fl;
    }
}

// Body integration at the end of a substep (the reference's "Integrate bodies" pass,
// src_mpm/pipeline.rs:510-520, gpu_rigid_impulses_update in rigid_impulses.rs:66-140):
// MPM integrates its own copy of every coupled body (gpu_rigid_impulses_update,
// rigid_impulses.rs:66-140).  The particles' impulses on the bodies are compiled out in
// the reference (src_mpm_shaders/lib.rs:38), so the velocity is not changed by them and
// never capped; the pose advances about the world centre of mass c = pose * local_com:
// t' = c + exp(w dt) (t - c) + v dt, q' = normalize(exp(w dt) q) (Velocity::integrate,
// src_rbd_shaders/dynamics/body.rs:253-268, with glam's from_scaled_axis, quaternion
// product and normalize); then gravity is added to v along the axes where the body's
// inverse mass is nonzero (a dynamic body coupled two ways; zero for the others).
__device__ inline void integrate_collider(ColliderDev &c, float dt, f3 g)
{
    const f3 t0 = mk3(c.t[0], c.t[1], c.t[2]);
    const f3 com = collider_com(c);
    const f3 sa = mk3(c.angvel[0] * dt, c.angvel[1] * dt, c.angvel[2] * dt);
    const float len = sqrtf(sa.x * sa.x + sa.y * sa.y + sa.z * sa.z);
    float dx = 0.0f, dy = 0.0f, dz = 0.0f, dw = 1.0f;
    if (len != 0.0f) {
        float sn, cs;
        sincosf(len * 0.5f, &sn, &cs);
        dx = sa.x / len * sn;
        dy = sa.y / len * sn;
        dz = sa.z / len * sn;
        dw = cs;
    }
    const f3 arm = quat_rotate(dx, dy, dz, dw, t0 - com);
    c.t[0] = (com.x + arm.x) + c.linvel[0] * dt;
    c.t[1] = (com.y + arm.y) + c.linvel[1] * dt;
    c.t[2] = (com.z + arm.z) + c.linvel[2] * dt;
    const float qx = c.q[0], qy = c.q[1], qz = c.q[2], qw = c.q[3];
    const float nx = dw * qx + dx * qw + dy * qz - dz * qy;
    const float ny = dw * qy - dx * qz + dy * qw + dz * qx;
    const float nz = dw * qz + dx * qy - dy * qx + dz * qw;
    const float nw = dw * qw - dx * qx - dy * qy - dz * qz;
    const float r = 1.0f / sqrtf(nx * nx + ny * ny + nz * nz + nw * nw);
    c.q[0] = nx * r;
    c.q[1] = ny * r;
    c.q[2] = nz * r;
    c.q[3] = nw * r;
    c.linvel[0] = c.linvel[0] + (c.inv_mass[0] != 0.0f ? g.x : 0.0f) * dt;
    c.linvel[1] = c.linvel[1] + (c.inv_mass[1] != 0.0f ? g.y : 0.0f) * dt;
    c.linvel[2] = c.linvel[2] + (c.inv_mass[2] != 0.0f ? g.z : 0.0f) * dt;
}

// Every collider of every active scene, 16 slots per scene, grid-stride.  Runs inside the
// substep's last kernel (nothing after it in the substep reads the colliders).
__device__ inline void integrate_colliders_part(const WorldDev &w)
{
    const uint32_t n = w.s_count * 16u;
    for (uint32_t t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        const SceneDev sc = w.scenes[t >> 4];
        const uint32_t c = t & 15u;
        if (!sc.active || c >= sc.col_count) continue;
        integrate_collider(w.cols[sc.col_start + c], sc.dt, mk3(sc.gx, sc.gy, sc.gz));
    }
}

__global__ void __launch_bounds__(256) k_integrate_colliders(WorldDev w) { integrate_colliders_part(w); }

// Contact particles of CPIC scenes: those with affinity bits (the others are compatible
// with every node and have a zero contact field, so the plain path is exact for them).
__device__ inline bool contact_particle(const WorldDev &w, uint32_t pos, uint32_t d)
{
    // breach is zero for the blocks of scenes without CPIC.
    return w.breach[d] && (w.out.uplane(PAFF)[pos] & 0xFFFFu) != 0u;
}

// ------------------------------------------------------------------------------------
// CFL timestep bound (timestep_bound.rs:61-111, models/default.rs:268-330,
// models/utils.rs:157-172) over the input arena: per enabled particle the smaller of the
// model's sound-speed bound and the velocity bound, as 1e-12 s units (saturating),
// min-reduced per scene.  Particles of one scene are contiguous, so a wave first
// reduces over the lanes that share a scene and issues one atomic per scene present.
__device__ inline uint32_t secs_to_int(float dt)
{
    const float x = dt * 1.0e12f;
    if (!(x > 0.0f)) return 0u;
    if (x >= 4294967295.0f) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

__device__ inline float particle_dt_bound(const ClassDev &cls, float h, float mass, float vol0, const f3 &v,
                                          const m3 &F, const m3 &A, float pdet)
{
    const uint32_t tag = cls.tag;
    float bulk, shear, cfl;
    if (tag <= 1u) { bulk = cls.w[0] + 2.0f * cls.w[1] / 3.0f; shear = cls.w[1]; cfl = cls.w[2]; }
    else if (tag <= 3u) { bulk = cls.w[10] + 2.0f * cls.w[11] / 3.0f; shear = cls.w[11]; cfl = cls.w[12]; }
    else if (tag == 4u) { bulk = cls.w[0] * cls.w[1]; shear = 0.0f; cfl = cls.w[3]; }
    else if (tag == 5u) {
        // Snow: the linear elastic backbone with its moduli scaled by the hardening factor.
        const float hard = expf(cls.w[3] * (1.0f - pdet));
        bulk = (cls.w[4] + 2.0f * cls.w[5] / 3.0f) * hard; shear = cls.w[5] * hard; cfl = cls.w[6];
    }
    else { bulk = shear = cfl = 0.0f; }
    float dt = 0.0f;
    if (tag <= 5u) {
        const float density0 = mass / vol0;
        const float density = density0 / fmaxf(det(F), 1.0e-6f);
        const float c = sqrtf((bulk + shear * 4.0f / 3.0f) / density);
        dt = cfl * h / fmaxf(length(v), c);
    }
    const float na2 = dot(A.c0, A.c0) + dot(A.c1, A.c1) + dot(A.c2, A.c2);
    const float norm_b = (h * h) / 4.0f * sqrtf(na2) / mass;
    const float apic_v = norm_b * 6.0f * sqrtf(3.0f) / h;
    return fminf(dt, h / (length(v) + apic_v));
}

// The bound of arena slot i (flags fl, scene sc) in 1e-12 s units.
__device__ inline uint32_t arena_dt_bound(const Arena &a, uint32_t i, uint32_t fl, const SceneDev &sc,
                                          const ClassDev *classes)
{
    m3 F, A;
    F.c0 = mk3(a.plane(PF0)[i], a.plane(PF1)[i], a.plane(PF2)[i]);
    F.c1 = mk3(a.plane(PF3)[i], a.plane(PF4)[i], a.plane(PF5)[i]);
    F.c2 = mk3(a.plane(PF6)[i], a.plane(PF7)[i], a.plane(PF8)[i]);
    A.c0 = mk3(a.plane(PC0)[i], a.plane(PC1)[i], a.plane(PC2)[i]);
    A.c1 = mk3(a.plane(PC3)[i], a.plane(PC4)[i], a.plane(PC5)[i]);
    A.c2 = mk3(a.plane(PC6)[i], a.plane(PC7)[i], a.plane(PC8)[i]);
    const f3 v = mk3(a.plane(PVX)[i], a.plane(PVY)[i], a.plane(PVZ)[i]);
    const ClassDev &cls = classes[sc.class_base + flags_class(fl)];
    const float pdet = cls.tag == 5u ? a.plane(PS0)[i] : 1.0f;
    return secs_to_int(particle_dt_bound(cls, sc.h, a.plane(PMASS)[i], a.plane(PVOL0)[i], v, F, A, pdet));
}

// Wave reduction per scene of the lanes' running minima (cur_s = NONE: none): one
// atomicMin into out[s * stride] per distinct scene held by the wave's lanes.  Called by
// every lane of the wave.
__device__ inline void wave_scene_min(uint32_t cur_s, uint32_t cur_min, uint32_t *out, uint32_t stride)
{
    const uint32_t l = lane_id();
    uint64_t pending = __ballot(cur_s != NONE);
    while (pending) {
        const uint32_t leader = (uint32_t)__ffsll((unsigned long long)pending) - 1u;
        const uint32_t s = __shfl(cur_s, (int)leader, WAVE);
        uint32_t v = (cur_s == s) ? cur_min : 0xFFFFFFFFu;
        for (int o = 32; o > 0; o >>= 1) v = min(v, (uint32_t)__shfl_xor(v, o, WAVE));
        if (l == leader) atomicMin(&out[(size_t)s * stride], v);
        pending &= ~__ballot(cur_s == s);
    }
}

// Adaptive scenes: the CFL bound of the state a G2P kernel has just written at sorted
// position pos (valid lanes; every lane of the wave calls), min-reduced per scene into the
// scene's clock for the next substep.
__device__ inline void next_dt_bound(const WorldDev &w, uint32_t pos, bool valid, uint32_t s)
{
    uint32_t cand = NONE, cs = NONE;
    if (valid) {
        const SceneDev sc = w.scenes[s];
        const uint32_t fl = w.out.uplane(PFLAGS)[pos];
        if (sc.adaptive && (fl & FLAG_ENABLED)) {
            cand = arena_dt_bound(w.out, pos, fl, sc, w.classes);
            cs = s;
        }
    }
    wave_scene_min(cs, cand, &w.clock[0].bound, (uint32_t)(sizeof(SceneClock) / 4));
}

// Start of a substep (first workgroup of its first scan): each adaptive scene's substep
// length becomes min(dt_max, the bound of its state), written into the scene table for
// the substep's later kernels; every scene's clock advances by its substep length.
__device__ inline void advance_clocks(const WorldDev &w)
{
    if (!w.clock) return;
    for (uint32_t s = threadIdx.x; s < w.s_count; s += blockDim.x) {
        SceneDev &sc = const_cast<SceneDev &>(w.scenes[s]);
        if (!sc.active) continue;
        SceneClock &c = w.clock[s];
        float dt = sc.dt;
        if (sc.adaptive) {
            dt = sc.dt_max;
            // The bound in seconds as the reference's host converts it (timestep_bound.rs:45).
            const float b = (float)c.bound / 1.0e12f;
            if (c.bound != NONE && b < dt) dt = b;
            if (!(dt > 0.0f)) dt = sc.dt_max;   // a zero bound (no finite one) keeps the maximum
            sc.dt = dt;
            sc.vlim = sc.h / dt;
        }
        c.bound = NONE;
        c.dt = dt;
        c.steps += 1u;
        c.t += (double)dt;
    }
}

// After this substep's scans: their look-back state and tickets start empty for the next
// substep, whose keys the G2P kernels insert into the hash (emptied by the grid update).
__device__ inline void prepare_next_sort(const WorldDev &w)
{
    const uint32_t gid = blockIdx.x * blockDim.x + threadIdx.x, gsz = gridDim.x * blockDim.x;
    for (uint32_t i = gid; i < 2u * w.scan_stride; i += gsz) w.scan_state[i] = 0ull;
    if (gid == 0) { w.g->ticket[0] = 0u; w.g->ticket[1] = 0u; }
    mark_overflow_slots(w);
}

// One thread per sorted particle, nodes read from the grid: the launch (g2p_groups) gives
// every workgroup at most g2p_threads positions, and a thread takes at most one, so no loop keeps
// the arenas' 60-odd plane addresses live in scalar registers (they spilled to vector
// lanes and were read back for every particle).  In CPIC worlds this kernel takes the
// plain particles and k_g2p_contact the contact ones, so that the contact code does not
// weigh on the register budget of the common path.  Then the new positions give
// the next substep's block keys and insertions (touch_keys): the output arena is the next
// input arena, in the same order, so the next substep starts at its scan.  The keys are
// taken by whole waves (lanes past the range are invalid) from the position and flags the
// thread has just written, still in its registers.
template <bool CPIC, bool ADAPT>
__global__ void __launch_bounds__(g2p_threads<CPIC>()) k_g2p_update(WorldDev w)
{
    uint32_t pbegin, pend;
    xcd_range(w.g->p_out, pbegin, pend);
    GatherGrid grid{w.bup, w.gridv, w.gridv_i, w.ncdf, NONE, NONE, NONE, NONE, NONE, NONE, NONE, NONE};
    const uint32_t pos = pbegin + threadIdx.x;
    bool valid = pos < pend;
    f3 xo = mk3(0, 0, 0);
    uint32_t flo = 0u;
    if (valid) {
        const uint32_t d = w.sblock[pos];
        if (CPIC && contact_particle(w, pos, d)) valid = false;   // keyed by k_g2p_contact
        else g2p_particle<GatherGrid, false>(w, pos, d, grid, xo, flo);
    }
    if (pbegin + (threadIdx.x & ~(uint32_t)(WAVE - 1)) < pend) {   // the wave holds positions
        const uint32_t s = valid ? flags_scene(flo) : 0u;
        touch_keys(w, pos, valid, s, xo.x, xo.y, xo.z);
        if (ADAPT) next_dt_bound(w, pos, valid, s);
    }
    prepare_next_sort(w);
    if (!CPIC) integrate_colliders_part(w);   // CPIC worlds: after k_g2p_contact, which reads them
}

// The contact particles listed by k_particle_cdf, then their keys (as in k_g2p_update).
template <bool ADAPT>
__global__ void __launch_bounds__(256) k_g2p_contact(WorldDev w)
{
    const uint32_t n = w.g->n_contact;
    GatherGrid grid{w.bup, w.gridv, w.gridv_i, w.ncdf, NONE, NONE, NONE, NONE, NONE, NONE, NONE, NONE};
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const uint32_t pos = w.clist[i];
        f3 xo;
        uint32_t flo;
        g2p_particle<GatherGrid, true>(w, pos, w.sblock[pos], grid, xo, flo);
    }
    const uint32_t l = lane_id();
    for (uint32_t base = blockIdx.x * blockDim.x + (threadIdx.x & ~(uint32_t)(WAVE - 1)); base < n;
         base += gridDim.x * blockDim.x) {
        const uint32_t i = base + l;
        const bool valid = i < n;
        uint32_t pos = 0, s = 0;
        float x = 0.0f, y = 0.0f, z = 0.0f;
        if (valid) {
            pos = w.clist[i];
            s = w.bscene[w.sblock[pos]];
            x = w.out.plane(PX)[pos];
            y = w.out.plane(PY)[pos];
            z = w.out.plane(PZ)[pos];
        }
        touch_keys(w, pos, valid, s, x, y, z);
        if (ADAPT) next_dt_bound(w, pos, valid, s);
    }
}

// ------------------------------------------------------------------------------------
// CPIC contact distance field.

// Per rigid particle of a CPIC scene's mesh colliders: gpu_p2g_cdf (p2g_cdf.rs:60-259)
// as a scatter.  The sample and its triangle take the body's pose of the substep's start
// (gpu_transform_sample_points / gpu_transform_shape_points, rigid_particle_update.rs:
// 10-55).  The nodes of the 3-cell range from the sample's associated cell
// round(x / h) - 1 that lie in active blocks get, where their projection onto the
// triangle's plane falls inside the triangle (edge tests with n = ab x ac), the
// collider's affinity bit with the side n . (x - a) < 0, and the candidate distance
// |n . (x - a)| / |n|.  Per node a 64-bit minimum of (distance bits << 32) | collider
// keeps the least distance, the lower collider on a tie (the distances are not
// negative, so their bits order like their values).  The reference reaches the same
// nodes through its rigid-particle sort (a primary block, extras in the +1 neighbours,
// grid/sort.rs:165-250, 493-680) and a per-block loop; here each block of the range is
// looked up in the scene's hash.
__global__ void __launch_bounds__(256) k_mesh_cdf(WorldDev w)
{
    for (uint32_t r = blockIdx.x * blockDim.x + threadIdx.x; r < w.n_rp; r += gridDim.x * blockDim.x) {
        const float4 rp = w.mesh_rp[r];
        const uint4 tri = w.mesh_t[__float_as_uint(rp.w)];
        const SceneDev sc = w.scenes[tri.w >> 4];
        if (!sc.active || !sc.cpic) continue;
        const uint32_t cid = tri.w & 15u;
        const ColliderDev &c = w.cols[sc.col_start + cid];
        const f3 t = mk3(c.t[0], c.t[1], c.t[2]);
        const float qx = c.q[0], qy = c.q[1], qz = c.q[2], qw = c.q[3];
        const float4 va = w.mesh_v[tri.x], vb = w.mesh_v[tri.y], vc = w.mesh_v[tri.z];
        const f3 p = quat_rotate(qx, qy, qz, qw, mk3(rp.x, rp.y, rp.z)) + t;
        const f3 a = quat_rotate(qx, qy, qz, qw, mk3(va.x, va.y, va.z)) + t;
        const f3 b = quat_rotate(qx, qy, qz, qw, mk3(vb.x, vb.y, vb.z)) + t;
        const f3 cc = quat_rotate(qx, qy, qz, qw, mk3(vc.x, vc.y, vc.z)) + t;
        const f3 ab = b - a, ac = cc - a, bc = cc - b;
        const f3 n = cross(ab, ac);
        const float nl = length(n);
        if (nl == 0.0f) continue;
        const f3 eab = cross(ab, n), ebc = cross(bc, n), eac = cross(ac, n);
        const float h = sc.h, ih = sc.inv_h;
        const float fx = round_div(p.x, h, ih) - 1.0f, fy = round_div(p.y, h, ih) - 1.0f,
                    fz = round_div(p.z, h, ih) - 1.0f;
        if (!(fabsf(fx) < 4000.0f && fabsf(fy) < 2000.0f && fabsf(fz) < 4000.0f)) continue;   // NaN too
        const int cx = (int)fx, cy = (int)fy, cz = (int)fz;
        for (int bz = cz >> 2; bz <= (cz + 2) >> 2; ++bz)
            for (int by = cy >> 2; by <= (cy + 2) >> 2; ++by)
                for (int bx = cx >> 2; bx <= (cx + 2) >> 2; ++bx) {
                    // Keys of blocks outside the packable range would alias others.
                    if (bx < -1023 || bx > 1023 || by < -511 || by > 511 || bz < -1023 || bz > 1023) continue;
                    const uint32_t d = find_block(w, sc, bx, by, bz);
                    if (d == NONE) continue;
                    const int x0 = max(cx, 4 * bx), x1 = min(cx + 2, 4 * bx + 3);
                    const int y0 = max(cy, 4 * by), y1 = min(cy + 2, 4 * by + 3);
                    const int z0 = max(cz, 4 * bz), z1 = min(cz + 2, 4 * bz + 3);
                    for (int z = z0; z <= z1; ++z)
                        for (int y = y0; y <= y1; ++y)
                            for (int x = x0; x <= x1; ++x) {
                                const f3 pt = mk3((float)x, (float)y, (float)z) * h;
                                const f3 ap = pt - a, bp = pt - b, cp = pt - cc;
                                if (!(dot(eab, ap) <= 0.0f && dot(ebc, bp) <= 0.0f && dot(eac, cp) >= 0.0f)) continue;
                                const float sd = dot(n, ap) / nl;
                                const size_t node = (size_t)d * NODES_PER_BLOCK + (size_t)((x - 4 * bx) + 4 * (y - 4 * by) + 16 * (z - 4 * bz));
                                atomicOr(&w.maff[node], (sd < 0.0f ? 0x00010001u : 0x00000001u) << cid);
                                atomicMin(&w.mkey[node], ((unsigned long long)__float_as_uint(fabsf(sd)) << 32) | cid);
                            }
                }
    }
}

// Per node of a CPIC scene's blocks: gpu_grid_update_cdf (grid_update_cdf.rs:60-132) for
// the cuboids, then the meshes' part left by k_mesh_cdf: its affinity bits are added and
// its closest collider replaces the cuboids' when strictly nearer (p2g_cdf.rs:241).  The
// mesh part is reset to empty for the next substep.
__global__ void __launch_bounds__(256) k_grid_cdf(WorldDev w)
{
    // One wave per block: lane = node; bany[d] records whether any node of block d is in
    // reach of a collider (the particle pass skips stencils without such nodes).
    const uint32_t waves = gridDim.x * (blockDim.x / WAVE);
    const uint32_t l = threadIdx.x % WAVE;
    if (blockIdx.x == 0 && threadIdx.x == 0) w.g->n_contact = 0u;   // contact list of this substep
    for (uint32_t d = blockIdx.x * (blockDim.x / WAVE) + threadIdx.x / WAVE; d < w.g->n_blocks; d += waves) {
        const SceneDev sc = w.scenes[w.bscene[d]];
        if (!sc.cpic) continue;
        const int4 v = w.bvid[d];
        float dist = MAX_FLT_REF;
        uint32_t aff = 0u, closest = NONE, bc = 0u;
        const uint32_t mask = v.x != OVERFLOW_VID ? w.bmask[d] : 0u;
        const size_t n = (size_t)d * NODES_PER_BLOCK + l;
        const int i = l & 3, j = (l >> 2) & 3, k = l >> 4;
        const int4 vb = v.x != OVERFLOW_VID ? v : make_int4(0, 0, 0, 0);   // no mesh reaches the overflow block
        const f3 pt = mk3((float)(vb.x * 4 + i), (float)(vb.y * 4 + j), (float)(vb.z * 4 + k)) * sc.h;
        if (mask) node_cdf(w.cols + sc.col_start, sc.col_count, mask, sc.h, pt, dist, aff, closest);
        if (w.n_rp) {
            const unsigned long long key = w.mkey[n];
            if (key != ~0ull) {
                w.mkey[n] = ~0ull;
                const float md = __uint_as_float((uint32_t)(key >> 32));
                if (md < dist) { dist = md; closest = (uint32_t)key & 15u; }
            }
            const uint32_t ma = w.maff[n];
            if (ma) { w.maff[n] = 0u; aff |= ma; }
        }
        if (closest != NONE) {
            // For G2P's incompatible nodes: the closest body's velocity at the node
            // (g2p.rs:214-231 evaluates it at the node position as dpt + x) and its
            // boundary condition.
            const ColliderDev &c = w.cols[sc.col_start + closest];
            const f3 bpv = body_point_velocity(c, pt);
            w.nbody[n] = make_float4(bpv.x, bpv.y, bpv.z, c.bc_friction);
            bc = c.bc_type;
        }
        w.ncdf[n] = make_uint4(__float_as_uint(dist), aff, closest, bc);
        const uint64_t any = __ballot(aff != 0u);
        if (l == 0) w.bany[d] = any != 0ull ? 1u : 0u;
        // The primary blocks whose particles' stencils include nodes of this block (block d
        // is their +e neighbour, e in {0,1}^3) have particles in reach of a collider.
        if (any && l < 8u) {
            const uint32_t src = w.bdn[8 * (size_t)d + l];
            if (src != NONE) w.breach[src] = 1u;
        }
    }
}

// Per particle of a CPIC scene: gpu_g2p_cdf (g2p_cdf.rs:150-331).  The affinity word is
// the union of the stencil nodes' affinities; a collider reached for the first time gets
// its side from the sign of the weighted signed node distances, a known one keeps its
// side.  Then the moving-least-squares fit d(x) = n . (x - x_p) + d_p over the nodes that
// share an affinity with the particle (distance negated where the node's side differs)
// gives the normal and the signed distance; when its matrix is singular (det <= 1e-8) the
// field is zero, affinity included.  Results go to the output arena at the sorted
// position, where P2G and G2P read them and where they stay as the next state.
__device__ __forceinline__ uint32_t particle_cdf_one(const WorldDev &w, uint32_t pos, uint32_t d, const SceneDev &sc)
{
    const uint32_t slot = w.perm[pos];
    const uint32_t prev = w.in.uplane(PAFF)[slot];
    uint32_t aff = prev;
    f3 nrm = mk3(0, 0, 0);
    float sd = 0.0f;
    const int4 bv = w.bvid[d];
    if (bv.x != OVERFLOW_VID) {
        const float px = w.in.plane(PX)[slot], py = w.in.plane(PY)[slot], pz = w.in.plane(PZ)[slot];
        const float h = sc.h, ih = sc.inv_h;
        const float cx = round_div(px, h, ih) - 1.0f, cy = round_div(py, h, ih) - 1.0f,
                    cz = round_div(pz, h, ih) - 1.0f;
        const int bx = (int)cx - 4 * bv.x, by = (int)cy - 4 * bv.y, bz = (int)cz - 4 * bv.z;
        if ((unsigned)bx < 4u && (unsigned)by < 4u && (unsigned)bz < 4u) {
            const f3 ref = mk3(cx * h - px, cy * h - py, cz * h - pz);
            float wx[3], wy[3], wz[3];
            {
                const float fx = -ref.x * ih, fy = -ref.y * ih, fz = -ref.z * ih;
                wx[0] = 0.5f * (1.5f - fx) * (1.5f - fx); wx[1] = 0.75f - (fx - 1.0f) * (fx - 1.0f); wx[2] = 0.5f * (fx - 0.5f) * (fx - 0.5f);
                wy[0] = 0.5f * (1.5f - fy) * (1.5f - fy); wy[1] = 0.75f - (fy - 1.0f) * (fy - 1.0f); wy[2] = 0.5f * (fy - 0.5f) * (fy - 0.5f);
                wz[0] = 0.5f * (1.5f - fz) * (1.5f - fz); wz[1] = 0.75f - (fz - 1.0f) * (fz - 1.0f); wz[2] = 0.5f * (fz - 0.5f) * (fz - 0.5f);
            }
            // No node of the 8 blocks of the stencil (block d and its +1 neighbours) is in
            // reach of a collider: every affinity is empty, the fit has no node and the
            // field is zero.
            if (!w.breach[d]) {
                w.out.uplane(PAFF)[pos] = 0u;
                return 0u;
            }
            GatherGrid gather{w.bup, w.gridv, w.gridv_i, w.ncdf, NONE, NONE, NONE, NONE, NONE, NONE, NONE, NONE};
            gather.prepare(d);
            float nd[27], wq[27];
            uint32_t na[27];
            uint32_t pa = 0u;
            #pragma unroll
            for (int q = 0; q < 27; ++q) {
                const int sx = NBH[q][0], sy = NBH[q][1], sz = NBH[q][2];
                const uint32_t ni = gather.index(bx + sx, by + sy, bz + sz);
                const uint4 nc = ni == NONE ? make_uint4(__float_as_uint(MAX_FLT_REF), 0u, NONE, 0u) : w.ncdf[ni];
                nd[q] = __uint_as_float(nc.x);
                na[q] = nc.y;
                wq[q] = wx[sx] * wy[sy] * wz[sz];
                pa |= nc.y & 0xFFFFu;
            }
            // Sides: kept for colliders known before, from the weighted signed
            // distances for newly reached ones.
            pa |= prev & ((prev & 0xFFFFu) << 16);
            // Only colliders some stencil node is affine to can give a side (for the others
            // every term is zero), and only those not known before need one.
            for (uint32_t todo = pa & ~prev & 0xFFFFu; todo != 0u; todo &= todo - 1u) {
                const uint32_t c = (uint32_t)__ffs(todo) - 1u;
                float acc = 0.0f;
                #pragma unroll
                for (int q = 0; q < 27; ++q) {
                    const float comp = (na[q] >> c) & 1u ? 1.0f : 0.0f;
                    const float sgn = (na[q] >> (16 + c)) & 1u ? -1.0f : 1.0f;
                    acc += comp * wq[q] * sgn * nd[q];
                }
                if (acc < 0.0f) pa |= 0x00010000u << c;
            }
            Sym4 m{0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
            float u[4] = {0.0f, 0.0f, 0.0f, 0.0f};
            #pragma unroll
            for (int q = 0; q < 27; ++q) {
                const uint32_t comb = na[q] & pa & 0xFFFFu;
                if (comb == 0u) continue;
                const uint32_t sdiff = ((na[q] >> 16) ^ (pa >> 16)) & comb;
                const float sx = (float)NBH[q][0] * h, sy = (float)NBH[q][1] * h, sz = (float)NBH[q][2] * h;
                const float dx = ref.x + sx, dy = ref.y + sy, dz = ref.z + sz;
                const float wgt = wq[q];
                const float dd = sdiff == 0u ? nd[q] : -nd[q];
                m.q00 += dx * dx * wgt; m.q01 += dx * dy * wgt; m.q02 += dx * dz * wgt; m.q03 += dx * wgt;
                m.q11 += dy * dy * wgt; m.q12 += dy * dz * wgt; m.q13 += dy * wgt;
                m.q22 += dz * dz * wgt; m.q23 += dz * wgt;
                m.q33 += wgt;
                u[0] += dx * (wgt * dd); u[1] += dy * (wgt * dd); u[2] += dz * (wgt * dd); u[3] += wgt * dd;
            }
            float res[4];
            if (sym4_solve(m, u, res)) {
                const f3 g = mk3(res[0], res[1], res[2]);
                const float len = length(g);
                nrm = len > 1.0e-6f ? g / len : mk3(0, 0, 0);
                sd = res[3];
                aff = pa;
            } else {
                aff = 0u;
            }
        }
    }
    // The normal and signed distance are only read where the affinity is not empty.
    w.out.uplane(PAFF)[pos] = aff;
    if (aff == 0u) return 0u;
    w.out.plane(PNX)[pos] = nrm.x;
    w.out.plane(PNY)[pos] = nrm.y;
    w.out.plane(PNZ)[pos] = nrm.z;
    w.out.plane(PSD)[pos] = sd;
    return aff;
}

// Appends pos to a list (one counter atomic per wave; the order does not matter).
__device__ inline void wave_append(uint32_t *list, uint32_t *count, bool take, uint32_t pos)
{
    const uint32_t l = lane_id();
    const uint64_t active = __ballot(1), m = __ballot(take);
    if (!m) return;
    const uint32_t leader = (uint32_t)__ffsll((unsigned long long)active) - 1u;
    uint32_t base = 0u;
    if (l == leader) base = atomicAdd(count, (uint32_t)__popcll(m));
    base = __shfl(base, (int)leader, WAVE);
    if (take) list[base + (uint32_t)__popcll(m & ((1ull << l) - 1ull))] = pos;
}

// Per particle of a CPIC scene: its contact field (particle_cdf_one); the particles left
// with affinity bits go to the contact list that k_g2p_contact processes densely.
__global__ void __launch_bounds__(256) k_particle_cdf(WorldDev w)
{
    uint32_t pbegin, pend;
    xcd_range(w.g->p_out, pbegin, pend);
    for (uint32_t pos = pbegin + threadIdx.x; pos < pend; pos += blockDim.x) {
        const uint32_t d = w.sblock[pos];
        const SceneDev sc = w.scenes[w.bscene[d]];
        const bool contact = sc.cpic && (particle_cdf_one(w, pos, d, sc) & 0xFFFFu) != 0u;
        wave_append(w.clist, &w.g->n_contact, contact, pos);
    }
}

__global__ void __launch_bounds__(256) k_timestep_bound(WorldDev w, uint32_t n, uint32_t *out, uint32_t stride)
{
    const uint32_t nr = (n + WAVE - 1) / WAVE * WAVE;
    uint32_t cur_s = NONE, cur_min = 0xFFFFFFFFu;     // per lane: running minimum of one scene
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < nr; i += gridDim.x * blockDim.x) {
        if (i >= n) continue;
        const uint32_t fl = w.in.uplane(PFLAGS)[i];
        const uint32_t s = flags_scene(fl);
        const SceneDev sc = w.scenes[s];
        if (!sc.active || !(fl & FLAG_ENABLED)) continue;
        const uint32_t cand = arena_dt_bound(w.in, i, fl, sc, w.classes);
        if (s != cur_s) {
            if (cur_s != NONE) atomicMin(&out[(size_t)cur_s * stride], cur_min);
            cur_s = s;
            cur_min = cand;
        } else {
            cur_min = min(cur_min, cand);
        }
    }
    wave_scene_min(cur_s, cur_min, out, stride);
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
        cell_of(prev.plane(PX)[slot], sc.h, sc.inv_h, b[0], loc[0]);
        cell_of(prev.plane(PY)[slot], sc.h, sc.inv_h, b[1], loc[1]);
        cell_of(prev.plane(PZ)[slot], sc.h, sc.inv_h, b[2], loc[2]);
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

// Workgroups of k_g2p_update: enough that xcd_range gives none more than `threads` (its
// workgroup size) of the at most p_hint sorted positions (one per thread).
static inline int g2p_groups(uint64_t p_hint, int threads)
{
    return grid_for(p_hint, threads, std::numeric_limits<int>::max() / 2);
}

void launch_substep(const WorldDev &w, const LaunchConfig &lc, hipStream_t st, hipEvent_t *ev, bool keys)
{
    const int T = 256;
    const int cap = lc.persistent_blocks;
    if (ev) (void)hipEventRecord(ev[0], st);
    if (keys) k_keys_touch<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    // Adaptive scenes after a host change: the bound of the input state (otherwise the
    // previous substep's G2P left it in the clocks).
    if (keys && w.any_adaptive && w.clock)
        k_timestep_bound<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w, (uint32_t)lc.p_hint, &w.clock[0].bound,
                                                                      (uint32_t)(sizeof(SceneClock) / 4));
    const uint32_t htile = SCAN_TILE * scan_chunks(w.h_total);
    const int hscan = std::min((int)((w.h_total + htile - 1) / htile), cap);
    k_scan<MODE_SLOTS><<<std::max(hscan, 1), 256, 0, st>>>(w);
    k_block_count<<<grid_for(std::max<uint64_t>(lc.p_hint, (uint64_t)w.b_total * 16), T, cap), T, 0, st>>>(w);
    const uint32_t btile = SCAN_TILE * scan_chunks(w.b_total);
    const int bscan = std::min((int)((w.b_total + btile - 1) / btile), cap);
    k_scan<MODE_COUNTS><<<std::max(bscan, 1), 256, 0, st>>>(w);
    k_scatter<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
    if (ev) (void)hipEventRecord(ev[1], st);
    if (w.any_cpic) {
        // CPIC worlds: the contact distance fields, then the CPIC variants of the
        // transfers (scenes without CPIC take the plain path inside them).
        if (w.n_rp) k_mesh_cdf<<<grid_for(w.n_rp, T, cap), T, 0, st>>>(w);
        k_grid_cdf<<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);   // one wave per block
        k_particle_cdf<<<grid_for(lc.p_hint, T, cap), T, 0, st>>>(w);
        if (ev) hipExtLaunchKernelGGL(k_p2g<true>, dim3(lc.p2g_blocks), dim3(64), 0, st, ev[5], ev[6], 0, w);
        else k_p2g<true><<<lc.p2g_blocks, 64, 0, st>>>(w);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (w.mesh_pass) k_grid_update<true, true><<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);
        else k_grid_update<true, false><<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);
        if (ev) (void)hipEventRecord(ev[3], st);
        // The contact particles' G2P touches other particles than the plain one and is
        // latency-bound on few waves: it runs beside it on the side stream when there is one.
        hipStream_t cs = st;
        if (lc.side && lc.fork && lc.join) {
            (void)hipEventRecord(lc.fork, st);
            (void)hipStreamWaitEvent(lc.side, lc.fork, 0);
            cs = lc.side;
        }
        // Adaptive worlds take the variants that also min-reduce the next substep's bound.
        if (w.any_adaptive) k_g2p_contact<true><<<grid_for(lc.p_hint, T, cap), T, 0, cs>>>(w);
        else k_g2p_contact<false><<<grid_for(lc.p_hint, T, cap), T, 0, cs>>>(w);
        auto g2p = w.any_adaptive ? k_g2p_update<true, true> : k_g2p_update<true, false>;
        constexpr int GT = g2p_threads<true>();
        if (ev) hipExtLaunchKernelGGL(g2p, dim3(g2p_groups(lc.p_hint, GT)), dim3(GT), 0, st, ev[7], ev[8], 0, w);
        else g2p<<<g2p_groups(lc.p_hint, GT), GT, 0, st>>>(w);
        if (cs != st) {
            (void)hipEventRecord(lc.join, cs);
            (void)hipStreamWaitEvent(st, lc.join, 0);
        }
        k_integrate_colliders<<<std::max(1, (int)((w.s_count * 16u + T - 1) / T)), T, 0, st>>>(w);
    } else {
        if (ev) hipExtLaunchKernelGGL(k_p2g<false>, dim3(lc.p2g_blocks), dim3(64), 0, st, ev[5], ev[6], 0, w);
        else k_p2g<false><<<lc.p2g_blocks, 64, 0, st>>>(w);
        if (ev) (void)hipEventRecord(ev[2], st);
        if (w.mesh_pass) k_grid_update<false, true><<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);
        else k_grid_update<false, false><<<grid_for((uint64_t)w.b_total * 64, T, cap), T, 0, st>>>(w);
        if (ev) (void)hipEventRecord(ev[3], st);
        auto g2p = w.any_adaptive ? k_g2p_update<false, true> : k_g2p_update<false, false>;
        constexpr int GT = g2p_threads<false>();
        if (ev) hipExtLaunchKernelGGL(g2p, dim3(g2p_groups(lc.p_hint, GT)), dim3(GT), 0, st, ev[7], ev[8], 0, w);
        else g2p<<<g2p_groups(lc.p_hint, GT), GT, 0, st>>>(w);
    }
    if (ev) (void)hipEventRecord(ev[4], st);
}

__global__ void k_set_particles(Globals *g, uint32_t p_in) { g->p_in = p_in; }

void launch_set_particles(Globals *g, uint32_t p_in, hipStream_t st)
{
    k_set_particles<<<1, 1, 0, st>>>(g, p_in);
}

void launch_clear(const WorldDev &w, const LaunchConfig &lc, hipStream_t st)
{
    const int T = 256;
    k_clear<<<grid_for(std::max(std::max(w.h_total, w.b_total), 2u * w.scan_stride), T, lc.persistent_blocks), T, 0, st>>>(w);
}

void launch_timestep_bound(const WorldDev &w, uint32_t n, uint32_t *out, uint32_t stride, int max_blocks, hipStream_t st)
{
    k_timestep_bound<<<grid_for(n, 256, max_blocks), 256, 0, st>>>(w, n, out, stride);
}

void launch_canonical(const WorldDev &w, const Arena &prev, uint32_t *pcell, uint32_t *buckets, uint32_t p_hint,
                      hipStream_t st)
{
    k_canonical<<<grid_for(p_hint, 256, 4096), 256, 0, st>>>(w, prev, pcell, buckets);
}

// ------------------------------------------------------------------------------------
// Render readback: gpu_prep_readback (src_mpm_shaders/solver/prep_readback.rs:245-300)
// over arena positions [start, start + count) of one scene, written at each particle's id
// as 24 floats (ReadbackData: colour, deformation as a column-major 4x4 with the 3x3 in
// its upper left, position, padding).  group / phase / radius are the particles' constant
// properties by id; cdf says whether the scene's contact field is live (CPIC).
__global__ void __launch_bounds__(256) k_prep_readback(Arena a, uint32_t start, uint32_t count, const uint32_t *group,
                                                       const float *phase, const float *radius, const float4 *palette,
                                                       uint32_t npal, uint32_t mode, float h, float dt, uint32_t cdf,
                                                       float *out)
{
    for (uint32_t i = blockIdx.x * blockDim.x + threadIdx.x; i < count; i += gridDim.x * blockDim.x) {
        const uint32_t slot = start + i;
        const uint32_t id = a.uplane(PID)[slot];
        const uint32_t fl = a.uplane(PFLAGS)[slot];
        m3 F;
        F.c0 = mk3(a.plane(PF0)[slot], a.plane(PF1)[slot], a.plane(PF2)[slot]);
        F.c1 = mk3(a.plane(PF3)[slot], a.plane(PF4)[slot], a.plane(PF5)[slot]);
        F.c2 = mk3(a.plane(PF6)[slot], a.plane(PF7)[slot], a.plane(PF8)[slot]);
        const float4 base = palette[group[id] % npal];
        // compute_deformation (:97-118): diag(2 r) times F with entries clamped to [-4, 4].
        const float d = radius[id] * 2.0f;
        float *o = out + 24 * (size_t)id;
        const f3 cols[3] = {F.c0, F.c1, F.c2};
        #pragma unroll
        for (int c = 0; c < 3; ++c) {
            o[4 + 4 * c + 0] = d * fminf(fmaxf(cols[c].x, -4.0f), 4.0f);
            o[4 + 4 * c + 1] = d * fminf(fmaxf(cols[c].y, -4.0f), 4.0f);
            o[4 + 4 * c + 2] = d * fminf(fmaxf(cols[c].z, -4.0f), 4.0f);
            o[4 + 4 * c + 3] = 0.0f;
        }
        o[16] = 0.0f; o[17] = 0.0f; o[18] = 0.0f; o[19] = 1.0f;
        // compute_color (:170-243).
        float4 col = base;
        const uint32_t aff = cdf ? a.uplane(PAFF)[slot] : 0u;
        if (mode == 2u) {           // velocity
            const f3 v = mk3(a.plane(PVX)[slot], a.plane(PVY)[slot], a.plane(PVZ)[slot]);
            col = make_float4(fabsf(v.x) * dt * 100.0f + 0.2f, fabsf(v.y) * dt * 100.0f + 0.2f,
                              fabsf(v.z) * dt * 100.0f + 0.2f, base.w);
        } else if (mode == 1u) {    // volume: singular values of F, largest first
            m3 U, V;
            f3 sv;
            svd3(F, U, sv, V);
            float s0 = sv.x, s1 = sv.y, s2 = sv.z, t;
            if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
            if (s1 < s2) { t = s1; s1 = s2; s2 = t; }
            if (s0 < s1) { t = s0; s0 = s1; s1 = t; }
            col = make_float4((1.0f - s0) / 0.005f + 0.2f, (1.0f - s1) / 0.005f + 0.2f, (1.0f - s2) / 0.005f + 0.2f,
                              base.w);
        } else if (mode == 3u) {    // phase
            const float ph = phase[id];
            col = make_float4(0.0f, 0.4f * ph, 0.4f * (1.0f - ph), base.w);
        } else if (mode == 4u) {    // contact normals
            const f3 n = aff ? mk3(a.plane(PNX)[slot], a.plane(PNY)[slot], a.plane(PNZ)[slot]) : mk3(0, 0, 0);
            col = n.x == 0.0f && n.y == 0.0f && n.z == 0.0f
                      ? make_float4(0.0f, 0.0f, 0.0f, base.w)
                      : make_float4((n.x + 1.0f) * 0.5f, (n.y + 1.0f) * 0.5f, (n.z + 1.0f) * 0.5f, base.w);
        } else if (mode == 5u) {    // contact signed distances
            const float sd = aff ? a.plane(PSD)[slot] : 0.0f;
            const float dd = sd / (h * 1.5f);
            col = dd > 0.0f ? make_float4(0.0f, fabsf(dd), 0.0f, base.w) : make_float4(fabsf(dd), 0.0f, 0.0f, base.w);
        } else if (mode == 6u) {    // contact sides
            const uint32_t inside = (aff >> 16) & (aff & 0xFFFFu);
            col = aff == 0u ? make_float4(0.0f, 0.0f, 0.0f, base.w)
                            : (inside == 0u ? make_float4(0.0f, 1.0f, 0.0f, base.w)
                                            : make_float4(1.0f, 0.0f, 0.0f, base.w));
        }
        if (!(fl & FLAG_ENABLED)) col = make_float4(1.0f, 0.0f, 0.0f, 1.0f);   // failed particles in red
        o[0] = col.x; o[1] = col.y; o[2] = col.z; o[3] = col.w;
        o[20] = a.plane(PX)[slot]; o[21] = a.plane(PY)[slot]; o[22] = a.plane(PZ)[slot]; o[23] = 0.0f;
    }
}

void launch_prep_readback(const Arena &a, uint32_t start, uint32_t count, const uint32_t *group, const float *phase,
                          const float *radius, const float4 *palette, uint32_t npal, uint32_t mode, float h, float dt,
                          bool cdf, float *out, int cap, hipStream_t st)
{
    if (count == 0) return;
    const int T = 256;
    k_prep_readback<<<grid_for(count, T, cap), T, 0, st>>>(a, start, count, group, phase, radius, palette, npal, mode,
                                                          h, dt, cdf ? 1u : 0u, out);
}

// ------------------------------------------------------------------------------------
// Substep without allocation: the world's tables and arenas live in a buffer the host
// carved up (the library's static scratch array); the particles come from and return to
// the caller's arrays in id order.

// Tables and counters (one workgroup).
__global__ void k_noalloc_tables(WorldDev w, NoallocTables t)
{
    const uint32_t tid = threadIdx.x;
    if (tid == 0) {
        *const_cast<SceneDev *>(w.scenes) = t.scene;
        w.ctr[0] = SceneCtr{0u, 0u, 0u, 0u};
        const_cast<uint32_t *>(w.region_start)[0] = 0u;
        const_cast<uint32_t *>(w.region_scene)[0] = 0u;
        w.g->p_in = t.n;
        w.g->n_contact = 0u;
    }
    if (tid < t.n_classes) const_cast<ClassDev *>(w.classes)[tid] = t.classes[tid];
    if (tid < t.n_cols) {
        ColliderDev c{};
        float *words = reinterpret_cast<float *>(&c);
        for (int k = 0; k < COLLIDER_BODY_WORDS; ++k) words[k] = t.cols[tid].body[k];
        c.shape = SHAPE_CUBOID;
        c.reach = t.cols[tid].reach;
        w.cols[tid] = c;
    }
}

// The input arena from the caller's arrays (after k_noalloc_tables: it may report a bad
// material index in the scene's error word).
__global__ void k_noalloc_pack(WorldDev w, NoallocTables t, NoallocParticles p)
{
    const uint32_t tid = threadIdx.x;
    for (uint32_t i = blockIdx.x * blockDim.x + tid; i < t.n; i += gridDim.x * blockDim.x) {
        const Arena a = w.in;
        for (int c = 0; c < 3; ++c) {
            a.plane(PX + c)[i] = p.x[3 * (size_t)i + c];
            a.plane(PVX + c)[i] = p.v[3 * (size_t)i + c];
        }
        for (int c = 0; c < 9; ++c) {
            a.plane(PC0 + c)[i] = p.C[9 * (size_t)i + c];
            a.plane(PF0 + c)[i] = p.F[9 * (size_t)i + c];
        }
        a.plane(PMASS)[i] = p.mass[i];
        a.plane(PVOL0)[i] = p.vol0[i];
        const uint32_t f = p.flags[i];
        uint32_t cls = flags_class(f);
        if (cls >= t.n_classes) {
            // A material index outside the table: reported, and the particle takes material 0.
            atomicOr(&w.ctr[0].err, ERR_MODEL);
            cls = 0u;
        }
        // Plastic state: the caller's, or the material's initial words (the sand tags'
        // three, snow's one), as an upload takes it from the particle's model words.
        const uint32_t tag = t.classes[cls].tag;
        const int pw = (tag == 2u || tag == 3u) ? 3 : tag == 5u ? 1 : 0;
        for (int c = 0; c < 3; ++c)
            a.plane(PS0 + c)[i] = c < pw ? (p.pstate ? p.pstate[3 * (size_t)i + c] : t.classes[cls].w[c]) : 0.0f;
        for (int c = PAFF; c <= PSD; ++c) a.plane(c)[i] = 0.0f;
        a.uplane(PFLAGS)[i] = make_flags((f & FLAG_ENABLED) != 0u, (f & FLAG_FIXED) != 0u, cls, 0u);
        a.uplane(PID)[i] = i;
    }
}

__global__ void k_noalloc_unpack(WorldDev w, NoallocParticles p, uint32_t n)
{
    const Arena a = w.out;
    for (uint32_t pos = blockIdx.x * blockDim.x + threadIdx.x; pos < n; pos += gridDim.x * blockDim.x) {
        const uint32_t id = a.uplane(PID)[pos];
        if (id >= n) continue;
        for (int c = 0; c < 3; ++c) {
            p.x[3 * (size_t)id + c] = a.plane(PX + c)[pos];
            p.v[3 * (size_t)id + c] = a.plane(PVX + c)[pos];
        }
        for (int c = 0; c < 9; ++c) {
            p.C[9 * (size_t)id + c] = a.plane(PC0 + c)[pos];
            p.F[9 * (size_t)id + c] = a.plane(PF0 + c)[pos];
        }
        p.mass[id] = a.plane(PMASS)[pos];
        if (p.pstate)
            for (int c = 0; c < 3; ++c) p.pstate[3 * (size_t)id + c] = a.plane(PS0 + c)[pos];
        const uint32_t f = a.uplane(PFLAGS)[pos];
        p.flags[id] = (f & (FLAG_ENABLED | FLAG_FIXED)) | (flags_class(f) << 2);
    }
    if (blockIdx.x == 0 && threadIdx.x == 0 && p.errors) p.errors[0] = w.ctr[0].err;
}

void launch_noalloc_setup(const WorldDev &w, const NoallocTables &t, const NoallocParticles &p, int cap, hipStream_t st)
{
    const int T = 256;
    k_noalloc_tables<<<1, 64, 0, st>>>(w, t);
    k_noalloc_pack<<<grid_for(std::max<uint32_t>(t.n, 1u), T, cap), T, 0, st>>>(w, t, p);
}

void launch_noalloc_unpack(const WorldDev &w, const NoallocParticles &p, uint32_t n, int cap, hipStream_t st)
{
    const int T = 256;
    k_noalloc_unpack<<<grid_for(std::max<uint32_t>(n, 1u), T, cap), T, 0, st>>>(w, p, n);
}

}  // namespace nxm
