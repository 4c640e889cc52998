// mpm_api.hip — host runtime and C ABI (include/nexus_mpm.h).
//
// A World owns the device state of a set of scenes stepped together: two particle
// arenas (read in the previous step's sorted order, written in this step's order),
// per-particle sort scratch, per-block tables, hash regions, and the scene / material /
// collider tables.  A standalone scene (the reference's MpmState) is a World holding
// one scene; a batch is a World holding many.  Scenes enter and leave a World between
// steps; every step re-packs the live scenes' particles contiguously in slot order.
#include <algorithm>
#include <atomic>
#include <array>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <unordered_set>
#include <vector>
#include <cstddef>

#include "../../include/nexus_mpm.h"
#include "mpm_kernels.h"

namespace nxm {

static thread_local std::string g_last_error;
// Uploads still to fail on purpose (nxm_debug_fail_uploads): reaches the error paths of a
// world's bookkeeping from the host.
static std::atomic<int> g_fail_uploads{0};

static nxm_status fail(nxm_status code, const std::string &msg)
{
    g_last_error = msg;
    return code;
}

static nxm_status hip_fail(hipError_t e, const char *what)
{
    nxm_status code = (e == hipErrorOutOfMemory) ? NXM_ERR_OUT_OF_MEMORY : NXM_ERR_DEVICE;
    return fail(code, std::string(what) + ": " + hipGetErrorString(e));
}

#define HIP_TRY(expr)                                        \
    do {                                                     \
        hipError_t e_ = (expr);                              \
        if (e_ != hipSuccess) return hip_fail(e_, #expr);    \
    } while (0)

static uint32_t next_pow2(uint32_t v)
{
    uint32_t p = 1;
    while (p < v && p < (1u << 31)) p <<= 1;
    return p;
}

template <typename T>
struct DevBuf {
    T *p = nullptr;
    size_t n = 0;
    ~DevBuf() { release(); }
    void release()
    {
        if (p) (void)hipFree(p);
        p = nullptr;
        n = 0;
    }
    // Grows (discarding contents) to at least `want` elements.
    hipError_t reserve(size_t want)
    {
        if (want <= n) return hipSuccess;
        release();
        size_t c = std::max<size_t>(want, 1);
        hipError_t e = hipMalloc(reinterpret_cast<void **>(&p), c * sizeof(T));
        if (e == hipSuccess) n = c;
        return e;
    }
};

}  // namespace nxm

using namespace nxm;

// Pinned host chunks shared by the threads of one context: a scene staged for upload on
// one thread (nxm_scene_stage_upload) holds a chunk until the thread stepping its world
// queues the copies; the chunk is reused once those copies have executed.
struct PinnedPool {
    struct Chunk {
        char *p = nullptr;
        size_t n = 0;
        hipEvent_t ev = nullptr;
        bool held = false;
    };
    std::mutex mu;
    std::vector<Chunk> chunks;

    // The smallest free chunk of at least `bytes` whose last copies have executed.
    hipError_t take(size_t bytes, int *idx, char **out)
    {
        std::lock_guard<std::mutex> g(mu);
        int best = -1;
        for (int k = 0; k < (int)chunks.size(); ++k) {
            Chunk &c = chunks[k];
            if (c.held || c.n < bytes) continue;
            if (best >= 0 && chunks[best].n <= c.n) continue;
            const hipError_t q = hipEventQuery(c.ev);
            if (q == hipErrorNotReady) continue;
            if (q != hipSuccess) return q;
            best = k;
        }
        if (best < 0) {
            Chunk c;
            c.n = std::max<size_t>(bytes, (size_t)1 << 20);
            hipError_t e = hipHostMalloc(reinterpret_cast<void **>(&c.p), c.n, hipHostMallocDefault);
            if (e != hipSuccess) return e;
            e = hipEventCreate(&c.ev);
            if (e != hipSuccess) { (void)hipHostFree(c.p); return e; }
            chunks.push_back(c);
            best = (int)chunks.size() - 1;
        }
        chunks[best].held = true;
        *idx = best;
        *out = chunks[best].p;
        return hipSuccess;
    }
    // Gives the chunk back; with a stream, reusable once the work queued so far has run.
    hipError_t give(int idx, hipStream_t st)
    {
        std::lock_guard<std::mutex> g(mu);
        Chunk &c = chunks[idx];
        hipError_t e = st ? hipEventRecord(c.ev, st) : hipSuccess;
        c.held = false;
        return e;
    }
    char *ptr(int idx)
    {
        std::lock_guard<std::mutex> g(mu);
        return chunks[idx].p;
    }
    ~PinnedPool()
    {
        for (auto &c : chunks) {
            (void)hipEventSynchronize(c.ev);
            (void)hipEventDestroy(c.ev);
            (void)hipHostFree(c.p);
        }
    }
};

struct nxm_ctx {
    int device = 0;
    int num_cus = 256;
    void *scratch = nullptr;     // this device's address of the no-allocation scratch array
    // Scenes and batches created from this context; nxm_ctx_destroy defers the delete
    // until the last of them is destroyed.
    std::atomic<int> refs{0};
    std::atomic<bool> closed{false};
    PinnedPool pool;
};

// Scratch of nxm_pipeline_step_noalloc: one per GPU (each device holds its own copy of
// this array at its own address).  Host code never uses the name itself (it is a host-side
// stand-in, not a device address): nxm_ctx_create looks the device address up once per
// device with hipGetSymbolAddress and every use goes through that pointer.
static __device__ __align__(256) unsigned char scratch[NXM_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static std::map<int, void *> g_scratch_ptr;   // device -> address of its scratch array

// The calling thread's current device must be `device`.
static hipError_t scratch_address(int device, void **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it != g_scratch_ptr.end()) {
        *out = it->second;
        return hipSuccess;
    }
    void *p = nullptr;
    hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr.emplace(device, p);
    *out = p;
    return hipSuccess;
}

static void ctx_release(nxm_ctx *ctx)
{
    if (ctx && ctx->refs.fetch_sub(1) == 1 && ctx->closed.load()) delete ctx;
}

namespace nxm {

class World;

// Leading model words that are per-particle plastic state rather than parameters: the
// sand tags' {det Fp, hardening q, log volume gain} (models/default.rs:60-76) and snow's
// plastic determinant (models/default.rs:78-110).
static int pstate_words(uint32_t tag) { return (tag == 2u || tag == 3u) ? 3 : tag == 5u ? 1 : 0; }

// Host copy of a scene.  Particle state lives here while the scene is not resident
// in a World; static per-particle properties always live here (indexed by id).
// ------------------------------------------------------------------------------------
// Rigid particles of a triangle mesh: sample_mesh (src_mpm/sampling/sample_trimesh.rs:
// 60-231), in f32 like the reference.  Every triangle is sampled in rows across its
// longest edge (the base) and up its height, every s = spacing / sqrt 2 along both, each
// row ending before the nearer of the two other edges; then each edge not seen before
// is sampled every s from its first vertex.  Neither vertices nor the base carry
// samples of the triangle pass; the edge pass covers the edges.

struct V3f {
    float x, y, z;
};
static inline V3f operator-(V3f a, V3f b) { return {a.x - b.x, a.y - b.y, a.z - b.z}; }
static inline V3f operator+(V3f a, V3f b) { return {a.x + b.x, a.y + b.y, a.z + b.z}; }
static inline V3f operator*(V3f a, float k) { return {a.x * k, a.y * k, a.z * k}; }
static inline V3f operator/(V3f a, float k) { return {a.x / k, a.y / k, a.z / k}; }
static inline float dot3(V3f a, V3f b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline float norm3(V3f a) { return std::sqrt(dot3(a, a)); }

struct MeshSamples {
    std::vector<float> pts;        // 3 per sample, body frame
    std::vector<uint32_t> tri;     // triangle of each sample
    void push(V3f p, uint32_t t)
    {
        pts.push_back(p.x);
        pts.push_back(p.y);
        pts.push_back(p.z);
        tri.push_back(t);
    }
};

static constexpr float SAMPLE_EPS = 1.0e-5f;   // sample_trimesh.rs:12

static void sample_triangle_rows(V3f a, V3f b, V3f c, float spacing, uint32_t t, MeshSamples &out)
{
    const float lab = norm3(b - a), lbc = norm3(c - b), lca = norm3(a - c);
    const float longest = std::max(std::max(lab, lbc), lca);
    V3f p0 = a, p1 = b, p2 = c;                 // p0 p1: the base
    if (longest == lbc) { p0 = b; p1 = c; p2 = a; }
    else if (longest == lca) { p0 = c; p1 = a; p2 = b; }
    const V3f base = p1 - p0, side = p2 - p0;
    const float base_len = norm3(base);
    const V3f dir = base / base_len;
    const float s = spacing / std::sqrt(2.0f);
    const float along = dot3(side, dir);        // foot of the apex on the base
    const float rest = base_len - along;
    if (along < SAMPLE_EPS || rest < SAMPLE_EPS || base_len < SAMPLE_EPS) return;
    const V3f up_vec = side - dir * along;
    const float up_len = norm3(up_vec);
    const V3f up = up_vec / up_len;
    const float slope0 = up_len / along, slope1 = up_len / rest;
    const uint32_t rows = (uint32_t)std::ceil(base_len / s);
    const V3f row_step = dir * s, col_step = up * s;
    for (uint32_t i = 1; i < rows; ++i) {
        const V3f foot = p0 + row_step * (float)i;
        const float height = std::min(slope0 * norm3(p0 - foot), slope1 * norm3(p1 - foot));
        const uint32_t cols = (uint32_t)std::ceil(height / s);
        for (uint32_t j = 1; j < cols; ++j) {
            const V3f q = foot + col_step * (float)j;
            if (std::isfinite(q.x) && std::isfinite(q.y) && std::isfinite(q.z)) out.push(q, t);
        }
    }
}

static void sample_edge_points(V3f a, V3f b, float spacing, uint32_t t, MeshSamples &out)
{
    const V3f e = b - a;
    const float len = norm3(e);
    if (!(len > SAMPLE_EPS)) return;
    const V3f dir = e / len;
    const float s = spacing / std::sqrt(2.0f);
    const size_t steps = (size_t)std::ceil(len / s);
    for (size_t i = 1; i < steps; ++i) out.push(a + dir * (s * (float)i), t);
}

static MeshSamples sample_mesh_surface(const float *v, const uint32_t *tri, uint32_t nt, float spacing)
{
    MeshSamples out;
    std::unordered_set<uint64_t> edges;
    auto vert = [&](uint32_t i) { return V3f{v[3 * i], v[3 * i + 1], v[3 * i + 2]}; };
    for (uint32_t t = 0; t < nt; ++t) {
        const uint32_t *ix = tri + 3 * t;
        sample_triangle_rows(vert(ix[0]), vert(ix[1]), vert(ix[2]), spacing, t, out);
        for (int e = 0; e < 3; ++e) {
            const uint32_t i0 = ix[e], i1 = ix[(e + 1) % 3];
            const uint64_t key = ((uint64_t)std::max(i0, i1) << 32) | std::min(i0, i1);
            if (edges.insert(key).second) sample_edge_points(vert(i0), vert(i1), spacing, t, out);
        }
    }
    return out;
}

// A triangle-mesh collider's geometry (body frame) and its rigid particles at the cell
// width they were sampled for.
// A mesh collider: geometry in the body frame, the pseudo-normals of its inside test, the
// rigid particles of CPIC scenes and the cell grid of the node pass (both for the cell
// width they were made at).
struct HostMesh {
    std::vector<float> v;
    std::vector<uint32_t> tri;
    std::vector<float> vn, en;    // per vertex 3, per triangle 9 (edges ab, bc, ca)
    MeshSamples samples;
    float sampled_h = 0.0f;
    float grid_h = 0.0f;
    float origin[3] = {0.0f, 0.0f, 0.0f};
    float cell = 1.0f;
    uint32_t dims[3] = {1, 1, 1};
    std::vector<uint32_t> start, ctri;   // per cell its list in ctri (start[c]..start[c + 1])
};

// Pseudo-normals of the mesh's inside test as parry3d 0.30 computes them
// (TriMesh::pseudo_normals, read by src_rbd_shaders/shapes/trimesh.rs:116-152): per vertex
// the sum of its faces' unit normals weighted by each face's angle at the vertex; per
// triangle edge (ab, bc, ca) the sum of the unit normals of the faces sharing it; faces of
// zero area add nothing; both normalized.
static void mesh_pseudo_normals(HostMesh &m)
{
    const size_t nv = m.v.size() / 3, nt = m.tri.size() / 3;
    std::vector<double> vn(3 * nv, 0.0), fn(3 * nt, 0.0);
    std::vector<std::pair<uint64_t, size_t>> edges(3 * nt);
    for (size_t t = 0; t < nt; ++t) {
        double p[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) p[k][c] = m.v[3 * (size_t)m.tri[3 * t + k] + c];
        const double e1[3] = {p[1][0] - p[0][0], p[1][1] - p[0][1], p[1][2] - p[0][2]};
        const double e2[3] = {p[2][0] - p[0][0], p[2][1] - p[0][1], p[2][2] - p[0][2]};
        const double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        const double l = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0.0) {
            for (int c = 0; c < 3; ++c) fn[3 * t + c] = n[c] / l;
            for (int k = 0; k < 3; ++k) {
                const double *o = p[k], *q1 = p[(k + 1) % 3], *q2 = p[(k + 2) % 3];
                const double u[3] = {q1[0] - o[0], q1[1] - o[1], q1[2] - o[2]};
                const double w[3] = {q2[0] - o[0], q2[1] - o[1], q2[2] - o[2]};
                const double cr[3] = {u[1] * w[2] - u[2] * w[1], u[2] * w[0] - u[0] * w[2], u[0] * w[1] - u[1] * w[0]};
                const double ang = std::atan2(std::sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]),
                                              u[0] * w[0] + u[1] * w[1] + u[2] * w[2]);
                for (int c = 0; c < 3; ++c) vn[3 * (size_t)m.tri[3 * t + k] + c] += fn[3 * t + c] * ang;
            }
        }
        for (int k = 0; k < 3; ++k) {
            const uint32_t i0 = m.tri[3 * t + k], i1 = m.tri[3 * t + (k + 1) % 3];
            edges[3 * t + k] = {((uint64_t)std::min(i0, i1) << 32) | std::max(i0, i1), 3 * t + k};
        }
    }
    auto unit = [](double *x) {
        const double l = std::sqrt(x[0] * x[0] + x[1] * x[1] + x[2] * x[2]);
        if (l > 0.0) { x[0] /= l; x[1] /= l; x[2] /= l; }
    };
    m.vn.assign(3 * nv, 0.0f);
    for (size_t q = 0; q < nv; ++q) {
        unit(&vn[3 * q]);
        for (int c = 0; c < 3; ++c) m.vn[3 * q + c] = (float)vn[3 * q + c];
    }
    std::sort(edges.begin(), edges.end());
    m.en.assign(9 * nt, 0.0f);
    for (size_t a = 0; a < edges.size();) {
        size_t b = a;
        double sum[3] = {0.0, 0.0, 0.0};
        for (; b < edges.size() && edges[b].first == edges[a].first; ++b)
            for (int c = 0; c < 3; ++c) sum[c] += fn[3 * (edges[b].second / 3) + c];
        unit(sum);
        for (size_t q = a; q < b; ++q)
            for (int c = 0; c < 3; ++c) m.en[3 * edges[q].second + c] = (float)sum[c];
        a = b;
    }
}

// The mesh's node-pass cell grid for cell width h: over the triangles' bounds widened by
// the 1.5 h query distance (and a margin against rounding), cells of at least 2 h and at
// most 64 per axis; each cell lists, in ascending order, the triangles whose widened
// bounds meet it.
static void mesh_grid(HostMesh &m, float h)
{
    if (m.grid_h == h && !m.start.empty()) return;
    const size_t nt = m.tri.size() / 3;
    const double r = 1.5 * (double)h * 1.001 + 1e-6;
    double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
    for (size_t q = 0; q < m.v.size() / 3; ++q)
        for (int c = 0; c < 3; ++c) {
            lo[c] = std::min(lo[c], (double)m.v[3 * q + c]);
            hi[c] = std::max(hi[c], (double)m.v[3 * q + c]);
        }
    if (nt == 0) { for (int c = 0; c < 3; ++c) lo[c] = hi[c] = 0.0; }
    double ext = 0.0;
    for (int c = 0; c < 3; ++c) {
        lo[c] -= r;
        hi[c] += r;
        ext = std::max(ext, hi[c] - lo[c]);
    }
    const double cell = std::max(2.0 * (double)h, ext / 64.0);
    for (int c = 0; c < 3; ++c) {
        m.origin[c] = (float)lo[c];
        m.dims[c] = (uint32_t)std::max(1.0, std::ceil((hi[c] - lo[c]) / cell));
    }
    m.cell = (float)cell;
    const size_t ncell = (size_t)m.dims[0] * m.dims[1] * m.dims[2];
    std::vector<uint32_t> count(ncell + 1, 0u);
    auto cells_of = [&](size_t t, uint32_t *c0, uint32_t *c1) {
        for (int c = 0; c < 3; ++c) {
            double a = 1e300, b = -1e300;
            for (int k = 0; k < 3; ++k) {
                const double x = m.v[3 * (size_t)m.tri[3 * t + k] + c];
                a = std::min(a, x);
                b = std::max(b, x);
            }
            const double fa = std::floor((a - r - (double)m.origin[c]) / (double)m.cell);
            const double fb = std::floor((b + r - (double)m.origin[c]) / (double)m.cell);
            c0[c] = (uint32_t)std::min(std::max(fa, 0.0), (double)(m.dims[c] - 1));
            c1[c] = (uint32_t)std::min(std::max(fb, 0.0), (double)(m.dims[c] - 1));
        }
    };
    for (int pass = 0; pass < 2; ++pass) {
        if (pass == 1) {
            for (size_t c = 0; c < ncell; ++c) count[c + 1] += count[c];
            m.start.assign(count.begin(), count.end());
            m.ctri.assign(count[ncell], 0u);
            std::fill(count.begin(), count.end(), 0u);
        }
        for (size_t t = 0; t < nt; ++t) {
            uint32_t c0[3], c1[3];
            cells_of(t, c0, c1);
            for (uint32_t z = c0[2]; z <= c1[2]; ++z)
                for (uint32_t y = c0[1]; y <= c1[1]; ++y)
                    for (uint32_t x = c0[0]; x <= c1[0]; ++x) {
                        const size_t c = x + (size_t)m.dims[0] * (y + (size_t)m.dims[1] * z);
                        if (pass == 0) ++count[c + 1];
                        else m.ctri[m.start[c] + count[c]++] = (uint32_t)t;
                    }
        }
    }
    m.grid_h = h;
}

struct HostParticles {
    size_t n = 0;
    // state
    std::vector<float> x, v, F, C, mass, vol0, pstate;
    std::vector<uint32_t> enabled, fixed;
    std::vector<uint32_t> aff;    // CPIC affinity word
    std::vector<float> cdf;       // CPIC contact normal and signed distance (4 per particle)
    // static
    std::vector<float> radius, phase, bfric, damping;
    std::vector<uint32_t> group, tag, words, cls;

    void resize(size_t m)
    {
        n = m;
        x.resize(3 * m); v.resize(3 * m); F.resize(9 * m); C.resize(9 * m);
        mass.resize(m); vol0.resize(m); pstate.resize(3 * m);
        enabled.resize(m); fixed.resize(m);
        aff.resize(m); cdf.resize(4 * m);
        radius.resize(m); phase.resize(m); bfric.resize(m); damping.resize(m);
        group.resize(m); tag.resize(m); words.resize(24 * m); cls.resize(m);
    }
};

// Copies every per-particle array entry of particle `from` to `to`.
static void move_particle(HostParticles &hp, size_t from, size_t to)
{
    auto mv = [&](auto &vec, size_t k) {
        for (size_t c = 0; c < k; ++c) vec[k * to + c] = vec[k * from + c];
    };
    mv(hp.x, 3); mv(hp.v, 3); mv(hp.F, 9); mv(hp.C, 9); mv(hp.mass, 1); mv(hp.vol0, 1); mv(hp.pstate, 3);
    mv(hp.enabled, 1); mv(hp.fixed, 1); mv(hp.radius, 1); mv(hp.phase, 1); mv(hp.bfric, 1); mv(hp.damping, 1);
    mv(hp.group, 1); mv(hp.tag, 1); mv(hp.words, 24); mv(hp.cls, 1);
    mv(hp.aff, 1); mv(hp.cdf, 4);
}

struct SceneRange {
    uint32_t start, count;
};

}  // namespace nxm

struct nxm_scene {
    nxm_ctx *ctx = nullptr;
    float base_dt = 1.0f / 60.0f;
    float dt = 1.0f / 60.0f;
    float h = 1.0f;
    float g[3] = {0.0f, 0.0f, 0.0f};
    uint32_t grid_cap = 32768;
    std::vector<ColliderDev> cols;
    std::vector<ClassDev> classes;
    std::map<std::array<uint32_t, 27>, uint32_t> class_index;
    std::array<uint32_t, 27> last_key{};   // the most recently interned class
    uint32_t last_class = 0xFFFFFFFFu;
    HostParticles hp;
    bool on_host = true;                // hp state authoritative
    World *world = nullptr;             // resident world (own or a batch's)
    uint32_t slot = 0;
    // Colliders: the world's table holds the integrated poses while the scene is resident
    // (dev_col_start); cols_host says the host copy is newer (set_colliders, or not resident).
    uint32_t dev_col_start = 0xFFFFFFFFu;
    bool cols_host = true;
    // Particles packed ahead of the next upload (nxm_scene_stage_upload): a chunk of the
    // context's pinned pool in arena-plane layout, flags without the world slot.
    int staged_chunk = -1;
    size_t staged_n = 0;
    std::vector<SceneRange> ranges;     // particle ranges in the world's input arena
    std::unique_ptr<World> own;         // standalone world
    nxm_batch *batch = nullptr;
    uint32_t error = 0;                 // latched device error bits
    bool debug_grid = false;
    bool graph = false;                 // standalone steps replay a captured graph
    bool cpic = false;                  // MpmState::use_cpic
    std::vector<HostMesh> meshes;       // mesh colliders: the last meshes.size() entries of cols
    // Adaptive substeps (nxm_scene_set_adaptive_dt) and the scene's clock: the world's copy
    // is authoritative while the scene is resident, clock_host says the host copy is newer.
    bool adaptive = false;
    nxm::SceneClock clock{0xFFFFFFFFu, 0.0f, 0u, 0u, 0.0};
    bool clock_host = true;
};

namespace nxm {

enum SlotState { SLOT_FREE = 0, SLOT_ACTIVE, SLOT_RETIRING };

struct Slot {
    nxm_scene *scene = nullptr;
    SlotState state = SLOT_FREE;
};

class World {
  public:
    explicit World(nxm_ctx *ctx) : ctx_(ctx) {}
    ~World()
    {
        if (stream_) {
            (void)hipStreamSynchronize(stream_);
            drop_graphs();
            for (auto &e : ev_) (void)hipEventDestroy(e);
            for (auto &e : tune_ev_) (void)hipEventDestroy(e);
            for (auto &c : pins_) {
                (void)hipEventDestroy(c.ev);
                (void)hipHostFree(c.p);
            }
            for (auto &e : ring_) (void)hipEventDestroy(e);
            if (side_) {
                (void)hipStreamSynchronize(side_);
                (void)hipEventDestroy(fork_);
                (void)hipEventDestroy(join_);
                (void)hipStreamDestroy(side_);
            }
            (void)hipStreamDestroy(stream_);
        }
    }

    nxm_status init()
    {
        HIP_TRY(hipSetDevice(ctx_->device));
        cfg_ = default_config();
        HIP_TRY(p2g_start_.reserve(max_p2g_groups() + 1));
        HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
        for (auto &e : ev_) HIP_TRY(hipEventCreate(&e));
        HIP_TRY(g_.reserve(1));
        HIP_TRY(hipMemsetAsync(g_.p, 0, sizeof(Globals), stream_));
        HIP_TRY(hipStreamSynchronize(stream_));
        return NXM_OK;
    }

    hipStream_t stream() const { return stream_; }
    uint32_t particles() const { return p_in_; }
    size_t scenes() const
    {
        size_t c = 0;
        for (auto &s : slots_) c += s.state == SLOT_ACTIVE;
        return c;
    }

    nxm_status sync()
    {
        HIP_TRY(hipSetDevice(ctx_->device));
        HIP_TRY(hipStreamSynchronize(stream_));
        return collect_errors();
    }

    // Adds a scene whose state is on the host; uploads its particles.
    nxm_status add(nxm_scene *sc)
    {
        uint32_t slot = NONE;
        for (uint32_t i = 0; i < slots_.size(); ++i)
            if (slots_[i].state == SLOT_FREE) { slot = i; break; }
        if (slot == NONE) {
            if (slots_.size() >= MAX_SCENES) return fail(NXM_ERR_INVALID_ARGUMENT, "too many scenes in one world");
            slot = (uint32_t)slots_.size();
            slots_.push_back(Slot{});
        }
        slots_[slot].scene = sc;
        slots_[slot].state = SLOT_ACTIVE;
        sc->world = this;
        sc->slot = slot;
        sc->ranges.clear();
        dirty_ = true;
        nxm_status st = upload(sc, 0, sc->hp.n);
        if (st != NXM_OK) {
            // Nothing of the scene entered the arena (the particle count advances only
            // after its copies are queued): the slot is freed and the scene stays on the
            // host, detached, so that no table, step or destroy reaches it through here.
            slots_[slot].scene = nullptr;
            slots_[slot].state = SLOT_FREE;
            sc->world = nullptr;
            sc->ranges.clear();
            return st;
        }
        sc->on_host = false;
        return NXM_OK;
    }

    bool holds(const nxm_scene *sc) const
    {
        return sc->world == this && sc->slot < slots_.size() && slots_[sc->slot].scene == sc;
    }

    // Appends particles [first, first + count) of sc's host arrays to the input arena of
    // the next step: gathered into pinned staging and copied on the world's stream, after
    // any step in flight (which reads the other arena), without waiting for it.
    nxm_status upload(nxm_scene *sc, size_t first, size_t count)
    {
        if (count == 0) return NXM_OK;
        if (g_fail_uploads.load() > 0 && g_fail_uploads.fetch_sub(1) > 0)
            return fail(NXM_ERR_DEVICE, "injected upload failure (nxm_debug_fail_uploads)");
        clean_ = false;   // the new particles have no keys yet
        if (p_in_ + count > p_cap_) {
            nxm_status st = sync();   // the arenas are reallocated: nothing may use them
            if (st != NXM_OK) return st;
            st = ensure_particles(p_in_ + (uint32_t)count);
            if (st != NXM_OK) return st;
        }
        const HostParticles &hp = sc->hp;
        constexpr int NUP = NUM_PLANES;
        if (sc->staged_chunk >= 0) {
            if (first == 0 && count == sc->staged_n) {
                // Packed by nxm_scene_stage_upload: add the world slot to the flags and queue.
                const int k = sc->staged_chunk;
                float *f = reinterpret_cast<float *>(sc->ctx->pool.ptr(k));
                uint32_t *uf = reinterpret_cast<uint32_t *>(f + (size_t)PFLAGS * count);
                const uint32_t sbits = sc->slot << 16;
                for (size_t i = 0; i < count; ++i) uf[i] |= sbits;
                Arena a = arena(cur_);
                const uint32_t at = p_in_;
                for (int p = 0; p < NUP; ++p)
                    HIP_TRY(hipMemcpyAsync(a.plane(p) + at, f + (size_t)p * count, count * 4, hipMemcpyHostToDevice, stream_));
                sc->staged_chunk = -1;
                HIP_TRY(sc->ctx->pool.give(k, stream_));
                sc->ranges.push_back(SceneRange{at, (uint32_t)count});
                p_in_ += (uint32_t)count;
                pin_dirty_ = true;
                return NXM_OK;
            }
            const int k = sc->staged_chunk;   // stale: the particles changed since
            sc->staged_chunk = -1;
            HIP_TRY(sc->ctx->pool.give(k, nullptr));
        }
        char *stage = nullptr;
        size_t ci = 0;
        nxm_status st = pinned(NUP * count * 4, &stage, &ci);
        if (st != NXM_OK) return st;
        float *f = reinterpret_cast<float *>(stage);
        auto put = [&](int p, const float *src, int stride, int comp) {
            float *dst = f + (size_t)p * count;
            for (size_t i = 0; i < count; ++i) dst[i] = src[(first + i) * stride + comp];
        };
        for (int c = 0; c < 3; ++c) put(PX + c, hp.x.data(), 3, c);
        for (int c = 0; c < 3; ++c) put(PVX + c, hp.v.data(), 3, c);
        for (int c = 0; c < 9; ++c) put(PC0 + c, hp.C.data(), 9, c);
        for (int c = 0; c < 9; ++c) put(PF0 + c, hp.F.data(), 9, c);
        put(PMASS, hp.mass.data(), 1, 0);
        put(PVOL0, hp.vol0.data(), 1, 0);
        for (int c = 0; c < 3; ++c) put(PS0 + c, hp.pstate.data(), 3, c);
        std::memcpy(f + (size_t)PAFF * count, hp.aff.data() + first, count * 4);
        for (int c = 0; c < 4; ++c) put(PNX + c, hp.cdf.data(), 4, c);
        uint32_t *uf = reinterpret_cast<uint32_t *>(f + (size_t)PFLAGS * count);
        uint32_t *ui = reinterpret_cast<uint32_t *>(f + (size_t)PID * count);
        for (size_t i = 0; i < count; ++i) {
            uf[i] = make_flags(hp.enabled[first + i] != 0, hp.fixed[first + i] != 0, hp.cls[first + i], sc->slot);
            ui[i] = (uint32_t)(first + i);
        }
        Arena a = arena(cur_);
        const uint32_t at = p_in_;
        for (int p = 0; p < NUP; ++p)
            HIP_TRY(hipMemcpyAsync(a.plane(p) + at, f + (size_t)p * count, count * 4, hipMemcpyHostToDevice, stream_));
        HIP_TRY(hipEventRecord(pins_[ci].ev, stream_));
        sc->ranges.push_back(SceneRange{at, (uint32_t)count});
        p_in_ += (uint32_t)count;
        pin_dirty_ = true;
        return NXM_OK;
    }

    // Detaches a scene whose state is discarded (the scene is being destroyed): no copy.
    void drop(nxm_scene *sc)
    {
        slots_[sc->slot].state = SLOT_RETIRING;
        slots_[sc->slot].scene = nullptr;
        sc->world = nullptr;
        sc->ranges.clear();
        sc->dev_col_start = NONE;
        sc->cols_host = true;
        sc->clock_host = true;
        dirty_ = true;
    }

    // Copies the scene's state back to its host arrays (id order) and detaches it.
    nxm_status remove(nxm_scene *sc)
    {
        nxm_status st = pull(sc);
        if (st != NXM_OK) return st;
        sc->dev_col_start = NONE;
        sc->cols_host = true;
        sc->clock_host = true;
        slots_[sc->slot].state = SLOT_RETIRING;
        slots_[sc->slot].scene = nullptr;
        sc->world = nullptr;
        sc->ranges.clear();
        sc->on_host = true;
        dirty_ = true;
        return NXM_OK;
    }

    void mark_dirty() { dirty_ = true; }

    // Synchronizes and copies the scene's integrated collider poses to its host copy.
    nxm_status pull_colliders(nxm_scene *sc)
    {
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        if (!sc->cols_host && sc->dev_col_start != NONE && !sc->cols.empty())
            HIP_TRY(hipMemcpy(sc->cols.data(), cols_[cols_cur_].p + sc->dev_col_start, sc->cols.size() * sizeof(ColliderDev),
                              hipMemcpyDeviceToHost));
        return NXM_OK;
    }

    // Synchronizes and copies the scene's clock to its host copy.
    nxm_status pull_clock(nxm_scene *sc)
    {
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        if (!sc->clock_host && sc->world == this && sc->slot < clock_.n)
            HIP_TRY(hipMemcpy(&sc->clock, clock_.p + sc->slot, sizeof(SceneClock), hipMemcpyDeviceToHost));
        return NXM_OK;
    }

    // Reads the scene's particle state into its host arrays (id order).
    nxm_status pull(nxm_scene *sc)
    {
        nxm_status st = pull_colliders(sc);
        if (st == NXM_OK) st = pull_clock(sc);
        if (st != NXM_OK) return st;
        HostParticles &hp = sc->hp;
        Arena a = arena(cur_);
        std::vector<float> buf;
        std::vector<uint32_t> ids, fl;
        for (const SceneRange &r : sc->ranges) {
            buf.resize(r.count);
            ids.resize(r.count);
            fl.resize(r.count);
            HIP_TRY(hipMemcpy(ids.data(), a.uplane(PID) + r.start, r.count * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(fl.data(), a.uplane(PFLAGS) + r.start, r.count * 4, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < r.count; ++i) {
                hp.enabled[ids[i]] = (fl[i] & FLAG_ENABLED) ? 1u : 0u;
                hp.fixed[ids[i]] = (fl[i] & FLAG_FIXED) ? 1u : 0u;
            }
            auto get = [&](int p, float *dst, int stride, int comp) -> hipError_t {
                hipError_t e = hipMemcpy(buf.data(), a.plane(p) + r.start, r.count * 4, hipMemcpyDeviceToHost);
                if (e != hipSuccess) return e;
                for (uint32_t i = 0; i < r.count; ++i) dst[(size_t)ids[i] * stride + comp] = buf[i];
                return hipSuccess;
            };
            for (int c = 0; c < 3; ++c) HIP_TRY(get(PX + c, hp.x.data(), 3, c));
            for (int c = 0; c < 3; ++c) HIP_TRY(get(PVX + c, hp.v.data(), 3, c));
            for (int c = 0; c < 9; ++c) HIP_TRY(get(PC0 + c, hp.C.data(), 9, c));
            for (int c = 0; c < 9; ++c) HIP_TRY(get(PF0 + c, hp.F.data(), 9, c));
            HIP_TRY(get(PMASS, hp.mass.data(), 1, 0));
            HIP_TRY(get(PVOL0, hp.vol0.data(), 1, 0));
            // plastic state: the sand tags carry three words, snow one
            std::vector<float> ps(3 * (size_t)r.count);
            for (int c = 0; c < 3; ++c) {
                HIP_TRY(hipMemcpy(buf.data(), a.plane(PS0 + c) + r.start, r.count * 4, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < r.count; ++i) {
                    uint32_t id = ids[i];
                    if (c < pstate_words(hp.tag[id])) hp.pstate[3 * (size_t)id + c] = buf[i];
                }
            }
            if (sc->cpic) {
                HIP_TRY(hipMemcpy(buf.data(), a.plane(PAFF) + r.start, r.count * 4, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < r.count; ++i) std::memcpy(&hp.aff[ids[i]], &buf[i], 4);
                for (int c = 0; c < 4; ++c) HIP_TRY(get(PNX + c, hp.cdf.data(), 4, c));
                // An empty affinity means a zero contact field; the device does not store it.
                for (uint32_t i = 0; i < r.count; ++i)
                    if (hp.aff[ids[i]] == 0u) std::fill(&hp.cdf[4 * (size_t)ids[i]], &hp.cdf[4 * (size_t)ids[i] + 4], 0.0f);
            }
        }
        return NXM_OK;
    }

    nxm_status positions(nxm_scene *sc, float *xyz)
    {
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        Arena a = arena(cur_);
        std::vector<float> buf;
        std::vector<uint32_t> ids;
        for (const SceneRange &r : sc->ranges) {
            buf.resize(r.count);
            ids.resize(r.count);
            HIP_TRY(hipMemcpy(ids.data(), a.uplane(PID) + r.start, r.count * 4, hipMemcpyDeviceToHost));
            for (int c = 0; c < 3; ++c) {
                HIP_TRY(hipMemcpy(buf.data(), a.plane(PX + c) + r.start, r.count * 4, hipMemcpyDeviceToHost));
                for (uint32_t i = 0; i < r.count; ++i) xyz[3 * (size_t)ids[i] + c] = buf[i];
            }
        }
        return NXM_OK;
    }

    // Enqueues one substep of every active scene.
    nxm_status step(nxm_stage_times *times)
    {
        HIP_TRY(hipSetDevice(ctx_->device));
        if (dirty_) {
            nxm_status st = rebuild_tables();
            if (st != NXM_OK) return st;
        }
        if (pin_dirty_) {
            launch_set_particles(g_.p, p_in_, stream_);
            HIP_TRY(hipGetLastError());
            pin_dirty_ = false;
        }
        if (p_in_ == 0 || b_total_ == 0 || n_regions_ == 0) return NXM_OK;
        if (tune_on_ && !tuning_ && tuned_p_ && (p_in_ > 2 * tuned_p_ || 2 * p_in_ < tuned_p_)) start_tuning();
        const nxm_launch_config cfg = tuning_ ? tune_candidate() : cfg_;
        WorldDev w = view();
        w.p2g_groups = cfg.p2g_groups;
        w.p2g_block_cost = cfg.p2g_block_cost;
        LaunchConfig lc{};
        lc.p_hint = p_in_;
        lc.persistent_blocks = (int)cfg.grid_groups;
        lc.p2g_blocks = (int)cfg.p2g_groups;
        lc.side = side_;
        lc.fork = fork_;
        lc.join = join_;
        hipEvent_t *evs = times ? ev_ : nullptr;
        if (!times && profiling_ && profile_seen_++ % profile_every_ == 0) {
            nxm_status st = ring_events(&evs);
            if (st != NXM_OK) return st;
        }
        hipEvent_t tev[2] = {nullptr, nullptr};
        const bool tuning_step = tuning_ && !evs;
        if (tuning_step) {
            nxm_status st = tune_events(tev);
            if (st != NXM_OK) return st;
            if (tev[0]) HIP_TRY(hipEventRecord(tev[0], stream_));
        }
        // clean_: the input arena's block keys and insertions are in place, left by the
        // previous substep's G2P, and the counters and scan state are empty.  Otherwise
        // (first substep, arena or tables changed on the host) everything is emptied and
        // the substep starts with the key pass.
        const bool keys = !clean_;
        if (keys) {
            launch_clear(w, lc, stream_);
            HIP_TRY(hipGetLastError());
        }
        clean_ = false;
        if (graph_on_ && !evs && !tuning_ && !keys) {
            // Replays the substep's launches as one graph; recaptured when the launch
            // shapes or any buffer address change.  The graph holds kernels only.
            const GraphKey k{p_in_, b_total_, h_total_, gen_, cfg.p2g_groups, cfg.p2g_block_cost, cfg.grid_groups};
            if (!gexec_[cur_] || !(gkey_[cur_] == k)) {
                if (gexec_[cur_]) { HIP_TRY(hipGraphExecDestroy(gexec_[cur_])); gexec_[cur_] = nullptr; }
                hipGraph_t graph = nullptr;
                HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
                launch_substep(w, lc, stream_, nullptr, false);
                hipError_t le = hipGetLastError();
                hipError_t ce = hipStreamEndCapture(stream_, &graph);
                if (le != hipSuccess) { if (graph) (void)hipGraphDestroy(graph); return hip_fail(le, "launch_substep (capture)"); }
                HIP_TRY(ce);
                hipError_t ie = hipGraphInstantiate(&gexec_[cur_], graph, nullptr, nullptr, 0);
                (void)hipGraphDestroy(graph);
                HIP_TRY(ie);
                gkey_[cur_] = k;
            }
            HIP_TRY(hipGraphLaunch(gexec_[cur_], stream_));
        } else {
            launch_substep(w, lc, stream_, evs, keys);
        }
        HIP_TRY(hipGetLastError());
        clean_ = true;
        if (tuning_step) {
            if (tev[1]) HIP_TRY(hipEventRecord(tev[1], stream_));
            ++tune_substeps_;
            nxm_status st = tune_advance();
            if (st != NXM_OK) return st;
        }
        // The output arena holds every live particle, scenes in slot order.
        cur_ ^= 1;
        uint32_t at = 0;
        for (uint32_t i = 0; i < slots_.size(); ++i) {
            if (slots_[i].state == SLOT_RETIRING) { slots_[i].state = SLOT_FREE; dirty_ = true; }
            if (slots_[i].state != SLOT_ACTIVE) continue;
            nxm_scene *sc = slots_[i].scene;
            uint32_t n = (uint32_t)sc->hp.n;
            sc->ranges.assign(1, SceneRange{at, n});
            at += n;
        }
        if (at != p_in_) {
            p_in_ = at;
            pin_dirty_ = true;
        }
        // p_in on the device: the step's p_out, which equals p_in unless scenes retired.
        if (pin_dirty_) {
            HIP_TRY(hipMemcpyAsync(&g_.p->p_in, &g_.p->p_out, sizeof(uint32_t), hipMemcpyDeviceToDevice, stream_));
            pin_dirty_ = false;
        }
        stepped_ = true;
        if (times) {
            HIP_TRY(hipEventSynchronize(ev_[4]));
            nxm_status st = stage_times(ev_, times);
            if (st != NXM_OK) return st;
            return collect_errors();
        }
        return NXM_OK;
    }

    // ---- per-stage timing without synchronization (GpuTimestamps analog) ----
    // Events are recorded on one substep in `every` (event records between the kernels
    // cost GPU time of their own).
    void profile_begin(uint32_t every = 1)
    {
        profiling_ = true;
        ring_used_ = 0;
        profile_every_ = every ? every : 1u;
        profile_seen_ = 0;
    }

    // Stage and kernel times of one substep from its SUBSTEP_EVENTS events.
    static nxm_status stage_times(hipEvent_t *e, nxm_stage_times *t)
    {
        float ms[6] = {};
        for (int i = 0; i < 4; ++i) HIP_TRY(hipEventElapsedTime(&ms[i], e[i], e[i + 1]));
        HIP_TRY(hipEventElapsedTime(&ms[4], e[5], e[6]));
        HIP_TRY(hipEventElapsedTime(&ms[5], e[7], e[8]));
        t->sort_ms = ms[0];
        t->p2g_ms = ms[1];
        t->grid_update_ms = ms[2];
        t->g2p_update_ms = ms[3];
        t->total_ms = ms[0] + ms[1] + ms[2] + ms[3];
        t->p2g_kernel_ms = ms[4];
        t->g2p_update_kernel_ms = ms[5];
        return NXM_OK;
    }

    nxm_status profile_end(nxm_stage_times *avg, uint32_t *steps)
    {
        profiling_ = false;
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        double acc[7] = {};
        for (size_t k = 0; k < ring_used_; ++k) {
            nxm_stage_times t{};
            st = stage_times(&ring_[SUBSTEP_EVENTS * k], &t);
            if (st != NXM_OK) return st;
            const float v[7] = {t.sort_ms, t.p2g_ms, t.grid_update_ms, t.g2p_update_ms, t.total_ms,
                                t.p2g_kernel_ms, t.g2p_update_kernel_ms};
            for (int i = 0; i < 7; ++i) acc[i] += v[i];
        }
        const double n = ring_used_ ? (double)ring_used_ : 1.0;
        if (avg) {
            avg->sort_ms = (float)(acc[0] / n);
            avg->p2g_ms = (float)(acc[1] / n);
            avg->grid_update_ms = (float)(acc[2] / n);
            avg->g2p_update_ms = (float)(acc[3] / n);
            avg->total_ms = (float)(acc[4] / n);
            avg->p2g_kernel_ms = (float)(acc[5] / n);
            avg->g2p_update_kernel_ms = (float)(acc[6] / n);
        }
        if (steps) *steps = (uint32_t)ring_used_;
        ring_used_ = 0;
        return NXM_OK;
    }

    // ---- CFL timestep bound of every scene of the world (in 1e-12 s units) ----
    nxm_status timestep_bound(nxm_scene *sc, uint32_t *units)
    {
        HIP_TRY(hipSetDevice(ctx_->device));
        if (dirty_) {
            nxm_status st = rebuild_tables();
            if (st != NXM_OK) return st;
        }
        *units = 0xFFFFFFFFu;
        if (p_in_ == 0) return NXM_OK;
        DevBuf<uint32_t> out;
        HIP_TRY(out.reserve(slots_.size()));
        HIP_TRY(hipMemsetD32Async(out.p, 0xFFFFFFFF, slots_.size(), stream_));
        launch_timestep_bound(view(), p_in_, out.p, 1u, ctx_->num_cus * 8, stream_);
        HIP_TRY(hipGetLastError());
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        HIP_TRY(hipMemcpy(units, out.p + sc->slot, sizeof(uint32_t), hipMemcpyDeviceToHost));
        return NXM_OK;
    }

    // Render data of the scene's particles (gpu_prep_readback), in id order.
    nxm_status prep_readback(nxm_scene *sc, uint32_t mode, const float *palette, uint32_t npal, float *out)
    {
        HIP_TRY(hipSetDevice(ctx_->device));
        if (dirty_) {
            nxm_status st = rebuild_tables();
            if (st != NXM_OK) return st;
        }
        const size_t n = sc->hp.n;
        if (n == 0) return NXM_OK;
        const HostParticles &hp = sc->hp;
        DevBuf<uint32_t> group;
        DevBuf<float> props, dout;
        DevBuf<float4> pal;
        HIP_TRY(group.reserve(n));
        HIP_TRY(props.reserve(2 * n));
        HIP_TRY(pal.reserve(npal));
        HIP_TRY(dout.reserve(24 * n));
        nxm_status st = put_async(group.p, hp.group.data(), n * 4);
        if (st == NXM_OK) st = put_async(props.p, hp.phase.data(), n * 4);
        if (st == NXM_OK) st = put_async(props.p + n, hp.radius.data(), n * 4);
        if (st == NXM_OK) st = put_async(pal.p, palette, (size_t)npal * 16);
        if (st != NXM_OK) return st;
        for (const SceneRange &r : sc->ranges)
            launch_prep_readback(arena(cur_), r.start, r.count, group.p, props.p, props.p + n, pal.p, npal, mode, sc->h,
                                 sc->dt, sc->cpic, dout.p, (int)ctx_->num_cus * 8, stream_);
        HIP_TRY(hipGetLastError());
        st = sync();
        if (st != NXM_OK) return st;
        HIP_TRY(hipMemcpy(out, dout.p, 24 * n * sizeof(float), hipMemcpyDeviceToHost));
        return NXM_OK;
    }

    // ---- inspection of the last step ----
    nxm_status sort_size(nxm_scene *sc, size_t *nb)
    {
        nxm_status st = sync();
        if (st != NXM_OK) return st;
        *nb = 0;
        if (!stepped_) return NXM_OK;
        SceneCtr c;
        HIP_TRY(hipMemcpy(&c, ctr_.p + sc->slot, sizeof(c), hipMemcpyDeviceToHost));
        *nb = c.b_count > 0 ? c.b_count - 1 : 0;   // the overflow block is not reported
        last_b_start_ = c.b_start;
        return NXM_OK;
    }

    nxm_status canonical(nxm_scene *sc, const nxm_sort_canonical *out)
    {
        size_t nb = 0;
        nxm_status st = sort_size(sc, &nb);
        if (st != NXM_OK) return st;
        if (!stepped_) return fail(NXM_ERR_INVALID_ARGUMENT, "no step has run");
        const uint32_t b0 = last_b_start_;
        DevBuf<uint32_t> pcell, buckets;
        HIP_TRY(pcell.reserve(p_cap_));
        HIP_TRY(buckets.reserve((size_t)10 * b_total_));
        HIP_TRY(hipMemsetAsync(buckets.p, 0, sizeof(uint32_t) * 10 * b_total_, stream_));
        WorldDev w = view();
        // The last step read arena cur_^1 (pre-step positions) and wrote arena cur_.
        launch_canonical(w, arena(cur_ ^ 1), pcell.p, buckets.p, p_cap_, stream_);
        HIP_TRY(hipGetLastError());
        HIP_TRY(hipStreamSynchronize(stream_));
        std::vector<int4> vid(nb);
        std::vector<uint32_t> np(nb), bk(10 * nb);
        if (nb) {
            HIP_TRY(hipMemcpy(vid.data(), bvid_.p + b0, nb * sizeof(int4), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(np.data(), bnp_.p + b0, nb * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(bk.data(), buckets.p + 10 * (size_t)b0, nb * 40, hipMemcpyDeviceToHost));
        }
        for (size_t b = 0; b < nb; ++b) {
            if (out->block_vid) { out->block_vid[3 * b] = vid[b].x; out->block_vid[3 * b + 1] = vid[b].y; out->block_vid[3 * b + 2] = vid[b].z; }
            if (out->block_np) out->block_np[b] = np[b];
            uint32_t tot = 0;
            for (int k = 0; k < 10; ++k) {
                tot += bk[10 * b + k];
                if (out->block_buckets) out->block_buckets[10 * b + k] = bk[10 * b + k];
            }
            if (out->block_npe) out->block_npe[b] = tot;
        }
        // particles of the scene: its output range
        for (const SceneRange &r : sc->ranges) {
            std::vector<uint32_t> ids(r.count), sb(r.count), pc(r.count);
            HIP_TRY(hipMemcpy(ids.data(), arena(cur_).uplane(PID) + r.start, r.count * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(sb.data(), sblock_.p + r.start, r.count * 4, hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(pc.data(), pcell.p + r.start, r.count * 4, hipMemcpyDeviceToHost));
            for (uint32_t i = 0; i < r.count; ++i) {
                if (out->particle_block) out->particle_block[ids[i]] = sb[i] - b0;
                if (out->particle_cell) out->particle_cell[ids[i]] = pc[i];
            }
        }
        return NXM_OK;
    }

    nxm_status grid(nxm_scene *sc, float *mass, float *mom, float *vel)
    {
        size_t nb = 0;
        nxm_status st = sort_size(sc, &nb);
        if (st != NXM_OK) return st;
        const uint32_t b0 = last_b_start_;
        std::vector<float4> gv(nb * 64);
        if (nb) HIP_TRY(hipMemcpy(gv.data(), gridv_.p + (size_t)b0 * 64, nb * 64 * sizeof(float4), hipMemcpyDeviceToHost));
        for (size_t i = 0; i < nb * 64; ++i) {
            if (mass) mass[i] = gv[i].w;
            if (vel) { vel[3 * i] = gv[i].x; vel[3 * i + 1] = gv[i].y; vel[3 * i + 2] = gv[i].z; }
        }
        if (mom) {
            if (!dbg_mom_.p) return fail(NXM_ERR_INVALID_ARGUMENT, "grid momentum needs nxm_scene_set_debug_grid before the step");
            HIP_TRY(hipMemcpy(mom, dbg_mom_.p + (size_t)b0 * 64 * 3, nb * 64 * 3 * sizeof(float), hipMemcpyDeviceToHost));
        }
        return NXM_OK;
    }

    nxm_status grid_cpic(nxm_scene *sc, float *dist, uint32_t *aff, uint32_t *closest, float *mass_i, float *vel_i)
    {
        if (!sc->cpic) return fail(NXM_ERR_INVALID_ARGUMENT, "the scene does not use CPIC");
        size_t nb = 0;
        nxm_status st = sort_size(sc, &nb);
        if (st != NXM_OK) return st;
        const uint32_t b0 = last_b_start_;
        std::vector<uint4> nc(nb * 64);
        std::vector<float4> gv(nb * 64);
        if (nb) {
            HIP_TRY(hipMemcpy(nc.data(), ncdf_.p + (size_t)b0 * 64, nb * 64 * sizeof(uint4), hipMemcpyDeviceToHost));
            HIP_TRY(hipMemcpy(gv.data(), gridv_i_.p + (size_t)b0 * 64, nb * 64 * sizeof(float4), hipMemcpyDeviceToHost));
        }
        for (size_t i = 0; i < nb * 64; ++i) {
            if (dist) std::memcpy(&dist[i], &nc[i].x, 4);
            if (aff) aff[i] = nc[i].y;
            if (closest) closest[i] = nc[i].z;
            if (mass_i) mass_i[i] = gv[i].w;
            if (vel_i) { vel_i[3 * i] = gv[i].x; vel_i[3 * i + 1] = gv[i].y; vel_i[3 * i + 2] = gv[i].z; }
        }
        return NXM_OK;
    }

  private:
    // Pinned host staging, reused once the copies that read it have executed.
    struct PinnedChunk {
        char *p = nullptr;
        size_t n = 0;
        hipEvent_t ev = nullptr;
    };

    nxm_status pinned(size_t bytes, char **out, size_t *idx)
    {
        for (size_t k = 0; k < pins_.size(); ++k) {
            if (pins_[k].n < bytes) continue;
            const hipError_t q = hipEventQuery(pins_[k].ev);
            if (q == hipErrorNotReady) continue;
            HIP_TRY(q);
            *out = pins_[k].p;
            *idx = k;
            return NXM_OK;
        }
        PinnedChunk c;
        c.n = std::max<size_t>(bytes, (size_t)1 << 20);
        HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&c.p), c.n, hipHostMallocDefault));
        hipError_t e = hipEventCreate(&c.ev);
        if (e != hipSuccess) { (void)hipHostFree(c.p); return hip_fail(e, "hipEventCreate"); }
        pins_.push_back(c);
        *out = c.p;
        *idx = pins_.size() - 1;
        return NXM_OK;
    }

    // Copies host bytes to a device buffer on the world's stream through pinned staging.
    nxm_status put_async(void *dst, const void *src, size_t bytes)
    {
        if (bytes == 0) return NXM_OK;
        char *stage = nullptr;
        size_t ci = 0;
        nxm_status st = pinned(bytes, &stage, &ci);
        if (st != NXM_OK) return st;
        std::memcpy(stage, src, bytes);
        HIP_TRY(hipMemcpyAsync(dst, stage, bytes, hipMemcpyHostToDevice, stream_));
        HIP_TRY(hipEventRecord(pins_[ci].ev, stream_));
        return NXM_OK;
    }

    nxm_status ring_events(hipEvent_t **out)
    {
        if (SUBSTEP_EVENTS * (ring_used_ + 1) > ring_.size()) {
            size_t add = std::max<size_t>(SUBSTEP_EVENTS * 64, ring_.size());
            for (size_t i = 0; i < add; ++i) {
                hipEvent_t e;
                HIP_TRY(hipEventCreate(&e));
                ring_.push_back(e);
            }
        }
        *out = &ring_[SUBSTEP_EVENTS * ring_used_];
        ++ring_used_;
        return NXM_OK;
    }

    Arena arena(int k) const { return Arena{arenas_[k].p, p_cap_}; }

    WorldDev view() const
    {
        WorldDev w{};
        w.in = arena(cur_);
        w.out = arena(cur_ ^ 1);
        w.pkey = pkey_.p; w.pdense = pdense_.p; w.prank = prank_.p; w.perm = perm_.p; w.sblock = sblock_.p;
        w.bscene = bscene_.p; w.bvid = bvid_.p; w.bnp = bnp_.p; w.bfirst = bfirst_.p;
        w.bup = bup_.p; w.bdn = bdn_.p; w.stage = stage_.p; w.gridv = gridv_.p;
        w.hkeys = hkeys_.p; w.hvals = hvals_.p; w.hmark = hmark_.p; w.bprim = bprim_.p; w.bmask = bmask_.p;
        w.region_start = region_start_.p; w.region_scene = region_scene_.p; w.n_regions = n_regions_;
        w.scenes = scenes_.p; w.ctr = ctr_.p; w.classes = classes_.p; w.cols = cols_[cols_cur_].p;
        w.g = g_.p; w.scan_state = scan_state_.p; w.scan_stride = scan_stride_;
        w.p2g_start = p2g_start_.p; w.p2g_groups = cfg_.p2g_groups; w.p2g_block_cost = cfg_.p2g_block_cost;
        w.dbg_mom = debug_ ? dbg_mom_.p : nullptr;
        w.ncdf = ncdf_.p; w.nbody = nbody_.p; w.binc = binc_.p; w.bany = bany_.p; w.breach = any_cpic_ ? breach_.p : nullptr; w.clist = clist_.p; w.stage_i = stage_i_.p; w.gridv_i = gridv_i_.p; w.any_cpic = any_cpic_ ? 1u : 0u;
        w.mesh_v = mesh_v_.p; w.mesh_t = mesh_t_.p; w.mesh_rp = mesh_rp_.p; w.mkey = mkey_.p; w.maff = maff_.p;
        w.mesh_vn = mesh_vn_.p; w.mesh_en = mesh_en_.p; w.mesh_grid = mesh_grid_.p; w.mesh_cell = mesh_cell_.p;
        w.mesh_ctri = mesh_ctri_.p; w.col_mesh = col_mesh_.p; w.mesh_pass = mesh_pass_ ? 1u : 0u;
        w.n_rp = n_rp_;
        w.p_cap = p_cap_; w.b_total = b_total_; w.h_total = h_total_; w.s_count = (uint32_t)slots_.size();
        w.clock = clock_.p; w.any_adaptive = any_adaptive_ ? 1u : 0u;
        return w;
    }

    nxm_status ensure_particles(uint32_t need)
    {
        if (need <= p_cap_) return NXM_OK;
        uint32_t cap = std::max<uint32_t>(need, p_cap_ + p_cap_ / 2);
        cap = (cap + 255u) & ~255u;
        DevBuf<float> na[2];
        HIP_TRY(na[0].reserve((size_t)NUM_PLANES * cap));
        HIP_TRY(na[1].reserve((size_t)NUM_PLANES * cap));
        if (p_in_) {
            for (int p = 0; p < NUM_PLANES; ++p)
                HIP_TRY(hipMemcpy(na[cur_].p + (size_t)p * cap, arenas_[cur_].p + (size_t)p * p_cap_, (size_t)p_in_ * 4,
                                  hipMemcpyDeviceToDevice));
        }
        for (int k = 0; k < 2; ++k) {
            arenas_[k].release();
            std::swap(arenas_[k].p, na[k].p);
            std::swap(arenas_[k].n, na[k].n);
        }
        p_cap_ = cap;
        ++gen_;
        HIP_TRY(pkey_.reserve(cap));
        HIP_TRY(pdense_.reserve(cap));
        HIP_TRY(prank_.reserve(cap));
        HIP_TRY(perm_.reserve(cap));
        HIP_TRY(sblock_.reserve(cap));
        HIP_TRY(clist_.reserve(cap));
        return NXM_OK;
    }

    // Lays out block ids, hash regions, classes and colliders of the active scenes
    // in slot order and uploads the tables.
    nxm_status rebuild_tables()
    {
        layout_.assign(slots_.size(), SceneDev{});
        std::vector<ClassDev> cls;
        uint32_t ncols = 0;
        std::vector<uint32_t> region_start, region_scene;
        uint32_t b = 0, hsum = 0;
        bool any_cpic = false, any_adaptive = false;
        std::vector<float4> mv, mrp, mvn, men;
        std::vector<uint4> mt;
        std::vector<MeshGridDev> mgrid;
        std::vector<uint32_t> mcell, mctri, col_mesh;
        for (uint32_t i = 0; i < slots_.size(); ++i) {
            SceneDev &L = layout_[i];
            if (slots_[i].state != SLOT_ACTIVE) continue;
            nxm_scene *sc = slots_[i].scene;
            L.h = sc->h;
            L.inv_h = 1.0f / sc->h;
            L.dt = sc->dt;
            L.vlim = sc->h / sc->dt;
            L.gx = sc->g[0]; L.gy = sc->g[1]; L.gz = sc->g[2];
            L.inv_d = 4.0f / (sc->h * sc->h);
            L.active = 1;
            L.b_cap = sc->grid_cap;
            b += sc->grid_cap + 1;
            L.h_base = hsum;
            L.h_cap = next_pow2(2 * sc->grid_cap);
            hsum += L.h_cap + 1;
            region_start.push_back(L.h_base);
            region_scene.push_back(i);
            L.col_start = ncols;
            L.col_count = (uint32_t)sc->cols.size();
            ncols += L.col_count;
            L.class_base = (uint32_t)cls.size();
            cls.insert(cls.end(), sc->classes.begin(), sc->classes.end());
            L.cpic = sc->cpic ? 1u : 0u;
            any_cpic |= sc->cpic;
            L.dt_max = sc->dt;
            L.adaptive = sc->adaptive ? 1u : 0u;
            any_adaptive |= sc->adaptive;
            // Mesh colliders: geometry appended to the world's tables; CPIC scenes add
            // their rigid particles sampled at the cell width (src_mpm/pipeline.rs:
            // 310-316), the others the pseudo-normals and cell grid of the node pass.
            const uint32_t nmesh = (uint32_t)sc->meshes.size();
            col_mesh.resize(ncols, NONE);
            for (uint32_t k = 0; k < nmesh; ++k) {
                HostMesh &m = sc->meshes[k];
                const uint32_t v0 = (uint32_t)mv.size(), t0 = (uint32_t)mt.size();
                const uint32_t cid = L.col_count - nmesh + k;
                for (size_t q = 0; q < m.v.size() / 3; ++q) mv.push_back(make_float4(m.v[3 * q], m.v[3 * q + 1], m.v[3 * q + 2], 0.0f));
                for (size_t q = 0; q < m.tri.size() / 3; ++q)
                    mt.push_back(make_uint4(v0 + m.tri[3 * q], v0 + m.tri[3 * q + 1], v0 + m.tri[3 * q + 2], (i << 4) | cid));
                if (sc->cpic) {
                    if (m.sampled_h != sc->h) {
                        m.samples = sample_mesh_surface(m.v.data(), m.tri.data(), (uint32_t)(m.tri.size() / 3), sc->h);
                        m.sampled_h = sc->h;
                    }
                    for (size_t q = 0; q < m.samples.tri.size(); ++q) {
                        const uint32_t tid = t0 + m.samples.tri[q];
                        float tf;
                        std::memcpy(&tf, &tid, 4);
                        mrp.push_back(make_float4(m.samples.pts[3 * q], m.samples.pts[3 * q + 1], m.samples.pts[3 * q + 2], tf));
                    }
                    continue;
                }
                mvn.resize(v0, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                men.resize(3 * (size_t)t0, make_float4(0.0f, 0.0f, 0.0f, 0.0f));
                for (size_t q = 0; q < m.v.size() / 3; ++q) mvn.push_back(make_float4(m.vn[3 * q], m.vn[3 * q + 1], m.vn[3 * q + 2], 0.0f));
                for (size_t q = 0; q < 3 * (m.tri.size() / 3); ++q) men.push_back(make_float4(m.en[3 * q], m.en[3 * q + 1], m.en[3 * q + 2], 0.0f));
                mesh_grid(m, sc->h);
                MeshGridDev gd;
                gd.ox = m.origin[0]; gd.oy = m.origin[1]; gd.oz = m.origin[2]; gd.inv_cell = 1.0f / m.cell;
                gd.nx = m.dims[0]; gd.ny = m.dims[1]; gd.nz = m.dims[2];
                gd.cell_base = (uint32_t)mcell.size();
                const uint32_t c0 = (uint32_t)mctri.size();
                for (uint32_t st0 : m.start) mcell.push_back(c0 + st0);
                for (uint32_t ct : m.ctri) mctri.push_back(t0 + ct);
                col_mesh[L.col_start + cid] = (uint32_t)mgrid.size();
                mgrid.push_back(gd);
            }
        }
        col_mesh.resize(ncols, NONE);
        const bool any_mesh = !mrp.empty();   // rigid particles of CPIC scenes' meshes
        const bool any_geom = !mt.empty();    // mesh geometry of any scene
        // Dense block ids are ranks of occupied hash slots: sizing the block tables by
        // the slot count keeps every rank (and each scene's overflow block) in range
        // even when a scene exceeds its capacity (reported as an error).
        (void)b;
        // Buffers a step in flight may still use are reallocated only after it finished;
        // otherwise the new tables are copied behind it on the stream.
        // One look-back word per scan tile (k_scan), for each of the two scans.
        const uint32_t scan_stride = (hsum + SCAN_TILE - 1u) / SCAN_TILE + 2u;
        const uint32_t nreg = (uint32_t)region_start.size();
        const bool grow = std::max<size_t>(slots_.size(), 1) > scenes_.n || std::max<size_t>(slots_.size(), 1) > ctr_.n ||
                          std::max<size_t>(cls.size(), 1) > classes_.n || std::max<size_t>(ncols, 1) > cols_[0].n || std::max<size_t>(ncols, 1) > cols_[1].n ||
                          hsum > bscene_.n || 2 * (size_t)scan_stride > scan_state_.n ||
                          std::max<uint32_t>(hsum, 1) > hkeys_.n || std::max<uint32_t>(nreg, 1) > region_start_.n ||
                          (debug_ && (size_t)NODES_PER_BLOCK * 3 * std::max<size_t>(hsum, 1) > dbg_mom_.n) ||
                          (any_cpic && (size_t)NODES_PER_BLOCK * std::max<size_t>(hsum, 1) > ncdf_.n) ||
                          (any_mesh && ((size_t)NODES_PER_BLOCK * std::max<size_t>(hsum, 1) > mkey_.n ||
                                        mrp.size() > mesh_rp_.n)) ||
                          (any_geom && (mv.size() > mesh_v_.n || mt.size() > mesh_t_.n || mvn.size() > mesh_vn_.n ||
                                        men.size() > mesh_en_.n || mgrid.size() > mesh_grid_.n ||
                                        mcell.size() > mesh_cell_.n || mctri.size() > mesh_ctri_.n ||
                                        col_mesh.size() > col_mesh_.n));
        if (grow) {
            nxm_status st = sync();
            if (st != NXM_OK) return st;
            // The collider tables may be reallocated: keep the integrated poses on the host.
            for (uint32_t i = 0; i < slots_.size(); ++i) {
                if (slots_[i].state != SLOT_ACTIVE) continue;
                nxm_scene *sc = slots_[i].scene;
                st = pull_colliders(sc);
                if (st != NXM_OK) return st;
                sc->cols_host = true;
            }
        }
        h_total_ = hsum;
        b_total_ = hsum;
        HIP_TRY(scenes_.reserve(std::max<size_t>(slots_.size(), 1)));
        HIP_TRY(ctr_.reserve(std::max<size_t>(slots_.size(), 1)));
        if (ctr_.n > ctr_zeroed_) {
            HIP_TRY(hipMemsetAsync(ctr_.p, 0, ctr_.n * sizeof(SceneCtr), stream_));
            ctr_zeroed_ = ctr_.n;
        }
        HIP_TRY(classes_.reserve(std::max<size_t>(cls.size(), 1)));
        HIP_TRY(cols_[0].reserve(std::max<size_t>(ncols, 1)));
        HIP_TRY(cols_[1].reserve(std::max<size_t>(ncols, 1)));
        nxm_status st = put_async(scenes_.p, layout_.data(), layout_.size() * sizeof(SceneDev));
        if (st == NXM_OK) st = put_async(classes_.p, cls.data(), cls.size() * sizeof(ClassDev));
        if (st != NXM_OK) return st;
        // The new collider table: scenes resident before keep their integrated poses (copied
        // on the stream from the current table), the others take their host copy.
        const int nc = cols_cur_ ^ 1;
        for (uint32_t i = 0; i < slots_.size(); ++i) {
            if (slots_[i].state != SLOT_ACTIVE) continue;
            nxm_scene *sc = slots_[i].scene;
            const SceneDev &L = layout_[i];
            if (L.col_count == 0) { sc->dev_col_start = L.col_start; sc->cols_host = false; continue; }
            ColliderDev *dst = cols_[nc].p + L.col_start;
            if (sc->cols_host || sc->dev_col_start == NONE) {
                st = put_async(dst, sc->cols.data(), L.col_count * sizeof(ColliderDev));
                if (st != NXM_OK) return st;
            } else {
                HIP_TRY(hipMemcpyAsync(dst, cols_[cols_cur_].p + sc->dev_col_start, L.col_count * sizeof(ColliderDev),
                                       hipMemcpyDeviceToDevice, stream_));
            }
            sc->dev_col_start = L.col_start;
            sc->cols_host = false;
        }
        cols_cur_ = nc;
        // Clocks stay at their slot while the scene is resident (a larger table takes the
        // old one's contents on the stream); scenes new to the world bring theirs.
        if (slots_.size() > clock_.n) {
            DevBuf<SceneClock> grown;
            HIP_TRY(grown.reserve(std::max<size_t>(2 * slots_.size(), 64)));
            if (clock_.n)
                HIP_TRY(hipMemcpyAsync(grown.p, clock_.p, clock_.n * sizeof(SceneClock), hipMemcpyDeviceToDevice, stream_));
            HIP_TRY(hipStreamSynchronize(stream_));   // the old table may still be in use
            std::swap(clock_.p, grown.p);
            std::swap(clock_.n, grown.n);
        }
        for (uint32_t i = 0; i < slots_.size(); ++i) {
            if (slots_[i].state != SLOT_ACTIVE) continue;
            nxm_scene *sc = slots_[i].scene;
            if (!sc->clock_host) continue;
            st = put_async(clock_.p + i, &sc->clock, sizeof(SceneClock));
            if (st != NXM_OK) return st;
            sc->clock_host = false;
        }
        any_adaptive_ = any_adaptive;
        if (b_total_ > bscene_.n) {
            size_t nb = (size_t)b_total_ + b_total_ / 4;
            HIP_TRY(bscene_.reserve(nb));
            HIP_TRY(bprim_.reserve(nb));
            HIP_TRY(bmask_.reserve(nb));
            HIP_TRY(bvid_.reserve(nb));
            HIP_TRY(bnp_.reserve(nb));
            HIP_TRY(bfirst_.reserve(nb));
            HIP_TRY(bup_.reserve(8 * nb));
            HIP_TRY(bdn_.reserve(8 * nb));
            HIP_TRY(stage_.reserve((size_t)4 * TILE_NODES * nb));
            HIP_TRY(gridv_.reserve((size_t)NODES_PER_BLOCK * nb));
        }
        scan_stride_ = scan_stride;
        HIP_TRY(scan_state_.reserve(2 * (size_t)scan_stride_));
        if (debug_) HIP_TRY(dbg_mom_.reserve((size_t)NODES_PER_BLOCK * 3 * std::max<size_t>(b_total_, 1)));
        if (any_cpic && (size_t)NODES_PER_BLOCK * std::max<size_t>(b_total_, 1) > ncdf_.n) {
            const size_t nb = std::max<size_t>(b_total_, 1) + b_total_ / 4;
            HIP_TRY(ncdf_.reserve((size_t)NODES_PER_BLOCK * nb));
            HIP_TRY(nbody_.reserve((size_t)NODES_PER_BLOCK * nb));
            HIP_TRY(binc_.reserve(nb));
            HIP_TRY(bany_.reserve(nb));
            HIP_TRY(breach_.reserve(nb));
            HIP_TRY(stage_i_.reserve((size_t)4 * TILE_NODES * nb));
            HIP_TRY(gridv_i_.reserve((size_t)NODES_PER_BLOCK * nb));
        }
        any_cpic_ = any_cpic;
        if (any_cpic && !side_) {
            // Created with the first CPIC scene only: streams are spread over the device's
            // hardware queues in creation order, and worlds without CPIC keep one stream each.
            HIP_TRY(hipStreamCreateWithFlags(&side_, hipStreamNonBlocking));
            HIP_TRY(hipEventCreateWithFlags(&fork_, hipEventDisableTiming));
            HIP_TRY(hipEventCreateWithFlags(&join_, hipEventDisableTiming));
        }
        if (any_mesh) {
            // k_grid_cdf empties every node k_mesh_cdf wrote in the same substep; fresh
            // allocations are emptied here, outside any captured graph.
            const size_t nodes = (size_t)NODES_PER_BLOCK * (std::max<size_t>(b_total_, 1) + b_total_ / 4);
            if (nodes > mkey_.n) {
                HIP_TRY(mkey_.reserve(nodes));
                HIP_TRY(maff_.reserve(nodes));
                mesh_nodes_clear_ = 0;
            }
            if (mesh_nodes_clear_ < mkey_.n) {
                HIP_TRY(hipMemsetAsync(mkey_.p, 0xFF, mkey_.n * sizeof(unsigned long long), stream_));
                HIP_TRY(hipMemsetAsync(maff_.p, 0, maff_.n * sizeof(uint32_t), stream_));
                mesh_nodes_clear_ = mkey_.n;
            }
            HIP_TRY(mesh_rp_.reserve(mrp.size()));
            nxm_status ms = put_async(mesh_rp_.p, mrp.data(), mrp.size() * sizeof(float4));
            if (ms != NXM_OK) return ms;
        }
        if (any_geom) {
            HIP_TRY(mesh_v_.reserve(mv.size()));
            HIP_TRY(mesh_t_.reserve(mt.size()));
            HIP_TRY(mesh_vn_.reserve(std::max<size_t>(mvn.size(), 1)));
            HIP_TRY(mesh_en_.reserve(std::max<size_t>(men.size(), 1)));
            HIP_TRY(mesh_grid_.reserve(std::max<size_t>(mgrid.size(), 1)));
            HIP_TRY(mesh_cell_.reserve(std::max<size_t>(mcell.size(), 1)));
            HIP_TRY(mesh_ctri_.reserve(std::max<size_t>(mctri.size(), 1)));
            HIP_TRY(col_mesh_.reserve(std::max<size_t>(col_mesh.size(), 1)));
            nxm_status ms = put_async(mesh_v_.p, mv.data(), mv.size() * sizeof(float4));
            if (ms == NXM_OK) ms = put_async(mesh_t_.p, mt.data(), mt.size() * sizeof(uint4));
            if (ms == NXM_OK && !mvn.empty()) ms = put_async(mesh_vn_.p, mvn.data(), mvn.size() * sizeof(float4));
            if (ms == NXM_OK && !men.empty()) ms = put_async(mesh_en_.p, men.data(), men.size() * sizeof(float4));
            if (ms == NXM_OK && !mgrid.empty()) ms = put_async(mesh_grid_.p, mgrid.data(), mgrid.size() * sizeof(MeshGridDev));
            if (ms == NXM_OK && !mcell.empty()) ms = put_async(mesh_cell_.p, mcell.data(), mcell.size() * sizeof(uint32_t));
            if (ms == NXM_OK && !mctri.empty()) ms = put_async(mesh_ctri_.p, mctri.data(), mctri.size() * sizeof(uint32_t));
            if (ms == NXM_OK && !col_mesh.empty()) ms = put_async(col_mesh_.p, col_mesh.data(), col_mesh.size() * sizeof(uint32_t));
            if (ms != NXM_OK) return ms;
        }
        n_rp_ = (uint32_t)mrp.size();
        mesh_pass_ = !mgrid.empty();
        HIP_TRY(hkeys_.reserve(std::max<uint32_t>(h_total_, 1)));
        HIP_TRY(hvals_.reserve(std::max<uint32_t>(h_total_, 1)));
        HIP_TRY(hmark_.reserve(std::max<uint32_t>(h_total_, 1)));
        n_regions_ = (uint32_t)region_start.size();
        HIP_TRY(region_start_.reserve(std::max<uint32_t>(n_regions_, 1)));
        HIP_TRY(region_scene_.reserve(std::max<uint32_t>(n_regions_, 1)));
        st = put_async(region_start_.p, region_start.data(), n_regions_ * 4);
        if (st == NXM_OK) st = put_async(region_scene_.p, region_scene.data(), n_regions_ * 4);
        if (st != NXM_OK) return st;
        dirty_ = false;
        clean_ = false;
        ++gen_;
        return NXM_OK;
    }

    nxm_status collect_errors()
    {
        // Counters exist for the slots of the last table rebuild; a scene added since
        // then has not run yet.
        const size_t nslots = std::min(slots_.size(), ctr_.n);
        if (nslots == 0 || !ctr_.p) return NXM_OK;
        std::vector<SceneCtr> c(nslots);
        HIP_TRY(hipMemcpy(c.data(), ctr_.p, c.size() * sizeof(SceneCtr), hipMemcpyDeviceToHost));
        nxm_status first = NXM_OK;
        bool any = false;
        for (uint32_t i = 0; i < nslots; ++i) {
            if (!c[i].err) continue;
            any = true;
            if (slots_[i].scene) slots_[i].scene->error |= c[i].err;
        }
        if (any) {
            HIP_TRY(hipMemset(ctr_.p, 0, c.size() * sizeof(SceneCtr)));
            for (uint32_t i = 0; i < nslots; ++i) {
                if (!c[i].err) continue;
                if (first == NXM_OK) {
                    if (c[i].err & ERR_GRID_CAPACITY)
                        first = fail(NXM_ERR_GRID_CAPACITY, "active blocks exceed the grid capacity (scene slot " + std::to_string(i) + ")");
                    else if (c[i].err & ERR_DOMAIN)
                        first = fail(NXM_ERR_DOMAIN, "particle block outside the packable domain (scene slot " + std::to_string(i) + ")");
                    else
                        first = fail(NXM_ERR_DEVICE, "device error bits " + std::to_string(c[i].err));
                }
            }
        }
        return first;
    }

  public:
    void set_debug(bool on)
    {
        if (on && !debug_) dirty_ = true;
        debug_ = on;
        ++gen_;
    }

    nxm_status set_graph(bool on)
    {
        if (!on) {
            nxm_status st = sync();
            if (st != NXM_OK) return st;
            drop_graphs();
        }
        graph_on_ = on;
        return NXM_OK;
    }

  private:
    struct GraphKey {
        uint32_t p_in, b_total, h_total;
        uint64_t gen;
        uint32_t p2g_groups, p2g_block_cost, grid_groups;
        bool operator==(const GraphKey &o) const
        {
            return p_in == o.p_in && b_total == o.b_total && h_total == o.h_total && gen == o.gen &&
                   p2g_groups == o.p2g_groups && p2g_block_cost == o.p2g_block_cost && grid_groups == o.grid_groups;
        }
    };

    // ---- online autotuner ----
    // Coordinate descent: stage 0 varies p2g_groups, stage 1 p2g_block_cost, stage 2
    // grid_groups, each from the best so far; a candidate runs TUNE_REPS timed substeps
    // (after one untimed substep) and scores its fastest.
    static constexpr int TUNE_REPS = 3;

    nxm_launch_config default_config() const
    {
        const uint32_t cu = (uint32_t)ctx_->num_cus;
        return nxm_launch_config{cu * 32u, 16u, cu * 8u};
    }
    uint32_t max_p2g_groups() const { return (uint32_t)ctx_->num_cus * 512u; }

    std::vector<nxm_launch_config> stage_candidates(int stage, const nxm_launch_config &b) const
    {
        const uint32_t cu = (uint32_t)ctx_->num_cus;
        std::vector<nxm_launch_config> c;
        if (stage == 0) for (uint32_t k : {128u, 256u, 512u}) c.push_back({cu * k, b.p2g_block_cost, b.grid_groups});
        if (stage == 1) for (uint32_t k : {4u, 16u, 64u}) c.push_back({b.p2g_groups, k, b.grid_groups});
        if (stage == 2) for (uint32_t k : {4u, 8u, 16u, 32u, 64u}) c.push_back({b.p2g_groups, b.p2g_block_cost, cu * k});
        return c;
    }

    void start_tuning()
    {
        tuning_ = true;
        tune_stage_ = 0;
        tune_best_ = cfg_;
        tune_cands_ = stage_candidates(0, tune_best_);
        tune_idx_ = 0;
        tune_rep_ = -1;   // one untimed substep first
        tune_times_.assign(tune_cands_.size(), 1e30f);
        tune_used_ = 0;
        tune_substeps_ = 0;
    }

    nxm_launch_config tune_candidate() const { return tune_cands_[tune_idx_]; }

    nxm_status tune_events(hipEvent_t *tev)
    {
        if (tune_rep_ < 0) return NXM_OK;   // untimed warm-up substep
        const size_t need = 2 * (tune_used_ + 1);
        while (tune_ev_.size() < need) {
            hipEvent_t e;
            HIP_TRY(hipEventCreate(&e));
            tune_ev_.push_back(e);
        }
        tev[0] = tune_ev_[2 * tune_used_];
        tev[1] = tune_ev_[2 * tune_used_ + 1];
        tune_slot_.resize(tune_used_ + 1);
        tune_slot_[tune_used_] = tune_idx_;
        ++tune_used_;
        return NXM_OK;
    }

    // Called after each tuning substep; at the end of a stage reads the events (one
    // synchronization per stage) and moves to the next stage or finishes.
    nxm_status tune_advance()
    {
        if (++tune_rep_ < TUNE_REPS) return NXM_OK;
        tune_rep_ = 0;
        if (++tune_idx_ < tune_cands_.size()) return NXM_OK;
        HIP_TRY(hipStreamSynchronize(stream_));
        for (size_t k = 0; k < tune_used_; ++k) {
            float ms = 0.0f;
            HIP_TRY(hipEventElapsedTime(&ms, tune_ev_[2 * k], tune_ev_[2 * k + 1]));
            tune_times_[tune_slot_[k]] = std::min(tune_times_[tune_slot_[k]], ms);
        }
        size_t best = 0;
        for (size_t k = 1; k < tune_cands_.size(); ++k)
            if (tune_times_[k] < tune_times_[best]) best = k;
        tune_best_ = tune_cands_[best];
        tune_used_ = 0;
        if (++tune_stage_ < 3) {
            tune_cands_ = stage_candidates(tune_stage_, tune_best_);
            tune_idx_ = 0;
            tune_rep_ = -1;
            tune_times_.assign(tune_cands_.size(), 1e30f);
            return NXM_OK;
        }
        cfg_ = tune_best_;
        tuning_ = false;
        tuned_p_ = p_in_;
        return NXM_OK;
    }

  public:
    nxm_status set_autotune(bool on)
    {
        tune_on_ = on;
        if (on) start_tuning();
        else tuning_ = false;
        return NXM_OK;
    }

    nxm_launch_config launch_config() const { return cfg_; }

    // Whether the tuner is still trying shapes, and the substeps of its current (or last) run.
    void autotune_state(int *tuning, uint32_t *substeps) const
    {
        if (tuning) *tuning = tuning_ ? 1 : 0;
        if (substeps) *substeps = tune_substeps_;
    }

    nxm_status set_launch_config(const nxm_launch_config &c)
    {
        if (c.p2g_groups == 0 || c.p2g_groups > max_p2g_groups() || c.p2g_groups % 8 != 0)
            return fail(NXM_ERR_INVALID_ARGUMENT, "p2g_groups must be a positive multiple of 8 up to 512 per CU");
        if (c.grid_groups == 0) return fail(NXM_ERR_INVALID_ARGUMENT, "grid_groups must be positive");
        tune_on_ = false;
        tuning_ = false;
        cfg_ = c;
        return NXM_OK;
    }

  private:

    void drop_graphs()
    {
        for (auto &g : gexec_) {
            if (g) (void)hipGraphExecDestroy(g);
            g = nullptr;
        }
    }

    nxm_ctx *ctx_;
    hipStream_t stream_ = nullptr;
    hipStream_t side_ = nullptr;       // CPIC: the contact particles' G2P (default priority)
    hipEvent_t fork_ = nullptr, join_ = nullptr;
    nxm_launch_config cfg_{};
    bool tune_on_ = false, tuning_ = false;
    int tune_stage_ = 0, tune_rep_ = 0;
    size_t tune_idx_ = 0, tune_used_ = 0;
    uint32_t tuned_p_ = 0, tune_substeps_ = 0;
    nxm_launch_config tune_best_{};
    std::vector<nxm_launch_config> tune_cands_;
    std::vector<float> tune_times_;
    std::vector<size_t> tune_slot_;
    std::vector<hipEvent_t> tune_ev_;
    std::vector<PinnedChunk> pins_;
    bool graph_on_ = false;
    uint64_t gen_ = 0;                 // bumped whenever a buffer address or table changes
    hipGraphExec_t gexec_[2] = {};     // one per arena parity
    GraphKey gkey_[2] = {};
    hipEvent_t ev_[SUBSTEP_EVENTS] = {};
    std::vector<Slot> slots_;
    std::vector<SceneDev> layout_;
    bool dirty_ = true, pin_dirty_ = true, stepped_ = false, debug_ = false, profiling_ = false, clean_ = false;
    std::vector<hipEvent_t> ring_;
    size_t ring_used_ = 0;
    uint32_t profile_every_ = 1, profile_seen_ = 0;
    int cur_ = 0;
    uint32_t p_in_ = 0, p_cap_ = 0, b_total_ = 0, h_total_ = 0;
    size_t ctr_zeroed_ = 0;
    DevBuf<float> arenas_[2];
    DevBuf<uint32_t> pkey_, pdense_, prank_, perm_, sblock_;
    DevBuf<uint32_t> bscene_, bprim_, bmask_, bnp_, bfirst_, bup_, bdn_, hkeys_, hvals_, hmark_;
    DevBuf<SceneClock> clock_;       // per slot
    bool any_adaptive_ = false;
    DevBuf<unsigned long long> scan_state_;
    DevBuf<uint32_t> p2g_start_;
    uint32_t scan_stride_ = 0;
    DevBuf<uint32_t> region_start_, region_scene_;
    uint32_t n_regions_ = 0, last_b_start_ = 0;
    DevBuf<int4> bvid_;
    DevBuf<float> stage_, dbg_mom_;
    DevBuf<float4> gridv_;
    DevBuf<uint4> ncdf_;               // CPIC buffers, allocated once a CPIC scene is present
    DevBuf<float4> nbody_;
    DevBuf<uint32_t> binc_, bany_, breach_, clist_;
    DevBuf<float> stage_i_;
    DevBuf<float4> gridv_i_;
    bool any_cpic_ = false;
    DevBuf<float4> mesh_v_, mesh_rp_;  // mesh colliders: geometry; rigid particles of CPIC scenes (k_mesh_cdf)
    DevBuf<uint4> mesh_t_;
    DevBuf<float4> mesh_vn_, mesh_en_; // node pass (scenes without CPIC): pseudo-normals,
    DevBuf<MeshGridDev> mesh_grid_;    // cell grids and lists, each world collider's mesh
    DevBuf<uint32_t> mesh_cell_, mesh_ctri_, col_mesh_;
    bool mesh_pass_ = false;           // mgrid non-empty: k_grid_update runs its mesh variant
    DevBuf<unsigned long long> mkey_;  // per node, kept empty between substeps
    DevBuf<uint32_t> maff_;
    size_t mesh_nodes_clear_ = 0;      // nodes of mkey_ / maff_ set to empty since allocation
    uint32_t n_rp_ = 0;
    DevBuf<SceneDev> scenes_;
    DevBuf<SceneCtr> ctr_;
    DevBuf<ClassDev> classes_;
    DevBuf<ColliderDev> cols_[2];      // collider tables (integrated on the device), swapped at each rebuild
    int cols_cur_ = 0;
    DevBuf<Globals> g_;
};

}  // namespace nxm

struct nxm_batch {
    nxm_ctx *ctx = nullptr;
    std::unique_ptr<World> world;
    std::vector<nxm_scene *> members;
};

// ======================================================================================
// Scene helpers

// Interns the material class of one particle: the model tag, its parameter words
// (without the per-particle plastic state) and the damping.
static nxm_status intern_class(nxm_scene *sc, uint32_t tag, const uint32_t *words, float damping, float bfric,
                               uint32_t *out)
{
    std::array<uint32_t, 27> key{};
    key[0] = tag;
    const int pw = pstate_words(tag);
    for (int i = 0; i < 24; ++i) key[1 + i] = i < pw ? 0u : words[i];
    std::memcpy(&key[25], &damping, 4);
    std::memcpy(&key[26], &bfric, 4);
    // Consecutive particles usually share their material: try the last class first.
    if (sc->last_class != NONE && key == sc->last_key) {
        *out = sc->last_class;
        return NXM_OK;
    }
    auto it = sc->class_index.find(key);
    if (it != sc->class_index.end()) {
        *out = it->second;
        sc->last_key = key;
        sc->last_class = it->second;
        return NXM_OK;
    }
    if (sc->classes.size() >= MAX_CLASSES) return fail(NXM_ERR_INVALID_ARGUMENT, "more than 16384 distinct materials in one scene");
    ClassDev c{};
    c.tag = tag;
    c.damping = damping;
    c.bfric = bfric;
    std::memcpy(c.w, words, sizeof(c.w));
    uint32_t id = (uint32_t)sc->classes.size();
    sc->classes.push_back(c);
    sc->class_index.emplace(key, id);
    sc->last_key = key;
    sc->last_class = id;
    *out = id;
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

static nxm_status ensure_own_world(nxm_scene *sc)
{
    if (sc->batch) return NXM_OK;
    if (!sc->own) {
        sc->own = std::make_unique<World>(sc->ctx);
        nxm_status st = sc->own->init();
        if (st != NXM_OK) { sc->own.reset(); return st; }
        if (sc->graph) (void)sc->own->set_graph(true);
        if (sc->debug_grid) sc->own->set_debug(true);
    }
    if (sc->world == nullptr) return sc->own->add(sc);
    return NXM_OK;
}

// Forgets particles packed by nxm_scene_stage_upload (the host particles are changing).
static void drop_staged(nxm_scene *sc)
{
    if (sc->staged_chunk < 0) return;
    (void)sc->ctx->pool.give(sc->staged_chunk, nullptr);
    sc->staged_chunk = -1;
}

static nxm_status append_host(nxm_scene *sc, const nxm_particles_soa *p, const nxm_particle *aos, size_t n)
{
    drop_staged(sc);
    HostParticles &hp = sc->hp;
    size_t base = hp.n;
    // state of already-resident particles is kept on the device; only the new
    // particles are staged on the host and uploaded
    hp.resize(base + n);
    for (size_t k = 0; k < n; ++k) {
        size_t i = base + k;
        float pos[3], vel[3] = {0, 0, 0}, F[9] = {1, 0, 0, 0, 1, 0, 0, 0, 1}, C[9] = {0};
        float vol, rad, mass, damp = 0.0f, phase = 1.0f, bfr = 1.0f;
        uint32_t grp = 0, en = 1, fx = 0, tag;
        const uint32_t *words;
        if (aos) {
            const nxm_particle &a = aos[k];
            std::memcpy(pos, a.position, 12);
            std::memcpy(vel, a.velocity, 12);
            std::memcpy(F, a.def_grad, 36);
            std::memcpy(C, a.affine, 36);
            vol = a.init_volume; rad = a.init_radius; mass = a.mass; damp = a.damping; phase = a.phase;
            bfr = a.boundary_friction; grp = a.group_id; en = a.enabled; fx = a.fixed; tag = a.model_tag;
            words = a.model_data;
        } else {
            if (!p->position || !p->init_volume || !p->mass || !p->model_tag || !p->model_data) {
                hp.resize(base);
                return fail(NXM_ERR_INVALID_ARGUMENT, "position, init_volume, mass, model_tag and model_data are required");
            }
            std::memcpy(pos, p->position + 3 * k, 12);
            if (p->velocity) std::memcpy(vel, p->velocity + 3 * k, 12);
            if (p->def_grad) std::memcpy(F, p->def_grad + 9 * k, 36);
            if (p->affine) std::memcpy(C, p->affine + 9 * k, 36);
            vol = p->init_volume[k];
            rad = p->init_radius ? p->init_radius[k] : 0.5f * std::cbrt(vol);
            mass = p->mass[k];
            if (p->damping) damp = p->damping[k];
            if (p->phase) phase = p->phase[k];
            if (p->boundary_friction) bfr = p->boundary_friction[k];
            if (p->group_id) grp = p->group_id[k];
            if (p->enabled) en = p->enabled[k];
            if (p->fixed) fx = p->fixed[k];
            tag = p->model_tag[k];
            words = p->model_data + 24 * k;
        }
        // The reference's model constructors only produce tags 0-5 (particle_model.rs:282-325).
        if (tag > NXM_MODEL_SNOW) {
            hp.resize(base);   // the scene keeps the particles it had
            return fail(NXM_ERR_UNSUPPORTED_MODEL, "unknown model tag");
        }
        std::memcpy(&hp.x[3 * i], pos, 12);
        std::memcpy(&hp.v[3 * i], vel, 12);
        std::memcpy(&hp.F[9 * i], F, 36);
        std::memcpy(&hp.C[9 * i], C, 36);
        hp.mass[i] = mass;
        hp.vol0[i] = vol;
        hp.radius[i] = rad;
        hp.damping[i] = damp;
        hp.phase[i] = phase;
        hp.bfric[i] = bfr;
        hp.group[i] = grp;
        hp.enabled[i] = en;
        hp.fixed[i] = fx;
        hp.tag[i] = tag;
        std::memcpy(&hp.words[24 * i], words, 96);
        hp.pstate[3 * i] = hp.pstate[3 * i + 1] = hp.pstate[3 * i + 2] = 0.0f;
        std::memcpy(&hp.pstate[3 * i], words, 4 * (size_t)pstate_words(tag));
        hp.aff[i] = 0u;
        hp.cdf[4 * i] = hp.cdf[4 * i + 1] = hp.cdf[4 * i + 2] = hp.cdf[4 * i + 3] = 0.0f;
        nxm_status st = intern_class(sc, tag, words, damp, bfr, &hp.cls[i]);
        if (st != NXM_OK) {
            hp.resize(base);
            return st;
        }
    }
    if (sc->world) {
        nxm_status st = sc->world->upload(sc, base, n);
        // Not uploaded: the scene keeps the particles it had, on both sides.
        if (st != NXM_OK) hp.resize(base);
        return st;
    }
    return NXM_OK;
}

namespace nxm {
// Device memory a scene occupies in a world: both particle arenas and the per-particle
// sort arrays, plus, per hash slot (each slot may hold a block), the hash entry, the
// block tables, the P2G staging tile and the grid nodes.
size_t scene_device_bytes(size_t n, uint32_t grid_cap)
{
    const size_t per_particle = 2 * NUM_PLANES * 4 + 5 * 4;
    const size_t per_block = 3 * 4                    // hkeys, hvals, hmark
                             + 4 * 4 + sizeof(int4)   // bscene, bprim, bmask, bnp/bfirst
                             + 4 + 2 * 8 * 4           // bfirst, bup, bdn
                             + 4 * TILE_NODES * 4      // staging tile
                             + NODES_PER_BLOCK * 16;   // grid nodes
    const size_t slots = (size_t)next_pow2(2 * grid_cap) + 1;
    return n * per_particle + slots * per_block + sizeof(SceneDev) + sizeof(SceneCtr);
}
}  // namespace nxm

// ======================================================================================
// C ABI

// Carves the no-allocation scratch: sizes and offsets of a one-scene world.
namespace {
struct Carver {
    size_t at = 0;
    template <typename T>
    size_t take(size_t count)
    {
        const size_t off = (at + 255u) & ~(size_t)255u;
        at = off + count * sizeof(T);
        return off;
    }
};
}  // namespace

extern "C" {

const char *nxm_last_error(void) { return g_last_error.c_str(); }
const char *nxm_version(void) { return "nexus_amd mpm 0.1 (gfx950)"; }

nxm_status nxm_ctx_create(int device, nxm_ctx **out)
{
    if (!out) return fail(NXM_ERR_INVALID_ARGUMENT, "out is null");
    int count = 0;
    HIP_TRY(hipGetDeviceCount(&count));
    if (device < 0 || device >= count) return fail(NXM_ERR_INVALID_ARGUMENT, "no such device");
    HIP_TRY(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIP_TRY(hipGetDeviceProperties(&prop, device));
    void *scr = nullptr;
    HIP_TRY(scratch_address(device, &scr));
    auto *c = new nxm_ctx();
    c->device = device;
    c->num_cus = prop.multiProcessorCount > 0 ? prop.multiProcessorCount : 256;
    c->scratch = scr;
    *out = c;
    return NXM_OK;
}

void nxm_ctx_destroy(nxm_ctx *ctx)
{
    if (!ctx) return;
    ctx->closed.store(true);
    if (ctx->refs.load() == 0) delete ctx;
}

nxm_status nxm_scene_create(nxm_ctx *ctx, const nxm_capacities *caps, nxm_scene **out)
{
    if (!ctx || !out) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    auto *sc = new nxm_scene();
    sc->ctx = ctx;
    ctx->refs.fetch_add(1);
    if (caps && caps->grid_size) sc->grid_cap = next_pow2(caps->grid_size);
    // The scene stays on the host until it is stepped or joins a batch.
    *out = sc;
    return NXM_OK;
}

void nxm_scene_destroy(nxm_scene *sc)
{
    if (!sc) return;
    if (sc->batch) {
        auto &m = sc->batch->members;
        m.erase(std::remove(m.begin(), m.end(), sc), m.end());
    }
    // Resident in a world it does not own (a batch's): the slot is dropped, the state
    // discarded with the scene.  Its own world goes with it below.
    if (sc->world && sc->world != sc->own.get() && sc->world->holds(sc)) sc->world->drop(sc);
    sc->world = nullptr;
    sc->own.reset();
    nxm_ctx *ctx = sc->ctx;
    if (sc->staged_chunk >= 0) (void)ctx->pool.give(sc->staged_chunk, nullptr);
    delete sc;
    ctx_release(ctx);
}

nxm_status nxm_scene_stage_upload(nxm_scene *sc)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (sc->world) return fail(NXM_ERR_BUSY, "the scene is resident: its particles are uploaded already");
    const HostParticles &hp = sc->hp;
    const size_t n = hp.n;
    if (sc->staged_chunk >= 0) {
        (void)sc->ctx->pool.give(sc->staged_chunk, nullptr);
        sc->staged_chunk = -1;
    }
    if (n == 0) return NXM_OK;
    int k = -1;
    char *buf = nullptr;
    HIP_TRY(sc->ctx->pool.take((size_t)NUM_PLANES * n * 4, &k, &buf));
    float *f = reinterpret_cast<float *>(buf);
    auto put = [&](int p, const float *src, int stride, int comp) {
        float *dst = f + (size_t)p * n;
        for (size_t i = 0; i < n; ++i) dst[i] = src[i * stride + comp];
    };
    for (int c = 0; c < 3; ++c) put(PX + c, hp.x.data(), 3, c);
    for (int c = 0; c < 3; ++c) put(PVX + c, hp.v.data(), 3, c);
    for (int c = 0; c < 9; ++c) put(PC0 + c, hp.C.data(), 9, c);
    for (int c = 0; c < 9; ++c) put(PF0 + c, hp.F.data(), 9, c);
    put(PMASS, hp.mass.data(), 1, 0);
    put(PVOL0, hp.vol0.data(), 1, 0);
    for (int c = 0; c < 3; ++c) put(PS0 + c, hp.pstate.data(), 3, c);
    std::memcpy(f + (size_t)PAFF * n, hp.aff.data(), n * 4);
    for (int c = 0; c < 4; ++c) put(PNX + c, hp.cdf.data(), 4, c);
    uint32_t *uf = reinterpret_cast<uint32_t *>(f + (size_t)PFLAGS * n);
    uint32_t *ui = reinterpret_cast<uint32_t *>(f + (size_t)PID * n);
    for (size_t i = 0; i < n; ++i) {
        uf[i] = make_flags(hp.enabled[i] != 0, hp.fixed[i] != 0, hp.cls[i], 0u);
        ui[i] = (uint32_t)i;
    }
    sc->staged_chunk = k;
    sc->staged_n = n;
    return NXM_OK;
}

nxm_status nxm_scene_set_simulation_params(nxm_scene *sc, const nxm_sim_params *p)
{
    if (!sc || !p) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!(p->dt > 0.0f)) return fail(NXM_ERR_INVALID_ARGUMENT, "dt must be positive");
    std::memcpy(sc->g, p->gravity, 12);
    sc->base_dt = p->dt;
    sc->dt = p->dt;
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

nxm_status nxm_scene_write_substep_params(nxm_scene *sc, uint32_t num_substeps)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    sc->dt = sc->base_dt / (float)std::max<uint32_t>(num_substeps, 1u);
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

nxm_status nxm_scene_set_cell_width(nxm_scene *sc, float h, uint32_t grid_capacity)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (!(h > 0.0f)) return fail(NXM_ERR_INVALID_ARGUMENT, "cell width must be positive");
    if (grid_capacity == 0 || grid_capacity > (1u << 26)) return fail(NXM_ERR_INVALID_ARGUMENT, "grid capacity out of range");
    sc->h = h;
    sc->grid_cap = next_pow2(grid_capacity);
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

nxm_status nxm_scene_add_particles(nxm_scene *sc, const nxm_particle *particles, size_t n)
{
    if (!sc || (!particles && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    return append_host(sc, nullptr, particles, n);
}

nxm_status nxm_scene_add_particles_soa(nxm_scene *sc, const nxm_particles_soa *p, size_t n)
{
    if (!sc || (!p && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    return append_host(sc, p, nullptr, n);
}

// The device record of a collider: the body, the shape and, for the analytic shapes,
// the radius about the body origin of a sphere holding the shape (the block masks'
// bound, widened by a relative 1e-5 against rounding).
static ColliderDev collider_record(const nxm_cuboid &c, uint32_t shape, const float *dims = nullptr)
{
    static_assert(offsetof(ColliderDev, shape) == sizeof(nxm_cuboid), "nxm_cuboid is a prefix of ColliderDev");
    static_assert(offsetof(ColliderDev, shape) == offsetof(nxm_collider, shape) &&
                  offsetof(ColliderDev, dims) == offsetof(nxm_collider, dims), "nxm_collider is a prefix of ColliderDev");
    ColliderDev d{};
    std::memcpy(&d, &c, sizeof(nxm_cuboid));
    d.shape = shape;
    if (dims) std::memcpy(d.dims, dims, sizeof(d.dims));
    if (shape != SHAPE_CUBOID) d.he[0] = d.he[1] = d.he[2] = 0.0f;
    auto norm = [](double x, double y, double z) { return std::sqrt(x * x + y * y + z * z); };
    double reach = 0.0;
    if (shape == SHAPE_CUBOID) reach = norm(c.half_extents[0], c.half_extents[1], c.half_extents[2]);
    else if (shape == SHAPE_BALL) reach = d.dims[0];
    else if (shape == SHAPE_CAPSULE)
        reach = std::max(norm(d.dims[0], d.dims[1], d.dims[2]), norm(d.dims[3], d.dims[4], d.dims[5])) + d.dims[6];
    else if (shape == SHAPE_CONE || shape == SHAPE_CYLINDER) reach = norm(d.dims[0], d.dims[1], 0.0);
    d.reach = (float)(reach * (1.0 + 1e-5));
    return d;
}

// Checks an analytic collider: a known kind with finite, non-negative dimensions (a
// capsule's segment ends may be anywhere finite).
static bool valid_shape(const nxm_collider &c, std::string &why)
{
    auto finite_nonneg = [](float x) { return std::isfinite(x) && x >= 0.0f; };
    switch (c.shape) {
    case NXM_SHAPE_BALL:
        if (!finite_nonneg(c.dims[0])) { why = "ball radius must be finite and non-negative"; return false; }
        return true;
    case NXM_SHAPE_CUBOID:
        for (int i = 0; i < 3; ++i)
            if (!finite_nonneg(c.body.half_extents[i])) { why = "cuboid half extents must be finite and non-negative"; return false; }
        return true;
    case NXM_SHAPE_CAPSULE:
        for (int i = 0; i < 6; ++i)
            if (!std::isfinite(c.dims[i])) { why = "capsule segment ends must be finite"; return false; }
        if (!finite_nonneg(c.dims[6])) { why = "capsule radius must be finite and non-negative"; return false; }
        return true;
    case NXM_SHAPE_CONE:
    case NXM_SHAPE_CYLINDER:
        if (!finite_nonneg(c.dims[0]) || !finite_nonneg(c.dims[1])) {
            why = "half height and radius must be finite and non-negative";
            return false;
        }
        return true;
    default:
        why = "unknown shape kind " + std::to_string(c.shape);
        return false;
    }
}

// Replaces the cuboids (ids 0..n-1) or the meshes (the ids after the cuboids) and keeps
// the other kind with its current poses.
static nxm_status replace_colliders(nxm_scene *sc, const std::vector<ColliderDev> &cub, const std::vector<ColliderDev> *mesh)
{
    if (sc->world) {
        nxm_status st = sc->world->pull_colliders(sc);
        if (st != NXM_OK) return st;
    }
    const size_t nmesh_old = sc->meshes.size();
    std::vector<ColliderDev> cols(cub);
    if (mesh) cols.insert(cols.end(), mesh->begin(), mesh->end());
    else cols.insert(cols.end(), sc->cols.end() - nmesh_old, sc->cols.end());
    sc->cols.swap(cols);
    sc->cols_host = true;
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

nxm_status nxm_scene_set_colliders(nxm_scene *sc, const nxm_cuboid *c, size_t n)
{
    if (!sc || (!c && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (n + sc->meshes.size() > NXM_MAX_COLLIDERS) return fail(NXM_ERR_TOO_MANY_COLLIDERS, "at most 16 colliders");
    std::vector<ColliderDev> cub(n);
    for (size_t i = 0; i < n; ++i) cub[i] = collider_record(c[i], SHAPE_CUBOID);
    return replace_colliders(sc, cub, nullptr);
}

nxm_status nxm_scene_set_shape_colliders(nxm_scene *sc, const nxm_collider *c, size_t n)
{
    if (!sc || (!c && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (n + sc->meshes.size() > NXM_MAX_COLLIDERS) return fail(NXM_ERR_TOO_MANY_COLLIDERS, "at most 16 colliders");
    std::vector<ColliderDev> cols(n);
    for (size_t i = 0; i < n; ++i) {
        std::string why;
        if (!valid_shape(c[i], why)) return fail(NXM_ERR_INVALID_ARGUMENT, "collider " + std::to_string(i) + ": " + why);
        cols[i] = collider_record(c[i].body, c[i].shape, c[i].dims);
    }
    return replace_colliders(sc, cols, nullptr);
}

nxm_status nxm_scene_set_trimesh_colliders(nxm_scene *sc, const nxm_trimesh *m, size_t n)
{
    if (!sc || (!m && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    const size_t ncub = sc->cols.size() - sc->meshes.size();
    if (ncub + n > NXM_MAX_COLLIDERS) return fail(NXM_ERR_TOO_MANY_COLLIDERS, "at most 16 colliders");
    std::vector<HostMesh> meshes(n);
    std::vector<ColliderDev> bodies(n);
    for (size_t k = 0; k < n; ++k) {
        if ((!m[k].vertices && m[k].num_vertices) || (!m[k].triangles && m[k].num_triangles))
            return fail(NXM_ERR_INVALID_ARGUMENT, "null mesh array");
        for (size_t q = 0; q < (size_t)m[k].num_triangles * 3; ++q)
            if (m[k].triangles[q] >= m[k].num_vertices)
                return fail(NXM_ERR_INVALID_ARGUMENT, "triangle vertex index out of range (mesh " + std::to_string(k) + ")");
        meshes[k].v.assign(m[k].vertices, m[k].vertices + (size_t)m[k].num_vertices * 3);
        meshes[k].tri.assign(m[k].triangles, m[k].triangles + (size_t)m[k].num_triangles * 3);
        for (float x : meshes[k].v)
            if (!std::isfinite(x)) return fail(NXM_ERR_INVALID_ARGUMENT, "non-finite mesh vertex (mesh " + std::to_string(k) + ")");
        mesh_pseudo_normals(meshes[k]);
        bodies[k] = collider_record(m[k].body, SHAPE_TRIMESH);
        double reach = 0.0;   // bounding sphere of the vertices about the body origin (block masks)
        for (size_t q = 0; q < meshes[k].v.size() / 3; ++q) {
            const double x = meshes[k].v[3 * q], y = meshes[k].v[3 * q + 1], z = meshes[k].v[3 * q + 2];
            reach = std::max(reach, std::sqrt(x * x + y * y + z * z));
        }
        bodies[k].reach = (float)(reach * (1.0 + 1e-5));
    }
    std::vector<ColliderDev> cub(sc->cols.begin(), sc->cols.begin() + ncub);
    // The cuboids' current poses first (replace_colliders pulls them), then the meshes.
    nxm_status st = NXM_OK;
    if (sc->world) {
        st = sc->world->pull_colliders(sc);
        if (st != NXM_OK) return st;
        cub.assign(sc->cols.begin(), sc->cols.begin() + ncub);
    }
    st = replace_colliders(sc, cub, &bodies);
    if (st != NXM_OK) return st;
    sc->meshes.swap(meshes);
    return NXM_OK;
}

nxm_status nxm_sample_trimesh(const float *vertices, uint32_t num_vertices, const uint32_t *triangles,
                              uint32_t num_triangles, float spacing, float *points, uint32_t *triangle_of,
                              size_t capacity, size_t *count)
{
    if (!count || (!vertices && num_vertices) || (!triangles && num_triangles) || (capacity && (!points || !triangle_of)))
        return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!(spacing > 0.0f)) return fail(NXM_ERR_INVALID_ARGUMENT, "spacing must be positive");
    for (size_t q = 0; q < (size_t)num_triangles * 3; ++q)
        if (triangles[q] >= num_vertices) return fail(NXM_ERR_INVALID_ARGUMENT, "triangle vertex index out of range");
    MeshSamples smp = sample_mesh_surface(vertices, triangles, num_triangles, spacing);
    *count = smp.tri.size();
    const size_t k = std::min(capacity, smp.tri.size());
    if (k) {
        std::memcpy(points, smp.pts.data(), k * 3 * sizeof(float));
        std::memcpy(triangle_of, smp.tri.data(), k * sizeof(uint32_t));
    }
    return NXM_OK;
}

static nxm_status latched(nxm_scene *sc, nxm_status st);

nxm_status nxm_scene_read_colliders(nxm_scene *sc, nxm_cuboid *out, size_t n)
{
    if (!sc || (!out && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (n < sc->cols.size()) return fail(NXM_ERR_INVALID_ARGUMENT, "output too small");
    if (sc->world) {
        nxm_status st = sc->world->pull_colliders(sc);
        if (st != NXM_OK) return st;
    }
    for (size_t i = 0; i < sc->cols.size(); ++i) std::memcpy(&out[i], &sc->cols[i], sizeof(nxm_cuboid));
    return latched(sc, NXM_OK);
}

size_t nxm_scene_len(const nxm_scene *sc) { return sc ? sc->hp.n : 0; }

size_t nxm_scene_device_bytes(const nxm_scene *sc)
{
    if (!sc) return 0;
    return nxm::scene_device_bytes(sc->hp.n, sc->grid_cap);
}

static nxm_status latched(nxm_scene *sc, nxm_status st)
{
    if (st != NXM_OK) return st;
    if (sc->error) {
        uint32_t e = sc->error;
        if (e & ERR_GRID_CAPACITY) return fail(NXM_ERR_GRID_CAPACITY, "active blocks exceeded the grid capacity");
        if (e & ERR_DOMAIN) return fail(NXM_ERR_DOMAIN, "a particle block is outside the packable domain");
        return fail(NXM_ERR_DEVICE, "device error");
    }
    return NXM_OK;
}

nxm_status nxm_scene_synchronize(nxm_scene *sc)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    nxm_status st = sc->world ? sc->world->sync() : NXM_OK;
    return latched(sc, st);
}

nxm_status nxm_scene_read_positions(nxm_scene *sc, float *xyz, size_t n)
{
    if (!sc || (!xyz && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (n < sc->hp.n) return fail(NXM_ERR_INVALID_ARGUMENT, "output too small");
    if (!sc->world) {
        std::memcpy(xyz, sc->hp.x.data(), sc->hp.n * 12);
        return latched(sc, NXM_OK);
    }
    return latched(sc, sc->world->positions(sc, xyz));
}

nxm_status nxm_scene_read_particles_soa(nxm_scene *sc, const nxm_particles_soa *o, size_t n)
{
    if (!sc || !o) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (n < sc->hp.n) return fail(NXM_ERR_INVALID_ARGUMENT, "output too small");
    if (sc->world) {
        nxm_status st = sc->world->pull(sc);
        if (st != NXM_OK) return st;
    }
    const HostParticles &hp = sc->hp;
    size_t m = hp.n;
    if (o->position) std::memcpy(o->position, hp.x.data(), m * 12);
    if (o->velocity) std::memcpy(o->velocity, hp.v.data(), m * 12);
    if (o->def_grad) std::memcpy(o->def_grad, hp.F.data(), m * 36);
    if (o->affine) std::memcpy(o->affine, hp.C.data(), m * 36);
    if (o->init_volume) std::memcpy(o->init_volume, hp.vol0.data(), m * 4);
    if (o->init_radius) std::memcpy(o->init_radius, hp.radius.data(), m * 4);
    if (o->mass) std::memcpy(o->mass, hp.mass.data(), m * 4);
    if (o->damping) std::memcpy(o->damping, hp.damping.data(), m * 4);
    if (o->phase) std::memcpy(o->phase, hp.phase.data(), m * 4);
    if (o->boundary_friction) std::memcpy(o->boundary_friction, hp.bfric.data(), m * 4);
    if (o->group_id) std::memcpy(o->group_id, hp.group.data(), m * 4);
    if (o->enabled) std::memcpy(o->enabled, hp.enabled.data(), m * 4);
    if (o->fixed) std::memcpy(o->fixed, hp.fixed.data(), m * 4);
    if (o->model_tag) std::memcpy(o->model_tag, hp.tag.data(), m * 4);
    if (o->model_data) {
        for (size_t i = 0; i < m; ++i) {
            std::memcpy(o->model_data + 24 * i, &hp.words[24 * i], 96);
            std::memcpy(o->model_data + 24 * i, &hp.pstate[3 * i], 4 * (size_t)pstate_words(hp.tag[i]));
        }
    }
    return latched(sc, NXM_OK);
}

nxm_status nxm_scene_swap_remove(nxm_scene *sc, const uint32_t *ids, size_t n, uint32_t *moved_from,
                                 uint32_t *moved_to, size_t *n_moved)
{
    if (!sc || (!ids && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    World *w = sc->world;
    if (w) {
        // Bring the state to the host and leave the world; the scene re-enters it below.
        nxm_status st = w->remove(sc);
        if (st != NXM_OK) return st;
    }
    drop_staged(sc);
    std::vector<uint32_t> targets(ids, ids + n);
    std::sort(targets.begin(), targets.end(), std::greater<uint32_t>());
    targets.erase(std::unique(targets.begin(), targets.end()), targets.end());
    HostParticles &hp = sc->hp;
    size_t len = hp.n, moved = 0;
    for (uint32_t id : targets) {
        if (id >= len) continue;
        const size_t last = len - 1;
        if (id != last) {
            move_particle(hp, last, id);
            if (moved_from) moved_from[moved] = (uint32_t)last;
            if (moved_to) moved_to[moved] = id;
            ++moved;
        }
        --len;
    }
    hp.resize(len);
    if (n_moved) *n_moved = moved;
    if (w) return w->add(sc);
    return NXM_OK;
}

nxm_status nxm_scene_timestep_bound(nxm_scene *sc, float *max_dt)
{
    if (!sc || !max_dt) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!sc->batch) {
        nxm_status st = ensure_own_world(sc);
        if (st != NXM_OK) return st;
    }
    uint32_t units = 0xFFFFFFFFu;
    nxm_status st = sc->world->timestep_bound(sc, &units);
    if (st != NXM_OK) return st;
    *max_dt = (float)units / 1.0e12f;
    return latched(sc, NXM_OK);
}

nxm_status nxm_scene_set_adaptive_dt(nxm_scene *sc, int enabled)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    sc->adaptive = enabled != 0;
    if (sc->world) sc->world->mark_dirty();
    return NXM_OK;
}

nxm_status nxm_scene_read_clock(nxm_scene *sc, double *time, float *last_dt, uint64_t *substeps)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (sc->world) {
        nxm_status st = sc->world->pull_clock(sc);   // synchronizes
        if (st != NXM_OK) return st;
    }
    if (time) *time = sc->clock.t;
    if (last_dt) *last_dt = sc->clock.dt;
    if (substeps) *substeps = sc->clock.steps;
    return latched(sc, NXM_OK);
}

// GpuReadbackData::DEFAULT_GROUP_COLORS (src_mpm/solver/prep_readback.rs:58-65): the
// palette used when the caller gives none.
static const float DEFAULT_GROUP_COLORS[6][4] = {
    {124.0f / 255.0f, 144.0f / 255.0f, 1.0f, 1.0f},         {8.0f / 255.0f, 144.0f / 255.0f, 1.0f, 1.0f},
    {124.0f / 255.0f, 7.0f / 255.0f, 1.0f, 1.0f},           {124.0f / 255.0f, 144.0f / 255.0f, 7.0f / 255.0f, 1.0f},
    {200.0f / 255.0f, 37.0f / 255.0f, 1.0f, 1.0f},          {124.0f / 255.0f, 230.0f / 255.0f, 25.0f / 255.0f, 1.0f},
};

nxm_status nxm_scene_prep_readback(nxm_scene *sc, uint32_t mode, const float *group_colors, uint32_t num_groups,
                                   nxm_readback_data *out, size_t n)
{
    if (!sc || (!out && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (num_groups && !group_colors) return fail(NXM_ERR_INVALID_ARGUMENT, "null palette");
    if (n < sc->hp.n) return fail(NXM_ERR_INVALID_ARGUMENT, "output too small");
    static_assert(sizeof(nxm_readback_data) == 96, "ReadbackData layout");
    const float *pal = num_groups ? group_colors : &DEFAULT_GROUP_COLORS[0][0];
    const uint32_t npal = num_groups ? num_groups : 6u;
    if (!sc->batch) {
        nxm_status st = ensure_own_world(sc);
        if (st != NXM_OK) return st;
    }
    nxm_status st = sc->world->prep_readback(sc, mode, pal, npal, reinterpret_cast<float *>(out));
    if (st != NXM_OK) return st;
    return latched(sc, NXM_OK);
}

nxm_status nxm_scene_prep_readback_rigid(nxm_scene *sc, const float *group_colors, uint32_t num_groups,
                                         nxm_readback_data *out, size_t n, size_t *count)
{
    if (!sc || !count || (!out && n)) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (num_groups && !group_colors) return fail(NXM_ERR_INVALID_ARGUMENT, "null palette");
    const float *pal = num_groups ? group_colors : &DEFAULT_GROUP_COLORS[0][0];
    const uint32_t npal = num_groups ? num_groups : 6u;
    if (sc->world) {
        nxm_status st = sc->world->pull_colliders(sc);
        if (st != NXM_OK) return st;
    }
    // The rigid particles, posed by their bodies (gpu_transform_sample_points), in the
    // order of the meshes and their samples: gpu_prep_readback_rigid (prep_readback.rs:
    // 429-469) with rigid particle i coloured by palette entry i mod the palette size.
    const size_t ncub = sc->cols.size() - sc->meshes.size();
    size_t total = 0;
    for (size_t k = 0; k < sc->meshes.size(); ++k) {
        HostMesh &m = sc->meshes[k];
        if (m.sampled_h != sc->h) {
            m.samples = sample_mesh_surface(m.v.data(), m.tri.data(), (uint32_t)(m.tri.size() / 3), sc->h);
            m.sampled_h = sc->h;
        }
        const ColliderDev &b = sc->cols[ncub + k];
        for (size_t q = 0; q < m.samples.tri.size(); ++q, ++total) {
            if (total >= n) continue;
            float *o = reinterpret_cast<float *>(out + total);
            const float px = m.samples.pts[3 * q], py = m.samples.pts[3 * q + 1], pz = m.samples.pts[3 * q + 2];
            // rotation * p + translation, the quaternion rotation v (w^2 - b.b) + 2 b (b.v) + 2 w (b x v)
            const float bx = b.q[0], by = b.q[1], bz = b.q[2], bw = b.q[3];
            const float b2 = bx * bx + by * by + bz * bz, bv = bx * px + by * py + bz * pz;
            const float k0 = bw * bw - b2, k2 = 2.0f * bw;
            o[20] = px * k0 + bx * (2.0f * bv) + (by * pz - bz * py) * k2 + b.t[0];
            o[21] = py * k0 + by * (2.0f * bv) + (bz * px - bx * pz) * k2 + b.t[1];
            o[22] = pz * k0 + bz * (2.0f * bv) + (bx * py - by * px) * k2 + b.t[2];
            o[23] = 0.0f;
            const float *c = pal + 4 * (total % npal);
            o[0] = c[0]; o[1] = c[1]; o[2] = c[2]; o[3] = c[3];
            const float sdiag = sc->h * 0.4f;
            for (int e = 0; e < 16; ++e) o[4 + e] = 0.0f;
            o[4] = sdiag; o[9] = sdiag; o[14] = sdiag; o[19] = 1.0f;
        }
    }
    *count = total;
    return NXM_OK;
}

nxm_status nxm_scene_sort_size(nxm_scene *sc, size_t *nb)
{
    if (!sc || !nb) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!sc->world) { *nb = 0; return NXM_OK; }
    return sc->world->sort_size(sc, nb);
}

nxm_status nxm_scene_read_sort_canonical(nxm_scene *sc, const nxm_sort_canonical *out)
{
    if (!sc || !out) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    return sc->world->canonical(sc, out);
}

nxm_status nxm_scene_read_grid(nxm_scene *sc, float *mass, float *momentum, float *velocity)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (!sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    return sc->world->grid(sc, mass, momentum, velocity);
}

nxm_status nxm_scene_read_grid_cpic(nxm_scene *sc, float *distance, uint32_t *affinity, uint32_t *closest,
                                    float *mass_incompatible, float *velocity_incompatible)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (!sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    return sc->world->grid_cpic(sc, distance, affinity, closest, mass_incompatible, velocity_incompatible);
}

nxm_status nxm_scene_set_cpic(nxm_scene *sc, int enabled)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    const bool on = enabled != 0;
    if (on == sc->cpic) return NXM_OK;
    World *w = sc->world;
    if (w) {
        // The world's tables and buffers change: bring the state to the host and re-enter.
        nxm_status st = w->remove(sc);
        if (st != NXM_OK) return st;
    }
    sc->cpic = on;
    drop_staged(sc);
    std::fill(sc->hp.aff.begin(), sc->hp.aff.end(), 0u);
    std::fill(sc->hp.cdf.begin(), sc->hp.cdf.end(), 0.0f);
    if (w) return w->add(sc);
    return NXM_OK;
}

nxm_status nxm_scene_read_cdf(nxm_scene *sc, uint32_t *affinity, float *normal, float *signed_distance, size_t n)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (n < sc->hp.n) return fail(NXM_ERR_INVALID_ARGUMENT, "output too small");
    if (sc->world) {
        nxm_status st = sc->world->pull(sc);
        if (st != NXM_OK) return st;
    }
    const HostParticles &hp = sc->hp;
    for (size_t i = 0; i < hp.n; ++i) {
        if (affinity) affinity[i] = hp.aff[i];
        if (normal) { normal[3 * i] = hp.cdf[4 * i]; normal[3 * i + 1] = hp.cdf[4 * i + 1]; normal[3 * i + 2] = hp.cdf[4 * i + 2]; }
        if (signed_distance) signed_distance[i] = hp.cdf[4 * i + 3];
    }
    return latched(sc, NXM_OK);
}

nxm_status nxm_debug_fail_uploads(uint32_t count)
{
    g_fail_uploads.store((int)std::min<uint32_t>(count, 1u << 30));
    return NXM_OK;
}

nxm_status nxm_scene_set_debug_grid(nxm_scene *sc, int enabled)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    sc->debug_grid = enabled != 0;
    if (sc->world) sc->world->set_debug(enabled != 0);
    return NXM_OK;
}

nxm_status nxm_pipeline_step(nxm_ctx *ctx, nxm_scene *sc, nxm_stage_times *times)
{
    if (!ctx || !sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (sc->batch) return fail(NXM_ERR_BUSY, "scene belongs to a batch");
    nxm_status st = ensure_own_world(sc);
    if (st != NXM_OK) return st;
    return sc->world->step(times);
}

nxm_status nxm_pipeline_run(nxm_ctx *ctx, nxm_scene *sc, uint32_t substeps)
{
    for (uint32_t i = 0; i < substeps; ++i) {
        nxm_status st = nxm_pipeline_step(ctx, sc, nullptr);
        if (st != NXM_OK) return st;
    }
    return NXM_OK;
}

nxm_status nxm_pipeline_step_noalloc(nxm_ctx *ctx, const nxm_noalloc_scene *scene,
                                     const nxm_device_particles *p, uint32_t n, uint32_t *errors, void *stream)
{
    using namespace nxm;
    if (!ctx || !scene || !p) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!ctx->scratch) return fail(NXM_ERR_DEVICE, "no scratch address for this device");
    if (!(scene->cell_width > 0.0f) || !std::isfinite(scene->cell_width))
        return fail(NXM_ERR_INVALID_ARGUMENT, "cell_width must be positive and finite");
    if (!(scene->params.dt > 0.0f) || !std::isfinite(scene->params.dt))
        return fail(NXM_ERR_INVALID_ARGUMENT, "dt must be positive and finite");
    if (scene->grid_capacity == 0 || scene->grid_capacity > (1u << 26))
        return fail(NXM_ERR_INVALID_ARGUMENT, "grid_capacity must be in [1, 2^26]");
    if (!scene->materials || scene->num_materials == 0 || scene->num_materials > (uint32_t)NOALLOC_MAX_CLASSES)
        return fail(NXM_ERR_INVALID_ARGUMENT, "1 to 16 materials");
    if (scene->num_colliders > (uint32_t)NOALLOC_MAX_COLLIDERS) return fail(NXM_ERR_TOO_MANY_COLLIDERS, "at most 16 colliders");
    if (scene->num_colliders && !scene->colliders) return fail(NXM_ERR_INVALID_ARGUMENT, "null colliders");
    for (uint32_t i = 0; i < scene->num_materials; ++i)
        if (scene->materials[i].model_tag > NXM_MODEL_SNOW)
            return fail(NXM_ERR_UNSUPPORTED_MODEL, "model tag above NXM_MODEL_SNOW (material " + std::to_string(i) + ")");
    if (n == 0) return NXM_OK;
    if (!p->position || !p->velocity || !p->def_grad || !p->affine || !p->mass || !p->init_volume || !p->flags)
        return fail(NXM_ERR_INVALID_ARGUMENT, "null particle array");

    // The world of one scene: the same tables and sizes a scene alone gets (its hash
    // region of 2 x capacity slots plus the overflow slot; block ids sized by slots).
    const uint32_t cap = next_pow2(scene->grid_capacity);
    const uint32_t h_cap = next_pow2(2 * cap);
    const uint32_t h_total = h_cap + 1u;
    const uint32_t b_total = h_total;
    const uint32_t p_cap = (n + 63u) & ~63u;
    const uint32_t cu = (uint32_t)ctx->num_cus;
    const uint32_t p2g_groups = cu * 32u;
    const uint32_t scan_stride = (h_total + SCAN_TILE - 1u) / SCAN_TILE + 2u;
    Carver c;
    const size_t o_arena0 = c.take<float>((size_t)NUM_PLANES * p_cap);
    const size_t o_arena1 = c.take<float>((size_t)NUM_PLANES * p_cap);
    const size_t o_pkey = c.take<uint32_t>(p_cap), o_pdense = c.take<uint32_t>(p_cap), o_prank = c.take<uint32_t>(p_cap);
    const size_t o_perm = c.take<uint32_t>(p_cap), o_sblock = c.take<uint32_t>(p_cap);
    const size_t o_bscene = c.take<uint32_t>(b_total), o_bvid = c.take<int4>(b_total), o_bnp = c.take<uint32_t>(b_total);
    const size_t o_bfirst = c.take<uint32_t>(b_total), o_bup = c.take<uint32_t>(8 * (size_t)b_total);
    const size_t o_bdn = c.take<uint32_t>(8 * (size_t)b_total);
    const size_t o_stage = c.take<float>((size_t)4 * TILE_NODES * b_total);
    const size_t o_gridv = c.take<float4>((size_t)NODES_PER_BLOCK * b_total);
    const size_t o_hkeys = c.take<uint32_t>(h_total), o_hvals = c.take<uint32_t>(h_total), o_hmark = c.take<uint32_t>(h_total);
    const size_t o_bprim = c.take<uint32_t>(b_total), o_bmask = c.take<uint32_t>(b_total);
    const size_t o_rstart = c.take<uint32_t>(1), o_rscene = c.take<uint32_t>(1);
    const size_t o_scenes = c.take<SceneDev>(1), o_ctr = c.take<SceneCtr>(1);
    const size_t o_classes = c.take<ClassDev>(NOALLOC_MAX_CLASSES), o_cols = c.take<ColliderDev>(NOALLOC_MAX_COLLIDERS);
    const size_t o_g = c.take<Globals>(1), o_p2g = c.take<uint32_t>((size_t)p2g_groups + 1);
    const size_t o_scan = c.take<unsigned long long>(2 * (size_t)scan_stride);
    if (c.at > (size_t)NXM_NOALLOC_SCRATCH_BYTES)
        return fail(NXM_ERR_INVALID_ARGUMENT, "the scene needs " + std::to_string(c.at) + " bytes of scratch; "
                    "nxm_pipeline_step_noalloc has " + std::to_string((size_t)NXM_NOALLOC_SCRATCH_BYTES) +
                    " (NXM_NOALLOC_SCRATCH_BYTES): use fewer particles or a smaller grid capacity");

    char *base = static_cast<char *>(ctx->scratch);
    auto at = [&](size_t off) { return static_cast<void *>(base + off); };
    WorldDev w{};
    w.in = Arena{static_cast<float *>(at(o_arena0)), p_cap};
    w.out = Arena{static_cast<float *>(at(o_arena1)), p_cap};
    w.pkey = static_cast<uint32_t *>(at(o_pkey)); w.pdense = static_cast<uint32_t *>(at(o_pdense));
    w.prank = static_cast<uint32_t *>(at(o_prank)); w.perm = static_cast<uint32_t *>(at(o_perm));
    w.sblock = static_cast<uint32_t *>(at(o_sblock));
    w.bscene = static_cast<uint32_t *>(at(o_bscene)); w.bvid = static_cast<int4 *>(at(o_bvid));
    w.bnp = static_cast<uint32_t *>(at(o_bnp)); w.bfirst = static_cast<uint32_t *>(at(o_bfirst));
    w.bup = static_cast<uint32_t *>(at(o_bup)); w.bdn = static_cast<uint32_t *>(at(o_bdn));
    w.stage = static_cast<float *>(at(o_stage)); w.gridv = static_cast<float4 *>(at(o_gridv));
    w.hkeys = static_cast<uint32_t *>(at(o_hkeys)); w.hvals = static_cast<uint32_t *>(at(o_hvals));
    w.hmark = static_cast<uint32_t *>(at(o_hmark));
    w.bprim = static_cast<uint32_t *>(at(o_bprim)); w.bmask = static_cast<uint32_t *>(at(o_bmask));
    w.region_start = static_cast<const uint32_t *>(at(o_rstart)); w.region_scene = static_cast<const uint32_t *>(at(o_rscene));
    w.n_regions = 1;
    w.scenes = static_cast<const SceneDev *>(at(o_scenes)); w.ctr = static_cast<SceneCtr *>(at(o_ctr));
    w.classes = static_cast<const ClassDev *>(at(o_classes)); w.cols = static_cast<ColliderDev *>(at(o_cols));
    w.g = static_cast<Globals *>(at(o_g));
    w.p2g_start = static_cast<uint32_t *>(at(o_p2g)); w.p2g_groups = p2g_groups; w.p2g_block_cost = 16u;
    w.scan_state = static_cast<unsigned long long *>(at(o_scan)); w.scan_stride = scan_stride;
    w.p_cap = p_cap; w.b_total = b_total; w.h_total = h_total; w.s_count = 1;

    NoallocTables t{};
    SceneDev &L = t.scene;
    const float h = scene->cell_width, dt = scene->params.dt;
    L.h = h; L.inv_h = 1.0f / h; L.dt = dt; L.vlim = h / dt;
    L.gx = scene->params.gravity[0]; L.gy = scene->params.gravity[1]; L.gz = scene->params.gravity[2];
    L.inv_d = 4.0f / (h * h);
    L.active = 1; L.b_cap = cap; L.h_base = 0; L.h_cap = h_cap;
    L.col_start = 0; L.col_count = scene->num_colliders; L.class_base = 0; L.cpic = 0;
    for (uint32_t i = 0; i < scene->num_materials; ++i) {
        const nxm_material &m = scene->materials[i];
        ClassDev &k = t.classes[i];
        k.tag = m.model_tag;
        k.damping = m.damping;
        k.bfric = m.boundary_friction;
        std::memcpy(k.w, m.model_data, sizeof(k.w));
    }
    for (uint32_t i = 0; i < scene->num_colliders; ++i) {
        const ColliderDev d = collider_record(scene->colliders[i], SHAPE_CUBOID);
        std::memcpy(t.cols[i].body, &d, sizeof(t.cols[i].body));
        t.cols[i].reach = d.reach;
    }
    t.n_classes = scene->num_materials;
    t.n_cols = scene->num_colliders;
    t.n = n;
    NoallocParticles dp{p->position, p->velocity, p->def_grad, p->affine, p->mass, p->init_volume, p->plastic, p->flags,
                        errors};

    HIP_TRY(hipSetDevice(ctx->device));
    hipStream_t st = static_cast<hipStream_t>(stream);
    LaunchConfig lc{};
    lc.p_hint = n;
    lc.persistent_blocks = (int)(cu * 8u);
    lc.p2g_blocks = (int)p2g_groups;
    launch_noalloc_setup(w, t, dp, lc.persistent_blocks, st);
    launch_clear(w, lc, st);
    launch_substep(w, lc, st, nullptr, true);
    launch_noalloc_unpack(w, dp, n, lc.persistent_blocks, st);
    HIP_TRY(hipGetLastError());
    return NXM_OK;
}

nxm_status nxm_scene_profile_begin(nxm_scene *sc)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (sc->batch) return fail(NXM_ERR_BUSY, "scene belongs to a batch");
    nxm_status st = ensure_own_world(sc);
    if (st != NXM_OK) return st;
    sc->world->profile_begin();
    return NXM_OK;
}

nxm_status nxm_scene_profile_sampled(nxm_scene *sc, uint32_t every)
{
    nxm_status st = nxm_scene_profile_begin(sc);
    if (st != NXM_OK) return st;
    sc->world->profile_begin(every);
    return NXM_OK;
}

nxm_status nxm_scene_profile_end(nxm_scene *sc, nxm_stage_times *avg, uint32_t *steps)
{
    if (!sc || !sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    return sc->world->profile_end(avg, steps);
}

nxm_status nxm_batch_profile_begin(nxm_batch *b)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    b->world->profile_begin();
    return NXM_OK;
}

nxm_status nxm_batch_profile_end(nxm_batch *b, nxm_stage_times *avg, uint32_t *steps)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    return b->world->profile_end(avg, steps);
}

nxm_status nxm_batch_create(nxm_ctx *ctx, nxm_batch **out)
{
    if (!ctx || !out) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    auto *b = new nxm_batch();
    b->ctx = ctx;
    ctx->refs.fetch_add(1);
    b->world = std::make_unique<World>(ctx);
    nxm_status st = b->world->init();
    if (st != NXM_OK) { delete b; ctx_release(ctx); return st; }
    *out = b;
    return NXM_OK;
}

void nxm_batch_destroy(nxm_batch *b)
{
    if (!b) return;
    std::vector<nxm_scene *> m = b->members;
    for (nxm_scene *sc : m) (void)nxm_batch_remove_scene(b, sc);
    nxm_ctx *ctx = b->ctx;
    delete b;
    ctx_release(ctx);
}

nxm_status nxm_batch_add_scene(nxm_batch *b, nxm_scene *sc)
{
    if (!b || !sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (sc->batch) return fail(NXM_ERR_BUSY, "scene already belongs to a batch");
    if (sc->world) {
        nxm_status st = sc->world->remove(sc);
        if (st != NXM_OK) return st;
    }
    sc->batch = b;
    nxm_status st = b->world->add(sc);
    if (st != NXM_OK) { sc->batch = nullptr; return st; }
    if (sc->debug_grid) b->world->set_debug(true);
    b->members.push_back(sc);
    return NXM_OK;
}

nxm_status nxm_batch_remove_scene(nxm_batch *b, nxm_scene *sc)
{
    if (!b || !sc || sc->batch != b) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not in this batch");
    // A member whose re-entry failed (nxm_scene_swap_remove, nxm_scene_set_cpic) is on the
    // host already and holds no slot.
    if (b->world->holds(sc)) {
        nxm_status st = b->world->remove(sc);
        if (st != NXM_OK) return st;
    }
    sc->batch = nullptr;
    auto &m = b->members;
    m.erase(std::remove(m.begin(), m.end(), sc), m.end());
    // The scene's state is back on the host; a later standalone step uploads it again.
    return NXM_OK;
}

nxm_status nxm_batch_step(nxm_batch *b, nxm_stage_times *times)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    return b->world->step(times);
}

nxm_status nxm_batch_run(nxm_batch *b, uint32_t substeps)
{
    for (uint32_t i = 0; i < substeps; ++i) {
        nxm_status st = nxm_batch_step(b, nullptr);
        if (st != NXM_OK) return st;
    }
    return NXM_OK;
}

nxm_status nxm_batch_synchronize(nxm_batch *b)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    return b->world->sync();
}

size_t nxm_batch_len(const nxm_batch *b) { return b ? b->members.size() : 0; }

nxm_status nxm_batch_set_graph(nxm_batch *b, int enabled)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    return b->world->set_graph(enabled != 0);
}

nxm_status nxm_batch_set_autotune(nxm_batch *b, int enabled)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    return b->world->set_autotune(enabled != 0);
}

nxm_status nxm_scene_set_autotune(nxm_scene *sc, int enabled)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (sc->batch) return fail(NXM_ERR_BUSY, "scene belongs to a batch");
    nxm_status st = ensure_own_world(sc);
    if (st != NXM_OK) return st;
    return sc->own->set_autotune(enabled != 0);
}

nxm_status nxm_batch_get_launch_config(nxm_batch *b, nxm_launch_config *out)
{
    if (!b || !out) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    *out = b->world->launch_config();
    return NXM_OK;
}

nxm_status nxm_batch_set_launch_config(nxm_batch *b, const nxm_launch_config *cfg)
{
    if (!b || !cfg) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    return b->world->set_launch_config(*cfg);
}

nxm_status nxm_scene_get_launch_config(nxm_scene *sc, nxm_launch_config *out)
{
    if (!sc || !out) return fail(NXM_ERR_INVALID_ARGUMENT, "null argument");
    if (!sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    *out = sc->world->launch_config();
    return NXM_OK;
}

nxm_status nxm_batch_get_autotune(nxm_batch *b, int *tuning, uint32_t *substeps)
{
    if (!b) return fail(NXM_ERR_INVALID_ARGUMENT, "null batch");
    b->world->autotune_state(tuning, substeps);
    return NXM_OK;
}

nxm_status nxm_scene_get_autotune(nxm_scene *sc, int *tuning, uint32_t *substeps)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (!sc->world) return fail(NXM_ERR_INVALID_ARGUMENT, "scene is not resident");
    sc->world->autotune_state(tuning, substeps);
    return NXM_OK;
}

nxm_status nxm_scene_set_graph(nxm_scene *sc, int enabled)
{
    if (!sc) return fail(NXM_ERR_INVALID_ARGUMENT, "null scene");
    if (sc->batch) return fail(NXM_ERR_BUSY, "scene belongs to a batch");
    sc->graph = enabled != 0;
    if (sc->own) return sc->own->set_graph(sc->graph);
    return NXM_OK;
}

}  // extern "C"
