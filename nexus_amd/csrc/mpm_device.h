// mpm_device.h — device-side data layout and per-particle / per-node math of the
// MI355X MLS-MPM substep.  Everything here is __device__ code shared by the kernels in
// mpm_world.hip.  Citations are file:line in the dimforge/nexus tree.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace nxm {

constexpr uint32_t NONE = 0xFFFFFFFFu;
constexpr uint32_t EMPTY_KEY = 0xFFFFFFFFu;
constexpr int WAVE = 64;
constexpr int TILE = 6;                   // 6^3 node tile: a block plus its +1 halo
constexpr int TILE_NODES = TILE * TILE * TILE;
constexpr int NODES_PER_BLOCK = 64;       // 4^3 (src_mpm_shaders/grid/grid.rs:25)
constexpr float MAX_FLT_REF = 1.0e18f;    // src_rbd_shaders/lib.rs:283
// Elements per tile of the single-pass scans (k_scan).  The look-back state array holds
// one word per tile for each of the two scans: the host sizes it by this constant.
constexpr int SCAN_TILE = 1024;

// Particle flags word: bit 0 enabled, bit 1 fixed, bits 2..15 material class,
// bits 16..31 scene slot.
constexpr uint32_t FLAG_ENABLED = 1u;
constexpr uint32_t FLAG_FIXED = 2u;
__host__ __device__ inline uint32_t flags_class(uint32_t f) { return (f >> 2) & 0x3FFFu; }
__host__ __device__ inline uint32_t flags_scene(uint32_t f) { return f >> 16; }
__host__ __device__ inline uint32_t make_flags(bool enabled, bool fixed, uint32_t cls, uint32_t scene)
{
    return (enabled ? FLAG_ENABLED : 0u) | (fixed ? FLAG_FIXED : 0u) | ((cls & 0x3FFFu) << 2) | (scene << 16);
}
constexpr uint32_t MAX_CLASSES = 1u << 14;
constexpr uint32_t MAX_SCENES = 1u << 16;

// Particle arena planes (structure of arrays, one f32/u32 plane per component).
enum Plane : int {
    PX = 0, PY, PZ,
    PVX, PVY, PVZ,
    PC0, PC1, PC2, PC3, PC4, PC5, PC6, PC7, PC8,   // affine, column-major
    PF0, PF1, PF2, PF3, PF4, PF5, PF6, PF7, PF8,   // deformation gradient, column-major
    PMASS, PVOL0,
    PS0, PS1, PS2,                                  // plastic state (sand: 3 words, snow: 1)
    PAFF, PNX, PNY, PNZ, PSD,                       // CPIC: affinity word, contact normal, signed distance
    PFLAGS, PID,
    NUM_PLANES
};

struct Arena {
    float *base;       // NUM_PLANES planes of `stride` elements
    uint32_t stride;
    __device__ __host__ float *plane(int p) const { return base + (size_t)p * stride; }
    __device__ __host__ uint32_t *uplane(int p) const { return reinterpret_cast<uint32_t *>(base + (size_t)p * stride); }
};

// One scene of a world.
struct SceneDev {
    float h, inv_h, dt, vlim;     // vlim = h / dt
    float gx, gy, gz, inv_d;      // inv_d = 4 / h^2
    uint32_t active;
    uint32_t b_cap;               // block capacity (the overflow block not included)
    uint32_t h_base, h_cap;       // hash region [h_base, h_base + h_cap) (h_cap a power of two), then one overflow slot
    uint32_t col_start, col_count;
    uint32_t class_base;          // first entry of the scene's material classes
    uint32_t cpic;                // compatible particle-in-cell transfers (MpmState::use_cpic)
    // Adaptive substeps: each substep's length is min(dt_max, the CFL bound of the state
    // at its start); dt and vlim above are then rewritten on the device each substep.
    float dt_max;                 // the scene's substep length (the maximum when adaptive)
    uint32_t adaptive;
};

// Per scene slot, on the device while the scene is resident: its simulated time, the
// length of its last substep and the CFL bound (1e-12 s units, NONE = none yet) that the
// next substep uses when adaptive.
struct SceneClock {
    uint32_t bound;
    float dt;
    uint32_t steps;
    uint32_t pad;
    double t;
};

struct SceneCtr {
    uint32_t b_start;  // first dense block id of the scene in the last step
    uint32_t b_count;  // its blocks, the overflow block (last) included
    uint32_t err;      // error bits
    uint32_t pad;
};
constexpr uint32_t ERR_GRID_CAPACITY = 1u;
constexpr uint32_t ERR_DOMAIN = 2u;
constexpr uint32_t ERR_MODEL = 4u;

struct Globals {
    uint32_t p_in;       // particles in the input arena
    uint32_t p_out;      // particles placed in the output arena (scan total)
    uint32_t n_blocks;   // dense block ids in use (scan total of occupied hash slots)
    uint32_t ticket[2];  // tile tickets of the two single-pass scans
    uint32_t n_contact;  // CPIC: contact particles listed for k_g2p_contact
    uint32_t pad[2];
};

// Material class: model tag and words (src_mpm_shaders/models/default.rs:41-44),
// plus the particle damping (ParticleProperties::damping, solver/particle.rs:166) and
// the boundary friction factor (Kinematics::boundary_friction, solver/particle.rs:143).
struct ClassDev {
    uint32_t tag;
    float damping;
    float w[24];
    float bfric;
    uint32_t pad;
};

// A collider on its body: the nxm_cuboid fields, then the shape kind and dimensions as
// nxm_collider holds them (a triangle mesh's geometry lives in the world's mesh tables;
// its half extents are zero).  Kinds are the reference's Shape tags
// (src_rbd_shaders/shapes/shape.rs:29-35).
constexpr uint32_t SHAPE_BALL = 0u;
constexpr uint32_t SHAPE_CUBOID = 1u;
constexpr uint32_t SHAPE_CAPSULE = 2u;
constexpr uint32_t SHAPE_CONE = 3u;
constexpr uint32_t SHAPE_CYLINDER = 4u;
constexpr uint32_t SHAPE_TRIMESH = 6u;
// A wave-uniform read of memory that no kernel running at the time writes, through the
// constant address space: the compiler then issues it as a scalar load.
template <class T>
__device__ __forceinline__ T ld_uniform(const T *p)
{
    static_assert(sizeof(T) % 4 == 0 && alignof(T) >= 4, "whole aligned dwords");
    using cword = const __attribute__((address_space(4))) uint32_t;
    T r;
    uint32_t *o = reinterpret_cast<uint32_t *>(&r);
    #pragma unroll
    for (uint32_t k = 0; k < sizeof(T) / 4; ++k) o[k] = ((cword *)p)[k];
    return r;
}

// ld_uniform where SCALAR holds (a kernel variant with scalar registers to spare), else a
// plain load.
template <bool SCALAR, class T>
__device__ __forceinline__ T ld_maybe_uniform(const T *p)
{
    if constexpr (SCALAR) return ld_uniform(p);
    else return *p;
}

struct ColliderDev {
    float he[3];
    float q[4];
    float t[3];
    float linvel[3];
    float angvel[3];
    uint32_t bc_type;
    float bc_friction;
    float inv_mass[3];
    float lcom[3];      // centre of mass in the body frame
    uint32_t shape;
    float dims[7];      // ball: r; capsule: a, b, r; cone, cylinder: half height, r
    float reach;        // radius about the body origin of a sphere holding the shape
    uint32_t pad;
};

// Block virtual id marker of a scene's overflow block (particles that could not be
// given a block: grid capacity exceeded or coordinates outside the packable range).
constexpr int32_t OVERFLOW_VID = INT32_MIN;

// ------------------------------------------------------------------------------------
// Keys: block of a position (src_mpm_shaders/solver/particle.rs:208-225,
// grid/grid.rs:122-134), packing (grid/grid.rs:102-106).

// Rust's saturating f32 -> i32 cast (NaN -> 0).
__device__ inline int32_t sat_i32(float f)
{
    if (f != f) return 0;
    if (f >= 2147483520.0f) return INT32_MAX;
    if (f <= -2147483648.0f) return INT32_MIN;
    return (int32_t)f;
}

// roundf(x / h) with the IEEE quotient, cheaply: x * inv_h is within 1.5 ulp of x / h,
// so the two round to the same integer unless the product lies within a few ulps of a
// half-integer; only then is the division done.  (roundf rounds half away from zero.)
__device__ inline float round_div(float x, float h, float inv_h)
{
    const float q = x * inv_h;
    const float f = fabsf(q - truncf(q));
    if (fabsf(f - 0.5f) > fabsf(q) * 0x1p-21f) return roundf(q);
    return roundf(x / h);
}

__device__ inline void cell_of(float x, float h, float inv_h, int32_t &blk, uint32_t &local)
{
    float cell = round_div(x, h, inv_h) - 1.0f;
    float b = floorf(cell / 4.0f);
    blk = sat_i32(b);
    float l = cell - b * 4.0f;
    local = (l != l) ? 0u : (uint32_t)l;
}

// Packable range: primaries keep their +1 neighbours inside the 11/10/11-bit key.
__device__ inline bool vid_in_range(int32_t x, int32_t y, int32_t z)
{
    return x >= -1023 && x <= 1022 && y >= -511 && y <= 510 && z >= -1023 && z <= 1022;
}

__device__ inline uint32_t pack_vid(int32_t x, int32_t y, int32_t z)
{
    return ((uint32_t)(x + 0x3FF) & 0x7FFu) | (((uint32_t)(y + 0x1FF) & 0x3FFu) << 11) |
           (((uint32_t)(z + 0x3FF) & 0x7FFu) << 21);
}

// Spreads the low 10 bits of v to every third bit.
__host__ __device__ inline uint32_t spread3(uint32_t v)
{
    v &= 0x3FFu;
    v = (v | (v << 16)) & 0x030000FFu;
    v = (v | (v << 8)) & 0x0300F00Fu;
    v = (v | (v << 4)) & 0x030C30C3u;
    v = (v | (v << 2)) & 0x09249249u;
    return v;
}

// First probe slot of a packed block key: the Morton code of its coordinates (low 10
// bits of each), so spatially close blocks land in close slots and the dense block ids,
// which are ranks of occupied slots, follow a Morton order.  Blocks whose codes agree
// modulo the region size collide and are resolved by linear probing.
__device__ inline uint32_t hash_key(uint32_t k)
{
    return spread3(k & 0x7FFu) | (spread3((k >> 11) & 0x3FFu) << 1) | (spread3(k >> 21) << 2);
}

// +1 neighbour offsets in the reference's blocks_associated_to_block order
// (grid/grid.rs:175-212); index 0 is the block itself.
__device__ __constant__ const int8_t ASSOC[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                                                     {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

// ------------------------------------------------------------------------------------
// Small linear algebra (column-major 3x3).

struct f3 { float x, y, z; };
struct m3 { f3 c0, c1, c2; };

__device__ inline f3 mk3(float x, float y, float z) { return f3{x, y, z}; }
__device__ inline f3 operator+(f3 a, f3 b) { return f3{a.x + b.x, a.y + b.y, a.z + b.z}; }
__device__ inline f3 operator-(f3 a, f3 b) { return f3{a.x - b.x, a.y - b.y, a.z - b.z}; }
__device__ inline f3 operator*(f3 a, float s) { return f3{a.x * s, a.y * s, a.z * s}; }
__device__ inline f3 operator/(f3 a, float s) { return f3{a.x / s, a.y / s, a.z / s}; }
__device__ inline float dot(f3 a, f3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ inline f3 cross(f3 a, f3 b) { return f3{a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y}; }
__device__ inline float length(f3 a) { return sqrtf(dot(a, a)); }

__device__ inline m3 mzero() { return m3{{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}; }
__device__ inline m3 mident() { return m3{{1, 0, 0}, {0, 1, 0}, {0, 0, 1}}; }
__device__ inline f3 mul(const m3 &m, f3 v) { return m.c0 * v.x + m.c1 * v.y + m.c2 * v.z; }
__device__ inline m3 mul(const m3 &a, const m3 &b) { return m3{mul(a, b.c0), mul(a, b.c1), mul(a, b.c2)}; }
__device__ inline m3 operator+(const m3 &a, const m3 &b) { return m3{a.c0 + b.c0, a.c1 + b.c1, a.c2 + b.c2}; }
__device__ inline m3 operator-(const m3 &a, const m3 &b) { return m3{a.c0 - b.c0, a.c1 - b.c1, a.c2 - b.c2}; }
__device__ inline m3 operator*(const m3 &a, float s) { return m3{a.c0 * s, a.c1 * s, a.c2 * s}; }
__device__ inline m3 transpose(const m3 &a)
{
    return m3{{a.c0.x, a.c1.x, a.c2.x}, {a.c0.y, a.c1.y, a.c2.y}, {a.c0.z, a.c1.z, a.c2.z}};
}
__device__ inline float det(const m3 &a) { return dot(a.c2, cross(a.c0, a.c1)); }
__device__ inline float trace(const m3 &a) { return a.c0.x + a.c1.y + a.c2.z; }

// ------------------------------------------------------------------------------------
// 3x3 SVD in registers: one-sided Jacobi (Hestenes' method) on the columns of F,
// cyclic pairs (0,1), (0,2), (1,2), at most 8 sweeps with early exit once every
// column pair is orthogonal to FP32 precision.  Returns F = U diag(s) V^T with U, V
// rotations; s is not sorted, and its smallest entry carries the sign of det F.
// The rotations use the hardware reciprocal, reciprocal square root and square root
// (about 1 ulp): the SVD needs its rotations orthogonal to FP32 rounding, not correctly
// rounded quotients.  Convergence is tested on (gamma^2 / (alpha beta)), the squared
// cosine between the columns, against (1e-7)^2.
__device__ inline float rcp_hw(float x) { return __builtin_amdgcn_rcpf(x); }
__device__ inline float rsq_hw(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ inline float sqrt_hw(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ inline void jacobi_pair(f3 &ap, f3 &aq, f3 &vp, f3 &vq, float &off)
{
    float alpha = dot(ap, ap), beta = dot(aq, aq), gamma = dot(ap, aq);
    float ab = alpha * beta;
    float rel2 = ab > 0.0f ? (gamma * gamma) * rcp_hw(ab) : 0.0f;
    off = fmaxf(off, rel2);
    if (!(rel2 > 1.0e-14f)) return;
    float zeta = (beta - alpha) * rcp_hw(2.0f * gamma);
    float az = fabsf(zeta);
    float t = az > 1.0e15f ? 0.5f * rcp_hw(zeta) : copysignf(rcp_hw(az + sqrt_hw(1.0f + zeta * zeta)), zeta);
    float c = rsq_hw(1.0f + t * t);
    float s = c * t;
    f3 a0 = ap, a1 = aq, w0 = vp, w1 = vq;
    ap = a0 * c - a1 * s;
    aq = a0 * s + a1 * c;
    vp = w0 * c - w1 * s;
    vq = w0 * s + w1 * c;
}

__device__ inline f3 any_orthogonal(f3 a)
{
    f3 e = fabsf(a.x) < 0.57f ? mk3(1, 0, 0) : (fabsf(a.y) < 0.57f ? mk3(0, 1, 0) : mk3(0, 0, 1));
    f3 c = cross(a, e);
    return c / length(c);
}

__device__ inline void svd3(const m3 &F, m3 &U, f3 &s, m3 &V)
{
    f3 a0 = F.c0, a1 = F.c1, a2 = F.c2;
    f3 w0 = mk3(1, 0, 0), w1 = mk3(0, 1, 0), w2 = mk3(0, 0, 1);
    for (int sweep = 0; sweep < 8; ++sweep) {
        float off = 0.0f;
        jacobi_pair(a0, a1, w0, w1, off);
        jacobi_pair(a0, a2, w0, w2, off);
        jacobi_pair(a1, a2, w1, w2, off);
        if (!(off > 1.0e-14f)) break;
    }
    float s0 = sqrt_hw(dot(a0, a0)), s1 = sqrt_hw(dot(a1, a1)), s2 = sqrt_hw(dot(a2, a2));
    float big = fmaxf(s0, fmaxf(s1, s2));
    float tiny = big * 1.0e-30f;
    f3 u0 = s0 > tiny ? a0 * rcp_hw(s0) : mk3(0, 0, 0);
    f3 u1 = s1 > tiny ? a1 * rcp_hw(s1) : mk3(0, 0, 0);
    f3 u2 = s2 > tiny ? a2 * rcp_hw(s2) : mk3(0, 0, 0);
    // Complete U where columns vanished (rank-deficient F).
    int nz = (s0 > tiny) + (s1 > tiny) + (s2 > tiny);
    if (nz < 3) {
        if (nz == 0) { u0 = mk3(1, 0, 0); u1 = mk3(0, 1, 0); u2 = mk3(0, 0, 1); }
        else {
            // move a valid column to u0 for the completion
            f3 ref = s0 > tiny ? u0 : (s1 > tiny ? u1 : u2);
            if (nz == 1) {
                f3 o1 = any_orthogonal(ref), o2 = cross(ref, o1);
                if (s0 > tiny) { u1 = o1; u2 = o2; }
                else if (s1 > tiny) { u0 = o1; u2 = o2; }
                else { u0 = o1; u1 = o2; }
            } else {
                if (!(s0 > tiny)) u0 = cross(u1, u2);
                else if (!(s1 > tiny)) u1 = cross(u2, u0);
                else u2 = cross(u0, u1);
            }
        }
    }
    U = m3{u0, u1, u2};
    V = m3{w0, w1, w2};
    s = mk3(s0, s1, s2);
    if (det(V) < 0.0f) { V.c2 = V.c2 * -1.0f; U.c2 = U.c2 * -1.0f; }
    if (det(U) < 0.0f) {
        // flip the column of the smallest singular value
        if (s.x <= s.y && s.x <= s.z) { U.c0 = U.c0 * -1.0f; s.x = -s.x; }
        else if (s.y <= s.z) { U.c1 = U.c1 * -1.0f; s.y = -s.y; }
        else { U.c2 = U.c2 * -1.0f; s.z = -s.z; }
    }
}

// U diag(d) U^T (symmetric), used by the corotated stress with a known SVD.
__device__ inline m3 u_diag_ut(const m3 &U, f3 d)
{
    m3 Ud = m3{U.c0 * d.x, U.c1 * d.y, U.c2 * d.z};
    return mul(Ud, transpose(U));
}

__device__ inline m3 recompose(const m3 &U, f3 s, const m3 &V)
{
    m3 Us = m3{U.c0 * s.x, U.c1 * s.y, U.c2 * s.z};
    return mul(Us, transpose(V));
}

// ------------------------------------------------------------------------------------
// Constitutive models (src_mpm_shaders/models/*).

// NeoHookeanModel::kirchoff_stress (neo_hookean_elasticity.rs:26-39)
__device__ inline m3 nh_stress(float lambda, float mu, const m3 &F)
{
    float j = fmaxf(det(F), 1.0e-10f);
    float dv = lambda * logf(j) - mu;
    m3 st = mul(F, transpose(F)) * mu;
    st.c0.x += dv;
    st.c1.y += dv;
    st.c2.z += dv;
    return st;
}

// LinearElasticModel::kirchoff_stress (linear_elasticity.rs:30-54) given the SVD of F:
// 2 mu U diag(s - 1) V^T F^T + lambda (J - 1) J I = 2 mu U diag((s - 1) s) U^T + ...
__device__ inline m3 linear_stress_svd(float lambda, float mu, const m3 &U, f3 s)
{
    float j = s.x * s.y * s.z;
    m3 st = u_diag_ut(U, mk3((s.x - 1.0f) * s.x, (s.y - 1.0f) * s.y, (s.z - 1.0f) * s.z)) * (2.0f * mu);
    float dv = lambda * (j - 1.0f) * j;
    st.c0.x += dv;
    st.c1.y += dv;
    st.c2.z += dv;
    return st;
}

// The Drucker-Prager return map runs on the hardware transcendentals (v_exp_f32,
// v_log_f32, v_sin_f32, about 1-2 ulp) and reciprocals: the reference leaves their
// precision to the shader compiler, and the parity bar on F and the plastic state is
// FP32 rounding, not correctly rounded libm results.
__device__ inline float dp_alpha(const float *dp, float q)
{
    float angle = dp[0] + (dp[1] * q - dp[3]) * __expf(-dp[2] * q);
    float sa = __sinf(angle);
    return 0.81649658f * (2.0f * sa) * rcp_hw(3.0f - sa);   // sqrt(2/3)
}

// DruckerPragerPlasticity::project (drucker_prager.rs:224-267 with :121-167).  Updates
// F and the plastic state {det Fp, q, log volume gain} in place; returns the SVD of
// the projected F in (U, s, V) for the elastic backbone.
__device__ inline void dp_project(const float *dp, float &pdet, float &q, float &lvg, m3 &F, m3 &U, f3 &s, m3 &V)
{
    svd3(F, U, s, V);
    if (dp[4] == 0.0f) return;
    float alpha = dp_alpha(dp, q);
    const float d = 3.0f, inv_d3 = 1.0f / 3.0f;
    float lg = lvg * inv_d3;
    f3 strain = mk3(__logf(s.x) + lg, __logf(s.y) + lg, __logf(s.z) + lg);
    float tr = strain.x + strain.y + strain.z;
    float m = tr * inv_d3;
    f3 dev = mk3(strain.x - m, strain.y - m, strain.z - m);
    float shifted = tr - dp[6];
    f3 ns;
    float dq;
    if (shifted > 0.0f || (dev.x == 0.0f && dev.y == 0.0f && dev.z == 0.0f)) {
        float e = __expf(dp[6] * inv_d3);
        ns = mk3(e, e, e);
        dq = length(strain);
    } else {
        float dn = length(dev);
        float gamma = dn + (d * dp[4] + 2.0f * dp[5]) * rcp_hw(2.0f * dp[5]) * shifted * alpha;
        if (gamma <= 0.0f) return;   // elastic: state and F unchanged
        float k = gamma * rcp_hw(dn);
        ns = mk3(__expf(strain.x - dev.x * k), __expf(strain.y - dev.y * k), __expf(strain.z - dev.z * k));
        dq = gamma;
    }
    float prev_det = s.x * s.y * s.z;
    float new_det = ns.x * ns.y * ns.z;
    pdet = pdet * prev_det * rcp_hw(new_det);
    lvg = lvg + __logf(prev_det) - __logf(new_det);
    q = q + dq;
    s = ns;
    F = recompose(U, s, V);
}

// SnowPlasticity::project and hardening_factor (snow.rs:77-113), sp = {critical
// compression, critical stretch, hardening}: the singular values are clamped to
// [1 - theta_c, 1 + theta_s], the removed volume change goes into the plastic
// determinant (kept in [0.1, 4]) and the elastic moduli scale by
// exp(hardening (1 - plastic det)).  Returns that factor; F and the plastic determinant
// are updated in place and (U, s, V) is the SVD of the returned F.
__device__ inline float snow_project(const float *sp, float &pdet, m3 &F, m3 &U, f3 &s, m3 &V)
{
    svd3(F, U, s, V);
    if (sp[0] <= 0.0f && sp[1] <= 0.0f) return __expf(sp[2] * (1.0f - pdet));
    const float lo = 1.0f - sp[0], hi = 1.0f + sp[1];
    const f3 c = mk3(fminf(fmaxf(s.x, lo), hi), fminf(fmaxf(s.y, lo), hi), fminf(fmaxf(s.z, lo), hi));
    const float prev_det = s.x * s.y * s.z;
    const float new_det = c.x * c.y * c.z;
    pdet = fminf(fmaxf(pdet * prev_det / new_det, 0.1f), 4.0f);
    s = c;
    F = recompose(U, s, V);
    return __expf(sp[2] * (1.0f - pdet));
}

// FluidModel (fluid.rs:54-84)
__device__ inline float fluid_pressure(const float *fl, float j)
{
    j = fmaxf(j, 1.0e-6f);
    if (j <= 1.0f) return fl[0] * (expf(-fl[1] * logf(j)) - 1.0f);
    return -fl[0] * fl[4] * (j - 1.0f);
}

__device__ inline m3 fluid_stress(const float *fl, const m3 &F, const m3 &gradv)
{
    float j = fmaxf(det(F), 1.0e-6f);
    m3 sr = (gradv + transpose(gradv)) * 0.5f;
    float sph = trace(sr) / 3.0f;
    sr.c0.x -= sph;
    sr.c1.y -= sph;
    sr.c2.z -= sph;
    m3 st = sr * (2.0f * fl[2] * j);
    float dv = -fluid_pressure(fl, j) * j;
    st.c0.x += dv;
    st.c1.y += dv;
    st.c2.z += dv;
    return st;
}

// ------------------------------------------------------------------------------------
// Cuboid collider boundary on grid nodes (solver/grid_update_collide.rs:30-183).

// v (w^2 - b.b) + b (2 b.v) + (b x v)(2 w): rotation of v by the quaternion (b, w).
__device__ inline f3 quat_rotate(float qx, float qy, float qz, float qw, f3 v)
{
    f3 b = mk3(qx, qy, qz);
    float b2 = dot(b, b);
    return v * (qw * qw - b2) + b * (dot(v, b) * 2.0f) + cross(b, v) * (qw * 2.0f);
}

// World centre of mass of the collider's body: pose * local centre of mass
// (LocalMassProperties::to_world, rigid_impulses.rs:163-185).
__device__ inline f3 collider_com(const ColliderDev &c)
{
    return quat_rotate(c.q[0], c.q[1], c.q[2], c.q[3], mk3(c.lcom[0], c.lcom[1], c.lcom[2])) + mk3(c.t[0], c.t[1], c.t[2]);
}

// Closest point on the boundary of a cuboid in its frame: outside, the clamp to the box;
// inside, the nearest face (first axis wins ties; the min face wins a same-axis tie),
// as parry3d 0.30's Cuboid::project_local_point with solid = false.
__device__ inline f3 project_cuboid_local(const float he[3], f3 local, bool &inside)
{
    float lp[3] = {local.x, local.y, local.z};
    float sh[3];
    bool zero = true;
    for (int i = 0; i < 3; ++i) {
        float mins_pt = -he[i] - lp[i];
        float pt_maxs = lp[i] - he[i];
        sh[i] = fmaxf(mins_pt, 0.0f) - fmaxf(pt_maxs, 0.0f);
        zero = zero && sh[i] == 0.0f;
    }
    if (!zero) {
        inside = false;
        for (int i = 0; i < 3; ++i) lp[i] += sh[i];
    } else {
        inside = true;
        float best = -3.0e38f;
        bool is_mins = false;
        int best_id = 0;
        for (int i = 0; i < 3; ++i) {
            float a = -he[i] - lp[i], b = lp[i] - he[i];
            if (a < b) {
                if (b > best) { best_id = i; is_mins = false; best = b; }
            } else if (a > best) {
                best_id = i; is_mins = true; best = a;
            }
        }
        for (int i = 0; i < 3; ++i)
            if (i == best_id) lp[i] += is_mins ? best : -best;
    }
    return mk3(lp[0], lp[1], lp[2]);
}

// parry3d 0.30's Ball::project_local_point with solid = false: the point scaled onto the
// sphere.  At the centre the scale is 0/0, a NaN projection that no collider pass
// picks, as in the reference.
__device__ inline f3 project_ball_local(float r, f3 p, bool &inside)
{
    const float d2 = dot(p, p);
    inside = d2 <= r * r;
    return p * (r / sqrtf(d2));
}

// Segment::project_local_point (src_rbd_shaders/shapes/segment.rs:43-60): the end a or b
// beyond the segment's slab, else a + ab (ab.ap / ab.ab).
__device__ inline f3 project_segment(f3 a, f3 b, f3 p)
{
    const f3 ab = b - a, ap = p - a;
    const float t = dot(ab, ap), n2 = dot(ab, ab);
    if (t <= 0.0f) return a;
    if (t >= n2) return b;
    return a + ab * (t / n2);
}

// Capsule::project_local_point_on_boundary (capsule.rs:62-83), d = {a, b, radius}:
// radially out from the nearest point of the axis segment; a point on the axis leaves
// along the first vector of orthonormal_basis3(axis direction) (capsule.rs:150-161).
__device__ inline f3 project_capsule_local(const float d[7], f3 p, bool &inside)
{
    const f3 a = mk3(d[0], d[1], d[2]), b = mk3(d[3], d[4], d[5]);
    const float r = d[6];
    const f3 on_axis = project_segment(a, b, p);
    const f3 dp = p - on_axis;
    const float dist = length(dp);
    if (dist > 0.0f) {
        inside = dist <= r;
        return on_axis + dp * (r / dist);
    }
    const f3 ax = b - a;
    const f3 u = ax / length(ax);
    const float sg = u.z >= 0.0f ? 1.0f : -1.0f;
    const float k = -1.0f / (sg + u.z);
    const float m = u.x * u.y * k;
    inside = true;
    return on_axis + mk3(1.0f + sg * u.x * u.x * k, sg * m, -sg * u.x) * r;
}

// Cylinder::project_local_point_on_boundary (cylinder.rs:90-148), axis y: inside, the
// nearest of the top, the bottom and the side (the side on ties); outside, the cap plane
// or its rim above or below, the side between.
__device__ inline f3 project_cylinder_local(float hh, float r, f3 p, bool &inside)
{
    const float planar = sqrtf(p.x * p.x + p.z * p.z);
    float ux = 1.0f, uz = 0.0f;
    if (planar > 0.0f) { ux = p.x / planar; uz = p.z / planar; }
    const float rx = ux * r, rz = uz * r;
    if (p.y >= -hh && p.y <= hh && planar <= r) {
        inside = true;
        const float top = hh - p.y, bottom = p.y - (-hh), side = r - planar;
        if (top < bottom && top < side) return mk3(p.x, hh, p.z);
        if (bottom < top && bottom < side) return mk3(p.x, -hh, p.z);
        return mk3(rx, p.y, rz);
    }
    inside = false;
    if (p.y > hh) return planar <= r ? mk3(p.x, hh, p.z) : mk3(rx, hh, rz);
    if (p.y < -hh) return planar <= r ? mk3(p.x, -hh, p.z) : mk3(rx, -hh, rz);
    return mk3(rx, p.y, rz);
}

// Cone::project_local_point_on_boundary (cone.rs:98-147), apex (0, hh, 0), base disc at
// y = -hh: below the disc, its plane; otherwise the segment from the apex to the base rim
// in the point's azimuth, or, inside, the nearer of that and the base plane.
__device__ inline f3 project_cone_local(float hh, float r, f3 p, bool &inside)
{
    const float planar = sqrtf(p.x * p.x + p.z * p.z);
    float ux = 1.0f, uz = 0.0f;
    if (planar > 0.0f) { ux = p.x / planar; uz = p.z / planar; }
    const f3 on_base = mk3(p.x, -hh, p.z);
    if (p.y < -hh && planar <= r) {
        inside = false;
        return on_base;
    }
    const f3 rim = mk3(ux * r, -hh, uz * r);
    const f3 apex = mk3(0.0f, hh, 0.0f);
    const f3 side = rim - apex;
    const f3 proj = project_segment(apex, rim, p);
    const f3 down = mk3(0.0f, -2.0f * hh, 0.0f);
    if (p.y >= -hh && p.y <= hh && dot(cross(side, p - apex), cross(side, down)) >= 0.0f) {
        inside = true;
        const f3 to_side = proj - p, to_base = on_base - p;
        return dot(to_side, to_side) > dot(to_base, to_base) ? on_base : proj;
    }
    inside = false;
    return proj;
}

// Shape::project_point_on_boundary (src_rbd_shaders/shapes/shape.rs:378-399) for a posed
// analytic collider: the point in the body frame (rotation^-1 (pt - translation)), its
// projection, back to the world.
__device__ inline f3 project_shape(const ColliderDev &c, f3 pt, bool &inside)
{
    const f3 t = mk3(c.t[0], c.t[1], c.t[2]);
    const f3 local = quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], pt - t);
    f3 lp;
    if (c.shape == SHAPE_BALL) lp = project_ball_local(c.dims[0], local, inside);
    else if (c.shape == SHAPE_CAPSULE) lp = project_capsule_local(c.dims, local, inside);
    else if (c.shape == SHAPE_CONE) lp = project_cone_local(c.dims[0], c.dims[1], local, inside);
    else if (c.shape == SHAPE_CYLINDER) lp = project_cylinder_local(c.dims[0], c.dims[1], local, inside);
    else lp = project_cuboid_local(c.he, local, inside);
    return quat_rotate(c.q[0], c.q[1], c.q[2], c.q[3], lp) + t;
}

// ------------------------------------------------------------------------------------
// Mesh colliders in the grid collider pass (CPIC off): grid_update_collide.rs:53-67 with
// TriMesh::project_local_point (src_rbd_shaders/shapes/trimesh.rs:155-207).

// A mesh's body-frame cell grid over the bounds of its triangles widened by the 1.5 h
// query distance: cell (ix, iy, iz) lists, in ascending order, every triangle whose
// widened bounds meet the cell, so a point's cell holds every triangle within 1.5 h of it.
struct MeshGridDev {
    float ox, oy, oz, inv_cell;
    uint32_t nx, ny, nz, cell_base;   // cell_base: the mesh's first entry of MeshSetDev::cell
};

// The world's mesh tables: vertices and triangles (global vertex ids), per vertex its
// pseudo-normal, per triangle those of its edges ab, bc, ca; per mesh its grid; per grid
// cell the start of its triangle list in ctri (cell[c + 1] its end); per world collider
// its mesh (NONE for the analytic shapes).
struct MeshSetDev {
    const float4 *v;
    const uint4 *t;
    const float4 *vn;
    const float4 *en;
    const MeshGridDev *grid;
    const uint32_t *cell;
    const uint32_t *ctri;
    const uint32_t *col_mesh;
};

// Triangle projection with feature location as the mesh query takes it (trimesh.rs:
// 237-391): the Voronoi regions of a, b, c, then of the edges ab, ac, bc (signed volumes),
// then the face; a degenerate face leaves the point itself (the solid case).  ft: 0
// vertex (fi 0 a, 1 b, 2 c), 1 edge (fi 0 ab, 1 bc, 2 ac), 2 face, 3 solid.
__device__ inline f3 tri_project(f3 a, f3 b, f3 c, f3 pt, int &ft, int &fi)
{
    const f3 ab = b - a, ac = c - a, ap = pt - a;
    const float ab_ap = dot(ab, ap), ac_ap = dot(ac, ap);
    if (ab_ap <= 0.0f && ac_ap <= 0.0f) { ft = 0; fi = 0; return a; }
    const f3 bp = pt - b;
    const float ab_bp = dot(ab, bp), ac_bp = dot(ac, bp);
    if (ab_bp >= 0.0f && ac_bp <= ab_bp) { ft = 0; fi = 1; return b; }
    const f3 cp = pt - c;
    const float ab_cp = dot(ab, cp), ac_cp = dot(ac, cp);
    if (ac_cp >= 0.0f && ab_cp <= ac_cp) { ft = 0; fi = 2; return c; }
    const f3 bc = c - b;
    const f3 n = cross(ab, ac);
    const float vc = dot(n, cross(ab, ap));
    if (vc < 0.0f && ab_ap >= 0.0f && ab_bp <= 0.0f) {
        ft = 1; fi = 0;
        return a + ab * (ab_ap / dot(ab, ab));
    }
    const float vb = -dot(n, cross(ac, cp));
    if (vb < 0.0f && ac_ap >= 0.0f && ac_cp <= 0.0f) {
        ft = 1; fi = 2;
        return a + ac * (ac_ap / dot(ac, ac));
    }
    const float va = dot(n, cross(bc, bp));
    if (va < 0.0f && ac_bp - ab_bp >= 0.0f && ab_cp - ac_cp >= 0.0f) {
        ft = 1; fi = 1;
        return b + bc * (dot(bc, bp) / dot(bc, bc));
    }
    if (va + vb + vc != 0.0f) {
        const float denom = 1.0f / (va + vb + vc);
        ft = 2; fi = 0;
        return a + ab * (vb * denom) + ac * (vc * denom);
    }
    ft = 3; fi = 0;
    return pt;
}

// The nearest triangle of mesh m within max_dist of the world point pt (a later triangle
// wins a tie, dist <= best), taken in the body frame by the inverse pose (R^-1 pt -
// R^-1 t, the reference's pose.inverse() * point); inside when the nearest feature's
// pseudo-normal points away from the point (a face's is ab x ac).  False when no
// triangle is that near.
__device__ inline bool project_trimesh(const MeshSetDev &ms, uint32_t m, const ColliderDev &c, float max_dist, f3 pt,
                                       f3 &proj, bool &inside)
{
    const MeshGridDev g = ms.grid[m];
    const f3 t = mk3(c.t[0], c.t[1], c.t[2]);
    const f3 lp = quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], pt) - quat_rotate(-c.q[0], -c.q[1], -c.q[2], c.q[3], t);
    const float fx = (lp.x - g.ox) * g.inv_cell, fy = (lp.y - g.oy) * g.inv_cell, fz = (lp.z - g.oz) * g.inv_cell;
    // outside the grid (or NaN): no triangle within reach
    if (!(fx >= 0.0f && fy >= 0.0f && fz >= 0.0f && fx < (float)g.nx && fy < (float)g.ny && fz < (float)g.nz)) return false;
    const uint32_t cell = g.cell_base + (uint32_t)fx + g.nx * ((uint32_t)fy + g.ny * (uint32_t)fz);
    const uint32_t kb = ms.cell[cell], ke = ms.cell[cell + 1];
    float best = max_dist;
    uint32_t bt = NONE;
    int bft = 0, bfi = 0;
    f3 bp = lp;
    for (uint32_t k = kb; k < ke; ++k) {
        const uint32_t tri = ms.ctri[k];
        const uint4 tv = ms.t[tri];
        const float4 A = ms.v[tv.x], B = ms.v[tv.y], C = ms.v[tv.z];
        int ft, fi;
        const f3 p = tri_project(mk3(A.x, A.y, A.z), mk3(B.x, B.y, B.z), mk3(C.x, C.y, C.z), lp, ft, fi);
        const float dist = length(p - lp);
        if (dist <= best) { best = dist; bt = tri; bft = ft; bfi = fi; bp = p; }
    }
    if (bt == NONE) return false;
    const uint4 tv = ms.t[bt];
    f3 pn;
    if (bft == 0) {
        const float4 q = ms.vn[bfi == 0 ? tv.x : (bfi == 1 ? tv.y : tv.z)];
        pn = mk3(q.x, q.y, q.z);
    } else if (bft == 1) {
        const float4 q = ms.en[3 * (size_t)bt + bfi];
        pn = mk3(q.x, q.y, q.z);
    } else {
        const float4 A = ms.v[tv.x], B = ms.v[tv.y], C = ms.v[tv.z];
        pn = cross(mk3(B.x - A.x, B.y - A.y, B.z - A.z), mk3(C.x - A.x, C.y - A.y, C.z - A.z));
    }
    inside = dot(pn, lp - bp) <= 0.0f;
    proj = quat_rotate(c.q[0], c.q[1], c.q[2], c.q[3], bp) + t;
    return true;
}

// BoundaryCondition::project_velocity (solver/boundary_condition.rs:77-114)
__device__ inline f3 bc_project(uint32_t ty, float friction, f3 vel, f3 n)
{
    if (ty == 0u) return mk3(0, 0, 0);
    if (ty == 1u) return vel - n * dot(vel, n);
    if (ty == 2u) {
        float nv = dot(vel, n);
        if (nv < 0.0f) {
            f3 tv = vel - n * nv;
            float tl = length(tv);
            f3 dir = tl > 1.0e-8f ? tv / tl : mk3(0, 0, 0);
            float pl = tl + friction * nv;
            return dir * (pl > 0.0f ? pl : 0.0f);
        }
        return vel;
    }
    return vel;
}

// mask: colliders that can reach the node's block (bit i for collider i); the others
// can be neither within 1.5 h of the node nor contain it, so skipping them leaves the
// reference's result unchanged.  MESH: mesh colliders (their mesh ms.col_mesh[col_base +
// i]) report only triangles within 1.5 h; without it the caller's scenes have none.  The
// colliders are read with ld_uniform: no kernel running beside the caller writes them.
template <bool MESH>
__device__ inline f3 collide_node(const ColliderDev *cols, uint32_t ncol, uint32_t mask, float h, float dt, f3 pt, f3 vel,
                                  const MeshSetDev &ms, uint32_t col_base)
{
    float cap = h * 1.5f;
    float best = MAX_FLT_REF;
    f3 normal = mk3(0, 0, 0);
    uint32_t best_id = 0;
    for (uint32_t m = ncol >= 32u ? mask : (mask & ((1u << ncol) - 1u)); m != 0u; m &= m - 1u) {
        const uint32_t i = (uint32_t)__ffs(m) - 1u;   // the masked colliders, ascending
        bool inside;
        f3 proj;
        const ColliderDev c = ld_uniform(cols + i);   // mask, and so i, is uniform over the wave
        if (c.shape == SHAPE_TRIMESH) {
            if (!MESH || !project_trimesh(ms, ms.col_mesh[col_base + i], c, cap, pt, proj, inside)) continue;
        } else {
            proj = project_shape(c, pt, inside);
        }
        f3 dpt = proj - pt;
        bool within = fabsf(dpt.x) <= cap && fabsf(dpt.y) <= cap && fabsf(dpt.z) <= cap;
        if (inside || within) {
            float sign = inside ? -1.0f : 1.0f;
            float dist = length(dpt);
            f3 n = dpt / (dist * -sign);
            float sd = sign * dist;
            if (sd < best) { best = sd; normal = n; best_id = i; }
        }
    }
    if (best != MAX_FLT_REF) {
        const ColliderDev &c = cols[best_id];
        f3 com = collider_com(c);
        f3 bv = mk3(c.linvel[0], c.linvel[1], c.linvel[2]) + cross(mk3(c.angvel[0], c.angvel[1], c.angvel[2]), pt - com);
        f3 delta = vel - bv;
        float nv = dot(delta, normal);
        if (best <= h) {
            vel = bv + bc_project(c.bc_type, c.bc_friction, delta, normal);
        } else if (-nv * dt > best - h) {
            vel = vel - normal * (nv + (best - h) / dt);
        }
    }
    return vel;
}

// ------------------------------------------------------------------------------------
// CPIC contact distance field (solver/grid_update_cdf.rs, g2p_cdf.rs; affinity words as
// grid/grid.rs:593-645: bit i = collider i within reach, bit 16 + i = its side).

__device__ inline bool aff_compatible(uint32_t a, uint32_t b)
{
    const uint32_t common = a & b & 0xFFFFu;
    return (((a ^ b) >> 16) & common) == 0u;
}

// gpu_grid_update_cdf's collide (grid_update_cdf.rs:18-60) for one node: among the
// colliders containing it or within 1.5 h on every axis, the strictly closest (its
// distance and id) and every such collider's affinity bit with its side.
__device__ inline void node_cdf(const ColliderDev *cols, uint32_t ncol, uint32_t mask, float h, f3 pt,
                                float &dist, uint32_t &aff, uint32_t &closest)
{
    const float cap = h * 1.5f;
    dist = MAX_FLT_REF;
    aff = 0u;
    closest = NONE;
    for (uint32_t m = ncol >= 32u ? mask : (mask & ((1u << ncol) - 1u)); m != 0u; m &= m - 1u) {
        const uint32_t i = (uint32_t)__ffs(m) - 1u;   // the masked colliders, ascending
        if (cols[i].shape == SHAPE_TRIMESH) continue;   // meshes: k_mesh_cdf
        bool inside;
        const f3 proj = project_shape(cols[i], pt, inside);
        const f3 dpt = proj - pt;
        const bool within = fabsf(dpt.x) <= cap && fabsf(dpt.y) <= cap && fabsf(dpt.z) <= cap;
        if (inside || within) {
            const float d = length(dpt);
            if (d < dist) { dist = d; closest = i; }
            aff |= (inside ? 0x00010001u : 0x00000001u) << i;
        }
    }
}

// Velocity::velocity_at_point (src_rbd_shaders/dynamics/body.rs:239-250) about the
// body's world centre of mass.
__device__ inline f3 body_point_velocity(const ColliderDev &c, f3 pt)
{
    const f3 com = collider_com(c);
    return mk3(c.linvel[0], c.linvel[1], c.linvel[2]) + cross(mk3(c.angvel[0], c.angvel[1], c.angvel[2]), pt - com);
}

// Symmetric 4x4 system of the particle's moving-least-squares distance fit: the
// determinant by the 2x2 minors of rows {0,1} and {2,3} (Laplace expansion) and, when it
// exceeds the reference's 1e-8 (g2p_cdf.rs:309), the solution adj(Q) u / det.
struct Sym4 {
    float q00, q01, q02, q03, q11, q12, q13, q22, q23, q33;
};

__device__ inline bool sym4_solve(const Sym4 &m, const float u[4], float out[4])
{
    const float a00 = m.q00, a01 = m.q01, a02 = m.q02, a03 = m.q03;
    const float a10 = m.q01, a11 = m.q11, a12 = m.q12, a13 = m.q13;
    const float a20 = m.q02, a21 = m.q12, a22 = m.q22, a23 = m.q23;
    const float a30 = m.q03, a31 = m.q13, a32 = m.q23, a33 = m.q33;
    const float s0 = a00 * a11 - a10 * a01, s1 = a00 * a12 - a10 * a02, s2 = a00 * a13 - a10 * a03;
    const float s3 = a01 * a12 - a11 * a02, s4 = a01 * a13 - a11 * a03, s5 = a02 * a13 - a12 * a03;
    const float c5 = a22 * a33 - a32 * a23, c4 = a21 * a33 - a31 * a23, c3 = a21 * a32 - a31 * a22;
    const float c2 = a20 * a33 - a30 * a23, c1 = a20 * a32 - a30 * a22, c0 = a20 * a31 - a30 * a21;
    const float det = s0 * c5 - s1 * c4 + s2 * c3 + s3 * c2 - s4 * c1 + s5 * c0;
    if (!(det > 1.0e-8f)) return false;
    const float b00 = a11 * c5 - a12 * c4 + a13 * c3, b01 = -a01 * c5 + a02 * c4 - a03 * c3;
    const float b02 = a31 * s5 - a32 * s4 + a33 * s3, b03 = -a21 * s5 + a22 * s4 - a23 * s3;
    const float b10 = -a10 * c5 + a12 * c2 - a13 * c1, b11 = a00 * c5 - a02 * c2 + a03 * c1;
    const float b12 = -a30 * s5 + a32 * s2 - a33 * s1, b13 = a20 * s5 - a22 * s2 + a23 * s1;
    const float b20 = a10 * c4 - a11 * c2 + a13 * c0, b21 = -a00 * c4 + a01 * c2 - a03 * c0;
    const float b22 = a30 * s4 - a31 * s2 + a33 * s0, b23 = -a20 * s4 + a21 * s2 - a23 * s0;
    const float b30 = -a10 * c3 + a11 * c1 - a12 * c0, b31 = a00 * c3 - a01 * c1 + a02 * c0;
    const float b32 = -a30 * s3 + a31 * s1 - a32 * s0, b33 = a20 * s3 - a21 * s1 + a22 * s0;
    const float r = 1.0f / det;
    out[0] = (b00 * u[0] + b01 * u[1] + b02 * u[2] + b03 * u[3]) * r;
    out[1] = (b10 * u[0] + b11 * u[1] + b12 * u[2] + b13 * u[3]) * r;
    out[2] = (b20 * u[0] + b21 * u[1] + b22 * u[2] + b23 * u[3]) * r;
    out[3] = (b30 * u[0] + b31 * u[1] + b32 * u[2] + b33 * u[3]) * r;
    return true;
}

}  // namespace nxm
