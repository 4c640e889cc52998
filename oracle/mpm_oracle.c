/*
 * mpm_oracle.c — TEST INFRASTRUCTURE ONLY (interface in mpm_oracle.h).
 *
 * A CPU restatement of dimforge/nexus's non-CPIC MLS-MPM substep, written from
 * the reference's shader sources.  It exists to check the HIP library; only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline may load it.
 * Citations are file:line in the dimforge/nexus tree.
 *
 * Built twice from this one file (oracle/Makefile):
 *   REAL=float   FP32 mirror: same operation order as the reference shaders,
 *                -ffp-contract=off, IEEE division, roundf (half away from zero).
 *   REAL=double  FP64 oracle: identical algorithm in double precision.
 *
 * Pass order of one substep (src_mpm/pipeline.rs:385-526, use_cpic = false):
 *   block activation (src_mpm_shaders/grid/grid.rs:447-578, grid/sort.rs:104-163)
 *   -> neighbour ids (grid/sort.rs:261-279) -> bucket counts (grid/sort.rs:287-349)
 *   -> exclusive scan + cursors (grid/sort.rs:357-420) -> scatter (grid/sort.rs:430-491)
 *   -> P2G (solver/p2g.rs:64-351) -> grid update (solver/grid_update.rs:21-88)
 *   -> collider boundary (solver/grid_update_collide.rs:30-183)
 *   -> G2P (solver/g2p.rs:42-347) -> particle update (solver/particle_update.rs:60-201).
 * The reference's atomics become a sequential emulation: header ids in first-touch
 * order, bucket lists in particle-id order.  That is one of the orders the
 * reference can produce; its canonical form (active block set, per-block bucket
 * counts, per-(block, bucket) id sets) is order independent.
 *
 * How each part is pinned (tests/test_oracle_*.py):
 *   block keys, buckets, spills   brute-force integer tables incl. x/h = k + 0.5 ties
 *   quadratic kernel              partition of unity, first and second moments
 *   P2G / G2P                     mass and momentum conservation, affine-field reproduction
 *   SVD                           numpy.linalg.svd (LAPACK) on random and degenerate F
 *   cuboid projection             brute-force closest point on the box surface
 *   NH / fluid / DP               closed forms (DP return map on singular values)
 *   whole substep                 free-fall closed form of the elastic cube
 * Parity unpinned (the reference cannot be built here and holds no vectors for this
 * path): the exact FP32 bits of glamx Mat3::svd (third-party, glamx 0.3), parry3d 0.30's
 * Cuboid::project_point tie-breaking between equidistant faces, and the precision of
 * the reference GPU's ln/exp/sin.  This file restates parry's published Aabb
 * projection (clamp outside; nearest face inside, first axis wins ties, min face wins
 * same-axis ties) and uses its own one-sided Jacobi SVD.
 */
#include "mpm_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#if defined(ORACLE_F64)
#define RSQRT sqrt
#define RLOG log
#define REXP exp
#define RSIN sin
#define RCOS cos
#define RFABS fabs
#define RROUND round
#define RFLOOR floor
#define RFMAX fmax
#define RFMIN fmin
#define JACOBI_EPS 1e-15
#else
#define RSQRT sqrtf
#define RLOG logf
#define REXP expf
#define RSIN sinf
#define RCOS cosf
#define RFABS fabsf
#define RROUND roundf
#define RFLOOR floorf
#define RFMAX fmaxf
#define RFMIN fminf
#define JACOBI_EPS 1e-7f
#endif

#define K(c) ((REAL)(c))
#define NONE 0xFFFFFFFFu
#define MAX_FLT_REF ((REAL)1.0e18f) /* src_rbd_shaders/lib.rs:283 */

/* ------------------------------------------------------------------------- */
/* Small vector / matrix algebra with the reference's (glam's) operation order. */

typedef struct { REAL x, y, z; } v3;
typedef struct { v3 c0, c1, c2; } m3; /* columns */

static inline v3 V3(REAL x, REAL y, REAL z) { v3 r; r.x = x; r.y = y; r.z = z; return r; }
static inline v3 vadd(v3 a, v3 b) { return V3(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 vsub(v3 a, v3 b) { return V3(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 vmul(v3 a, REAL s) { return V3(a.x * s, a.y * s, a.z * s); }
static inline v3 vdiv(v3 a, REAL s) { return V3(a.x / s, a.y / s, a.z / s); }
static inline REAL vdot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline v3 vcross(v3 a, v3 b)
{
    return V3(a.y * b.z - b.y * a.z, a.z * b.x - b.z * a.x, a.x * b.y - b.x * a.y);
}
static inline REAL vlen(v3 a) { return RSQRT(vdot(a, a)); }
static inline REAL vget(v3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
static inline void vset(v3 *a, int i, REAL s)
{
    if (i == 0) a->x = s; else if (i == 1) a->y = s; else a->z = s;
}

static inline m3 M3(v3 c0, v3 c1, v3 c2) { m3 r; r.c0 = c0; r.c1 = c1; r.c2 = c2; return r; }
static inline m3 mzero(void) { return M3(V3(0, 0, 0), V3(0, 0, 0), V3(0, 0, 0)); }
static inline m3 mident(void) { return M3(V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)); }
static inline m3 mdiag(v3 d) { return M3(V3(d.x, 0, 0), V3(0, d.y, 0), V3(0, 0, d.z)); }
static inline v3 mcol(m3 m, int j) { return j == 0 ? m.c0 : (j == 1 ? m.c1 : m.c2); }
/* glam Mat3 * Vec3: (c0*v.x + c1*v.y) + c2*v.z */
static inline v3 mvmul(m3 m, v3 v)
{
    return vadd(vadd(vmul(m.c0, v.x), vmul(m.c1, v.y)), vmul(m.c2, v.z));
}
static inline m3 mmul(m3 a, m3 b) { return M3(mvmul(a, b.c0), mvmul(a, b.c1), mvmul(a, b.c2)); }
static inline m3 madd(m3 a, m3 b) { return M3(vadd(a.c0, b.c0), vadd(a.c1, b.c1), vadd(a.c2, b.c2)); }
static inline m3 msub(m3 a, m3 b) { return M3(vsub(a.c0, b.c0), vsub(a.c1, b.c1), vsub(a.c2, b.c2)); }
static inline m3 mscale(m3 a, REAL s) { return M3(vmul(a.c0, s), vmul(a.c1, s), vmul(a.c2, s)); }
static inline m3 mtrans(m3 a)
{
    return M3(V3(a.c0.x, a.c1.x, a.c2.x), V3(a.c0.y, a.c1.y, a.c2.y), V3(a.c0.z, a.c1.z, a.c2.z));
}
/* glam Mat3::determinant: z_axis . (x_axis x y_axis) */
static inline REAL mdet(m3 a) { return vdot(a.c2, vcross(a.c0, a.c1)); }
static inline REAL mtrace(m3 a) { return a.c0.x + a.c1.y + a.c2.z; }
/* outer_product(a, b) = from_cols(a*b.x, a*b.y, a*b.z)   (solver/g2p.rs:378-381) */
static inline m3 mouter(v3 a, v3 b) { return M3(vmul(a, b.x), vmul(a, b.y), vmul(a, b.z)); }
static inline m3 mload(const REAL *p)
{
    return M3(V3(p[0], p[1], p[2]), V3(p[3], p[4], p[5]), V3(p[6], p[7], p[8]));
}
static inline void mstore(REAL *p, m3 a)
{
    p[0] = a.c0.x; p[1] = a.c0.y; p[2] = a.c0.z;
    p[3] = a.c1.x; p[4] = a.c1.y; p[5] = a.c1.z;
    p[6] = a.c2.x; p[7] = a.c2.y; p[8] = a.c2.z;
}
static inline v3 vload(const REAL *p) { return V3(p[0], p[1], p[2]); }
static inline void vstore(REAL *p, v3 a) { p[0] = a.x; p[1] = a.y; p[2] = a.z; }
static inline REAL rclamp(REAL x, REAL lo, REAL hi)
{
    /* Rust f32::clamp: NaN propagates. */
    if (x < lo) x = lo;
    if (x > hi) x = hi;
    return x;
}
static inline REAL wordf(const float *params, int i) { return (REAL)params[i]; }

/* ------------------------------------------------------------------------- */
/* Keys, blocks and buckets. */

/* Position::associated_cell_index_in_block_off_by_one + block_associated_to_point
 * (solver/particle.rs:208-225, grid/grid.rs:122-134). */
static void particle_block(v3 pt, REAL h, int32_t vid[3], uint32_t local[3])
{
    REAL p[3] = {pt.x, pt.y, pt.z};
    for (int d = 0; d < 3; ++d) {
        REAL cell = RROUND(p[d] / h) - K(1.0);
        REAL blk = RFLOOR(cell / K(4.0));
        vid[d] = (int32_t)blk;
        local[d] = (uint32_t)(cell - blk * K(4.0));
    }
}

/* BlockVirtualId::pack (grid/grid.rs:102-106): 11/10/11 bits. */
static uint32_t pack_vid(const int32_t v[3])
{
    return ((uint32_t)(v[0] + 0x3FF) & 0x7FFu) | (((uint32_t)(v[1] + 0x1FF) & 0x3FFu) << 11) |
           (((uint32_t)(v[2] + 0x3FF) & 0x7FFu) << 21);
}

/* hash (grid/grid.rs:430-436) */
static uint32_t hash_key(uint32_t k)
{
    k *= 0xCC9E2D51u;
    k = (k << 15) | (k >> 17);
    k *= 0x1B873593u;
    return k;
}

/* blocks_associated_to_block order (grid/grid.rs:175-212). */
static const int ASSOC_SHIFT[8][3] = {{0, 0, 0}, {0, 0, 1}, {0, 1, 0}, {0, 1, 1},
                                      {1, 0, 0}, {1, 0, 1}, {1, 1, 0}, {1, 1, 1}};

/* NBH_SHIFTS (grid/kernel.rs:42-70): the G2P summation order. */
static const int NBH_SHIFTS[27][3] = {
    {2, 2, 2}, {2, 0, 2}, {2, 1, 2}, {0, 2, 2}, {0, 0, 2}, {0, 1, 2}, {1, 2, 2}, {1, 0, 2}, {1, 1, 2},
    {2, 2, 0}, {2, 0, 0}, {2, 1, 0}, {0, 2, 0}, {0, 0, 0}, {0, 1, 0}, {1, 2, 0}, {1, 0, 0}, {1, 1, 0},
    {2, 2, 1}, {2, 0, 1}, {2, 1, 1}, {0, 2, 1}, {0, 0, 1}, {0, 1, 1}, {1, 2, 1}, {1, 0, 1}, {1, 1, 1}};

#define NUM_BUCKETS 10
#define NUM_PRIMARY_BUCKETS 4
#define EXTRA_MIN_SHIFT 2u /* grid/sort.rs:35 */

/* extra_sort_bucket (grid/sort.rs:56-62) */
static int extra_bucket(const uint32_t local[3], const int bshift[3])
{
    int l = (int)local[2] - bshift[2] * 4;
    if (l < -2) l = -2;
    return NUM_PRIMARY_BUCKETS + (l + 2);
}

/* spill rule (grid/sort.rs:316-328) */
static int spills(const uint32_t local[3], const int bshift[3])
{
    for (int d = 0; d < 3; ++d)
        if (bshift[d] != 0 && local[d] < EXTRA_MIN_SHIFT) return 0;
    return 1;
}

typedef struct {
    uint32_t *keys, *vals;
    uint32_t cap;
    uint32_t len;
} hmap;

static void hmap_init(hmap *m, uint32_t cap)
{
    m->cap = cap;
    m->len = 0;
    m->keys = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    m->vals = (uint32_t *)malloc(sizeof(uint32_t) * cap);
    memset(m->keys, 0xFF, sizeof(uint32_t) * cap);
}
static void hmap_free(hmap *m) { free(m->keys); free(m->vals); }
static uint32_t hmap_find(const hmap *m, uint32_t key)
{
    uint32_t s = hash_key(key) & (m->cap - 1);
    for (uint32_t i = 0; i < m->cap; ++i) {
        if (m->keys[s] == key) return m->vals[s];
        if (m->keys[s] == NONE) return NONE;
        s = (s + 1) & (m->cap - 1);
    }
    return NONE;
}
static int hmap_insert(hmap *m, uint32_t key, uint32_t val);
static void hmap_grow(hmap *m)
{
    hmap bigger;
    hmap_init(&bigger, m->cap * 2);
    for (uint32_t i = 0; i < m->cap; ++i)
        if (m->keys[i] != NONE) hmap_insert(&bigger, m->keys[i], m->vals[i]);
    hmap_free(m);
    *m = bigger;
}
/* Returns 1 if inserted (new), 0 if present. */
static int hmap_insert(hmap *m, uint32_t key, uint32_t val)
{
    if (2 * (m->len + 1) > m->cap) hmap_grow(m);
    uint32_t s = hash_key(key) & (m->cap - 1);
    for (;;) {
        if (m->keys[s] == key) return 0;
        if (m->keys[s] == NONE) {
            m->keys[s] = key;
            m->vals[s] = val;
            m->len++;
            return 1;
        }
        s = (s + 1) & (m->cap - 1);
    }
}

/* ------------------------------------------------------------------------- */
/* Quadratic B-spline (grid/kernel.rs:103-180). */

static inline REAL qk_eval(REAL x)
{
    REAL xa = RFABS(x);
    REAL part1 = K(0.75) - xa * xa;
    REAL part2 = K(0.5) * (K(1.5) - xa) * (K(1.5) - xa);
    REAL lt05 = (xa < K(0.5)) ? K(1.0) : K(0.0);
    REAL lt15 = (xa < K(1.5)) ? K(1.0) : K(0.0);
    return lt05 * part1 + (K(1.0) - lt05) * lt15 * part2;
}

static inline v3 qk_eval_all(REAL x)
{
    return V3(K(0.5) * (K(1.5) - x) * (K(1.5) - x), K(0.75) - (x - K(1.0)) * (x - K(1.0)),
              K(0.5) * (x - K(0.5)) * (x - K(0.5)));
}

static inline REAL qk_inv_d(REAL h) { return K(4.0) / (h * h); }

/* ------------------------------------------------------------------------- */
/* 3x3 SVD: one-sided Jacobi (Hestenes).  F = U diag(s) V^T with U, V proper
 * rotations; the smallest singular value carries the sign of det F. */

static void jacobi_rotate(v3 *ap, v3 *aq, v3 *vp, v3 *vq, REAL *off)
{
    REAL alpha = vdot(*ap, *ap), beta = vdot(*aq, *aq), gamma = vdot(*ap, *aq);
    REAL scale = RSQRT(alpha * beta);
    if (!(scale > K(0.0))) return;
    REAL rel = RFABS(gamma) / scale;
    if (rel > *off) *off = rel;
    if (rel <= JACOBI_EPS) return;
    REAL zeta = (beta - alpha) / (K(2.0) * gamma);
    REAL t;
    if (RFABS(zeta) > K(1.0e15)) {
        t = K(0.5) / zeta;
    } else {
        REAL sgn = zeta >= K(0.0) ? K(1.0) : K(-1.0);
        t = sgn / (RFABS(zeta) + RSQRT(K(1.0) + zeta * zeta));
    }
    REAL c = K(1.0) / RSQRT(K(1.0) + t * t);
    REAL s = c * t;
    v3 a0 = *ap, a1 = *aq, w0 = *vp, w1 = *vq;
    *ap = vsub(vmul(a0, c), vmul(a1, s));
    *aq = vadd(vmul(a0, s), vmul(a1, c));
    *vp = vsub(vmul(w0, c), vmul(w1, s));
    *vq = vadd(vmul(w0, s), vmul(w1, c));
}

static v3 any_orthogonal(v3 a)
{
    v3 e = RFABS(a.x) < K(0.57) ? V3(1, 0, 0) : (RFABS(a.y) < K(0.57) ? V3(0, 1, 0) : V3(0, 0, 1));
    v3 c = vcross(a, e);
    return vdiv(c, vlen(c));
}

static void svd3_impl(m3 F, m3 *U, v3 *s, m3 *V)
{
    v3 a[3] = {F.c0, F.c1, F.c2};
    v3 w[3] = {V3(1, 0, 0), V3(0, 1, 0), V3(0, 0, 1)};
    for (int sweep = 0; sweep < 40; ++sweep) {
        REAL off = K(0.0);
        jacobi_rotate(&a[0], &a[1], &w[0], &w[1], &off);
        jacobi_rotate(&a[0], &a[2], &w[0], &w[2], &off);
        jacobi_rotate(&a[1], &a[2], &w[1], &w[2], &off);
        if (off <= JACOBI_EPS) break;
    }
    REAL sig[3];
    v3 u[3];
    for (int i = 0; i < 3; ++i) sig[i] = vlen(a[i]);
    /* Order by decreasing singular value. */
    for (int i = 0; i < 2; ++i)
        for (int j = i + 1; j < 3; ++j)
            if (sig[j] > sig[i]) {
                REAL ts = sig[i]; sig[i] = sig[j]; sig[j] = ts;
                v3 ta = a[i]; a[i] = a[j]; a[j] = ta;
                v3 tw = w[i]; w[i] = w[j]; w[j] = tw;
            }
    REAL tiny = sig[0] * K(1.0e-30);
    for (int i = 0; i < 3; ++i) u[i] = sig[i] > tiny && sig[i] > K(0.0) ? vdiv(a[i], sig[i]) : V3(0, 0, 0);
    if (!(sig[0] > K(0.0))) { u[0] = V3(1, 0, 0); u[1] = V3(0, 1, 0); u[2] = V3(0, 0, 1); }
    else if (!(sig[1] > tiny)) { u[1] = any_orthogonal(u[0]); u[2] = vcross(u[0], u[1]); }
    else if (!(sig[2] > tiny)) { u[2] = vcross(u[0], u[1]); }
    m3 Um = M3(u[0], u[1], u[2]);
    m3 Vm = M3(w[0], w[1], w[2]);
    if (mdet(Vm) < K(0.0)) { Vm.c2 = vmul(Vm.c2, K(-1.0)); Um.c2 = vmul(Um.c2, K(-1.0)); }
    if (mdet(Um) < K(0.0)) { Um.c2 = vmul(Um.c2, K(-1.0)); sig[2] = -sig[2]; }
    *U = Um;
    *V = Vm;
    *s = V3(sig[0], sig[1], sig[2]);
}

/* ------------------------------------------------------------------------- */
/* Constitutive models. */

/* NeoHookeanModel::kirchoff_stress (models/neo_hookean_elasticity.rs:26-39) */
static m3 nh_stress(REAL lambda, REAL mu, m3 F)
{
    REAL j = RFMAX(mdet(F), K(1.0e-10));
    REAL diag_val = lambda * RLOG(j) - mu;
    m3 st = mscale(mmul(F, mtrans(F)), mu);
    st.c0.x += diag_val;
    st.c1.y += diag_val;
    st.c2.z += diag_val;
    return st;
}

/* LinearElasticModel::kirchoff_stress (models/linear_elasticity.rs:30-54) */
static m3 linear_stress(REAL lambda, REAL mu, m3 F)
{
    m3 U, V;
    v3 s;
    svd3_impl(F, &U, &s, &V);
    REAL j = s.x * s.y * s.z;
    v3 ms = vsub(s, V3(1, 1, 1));
    m3 rec = mmul(mmul(U, mdiag(ms)), mtrans(V));
    REAL diag_val = lambda * (j - K(1.0)) * j;
    m3 r = mscale(mmul(rec, mtrans(F)), K(2.0) * mu);
    r.c0.x += diag_val;
    r.c1.y += diag_val;
    r.c2.z += diag_val;
    return r;
}

/* DruckerPragerPlasticity::alpha (models/drucker_prager.rs:64-68) */
static REAL dp_alpha(const REAL dp[7], REAL q)
{
    REAL angle = dp[0] + (dp[1] * q - dp[3]) * REXP(-dp[2] * q);
    REAL sa = RSIN(angle);
    return RSQRT(K(2.0) / K(3.0)) * (K(2.0) * sa) / (K(3.0) - sa);
}

/* DruckerPragerPlasticity::project (models/drucker_prager.rs:224-267) with
 * project_deformation_gradient (:121-167).  dp = {ha, hb, hc, hd, lambda, mu, cohesion}. */
static m3 dp_project(const REAL dp[7], REAL st[3], m3 F)
{
    if (dp[4] == K(0.0)) return F;
    m3 U, V;
    v3 s;
    svd3_impl(F, &U, &s, &V);
    REAL alpha = dp_alpha(dp, st[1]);
    const REAL d = K(3.0);
    v3 strain = vadd(V3(RLOG(s.x), RLOG(s.y), RLOG(s.z)), V3(st[2] / d, st[2] / d, st[2] / d));
    REAL tr = strain.x + strain.y + strain.z;
    v3 dev = vsub(strain, V3(tr / d, tr / d, tr / d));
    REAL shifted = tr - dp[6];
    v3 ns;
    REAL dq;
    if (shifted > K(0.0) || (dev.x == K(0.0) && dev.y == K(0.0) && dev.z == K(0.0))) {
        REAL e = REXP(dp[6] / d);
        ns = V3(e, e, e);
        dq = vlen(strain);
    } else {
        REAL dn = vlen(dev);
        REAL gamma = dn + (d * dp[4] + K(2.0) * dp[5]) / (K(2.0) * dp[5]) * shifted * alpha;
        if (gamma <= K(0.0)) return F; /* elastic: state and F unchanged */
        v3 hh = vsub(strain, vmul(dev, gamma / dn));
        ns = V3(REXP(hh.x), REXP(hh.y), REXP(hh.z));
        dq = gamma;
    }
    REAL prev_det = s.x * s.y * s.z;
    REAL new_det = ns.x * ns.y * ns.z;
    st[0] = st[0] * prev_det / new_det;
    st[2] = st[2] + RLOG(prev_det) - RLOG(new_det);
    st[1] = st[1] + dq;
    return mmul(mmul(U, mdiag(ns)), mtrans(V));
}

/* SnowPlasticity::hardening_factor (models/snow.rs:77-79): exp(xi (1 - det Fp)). */
static REAL snow_hardening(REAL xi, REAL pdet) { return REXP(xi * (K(1.0) - pdet)); }

/* SnowPlasticity::project (models/snow.rs:83-113), sp = {theta_c, theta_s, xi}: clamp the
 * singular values to [1 - theta_c, 1 + theta_s]; det Fp absorbs the removed volume change
 * and stays in [0.1, 4].  st[0] = det Fp; returns the projected F, *hard the hardening. */
static m3 snow_project(const REAL sp[3], REAL st[3], m3 F, REAL *hard)
{
    if (sp[0] <= K(0.0) && sp[1] <= K(0.0)) {
        *hard = snow_hardening(sp[2], st[0]);
        return F;
    }
    m3 U, V;
    v3 s;
    svd3_impl(F, &U, &s, &V);
    const REAL lo = K(1.0) - sp[0], hi = K(1.0) + sp[1];
    const v3 c = V3(rclamp(s.x, lo, hi), rclamp(s.y, lo, hi), rclamp(s.z, lo, hi));
    const REAL prev_det = s.x * s.y * s.z;
    const REAL new_det = c.x * c.y * c.z;
    st[0] = rclamp(st[0] * prev_det / new_det, K(0.1), K(4.0));
    *hard = snow_hardening(sp[2], st[0]);
    return mmul(mmul(U, mdiag(c)), mtrans(V));
}

/* FluidModel::pressure / kirchoff_stress (models/fluid.rs:54-84) */
static REAL fluid_pressure(const REAL fl[5], REAL j)
{
    j = RFMAX(j, K(1.0e-6));
    if (j <= K(1.0)) {
        REAL ratio = REXP(-fl[1] * RLOG(j));
        return fl[0] * (ratio - K(1.0));
    }
    return -fl[0] * fl[4] * (j - K(1.0));
}

static m3 fluid_stress(const REAL fl[5], m3 F, m3 gradv)
{
    REAL j = RFMAX(mdet(F), K(1.0e-6));
    m3 sr = mscale(madd(gradv, mtrans(gradv)), K(0.5)); /* ParticleUpdateData::strain_rate */
    REAL sph = mtrace(sr) / K(3.0);                       /* models/utils.rs:60-110 */
    sr.c0.x -= sph;
    sr.c1.y -= sph;
    sr.c2.z -= sph;
    m3 stress = mscale(sr, K(2.0) * fl[2] * j);
    REAL diag_val = -fluid_pressure(fl, j) * j;
    stress.c0.x += diag_val;
    stress.c1.y += diag_val;
    stress.c2.z += diag_val;
    return stress;
}

/* DefaultParticleModel::update (models/default.rs:208-264).  Returns 0; F and the
 * per-particle plastic state st are updated in place (unknown tags: zero stress). */
static int model_update(uint32_t tag, const float *w, REAL st[3], m3 *F, m3 gradv, m3 *tau)
{
    switch (tag) {
    case 0:
        *tau = linear_stress(wordf(w, 0), wordf(w, 1), *F);
        return 0;
    case 1:
        *tau = nh_stress(wordf(w, 0), wordf(w, 1), *F);
        return 0;
    case 2:
    case 3: {
        REAL dp[7];
        for (int i = 0; i < 7; ++i) dp[i] = wordf(w, 3 + i);
        *F = dp_project(dp, st, *F);
        if (tag == 2) *tau = linear_stress(wordf(w, 10), wordf(w, 11), *F);
        else *tau = nh_stress(wordf(w, 10), wordf(w, 11), *F);
        return 0;
    }
    case 4: {
        REAL fl[5];
        for (int i = 0; i < 5; ++i) fl[i] = wordf(w, i);
        *tau = fluid_stress(fl, *F, gradv);
        return 0;
    }
    case 5: {
        /* SnowModel: words 0 = det Fp (state), 1-3 plasticity, 4-6 linear elastic backbone
         * (models/default.rs:78-120, 254-261). */
        REAL sp[3] = {wordf(w, 1), wordf(w, 2), wordf(w, 3)};
        REAL hard;
        *F = snow_project(sp, st, *F, &hard);
        *tau = linear_stress(wordf(w, 4) * hard, wordf(w, 5) * hard, *F);
        return 0;
    }
    default:
        *tau = mzero();
        return 0;
    }
}

/* ------------------------------------------------------------------------- */
/* Cuboid colliders. */

/* Rotation of v by quaternion (b, w): v (w^2 - b.b) + b (2 b.v) + (b x v)(2 w). */
static v3 quat_rotate(REAL qx, REAL qy, REAL qz, REAL qw, v3 v)
{
    v3 b = V3(qx, qy, qz);
    REAL b2 = vdot(b, b);
    return vadd(vadd(vmul(v, qw * qw - b2), vmul(b, vdot(v, b) * K(2.0))),
                vmul(vcross(b, v), qw * K(2.0)));
}

/* World centre of mass of a collider's body: pose * local_com
 * (LocalMassProperties::to_world, src_mpm_shaders/solver/rigid_impulses.rs:163-185). */
static v3 collider_com(const orc_cuboid *c)
{
    v3 lc = V3(c->local_com[0], c->local_com[1], c->local_com[2]);
    return vadd(quat_rotate(c->rotation[0], c->rotation[1], c->rotation[2], c->rotation[3], lc),
                V3(c->translation[0], c->translation[1], c->translation[2]));
}

/* parry3d 0.30 Cuboid::project_local_point(pt, solid = false), reached through
 * Shape::project_point_on_boundary (src_rbd_shaders/shapes/shape.rs:378-384): outside,
 * the clamp to the box; inside, the nearest face (first axis wins ties, the min face a
 * same-axis tie). */
static v3 project_cuboid_local(const orc_cuboid *c, v3 local, int *inside)
{
    v3 he = V3(c->half_extents[0], c->half_extents[1], c->half_extents[2]);
    v3 mins_pt = vsub(vmul(he, K(-1.0)), local);
    v3 pt_maxs = vsub(local, he);
    v3 shift = vsub(V3(RFMAX(mins_pt.x, K(0.0)), RFMAX(mins_pt.y, K(0.0)), RFMAX(mins_pt.z, K(0.0))),
                    V3(RFMAX(pt_maxs.x, K(0.0)), RFMAX(pt_maxs.y, K(0.0)), RFMAX(pt_maxs.z, K(0.0))));
    if (!(shift.x == K(0.0) && shift.y == K(0.0) && shift.z == K(0.0))) {
        *inside = 0;
        return vadd(local, shift);
    }
    *inside = 1;
    REAL best = -MAX_FLT_REF * K(1.0e20);
    int is_mins = 0, best_id = 0;
    for (int i = 0; i < 3; ++i) {
        REAL a = vget(mins_pt, i), b = vget(pt_maxs, i);
        if (a < b) {
            if (b > best) { best_id = i; is_mins = 0; best = b; }
        } else if (a > best) {
            best_id = i; is_mins = 1; best = a;
        }
    }
    v3 sh = V3(0, 0, 0);
    vset(&sh, best_id, is_mins ? best : -best);
    return vadd(local, sh);
}

/* parry3d 0.30 Ball::project_local_point(pt, solid = false): pt * (r / |pt|), inside
 * when |pt|^2 <= r^2 (NaN at the centre). */
static v3 project_ball_local(REAL r, v3 p, int *inside)
{
    REAL d2 = vdot(p, p);
    *inside = d2 <= r * r;
    return vmul(p, r / RSQRT(d2));
}

/* Segment::project_local_point (src_rbd_shaders/shapes/segment.rs:43-60). */
static v3 project_segment(v3 a, v3 b, v3 p)
{
    v3 ab = vsub(b, a), ap = vsub(p, a);
    REAL t = vdot(ab, ap), n2 = vdot(ab, ab);
    if (t <= K(0.0)) return a;
    if (t >= n2) return b;
    return vadd(a, vmul(ab, t / n2));
}

/* Capsule::project_local_point_on_boundary (capsule.rs:62-83) with orthonormal_basis3
 * (capsule.rs:150-161) for a point on the axis. */
static v3 project_capsule_local(const float *d, v3 p, int *inside)
{
    v3 a = V3(d[0], d[1], d[2]), b = V3(d[3], d[4], d[5]);
    REAL r = d[6];
    v3 on_axis = project_segment(a, b, p);
    v3 dp = vsub(p, on_axis);
    REAL dist = vlen(dp);
    if (dist > K(0.0)) {
        *inside = dist <= r;
        return vadd(on_axis, vmul(dp, r / dist));
    }
    v3 ax = vsub(b, a);
    v3 u = vdiv(ax, vlen(ax));
    REAL sg = u.z >= K(0.0) ? K(1.0) : K(-1.0);
    REAL k = K(-1.0) / (sg + u.z);
    REAL m = u.x * u.y * k;
    *inside = 1;
    return vadd(on_axis, vmul(V3(K(1.0) + sg * u.x * u.x * k, sg * m, -sg * u.x), r));
}

/* Cylinder::project_local_point_on_boundary (cylinder.rs:90-148), axis y. */
static v3 project_cylinder_local(REAL hh, REAL r, v3 p, int *inside)
{
    REAL planar = RSQRT(p.x * p.x + p.z * p.z);
    REAL ux = K(1.0), uz = K(0.0);
    if (planar > K(0.0)) { ux = p.x / planar; uz = p.z / planar; }
    REAL rx = ux * r, rz = uz * r;
    if (p.y >= -hh && p.y <= hh && planar <= r) {
        *inside = 1;
        REAL top = hh - p.y, bottom = p.y - (-hh), side = r - planar;
        if (top < bottom && top < side) return V3(p.x, hh, p.z);
        if (bottom < top && bottom < side) return V3(p.x, -hh, p.z);
        return V3(rx, p.y, rz);
    }
    *inside = 0;
    if (p.y > hh) return planar <= r ? V3(p.x, hh, p.z) : V3(rx, hh, rz);
    if (p.y < -hh) return planar <= r ? V3(p.x, -hh, p.z) : V3(rx, -hh, rz);
    return V3(rx, p.y, rz);
}

/* Cone::project_local_point_on_boundary (cone.rs:98-147): apex (0, hh, 0), base disc at
 * y = -hh. */
static v3 project_cone_local(REAL hh, REAL r, v3 p, int *inside)
{
    REAL planar = RSQRT(p.x * p.x + p.z * p.z);
    REAL ux = K(1.0), uz = K(0.0);
    if (planar > K(0.0)) { ux = p.x / planar; uz = p.z / planar; }
    v3 on_base = V3(p.x, -hh, p.z);
    if (p.y < -hh && planar <= r) {
        *inside = 0;
        return on_base;
    }
    v3 rim = V3(ux * r, -hh, uz * r);
    v3 apex = V3(K(0.0), hh, K(0.0));
    v3 side = vsub(rim, apex);
    v3 proj = project_segment(apex, rim, p);
    v3 down = V3(K(0.0), K(-2.0) * hh, K(0.0));
    if (p.y >= -hh && p.y <= hh && vdot(vcross(side, vsub(p, apex)), vcross(side, down)) >= K(0.0)) {
        *inside = 1;
        v3 to_side = vsub(proj, p), to_base = vsub(on_base, p);
        return vdot(to_side, to_side) > vdot(to_base, to_base) ? on_base : proj;
    }
    *inside = 0;
    return proj;
}

/* Shape::project_point_on_boundary (src_rbd_shaders/shapes/shape.rs:378-399): the point
 * in the body frame, rotation^-1 (pt - translation), projected, and back. */
static v3 project_shape(const orc_cuboid *c, v3 pt, int *inside)
{
    REAL qx = c->rotation[0], qy = c->rotation[1], qz = c->rotation[2], qw = c->rotation[3];
    v3 t = V3(c->translation[0], c->translation[1], c->translation[2]);
    v3 local = quat_rotate(-qx, -qy, -qz, qw, vsub(pt, t));
    v3 lp;
    switch (c->shape) {
    case ORC_SHAPE_BALL: lp = project_ball_local(c->dims[0], local, inside); break;
    case ORC_SHAPE_CAPSULE: lp = project_capsule_local(c->dims, local, inside); break;
    case ORC_SHAPE_CONE: lp = project_cone_local(c->dims[0], c->dims[1], local, inside); break;
    case ORC_SHAPE_CYLINDER: lp = project_cylinder_local(c->dims[0], c->dims[1], local, inside); break;
    default: lp = project_cuboid_local(c, local, inside); break;
    }
    return vadd(quat_rotate(qx, qy, qz, qw, lp), t);
}

/* BoundaryCondition::project_velocity (solver/boundary_condition.rs:77-114) */
static v3 bc_project(uint32_t ty, REAL friction, v3 vel, v3 n)
{
    if (ty == 0) return V3(0, 0, 0);
    if (ty == 1) {
        REAL nv = vdot(vel, n);
        return vsub(vel, vmul(n, nv));
    }
    if (ty == 2) {
        REAL nv = vdot(vel, n);
        if (nv < K(0.0)) {
            v3 tv = vsub(vel, vmul(n, nv));
            REAL tl = vlen(tv);
            v3 dir = tl > K(1.0e-8) ? vdiv(tv, tl) : V3(0, 0, 0);
            REAL pl = tl + friction * nv;
            pl = pl > K(0.0) ? pl : K(0.0);
            return vmul(dir, pl);
        }
        return vel;
    }
    return vel;
}

/* ------------------------------------------------------------------------- */
/* Mesh colliders in the grid collider pass (CPIC off). */

/* Triangle::project_local_point_and_get_location as the mesh query uses it
 * (src_rbd_shaders/shapes/trimesh.rs:237-391: Ericson's Voronoi regions, a then b then c,
 * then the edges ab, ac, bc by the signed volumes of stable_check_edges_voronoi, then the
 * face; a degenerate face falls through to the solid case, the point itself).  Feature
 * types: 0 vertex (id 0 a, 1 b, 2 c), 1 edge (id 0 ab, 1 bc, 2 ac), 2 face, 3 solid. */
static v3 tri_project(v3 a, v3 b, v3 c, v3 pt, int *ftype, int *fid)
{
    v3 ab = vsub(b, a), ac = vsub(c, a), ap = vsub(pt, a);
    REAL ab_ap = vdot(ab, ap), ac_ap = vdot(ac, ap);
    if (ab_ap <= K(0.0) && ac_ap <= K(0.0)) { *ftype = 0; *fid = 0; return a; }
    v3 bp = vsub(pt, b);
    REAL ab_bp = vdot(ab, bp), ac_bp = vdot(ac, bp);
    if (ab_bp >= K(0.0) && ac_bp <= ab_bp) { *ftype = 0; *fid = 1; return b; }
    v3 cp = vsub(pt, c);
    REAL ab_cp = vdot(ab, cp), ac_cp = vdot(ac, cp);
    if (ac_cp >= K(0.0) && ab_cp <= ac_cp) { *ftype = 0; *fid = 2; return c; }
    v3 bc = vsub(c, b);
    v3 n = vcross(ab, ac);
    REAL vc = vdot(n, vcross(ab, ap));
    if (vc < K(0.0) && ab_ap >= K(0.0) && ab_bp <= K(0.0)) {
        REAL v = ab_ap / vdot(ab, ab);
        *ftype = 1; *fid = 0;
        return vadd(a, vmul(ab, v));
    }
    REAL vb = -vdot(n, vcross(ac, cp));
    if (vb < K(0.0) && ac_ap >= K(0.0) && ac_cp <= K(0.0)) {
        REAL w = ac_ap / vdot(ac, ac);
        *ftype = 1; *fid = 2;
        return vadd(a, vmul(ac, w));
    }
    REAL va = vdot(n, vcross(bc, bp));
    if (va < K(0.0) && ac_bp - ab_bp >= K(0.0) && ab_cp - ac_cp >= K(0.0)) {
        REAL w = vdot(bc, bp) / vdot(bc, bc);
        *ftype = 1; *fid = 1;
        return vadd(b, vmul(bc, w));
    }
    if (va + vb + vc != K(0.0)) {
        REAL denom = K(1.0) / (va + vb + vc);
        REAL v = vb * denom, w = vc * denom;
        *ftype = 2; *fid = 0;
        return vadd(vadd(a, vmul(ab, v)), vmul(ac, w));
    }
    *ftype = 3; *fid = 0;
    return pt;
}

/* A mesh with the pseudo-normals of its inside test (parry3d 0.30's
 * TriMesh::pseudo_normals, read by trimesh.rs:116-152): per vertex the sum of its faces'
 * unit normals weighted by each face's angle at the vertex; per triangle edge (ab, bc,
 * ca) the sum of the unit normals of the faces sharing it (the angle at an edge is pi for
 * every face); faces without a normal (zero area) add nothing.  Both normalized. */
typedef struct {
    const orc_trimesh *m;
    double *vn;   /* nv*3 */
    double *en;   /* nt*9 */
} mesh_pn;

typedef struct {
    uint64_t key;   /* (min vertex << 32) | max vertex */
    int64_t slot;   /* 3 triangle + edge */
} edge_ref;

static int edge_ref_cmp(const void *x, const void *y)
{
    const edge_ref *a = (const edge_ref *)x, *b = (const edge_ref *)y;
    if (a->key != b->key) return a->key < b->key ? -1 : 1;
    return a->slot < b->slot ? -1 : (a->slot > b->slot);
}

static void dnormalize(double *v)
{
    double l = sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    if (l > 0.0) { v[0] /= l; v[1] /= l; v[2] /= l; }
}

static void mesh_pseudo_normals(mesh_pn *mp)
{
    const orc_trimesh *m = mp->m;
    mp->vn = (double *)calloc((size_t)(3 * (m->nv > 0 ? m->nv : 1)), sizeof(double));
    mp->en = (double *)calloc((size_t)(9 * (m->nt > 0 ? m->nt : 1)), sizeof(double));
    double *fn = (double *)calloc((size_t)(3 * (m->nt > 0 ? m->nt : 1)), sizeof(double));
    edge_ref *er = (edge_ref *)malloc(sizeof(edge_ref) * (size_t)(3 * (m->nt > 0 ? m->nt : 1)));
    for (int64_t t = 0; t < m->nt; ++t) {
        const uint32_t *id = m->triangles + 3 * t;
        double p[3][3];
        for (int k = 0; k < 3; ++k)
            for (int c = 0; c < 3; ++c) p[k][c] = m->vertices[3 * id[k] + c];
        double e1[3], e2[3];
        for (int c = 0; c < 3; ++c) { e1[c] = p[1][c] - p[0][c]; e2[c] = p[2][c] - p[0][c]; }
        double n[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2], e1[0] * e2[1] - e1[1] * e2[0]};
        double l = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (l > 0.0) {
            for (int c = 0; c < 3; ++c) fn[3 * t + c] = n[c] / l;
            for (int k = 0; k < 3; ++k) {
                const double *o = p[k], *q1 = p[(k + 1) % 3], *q2 = p[(k + 2) % 3];
                double u[3], v[3];
                for (int c = 0; c < 3; ++c) { u[c] = q1[c] - o[c]; v[c] = q2[c] - o[c]; }
                double cr[3] = {u[1] * v[2] - u[2] * v[1], u[2] * v[0] - u[0] * v[2], u[0] * v[1] - u[1] * v[0]};
                double ang = atan2(sqrt(cr[0] * cr[0] + cr[1] * cr[1] + cr[2] * cr[2]), u[0] * v[0] + u[1] * v[1] + u[2] * v[2]);
                for (int c = 0; c < 3; ++c) mp->vn[3 * id[k] + c] += fn[3 * t + c] * ang;
            }
        }
        for (int k = 0; k < 3; ++k) {
            uint32_t i0 = id[k], i1 = id[(k + 1) % 3];
            uint32_t lo = i0 < i1 ? i0 : i1, hi = i0 < i1 ? i1 : i0;
            er[3 * t + k].key = ((uint64_t)lo << 32) | hi;
            er[3 * t + k].slot = 3 * t + k;
        }
    }
    for (int64_t q = 0; q < m->nv; ++q) dnormalize(mp->vn + 3 * q);
    qsort(er, (size_t)(3 * m->nt), sizeof(edge_ref), edge_ref_cmp);
    for (int64_t a = 0; a < 3 * m->nt;) {
        int64_t b = a;
        double sum[3] = {0.0, 0.0, 0.0};
        while (b < 3 * m->nt && er[b].key == er[a].key) {
            for (int c = 0; c < 3; ++c) sum[c] += fn[3 * (er[b].slot / 3) + c];
            ++b;
        }
        dnormalize(sum);
        for (int64_t q = a; q < b; ++q)
            for (int c = 0; c < 3; ++c) mp->en[3 * er[q].slot + c] = sum[c];
        a = b;
    }
    free(fn);
    free(er);
}

static void mesh_pn_free(mesh_pn *mp)
{
    free(mp->vn);
    free(mp->en);
    mp->vn = mp->en = NULL;
}

/* TriMesh::project_local_point (trimesh.rs:155-207) without its BVH: the nearest triangle
 * within max_dist (a later triangle wins a tie, dist <= best), inside when the feature's
 * pseudo-normal points away from the point (pn . (pt - proj) <= 0; a face's is ab x ac).
 * Returns 0 when no triangle is that near. */
static int project_trimesh_local(const mesh_pn *mp, v3 lp, REAL max_dist, v3 *proj, int *inside)
{
    const orc_trimesh *m = mp->m;
    REAL best = max_dist;
    int64_t best_t = -1;
    int bt = 0, bid = 0;
    v3 bp = lp, ba = lp, bb = lp, bc = lp;
    for (int64_t t = 0; t < m->nt; ++t) {
        const uint32_t *id = m->triangles + 3 * t;
        v3 a = V3(m->vertices[3 * id[0]], m->vertices[3 * id[0] + 1], m->vertices[3 * id[0] + 2]);
        v3 b = V3(m->vertices[3 * id[1]], m->vertices[3 * id[1] + 1], m->vertices[3 * id[1] + 2]);
        v3 c = V3(m->vertices[3 * id[2]], m->vertices[3 * id[2] + 1], m->vertices[3 * id[2] + 2]);
        int ft, fi;
        v3 p = tri_project(a, b, c, lp, &ft, &fi);
        REAL dist = vlen(vsub(p, lp));
        if (dist <= best) {
            best = dist; best_t = t; bp = p; bt = ft; bid = fi;
            ba = a; bb = b; bc = c;
        }
    }
    if (best_t < 0) return 0;
    v3 pn;
    if (bt == 0) {
        const double *q = mp->vn + 3 * m->triangles[3 * best_t + bid];
        pn = V3((REAL)q[0], (REAL)q[1], (REAL)q[2]);
    } else if (bt == 1) {
        /* edge ids: 0 ab, 1 bc, 2 ac = the pseudo-normal slots ab, bc, ca */
        const double *q = mp->en + 9 * best_t + 3 * bid;
        pn = V3((REAL)q[0], (REAL)q[1], (REAL)q[2]);
    } else {
        pn = vcross(vsub(bb, ba), vsub(bc, ba));
    }
    *inside = vdot(pn, vsub(lp, bp)) <= K(0.0);
    *proj = bp;
    return 1;
}

/* The mesh collider's projection in the world (grid_update_collide.rs:53-67): the point
 * in the body frame by the inverse pose (R^-1 p - R^-1 t), the projection back by the
 * pose. */
static int project_trimesh(const mesh_pn *mp, const orc_cuboid *c, REAL max_dist, v3 pt, v3 *proj, int *inside)
{
    REAL qx = c->rotation[0], qy = c->rotation[1], qz = c->rotation[2], qw = c->rotation[3];
    v3 t = V3(c->translation[0], c->translation[1], c->translation[2]);
    v3 lp = vsub(quat_rotate(-qx, -qy, -qz, qw, pt), quat_rotate(-qx, -qy, -qz, qw, t));
    v3 lproj;
    if (!project_trimesh_local(mp, lp, max_dist, &lproj, inside)) return 0;
    *proj = vadd(quat_rotate(qx, qy, qz, qw, lproj), t);
    return 1;
}

/* The meshes of the colliders (the last nmesh of them) for the node pass. */
typedef struct {
    mesh_pn *mp;
    int nmesh;
} mesh_set;

/* gpu_grid_update_collide body (solver/grid_update_collide.rs:30-99, 105-183).
 * Velocity::velocity_at_point (src_rbd_shaders/dynamics/body.rs:239-250) with the
 * collider's world centre of mass at its pose translation. */
static v3 collide_node(const orc_cuboid *cols, int ncol, REAL h, REAL dt, v3 pt, v3 vel, const mesh_set *ms)
{
    REAL cap = h * K(1.5);
    REAL best = MAX_FLT_REF;
    v3 normal = V3(0, 0, 0);
    int best_id = 0;
    for (int i = 0; i < ncol; ++i) {
        int inside;
        v3 proj;
        if (cols[i].shape == ORC_SHAPE_TRIMESH) {
            /* a mesh only reports triangles within 1.5 h (its max_dist) */
            const int k = i - (ncol - (ms ? ms->nmesh : 0));
            if (!ms || k < 0 || !project_trimesh(&ms->mp[k], &cols[i], cap, pt, &proj, &inside)) continue;
        } else {
            proj = project_shape(&cols[i], pt, &inside);
        }
        v3 dpt = vsub(proj, pt);
        int within = RFABS(dpt.x) <= cap && RFABS(dpt.y) <= cap && RFABS(dpt.z) <= cap;
        if (inside || within) {
            REAL sign = inside ? K(-1.0) : K(1.0);
            REAL dist = vlen(dpt);
            v3 n = vdiv(dpt, dist * -sign);
            REAL sd = sign * dist;
            if (sd < best) { best = sd; normal = n; best_id = i; }
        }
    }
    if (best != MAX_FLT_REF) {
        const orc_cuboid *c = &cols[best_id];
        v3 com = collider_com(c);
        v3 lin = V3(c->linvel[0], c->linvel[1], c->linvel[2]);
        v3 ang = V3(c->angvel[0], c->angvel[1], c->angvel[2]);
        v3 bv = vadd(lin, vcross(ang, vsub(pt, com)));
        v3 delta = vsub(vel, bv);
        REAL nv = vdot(delta, normal);
        REAL margin = h;
        if (best <= margin) {
            vel = vadd(bv, bc_project(c->bc_type, (REAL)c->bc_friction, delta, normal));
        } else if (-nv * dt > best - margin) {
            v3 excess = vmul(normal, nv + (best - margin) / dt);
            vel = vsub(vel, excess);
        }
    }
    return vel;
}

/* ------------------------------------------------------------------------- */
/* CPIC: contact distance field (solver/grid_update_cdf.rs, g2p_cdf.rs, p2g.rs,
 * g2p.rs and particle_update.rs with use_cpic).  Affinity words: bit i = collider i is
 * within reach, bit 16 + i = the side of collider i (grid/grid.rs:593-645). */

#define AFF_MASK 0x0000FFFFu
#define CPIC_MAX_COLLIDERS 16

static int aff_compatible(uint32_t a, uint32_t b)
{
    uint32_t common = a & b & AFF_MASK;
    return ((a >> 16) & common) == ((b >> 16) & common);
}

/* gpu_grid_update_cdf's collide (grid_update_cdf.rs:18-60): over the colliders, the
 * closest (strictly smaller distance wins) of those containing the node or within
 * 1.5 h on every axis, and their affinity bits with the side (inside) as sign. */
static void node_cdf(const orc_cuboid *cols, int ncol, REAL h, v3 pt, REAL *dist, uint32_t *aff,
                     uint32_t *closest)
{
    REAL best = MAX_FLT_REF;
    uint32_t a = 0, cid = NONE;
    const REAL cap = h * K(1.5);
    for (int i = 0; i < ncol; ++i) {
        int inside;
        v3 proj = project_shape(&cols[i], pt, &inside);
        v3 dpt = vsub(proj, pt);
        int within = RFABS(dpt.x) <= cap && RFABS(dpt.y) <= cap && RFABS(dpt.z) <= cap;
        if (inside || within) {
            REAL d = vlen(dpt);
            if (d < best) { cid = (uint32_t)i; best = d; }
            a |= (inside ? 0x00010001u : 0x00000001u) << i;
        }
    }
    *dist = best;
    *aff = a;
    *closest = cid;
}

/* ------------------------------------------------------------------------- */
/* Mesh colliders: rigid particles (surface samples) and p2g_cdf. */

/* f32 vector helpers of the sampler: the reference samples in f32 whatever REAL is. */
typedef struct { float x, y, z; } f3s;
static f3s fs_sub(f3s a, f3s b) { f3s r = {a.x - b.x, a.y - b.y, a.z - b.z}; return r; }
static f3s fs_add(f3s a, f3s b) { f3s r = {a.x + b.x, a.y + b.y, a.z + b.z}; return r; }
static f3s fs_mul(f3s a, float k) { f3s r = {a.x * k, a.y * k, a.z * k}; return r; }
static f3s fs_div(f3s a, float k) { f3s r = {a.x / k, a.y / k, a.z / k}; return r; }
static float fs_dot(f3s a, f3s b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static float fs_len(f3s a) { return sqrtf(fs_dot(a, a)); }
static f3s fs_load(const float *v, uint32_t i) { f3s r = {v[3 * i], v[3 * i + 1], v[3 * i + 2]}; return r; }

#define SAMPLE_EPS 1.0e-5f /* sample_trimesh.rs:12 */

typedef struct { float *pts; uint32_t *tri; int64_t n, cap; } sample_out;

static void sample_push(sample_out *o, f3s p, uint32_t tri)
{
    if (o->n < o->cap) {
        o->pts[3 * o->n] = p.x;
        o->pts[3 * o->n + 1] = p.y;
        o->pts[3 * o->n + 2] = p.z;
        o->tri[o->n] = tri;
    }
    o->n++;
}

/* sample_edge (sample_trimesh.rs:111-131): points a + dir (s i), i = 1 .. ceil(len / s) - 1,
 * s = spacing / sqrt 2. */
static void sample_edge(f3s a, f3s b, float spacing, uint32_t tri, sample_out *o)
{
    f3s ab = fs_sub(b, a);
    float len = fs_len(ab);
    if (!(len > SAMPLE_EPS)) return;
    f3s dir = fs_div(ab, len);
    float s = spacing / sqrtf(2.0f);
    int64_t nsteps = (int64_t)ceilf(len / s);
    for (int64_t i = 1; i < nsteps; ++i) sample_push(o, fs_add(a, fs_mul(dir, s * (float)i)), tri);
}

/* sample_triangle (sample_trimesh.rs:133-231): the longest edge is the base AB; rows of
 * points every s along the base, each row running up the height direction every s, as
 * far as the nearer of the lines AC and BC; neither the base nor the vertices are
 * sampled. */
static void sample_triangle(f3s a, f3s b, f3s c, float spacing, uint32_t tri, sample_out *o)
{
    float dab = fs_len(fs_sub(b, a)), dbc = fs_len(fs_sub(c, b)), dca = fs_len(fs_sub(a, c));
    float mx = fmaxf(fmaxf(dab, dbc), dca);
    f3s A = a, B = b, Cc = c;
    if (mx == dbc) { A = b; B = c; Cc = a; }
    else if (mx == dca) { A = c; B = a; Cc = b; }
    f3s ac = fs_sub(Cc, A), base = fs_sub(B, A);
    float base_len = fs_len(base);
    f3s base_dir = fs_div(base, base_len);
    float s = spacing / sqrtf(2.0f);
    float base_steps = ceilf(base_len / s);
    f3s base_step = fs_mul(base_dir, s);
    float ac_off = fs_dot(ac, base_dir);
    float bc_off = base_len - ac_off;
    if (ac_off < SAMPLE_EPS || bc_off < SAMPLE_EPS || base_len < SAMPLE_EPS) return;
    f3s height = fs_sub(ac, fs_mul(base_dir, ac_off));
    float height_len = fs_len(height);
    f3s height_dir = fs_div(height, height_len);
    float tan_a = height_len / ac_off, tan_b = height_len / bc_off;
    uint32_t nb = (uint32_t)base_steps;
    for (uint32_t i = 1; i < nb; ++i) {
        f3s bp = fs_add(A, fs_mul(base_step, (float)i));
        float h_ac = tan_a * fs_len(fs_sub(A, bp));
        float h_bc = tan_b * fs_len(fs_sub(B, bp));
        float hl = fminf(h_ac, h_bc);
        uint32_t nh = (uint32_t)ceilf(hl / s);
        f3s hstep = fs_mul(height_dir, s);
        for (uint32_t j = 1; j < nh; ++j) {
            f3s p = fs_add(bp, fs_mul(hstep, (float)j));
            if (!isfinite(p.x) || !isfinite(p.y) || !isfinite(p.z)) continue;
            sample_push(o, p, tri);
        }
    }
}

int64_t orc_sample_trimesh(const float *vertices, int64_t nv, const uint32_t *triangles, int64_t nt,
                           float spacing, float *points, uint32_t *triangle_of, int64_t cap)
{
    (void)nv;
    sample_out o = {points, triangle_of, 0, cap};
    /* edges seen so far, keyed (max, min) vertex index (sample_trimesh.rs:68-77) */
    int64_t hcap = 16;
    while (hcap < 6 * nt + 16) hcap *= 2;
    uint64_t *seen = (uint64_t *)malloc(sizeof(uint64_t) * (size_t)hcap);
    for (int64_t i = 0; i < hcap; ++i) seen[i] = ~0ull;
    for (int64_t t = 0; t < nt; ++t) {
        const uint32_t *ix = triangles + 3 * t;
        f3s v0 = fs_load(vertices, ix[0]), v1 = fs_load(vertices, ix[1]), v2 = fs_load(vertices, ix[2]);
        sample_triangle(v0, v1, v2, spacing, (uint32_t)t, &o);
        const uint32_t ea[3] = {ix[0], ix[1], ix[2]}, eb[3] = {ix[1], ix[2], ix[0]};
        for (int e = 0; e < 3; ++e) {
            uint32_t hi = ea[e] > eb[e] ? ea[e] : eb[e], lo = ea[e] > eb[e] ? eb[e] : ea[e];
            uint64_t key = ((uint64_t)hi << 32) | lo;
            uint64_t slot = (key * 0x9E3779B97F4A7C15ull) >> 20;
            int fresh = 0;
            for (;;) {
                slot &= (uint64_t)(hcap - 1);
                if (seen[slot] == key) break;
                if (seen[slot] == ~0ull) { seen[slot] = key; fresh = 1; break; }
                slot++;
            }
            if (fresh) sample_edge(fs_load(vertices, ea[e]), fs_load(vertices, eb[e]), spacing, (uint32_t)t, &o);
        }
    }
    free(seen);
    return o.n;
}

/* Pose of a collider's body applied to a body-frame point: rotation * p + translation. */
static v3 body_to_world(const orc_cuboid *c, const float *p)
{
    return vadd(quat_rotate(c->rotation[0], c->rotation[1], c->rotation[2], c->rotation[3], V3(p[0], p[1], p[2])),
                V3(c->translation[0], c->translation[1], c->translation[2]));
}

/* gpu_p2g_cdf (p2g_cdf.rs:60-259) as a scatter from the rigid particles: each sample's
 * triangle, posed, reaches the nodes of the 3-cell range from its associated cell
 * round(x / h) - 1; a node whose projection onto the triangle's plane falls inside the
 * triangle (edge tests with n = ab x ac) takes the unsigned plane distance and the
 * collider's affinity bit, with the side n . (x - a) < 0 as sign.  Per node the
 * closest mesh collider is the least (distance, collider id); it replaces the
 * analytic shapes' closest only when strictly nearer (p2g_cdf.rs:241). */
static void mesh_cdf(const orc_cpic *cp, const orc_cuboid *cols, int ncol, REAL h, const hmap *hm,
                     const int32_t *bvid, int64_t nb, REAL *ndist, uint32_t *naff, uint32_t *nclosest)
{
    (void)bvid;
    REAL *mdist = (REAL *)malloc(sizeof(REAL) * 64 * (size_t)(nb > 0 ? nb : 1));
    uint32_t *mid = (uint32_t *)malloc(sizeof(uint32_t) * 64 * (size_t)(nb > 0 ? nb : 1));
    for (int64_t i = 0; i < 64 * nb; ++i) { mdist[i] = MAX_FLT_REF; mid[i] = NONE; }
    for (int k = 0; k < cp->nmesh; ++k) {
        const orc_trimesh *m = &cp->meshes[k];
        const uint32_t cid = (uint32_t)(ncol - cp->nmesh + k);
        const orc_cuboid *body = &cols[cid];
        for (int64_t s = 0; s < m->ns; ++s) {
            v3 p = body_to_world(body, m->samples + 3 * s);
            const uint32_t *ix = m->triangles + 3 * m->sample_triangle[s];
            v3 a = body_to_world(body, m->vertices + 3 * ix[0]);
            v3 b = body_to_world(body, m->vertices + 3 * ix[1]);
            v3 c = body_to_world(body, m->vertices + 3 * ix[2]);
            REAL pc[3] = {p.x, p.y, p.z};
            int64_t base[3];
            int ok = 1;
            for (int d = 0; d < 3; ++d) {
                REAL cell = RROUND(pc[d] / h) - K(1.0);
                if (!(cell > K(-4000.0) && cell < K(4000.0))) ok = 0;
                base[d] = ok ? (int64_t)cell : 0;
            }
            if (!ok) continue;
            v3 ab = vsub(b, a), ac = vsub(c, a), bc = vsub(c, b);
            v3 n = vcross(ab, ac);
            REAL nl = vlen(n);
            if (nl == K(0.0)) continue;
            for (int q = 0; q < 27; ++q) {
                int64_t cx = base[0] + q % 3, cy = base[1] + (q / 3) % 3, cz = base[2] + q / 9;
                int32_t vid[3] = {(int32_t)((cx - ((cx % 4 + 4) % 4)) / 4), (int32_t)((cy - ((cy % 4 + 4) % 4)) / 4),
                                  (int32_t)((cz - ((cz % 4 + 4) % 4)) / 4)};
                if (vid[0] < -1023 || vid[0] > 1023 || vid[1] < -511 || vid[1] > 511 || vid[2] < -1023 || vid[2] > 1023)
                    continue;
                uint32_t blk = hmap_find(hm, pack_vid(vid));
                if (blk == NONE) continue;
                int64_t node = 64 * (int64_t)blk + ((cx % 4 + 4) % 4) + 4 * ((cy % 4 + 4) % 4) + 16 * ((cz % 4 + 4) % 4);
                v3 x = vmul(V3((REAL)cx, (REAL)cy, (REAL)cz), h);
                v3 ap = vsub(x, a), bp = vsub(x, b), cpv = vsub(x, c);
                if (vdot(vcross(ab, n), ap) <= K(0.0) && vdot(vcross(bc, n), bp) <= K(0.0) &&
                    vdot(vcross(ac, n), cpv) >= K(0.0)) {
                    REAL sd = vdot(n, ap) / nl;
                    REAL d = RFABS(sd);
                    naff[node] |= (sd < K(0.0) ? 0x00010001u : 0x00000001u) << cid;
                    if (d < mdist[node] || (d == mdist[node] && cid < mid[node])) { mdist[node] = d; mid[node] = cid; }
                }
            }
        }
    }
    for (int64_t i = 0; i < 64 * nb; ++i)
        if (mid[i] != NONE && mdist[i] < ndist[i]) { ndist[i] = mdist[i]; nclosest[i] = mid[i]; }
    free(mdist);
    free(mid);
}

/* 4x4 determinant and inverse through the 2x2 minors of the first two and the last two
 * rows (Laplace expansion); m is row-major here (symmetric inputs only). */
static REAL m4_minors(const REAL m[16], REAL s[6], REAL c[6])
{
    s[0] = m[0] * m[5] - m[4] * m[1];
    s[1] = m[0] * m[6] - m[4] * m[2];
    s[2] = m[0] * m[7] - m[4] * m[3];
    s[3] = m[1] * m[6] - m[5] * m[2];
    s[4] = m[1] * m[7] - m[5] * m[3];
    s[5] = m[2] * m[7] - m[6] * m[3];
    c[5] = m[10] * m[15] - m[14] * m[11];
    c[4] = m[9] * m[15] - m[13] * m[11];
    c[3] = m[9] * m[14] - m[13] * m[10];
    c[2] = m[8] * m[15] - m[12] * m[11];
    c[1] = m[8] * m[14] - m[12] * m[10];
    c[0] = m[8] * m[13] - m[12] * m[9];
    return s[0] * c[5] - s[1] * c[4] + s[2] * c[3] + s[3] * c[2] - s[4] * c[1] + s[5] * c[0];
}

/* adj(m) b / det, with the minors of m4_minors. */
static void m4_solve(const REAL m[16], const REAL s[6], const REAL c[6], REAL det, const REAL b[4], REAL out[4])
{
    REAL inv[16];
    inv[0] = m[5] * c[5] - m[6] * c[4] + m[7] * c[3];
    inv[1] = -m[1] * c[5] + m[2] * c[4] - m[3] * c[3];
    inv[2] = m[13] * s[5] - m[14] * s[4] + m[15] * s[3];
    inv[3] = -m[9] * s[5] + m[10] * s[4] - m[11] * s[3];
    inv[4] = -m[4] * c[5] + m[6] * c[2] - m[7] * c[1];
    inv[5] = m[0] * c[5] - m[2] * c[2] + m[3] * c[1];
    inv[6] = -m[12] * s[5] + m[14] * s[2] - m[15] * s[1];
    inv[7] = m[8] * s[5] - m[10] * s[2] + m[11] * s[1];
    inv[8] = m[4] * c[4] - m[5] * c[2] + m[7] * c[0];
    inv[9] = -m[0] * c[4] + m[1] * c[2] - m[3] * c[0];
    inv[10] = m[12] * s[4] - m[13] * s[2] + m[15] * s[0];
    inv[11] = -m[8] * s[4] + m[9] * s[2] - m[11] * s[0];
    inv[12] = -m[4] * c[3] + m[5] * c[1] - m[6] * c[0];
    inv[13] = m[0] * c[3] - m[1] * c[1] + m[2] * c[0];
    inv[14] = -m[12] * s[3] + m[13] * s[1] - m[14] * s[0];
    inv[15] = m[8] * s[3] - m[9] * s[1] + m[10] * s[0];
    REAL r = K(1.0) / det;
    for (int i = 0; i < 4; ++i)
        out[i] = (inv[4 * i] * b[0] + inv[4 * i + 1] * b[1] + inv[4 * i + 2] * b[2] + inv[4 * i + 3] * b[3]) * r;
}

/* One particle's contact distance field (g2p_cdf.rs:150-331) from the node CDFs of its
 * 3x3x3 stencil: the affinity bits (the union of the nodes', with the side of each newly
 * reached collider from the weighted signed distances and the previous side kept
 * otherwise), then the moving-least-squares fit d(x) = n . (x - x_p) + d_p over the
 * nodes sharing an affinity, signed by whether the node is on the particle's side. */
static void particle_cdf(v3 refp, REAL h, const REAL *ndist[27], const uint32_t *naff[27], uint32_t prev,
                         uint32_t *aff, v3 *normal, REAL *sd)
{
    v3 w[3] = {qk_eval_all(-refp.x / h), qk_eval_all(-refp.y / h), qk_eval_all(-refp.z / h)};
    uint32_t pa = 0;
    REAL signs[CPIC_MAX_COLLIDERS];
    for (int c = 0; c < CPIC_MAX_COLLIDERS; ++c) signs[c] = K(0.0);
    for (int i = 0; i < 27; ++i) {
        int sx = NBH_SHIFTS[i][0], sy = NBH_SHIFTS[i][1], sz = NBH_SHIFTS[i][2];
        uint32_t a = *naff[i];
        REAL d = *ndist[i];
        pa |= a & AFF_MASK;
        REAL weight = vget(w[0], sx) * vget(w[1], sy) * vget(w[2], sz);
        for (int c = 0; c < CPIC_MAX_COLLIDERS; ++c) {
            REAL compatible = (a >> c) & 1u ? K(1.0) : K(0.0);
            REAL sign = (a >> (16 + c)) & 1u ? K(-1.0) : K(1.0);
            signs[c] += compatible * weight * sign * d;
        }
    }
    for (int c = 0; c < CPIC_MAX_COLLIDERS; ++c) {
        if (!((prev >> c) & 1u)) {
            if (signs[c] < K(0.0)) pa |= 0x00010000u << c;
        } else {
            pa |= prev & (0x00010000u << c);
        }
    }
    REAL q[16], u[4];
    for (int k = 0; k < 16; ++k) q[k] = K(0.0);
    for (int k = 0; k < 4; ++k) u[k] = K(0.0);
    for (int i = 0; i < 27; ++i) {
        int sx = NBH_SHIFTS[i][0], sy = NBH_SHIFTS[i][1], sz = NBH_SHIFTS[i][2];
        uint32_t a = *naff[i];
        v3 dpt = vadd(refp, vmul(V3((REAL)sx, (REAL)sy, (REAL)sz), h));
        REAL weight = vget(w[0], sx) * vget(w[1], sy) * vget(w[2], sz);
        uint32_t comb = a & pa & AFF_MASK;
        uint32_t sdiff = ((a >> 16) ^ (pa >> 16)) & comb;
        if (comb != 0u) {
            REAL pv[4] = {dpt.x, dpt.y, dpt.z, K(1.0)};
            REAL d = sdiff == 0u ? *ndist[i] : -*ndist[i];
            for (int r = 0; r < 4; ++r) {
                for (int cc = 0; cc < 4; ++cc) q[4 * r + cc] += pv[r] * pv[cc] * weight;
                u[r] += pv[r] * (weight * d);
            }
        }
    }
    REAL sm[6], cm[6];
    REAL det = m4_minors(q, sm, cm);
    if (det > K(1.0e-8)) {
        REAL res[4];
        m4_solve(q, sm, cm, det, u, res);
        v3 nv = V3(res[0], res[1], res[2]);
        REAL len = vlen(nv);
        *normal = len > K(1.0e-6) ? vdiv(nv, len) : V3(0, 0, 0);
        *sd = res[3];
        *aff = pa;
    } else {
        *normal = V3(0, 0, 0);
        *sd = K(0.0);
        *aff = 0u;
    }
}

/* gpu_rigid_impulses_update (rigid_impulses.rs:66-140) with the particles' impulses
 * compiled out (zero, so no velocity cap): Velocity::integrate about the world centre
 * of mass c (src_rbd_shaders/dynamics/body.rs:253-268): t' = c + exp(w dt) (t - c) + v dt,
 * q' = normalize(from_scaled_axis(w dt) q) (glam's from_scaled_axis: identity for a zero
 * vector, else (axis sin(a/2), cos(a/2)); Hamilton product; normalize = q / |q|); then
 * v += g dt along the axes with a nonzero inverse mass.  The pose and velocity are
 * stored in f32, as the reference's. */
static void integrate_cuboid(orc_cuboid *c, float dt, const float *g)
{
    v3 t0 = V3(c->translation[0], c->translation[1], c->translation[2]);
    v3 com = collider_com(c);
    REAL sx = (REAL)c->angvel[0] * (REAL)dt, sy = (REAL)c->angvel[1] * (REAL)dt, sz = (REAL)c->angvel[2] * (REAL)dt;
    REAL len = RSQRT(sx * sx + sy * sy + sz * sz);
    REAL dx = 0, dy = 0, dz = 0, dw = 1;
    if (len != K(0.0)) {
        REAL half = len * K(0.5);
        REAL sn = RSIN(half), cs = RCOS(half);
        dx = sx / len * sn;
        dy = sy / len * sn;
        dz = sz / len * sn;
        dw = cs;
    }
    v3 arm = quat_rotate(dx, dy, dz, dw, vsub(t0, com));
    v3 nt = vadd(vadd(com, arm), V3((REAL)c->linvel[0] * (REAL)dt, (REAL)c->linvel[1] * (REAL)dt, (REAL)c->linvel[2] * (REAL)dt));
    c->translation[0] = (float)nt.x;
    c->translation[1] = (float)nt.y;
    c->translation[2] = (float)nt.z;
    REAL qx = c->rotation[0], qy = c->rotation[1], qz = c->rotation[2], qw = c->rotation[3];
    REAL nx = dw * qx + dx * qw + dy * qz - dz * qy;
    REAL ny = dw * qy - dx * qz + dy * qw + dz * qx;
    REAL nz = dw * qz + dx * qy - dy * qx + dz * qw;
    REAL nw = dw * qw - dx * qx - dy * qy - dz * qz;
    REAL r = K(1.0) / RSQRT(nx * nx + ny * ny + nz * nz + nw * nw);
    c->rotation[0] = (float)(nx * r);
    c->rotation[1] = (float)(ny * r);
    c->rotation[2] = (float)(nz * r);
    c->rotation[3] = (float)(nw * r);
    for (int k = 0; k < 3; ++k)
        c->linvel[k] = (float)((REAL)c->linvel[k] + (c->inv_mass[k] != 0.0f ? (REAL)g[k] : K(0.0)) * (REAL)dt);
}

static v3 body_point_velocity(const orc_cuboid *c, v3 pt)
{
    v3 com = collider_com(c);
    v3 lin = V3(c->linvel[0], c->linvel[1], c->linvel[2]);
    v3 ang = V3(c->angvel[0], c->angvel[1], c->angvel[2]);
    return vadd(lin, vcross(ang, vsub(pt, com)));
}

/* gpu_particle_update for one particle (solver/particle_update.rs:60-201): from the G2P
 * velocity gv, velocity gradient gg and divergence div, and the contact field (signed
 * distance psd, normal pn, the bodies' velocity prv; psd = 0 outside CPIC).  Reads and
 * writes x, v, C, F, the plastic state ps (NULL for tags without one), mass and enabled.
 * Returns nonzero for an unknown model tag. */
static int particle_update_one(uint32_t tag, const float *params, REAL h, REAL dt, REAL damping, uint32_t fixed,
                               REAL vol0, const REAL *gv, const REAL *gg, REAL div, REAL psd, const REAL *pn,
                               const REAL *prv, REAL *x, REAL *v, REAL *C, REAL *F, REAL *ps, REAL *mass,
                               uint32_t *enabled)
{
    v3 vel = vload(gv);
    m3 gradv = mload(gg);
    const REAL inv_d = qk_inv_d(h);
    /* CPIC reprojection of a penetrating particle (particle_update.rs:92-102) */
    const int penetrating = psd < K(-0.05) * h;
    if (penetrating) {
        v3 rv = vload(prv);
        vel = vadd(rv, bc_project(1u, K(0.0), vsub(vel, rv), vload(pn)));
    }
    REAL vl = vlen(vel);
    if (vl > h / dt) vel = vdiv(vmul(vdiv(vel, vl), h), dt);
    vel = vdiv(vel, K(1.0) + damping * dt);
    if (fixed) vel = V3(0, 0, 0);
    v3 pos = vload(x);
    v3 npos = vadd(pos, vmul(vel, dt));
    if (penetrating) {
        /* penalty impulse (particle_update.rs:126-136) */
        REAL corrected = RFMAX(psd, K(-0.3) * h);
        vel = vadd(vel, vmul(vload(pn), dt * -corrected * K(1.0e3)));
    }
    m3 Fm = mload(F);
    if (tag != 4) {
        Fm = madd(Fm, mmul(mscale(gradv, dt), Fm));
    } else {
        REAL f0 = Fm.c0.x;
        REAL rate = div * dt / K(3.0);
        rate = rclamp(rate, K(-0.1), K(0.1));
        REAL nd = rclamp(f0 + rate * f0, K(0.5), K(2.0));
        Fm = mdiag(V3(nd, nd, nd));
    }
    REAL st[3] = {0, 0, 0};
    if (ps) { st[0] = ps[0]; st[1] = ps[1]; st[2] = ps[2]; }
    m3 tau;
    const int bad = model_update(tag, params, st, &Fm, gradv, &tau) != 0;
    if (ps) { ps[0] = st[0]; ps[1] = st[1]; ps[2] = st[2]; }
    m3 aff = msub(mscale(gradv, *mass), mscale(tau, vol0 * inv_d * dt));
    int nan_pos = npos.x != npos.x || npos.y != npos.y || npos.z != npos.z;
    if (!nan_pos && mdet(Fm) > K(0.0)) {
        vstore(x, npos);
        mstore(C, aff);
    } else {
        *enabled = 0;
        vel = V3(0, 0, 0);
        Fm = mident();
        mstore(C, mzero());
        *mass = K(0.0);
    }
    vstore(v, vel);
    mstore(F, Fm);
    return bad;
}

/* ------------------------------------------------------------------------- */
/* Workspace. */

struct orc_ws {
    int64_t nb, nbase, nsorted;
    int32_t *vid;      /* nb*3 */
    uint32_t *np, *npe, *first, *buckets, *nbh; /* nb, nb, nb, nb*10, nb*7 */
    uint32_t *sorted;  /* nsorted */
    REAL *mv, *m, *vel;
    int has_cpic;
    REAL *ndist, *mv_i, *m_i, *vel_i;   /* CPIC: node CDF distance, incompatible fields */
    uint32_t *naff, *nclosest;
    int64_t cap_b, cap_s;
    mesh_set meshes;   /* mesh colliders of the node pass (orc_ws_set_meshes) */
};

orc_ws *orc_ws_create(void)
{
    orc_ws *ws = (orc_ws *)calloc(1, sizeof(orc_ws));
    return ws;
}

void orc_ws_destroy(orc_ws *ws)
{
    if (!ws) return;
    free(ws->vid); free(ws->np); free(ws->npe); free(ws->first); free(ws->buckets);
    free(ws->nbh); free(ws->sorted); free(ws->mv); free(ws->m); free(ws->vel);
    free(ws->ndist); free(ws->mv_i); free(ws->m_i); free(ws->vel_i); free(ws->naff); free(ws->nclosest);
    orc_ws_set_meshes(ws, NULL, 0);
    free(ws);
}

int orc_ws_set_meshes(orc_ws *ws, const orc_trimesh *meshes, int nmesh)
{
    for (int k = 0; k < ws->meshes.nmesh; ++k) mesh_pn_free(&ws->meshes.mp[k]);
    free(ws->meshes.mp);
    ws->meshes.mp = NULL;
    ws->meshes.nmesh = 0;
    if (nmesh <= 0) return 0;
    ws->meshes.mp = (mesh_pn *)calloc((size_t)nmesh, sizeof(mesh_pn));
    for (int k = 0; k < nmesh; ++k) {
        ws->meshes.mp[k].m = &meshes[k];
        mesh_pseudo_normals(&ws->meshes.mp[k]);
    }
    ws->meshes.nmesh = nmesh;
    return 0;
}

int orc_project_trimesh(const orc_trimesh *m, const orc_cuboid *body, REAL max_dist, const REAL *pt, REAL *proj,
                        int *inside)
{
    mesh_pn mp;
    mp.m = m;
    mesh_pseudo_normals(&mp);
    v3 p;
    int ok = project_trimesh(&mp, body, max_dist, vload(pt), &p, inside);
    if (ok) vstore(proj, p);
    mesh_pn_free(&mp);
    return ok;
}

static void ws_reserve(orc_ws *ws, int64_t nb, int64_t ns)
{
    if (nb > ws->cap_b) {
        int64_t c = nb * 2;
        ws->vid = (int32_t *)realloc(ws->vid, sizeof(int32_t) * 3 * c);
        ws->np = (uint32_t *)realloc(ws->np, sizeof(uint32_t) * c);
        ws->npe = (uint32_t *)realloc(ws->npe, sizeof(uint32_t) * c);
        ws->first = (uint32_t *)realloc(ws->first, sizeof(uint32_t) * c);
        ws->buckets = (uint32_t *)realloc(ws->buckets, sizeof(uint32_t) * NUM_BUCKETS * c);
        ws->nbh = (uint32_t *)realloc(ws->nbh, sizeof(uint32_t) * 7 * c);
        ws->mv = (REAL *)realloc(ws->mv, sizeof(REAL) * 64 * 3 * c);
        ws->m = (REAL *)realloc(ws->m, sizeof(REAL) * 64 * c);
        ws->vel = (REAL *)realloc(ws->vel, sizeof(REAL) * 64 * 3 * c);
        ws->ndist = (REAL *)realloc(ws->ndist, sizeof(REAL) * 64 * c);
        ws->mv_i = (REAL *)realloc(ws->mv_i, sizeof(REAL) * 64 * 3 * c);
        ws->m_i = (REAL *)realloc(ws->m_i, sizeof(REAL) * 64 * c);
        ws->vel_i = (REAL *)realloc(ws->vel_i, sizeof(REAL) * 64 * 3 * c);
        ws->naff = (uint32_t *)realloc(ws->naff, sizeof(uint32_t) * 64 * c);
        ws->nclosest = (uint32_t *)realloc(ws->nclosest, sizeof(uint32_t) * 64 * c);
        ws->cap_b = c;
    }
    if (ns > ws->cap_s) {
        int64_t c = ns * 2;
        ws->sorted = (uint32_t *)realloc(ws->sorted, sizeof(uint32_t) * c);
        ws->cap_s = c;
    }
}

int64_t orc_ws_nblocks(const orc_ws *ws) { return ws->nb; }
int64_t orc_ws_nprimary_blocks(const orc_ws *ws) { return ws->nbase; }
int64_t orc_ws_nsorted(const orc_ws *ws) { return ws->nsorted; }

int orc_ws_blocks(const orc_ws *ws, int32_t *vid, uint32_t *np, uint32_t *npe, uint32_t *buckets,
                  uint32_t *first)
{
    memcpy(vid, ws->vid, sizeof(int32_t) * 3 * ws->nb);
    memcpy(np, ws->np, sizeof(uint32_t) * ws->nb);
    memcpy(npe, ws->npe, sizeof(uint32_t) * ws->nb);
    memcpy(buckets, ws->buckets, sizeof(uint32_t) * NUM_BUCKETS * ws->nb);
    memcpy(first, ws->first, sizeof(uint32_t) * ws->nb);
    return 0;
}

int orc_ws_sorted_ids(const orc_ws *ws, uint32_t *ids)
{
    memcpy(ids, ws->sorted, sizeof(uint32_t) * ws->nsorted);
    return 0;
}

int orc_ws_grid(const orc_ws *ws, REAL *mv, REAL *m, REAL *vel)
{
    memcpy(mv, ws->mv, sizeof(REAL) * 64 * 3 * ws->nb);
    memcpy(m, ws->m, sizeof(REAL) * 64 * ws->nb);
    memcpy(vel, ws->vel, sizeof(REAL) * 64 * 3 * ws->nb);
    return 0;
}

int orc_ws_grid_cpic(const orc_ws *ws, REAL *dist, uint32_t *aff, uint32_t *closest, REAL *mv_i, REAL *m_i,
                     REAL *vel_i)
{
    if (!ws->has_cpic) return -1;
    memcpy(dist, ws->ndist, sizeof(REAL) * 64 * ws->nb);
    memcpy(aff, ws->naff, sizeof(uint32_t) * 64 * ws->nb);
    memcpy(closest, ws->nclosest, sizeof(uint32_t) * 64 * ws->nb);
    memcpy(mv_i, ws->mv_i, sizeof(REAL) * 64 * 3 * ws->nb);
    memcpy(m_i, ws->m_i, sizeof(REAL) * 64 * ws->nb);
    memcpy(vel_i, ws->vel_i, sizeof(REAL) * 64 * 3 * ws->nb);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* One substep. */

int orc_step(orc_ws *ws, int64_t n, REAL *x, REAL *v, REAL *C, REAL *F, REAL *mass,
             const REAL *vol0, const REAL *damping, REAL *pstate, uint32_t *enabled,
             const uint32_t *fixed, const uint32_t *tag, const float *params, const float *sim,
             orc_cuboid *cols, int ncol, int keep_debug)
{
    return orc_step_ex(ws, n, x, v, C, F, mass, vol0, damping, pstate, enabled, fixed, tag, params, sim, cols,
                       ncol, keep_debug, NULL);
}

int orc_step_ex(orc_ws *ws, int64_t n, REAL *x, REAL *v, REAL *C, REAL *F, REAL *mass,
                const REAL *vol0, const REAL *damping, REAL *pstate, uint32_t *enabled,
                const uint32_t *fixed, const uint32_t *tag, const float *params, const float *sim,
                orc_cuboid *cols, int ncol, int keep_debug, const orc_cpic *cp)
{
    (void)keep_debug;
    if (cp && ncol > CPIC_MAX_COLLIDERS) return -3;
    const v3 g = V3((REAL)sim[0], (REAL)sim[1], (REAL)sim[2]);
    const REAL dt = (REAL)sim[3];
    const REAL h = (REAL)sim[4];
    int rc = 0;

    /* --- keys --- */
    int32_t *pvid = (int32_t *)malloc(sizeof(int32_t) * 3 * (n > 0 ? n : 1));
    uint32_t *ploc = (uint32_t *)malloc(sizeof(uint32_t) * 3 * (n > 0 ? n : 1));
    uint32_t *pblk = (uint32_t *)malloc(sizeof(uint32_t) * (n > 0 ? n : 1));
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < n; ++p) particle_block(vload(x + 3 * p), h, pvid + 3 * p, ploc + 3 * p);

    /* --- block activation: primary touch in particle order, then the +1
     *     neighbours of each base block in header order --- */
    hmap hm;
    hmap_init(&hm, 4096);
    int64_t nb = 0, capb = 1024;
    int32_t *bvid = (int32_t *)malloc(sizeof(int32_t) * 3 * capb);
    for (int64_t p = 0; p < n; ++p) {
        uint32_t key = pack_vid(pvid + 3 * p);
        if (hmap_insert(&hm, key, (uint32_t)nb)) {
            if (nb == capb) { capb *= 2; bvid = (int32_t *)realloc(bvid, sizeof(int32_t) * 3 * capb); }
            memcpy(bvid + 3 * nb, pvid + 3 * p, sizeof(int32_t) * 3);
            nb++;
        }
    }
    int64_t nbase = nb;
    for (int64_t b = 0; b < nbase; ++b) {
        for (int i = 1; i < 8; ++i) {
            int32_t nv[3] = {bvid[3 * b] + ASSOC_SHIFT[i][0], bvid[3 * b + 1] + ASSOC_SHIFT[i][1],
                             bvid[3 * b + 2] + ASSOC_SHIFT[i][2]};
            if (hmap_insert(&hm, pack_vid(nv), (uint32_t)nb)) {
                if (nb == capb) { capb *= 2; bvid = (int32_t *)realloc(bvid, sizeof(int32_t) * 3 * capb); }
                memcpy(bvid + 3 * nb, nv, sizeof(nv));
                nb++;
            }
        }
    }
    for (int64_t p = 0; p < n; ++p) pblk[p] = hmap_find(&hm, pack_vid(pvid + 3 * p));

    /* --- neighbour ids, bucket counts, scan, cursors, scatter --- */
    uint32_t *nbh = (uint32_t *)malloc(sizeof(uint32_t) * 7 * nb);
    for (int64_t b = 0; b < nb; ++b)
        for (int i = 1; i < 8; ++i) {
            int32_t nv[3] = {bvid[3 * b] + ASSOC_SHIFT[i][0], bvid[3 * b + 1] + ASSOC_SHIFT[i][1],
                             bvid[3 * b + 2] + ASSOC_SHIFT[i][2]};
            nbh[7 * b + i - 1] = hmap_find(&hm, pack_vid(nv));
        }
    uint32_t *cnt = (uint32_t *)calloc((size_t)(NUM_BUCKETS * nb + 1), sizeof(uint32_t));
    for (int64_t p = 0; p < n; ++p) {
        uint32_t b0 = pblk[p];
        const uint32_t *loc = ploc + 3 * p;
        cnt[NUM_BUCKETS * b0 + loc[2]]++;
        for (int i = 1; i < 8; ++i) {
            if (spills(loc, ASSOC_SHIFT[i])) {
                uint32_t bi = nbh[7 * b0 + i - 1];
                cnt[NUM_BUCKETS * bi + extra_bucket(loc, ASSOC_SHIFT[i])]++;
            }
        }
    }
    uint32_t *first = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
    uint32_t *npb = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
    uint32_t *npe = (uint32_t *)malloc(sizeof(uint32_t) * (nb + 1));
    uint32_t *cur = (uint32_t *)malloc(sizeof(uint32_t) * (NUM_BUCKETS * nb + 1));
    uint64_t running_total = 0;
    for (int64_t b = 0; b < nb; ++b) {
        first[b] = (uint32_t)running_total;
        uint32_t run = first[b];
        for (int k = 0; k < NUM_BUCKETS; ++k) {
            if (k == NUM_PRIMARY_BUCKETS) npb[b] = run - first[b];
            cur[NUM_BUCKETS * b + k] = run;
            run += cnt[NUM_BUCKETS * b + k];
        }
        npe[b] = run - first[b];
        running_total += npe[b];
    }
    int64_t nsorted = (int64_t)running_total;
    uint32_t *sorted = (uint32_t *)malloc(sizeof(uint32_t) * (nsorted > 0 ? nsorted : 1));
    for (int64_t p = 0; p < n; ++p) {
        uint32_t b0 = pblk[p];
        const uint32_t *loc = ploc + 3 * p;
        sorted[cur[NUM_BUCKETS * b0 + loc[2]]++] = (uint32_t)p;
        for (int i = 1; i < 8; ++i) {
            if (spills(loc, ASSOC_SHIFT[i])) {
                uint32_t bi = nbh[7 * b0 + i - 1];
                sorted[cur[NUM_BUCKETS * bi + extra_bucket(loc, ASSOC_SHIFT[i])]++] = (uint32_t)p;
            }
        }
    }

    /* --- CPIC: node CDFs (grid_update_cdf.rs:60-132 for the cuboids, p2g_cdf.rs:60-259
     *     for the meshes), then each primary particle's CDF (g2p_cdf.rs:331-407) --- */
    const int64_t nbn = nb > 0 ? nb : 1;
    REAL *ndist = NULL;
    uint32_t *naff = NULL, *nclosest = NULL;
    if (cp) {
        ndist = (REAL *)malloc(sizeof(REAL) * 64 * nbn);
        naff = (uint32_t *)malloc(sizeof(uint32_t) * 64 * nbn);
        nclosest = (uint32_t *)malloc(sizeof(uint32_t) * 64 * nbn);
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < 64 * nb; ++i) {
            int64_t b = i / 64;
            int t = (int)(i % 64);
            const int32_t *vid = bvid + 3 * b;
            v3 pt = vmul(V3((REAL)(vid[0] * 4 + (t & 3)), (REAL)(vid[1] * 4 + ((t >> 2) & 3)), (REAL)(vid[2] * 4 + (t >> 4))), h);
            node_cdf(cols, ncol - cp->nmesh, h, pt, ndist + i, naff + i, nclosest + i);
        }
        if (cp->nmesh > 0) mesh_cdf(cp, cols, ncol, h, &hm, bvid, nb, ndist, naff, nclosest);
        const REAL none_dist = MAX_FLT_REF;
        const uint32_t none_aff = 0u;
#pragma omp parallel
        {
            const REAL *tdist[216];
            const uint32_t *taff[216];
#pragma omp for schedule(dynamic, 4)
            for (int64_t b = 0; b < nb; ++b) {
                if (npb[b] == 0) continue;
                const int32_t *vid = bvid + 3 * b;
                for (int o = 0; o < 8; ++o) {
                    int ox = ASSOC_SHIFT[o][0], oy = ASSOC_SHIFT[o][1], oz = ASSOC_SHIFT[o][2];
                    int32_t nv[3] = {vid[0] + ox, vid[1] + oy, vid[2] + oz};
                    uint32_t hid = hmap_find(&hm, pack_vid(nv));
                    for (int t = 0; t < 64; ++t) {
                        int tx = t & 3, ty = (t >> 2) & 3, tz = t >> 4;
                        if ((ox && tx > 1) || (oy && ty > 1) || (oz && tz > 1)) continue;
                        int si = (ox * 4 + tx) + 6 * (oy * 4 + ty) + 36 * (oz * 4 + tz);
                        tdist[si] = hid != NONE ? ndist + 64 * (int64_t)hid + t : &none_dist;
                        taff[si] = hid != NONE ? naff + 64 * (int64_t)hid + t : &none_aff;
                    }
                }
                for (uint32_t k = first[b]; k < first[b] + npb[b]; ++k) {
                    uint32_t pid = sorted[k];
                    v3 pos = vload(x + 3 * pid);
                    v3 refp = vsub(vmul(vsub(V3(RROUND(pos.x / h), RROUND(pos.y / h), RROUND(pos.z / h)), V3(1, 1, 1)), h), pos);
                    const uint32_t *loc = ploc + 3 * pid;
                    int packed = (int)(loc[0] + 6 * loc[1] + 36 * loc[2]);
                    const REAL *sd27[27];
                    const uint32_t *af27[27];
                    for (int i = 0; i < 27; ++i) {
                        int si = packed + NBH_SHIFTS[i][0] + 6 * NBH_SHIFTS[i][1] + 36 * NBH_SHIFTS[i][2];
                        sd27[i] = tdist[si];
                        af27[i] = taff[si];
                    }
                    v3 nrm;
                    particle_cdf(refp, h, sd27, af27, cp->affinity[pid], &cp->affinity[pid], &nrm, &cp->signed_distance[pid]);
                    vstore(cp->normal + 3 * pid, nrm);
                }
            }
        }
    }

    /* --- P2G: one gather per node over the block's sorted list (p2g.rs:64-351); with
     *     CPIC, contributions of particles not compatible with the node go to the
     *     node's second (incompatible) momentum and mass --- */
    REAL *gmv = (REAL *)malloc(sizeof(REAL) * 64 * 3 * (nb > 0 ? nb : 1));
    REAL *gm = (REAL *)malloc(sizeof(REAL) * 64 * (nb > 0 ? nb : 1));
    REAL *gvel = (REAL *)malloc(sizeof(REAL) * 64 * 3 * (nb > 0 ? nb : 1));
    REAL *gmv_i = cp ? (REAL *)malloc(sizeof(REAL) * 64 * 3 * nbn) : NULL;
    REAL *gm_i = cp ? (REAL *)malloc(sizeof(REAL) * 64 * nbn) : NULL;
    REAL *gvel_i = cp ? (REAL *)malloc(sizeof(REAL) * 64 * 3 * nbn) : NULL;
    const REAL inv_h = K(1.0) / h;
#pragma omp parallel
    {
        int32_t zkey[64];
#pragma omp for schedule(dynamic, 4)
        for (int64_t b = 0; b < nb; ++b) {
            const int32_t *vid = bvid + 3 * b;
            uint32_t fst = first[b], last = first[b] + npe[b], pend = first[b] + npb[b];
            uint32_t num_chunks = (npe[b] + 63) / 64;
            for (int t = 0; t < 64; ++t) {
                int tx = t & 3, ty = (t >> 2) & 3, tz = t >> 4;
                v3 cell_pos = vmul(V3((REAL)(vid[0] * 4 + tx), (REAL)(vid[1] * 4 + ty), (REAL)(vid[2] * 4 + tz)), h);
                v3 acc_mv = V3(0, 0, 0), acc_mv_i = V3(0, 0, 0);
                REAL acc_m = K(0.0), acc_m_i = K(0.0);
                const uint32_t node_aff = cp ? naff[64 * b + t] : 0u;
                for (uint32_t ch = 0; ch < num_chunks; ++ch) {
                    uint32_t base = fst + ch * 64;
                    uint32_t len = last - base < 64 ? last - base : 64;
                    for (uint32_t k = 0; k < len; ++k) {
                        uint32_t pid = sorted[base + k];
                        REAL zc = RROUND(x[3 * pid + 2] / h) - K(1.0);
                        int32_t zk = (int32_t)zc - vid[2] * 4;
                        zkey[k] = zk < -2 ? -2 : zk;
                    }
                    int32_t zmin = zkey[0], zmax = zkey[len - 1];
                    if (pend > base && pend < base + len) {
                        uint32_t bb = pend - base;
                        if (zkey[bb] < zmin) zmin = zkey[bb];
                        if (zkey[bb - 1] > zmax) zmax = zkey[bb - 1];
                    }
                    if (!(tz >= zmin && tz <= zmax + 2)) continue;
                    for (uint32_t k = 0; k < len; ++k) {
                        uint32_t pid = sorted[base + k];
                        if (!enabled[pid]) continue; /* contributes exactly +0 (p2g.rs:202-211) */
                        v3 pos = vload(x + 3 * pid);
                        REAL pm = mass[pid];
                        v3 mom = vmul(vload(v + 3 * pid), pm);
                        m3 aff = mload(C + 9 * pid);
                        v3 dpt = vsub(cell_pos, pos);
                        REAL w = qk_eval(dpt.x * inv_h) * qk_eval(dpt.y * inv_h) * qk_eval(dpt.z * inv_h);
                        if (w != K(0.0)) {
                            v3 vc = vmul(vadd(mvmul(aff, dpt), mom), w);
                            if (cp && !aff_compatible(cp->affinity[pid], node_aff)) {
                                acc_mv_i = vadd(acc_mv_i, vc);
                                acc_m_i = acc_m_i + pm * w;
                            } else {
                                acc_mv = vadd(acc_mv, vc);
                                acc_m = acc_m + pm * w;
                            }
                        }
                    }
                }
                vstore(gmv + 3 * (64 * b + t), acc_mv);
                gm[64 * b + t] = acc_m;
                if (cp) {
                    vstore(gmv_i + 3 * (64 * b + t), acc_mv_i);
                    gm_i[64 * b + t] = acc_m_i;
                }
            }
        }
    }

    /* --- grid update + collider boundary (grid_update.rs:21-88, grid_update_collide.rs) --- */
    const REAL vlim = h / dt;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 64 * nb; ++i) {
        int64_t b = i / 64;
        int t = (int)(i % 64);
        int tx = t & 3, ty = (t >> 2) & 3, tz = t >> 4;
        const int32_t *vid = bvid + 3 * b;
        v3 mom = vload(gmv + 3 * i);
        REAL m = gm[i];
        REAL inv_mass = m > K(0.0) ? K(1.0) / m : K(0.0);
        v3 vel = vmul(vadd(mom, vmul(g, m * dt)), inv_mass);
        vel = V3(RFMIN(RFMAX(vel.x, -vlim), vlim), RFMIN(RFMAX(vel.y, -vlim), vlim),
                 RFMIN(RFMAX(vel.z, -vlim), vlim));
        if (cp) {
            /* CPIC: both velocity fields, no collider pass (grid_update.rs:55-66,
             * solver/grid_update.rs launch) */
            v3 mom_i = vload(gmv_i + 3 * i);
            REAL mi = gm_i[i];
            REAL inv_mi = mi > K(0.0) ? K(1.0) / mi : K(0.0);
            v3 vi = vmul(vadd(mom_i, vmul(g, mi * dt)), inv_mi);
            vi = V3(RFMIN(RFMAX(vi.x, -vlim), vlim), RFMIN(RFMAX(vi.y, -vlim), vlim), RFMIN(RFMAX(vi.z, -vlim), vlim));
            vstore(gvel_i + 3 * i, vi);
        } else {
            v3 pt = vmul(V3((REAL)(vid[0] * 4 + tx), (REAL)(vid[1] * 4 + ty), (REAL)(vid[2] * 4 + tz)), h);
            vel = collide_node(cols, ncol, h, dt, pt, vel, ws->meshes.nmesh ? &ws->meshes : NULL);
        }
        vstore(gvel + 3 * i, vel);
    }

    /* --- G2P over each block's primaries (g2p.rs:42-347) --- */
    REAL *pv = (REAL *)malloc(sizeof(REAL) * 3 * (n > 0 ? n : 1));
    REAL *pg = (REAL *)malloc(sizeof(REAL) * 9 * (n > 0 ? n : 1));
    REAL *pdiv = (REAL *)malloc(sizeof(REAL) * (n > 0 ? n : 1));
    const REAL inv_d = qk_inv_d(h);
#pragma omp parallel
    {
        v3 tile[216], tile_i[216];
        uint32_t tile_aff[216], tile_cid[216];
#pragma omp for schedule(dynamic, 4)
        for (int64_t b = 0; b < nb; ++b) {
            if (npb[b] == 0) continue;
            const int32_t *vid = bvid + 3 * b;
            for (int o = 0; o < 8; ++o) {
                int ox = ASSOC_SHIFT[o][0], oy = ASSOC_SHIFT[o][1], oz = ASSOC_SHIFT[o][2];
                int32_t nv[3] = {vid[0] + ox, vid[1] + oy, vid[2] + oz};
                uint32_t hid = hmap_find(&hm, pack_vid(nv));
                for (int t = 0; t < 64; ++t) {
                    int tx = t & 3, ty = (t >> 2) & 3, tz = t >> 4;
                    if ((ox && tx > 1) || (oy && ty > 1) || (oz && tz > 1)) continue;
                    int si = (ox * 4 + tx) + 6 * (oy * 4 + ty) + 36 * (oz * 4 + tz);
                    tile[si] = hid != NONE ? vload(gvel + 3 * (64 * (int64_t)hid + t)) : V3(0, 0, 0);
                    if (cp) {
                        tile_i[si] = hid != NONE ? vload(gvel_i + 3 * (64 * (int64_t)hid + t)) : V3(0, 0, 0);
                        tile_aff[si] = hid != NONE ? naff[64 * (int64_t)hid + t] : 0u;
                        tile_cid[si] = hid != NONE ? nclosest[64 * (int64_t)hid + t] : NONE;
                    }
                }
            }
            for (uint32_t k = first[b]; k < first[b] + npb[b]; ++k) {
                uint32_t pid = sorted[k];
                v3 vel = V3(0, 0, 0);
                m3 grad = mzero();
                REAL div = K(0.0);
                if (enabled[pid]) {
                    v3 pos = vload(x + 3 * pid);
                    v3 refp = vsub(vmul(vsub(V3(RROUND(pos.x / h), RROUND(pos.y / h), RROUND(pos.z / h)), V3(1, 1, 1)), h), pos);
                    v3 w[3] = {qk_eval_all(-refp.x / h), qk_eval_all(-refp.y / h), qk_eval_all(-refp.z / h)};
                    const uint32_t *loc = ploc + 3 * pid;
                    int packed = (int)(loc[0] + 6 * loc[1] + 36 * loc[2]);
                    for (int i = 0; i < 27; ++i) {
                        int sx = NBH_SHIFTS[i][0], sy = NBH_SHIFTS[i][1], sz = NBH_SHIFTS[i][2];
                        const int si = packed + sx + 6 * sy + 36 * sz;
                        v3 cv = tile[si];
                        v3 dpt = vadd(refp, vmul(V3((REAL)sx, (REAL)sy, (REAL)sz), h));
                        if (cp && !aff_compatible(cp->affinity[pid], tile_aff[si])) {
                            /* g2p.rs:206-233: the incompatible field, seen through the
                             * closest collider's boundary condition (its friction scaled by
                             * the particle's) relative to the body's velocity there */
                            cv = tile_i[si];
                            uint32_t cid = tile_cid[si];
                            if (cid != NONE) {
                                const orc_cuboid *c = &cols[cid];
                                v3 bpv = body_point_velocity(c, vadd(dpt, pos));
                                v3 nrm = vload(cp->normal + 3 * pid);
                                cv = vadd(bpv, bc_project(c->bc_type, (REAL)c->bc_friction * cp->boundary_friction[pid],
                                                          vsub(cv, bpv), nrm));
                            }
                        }
                        REAL weight = vget(w[0], sx) * vget(w[1], sy) * vget(w[2], sz);
                        vel = vadd(vel, vmul(cv, weight));
                        grad = madd(grad, mscale(mouter(cv, dpt), weight * inv_d));
                        div = div + weight * inv_d * vdot(cv, dpt);
                    }
                    if (cp) {
                        /* g2p.rs:248-258: the bodies' velocities at the particle, summed over
                         * its affinities */
                        v3 rv = V3(0, 0, 0);
                        for (int c = 0; c < ncol; ++c)
                            if ((cp->affinity[pid] >> c) & 1u) rv = vadd(rv, body_point_velocity(&cols[c], pos));
                        vstore(cp->rigid_vel + 3 * pid, rv);
                    }
                } else if (cp) {
                    vstore(cp->rigid_vel + 3 * pid, V3(0, 0, 0));
                }
                vstore(pv + 3 * pid, vel);
                mstore(pg + 9 * pid, grad);
                pdiv[pid] = div;
            }
        }
    }

    /* --- particle update (particle_update.rs:60-201); the CDF terms only with CPIC --- */
    int bad_model = 0;
#pragma omp parallel for schedule(static) reduction(| : bad_model)
    for (int64_t p = 0; p < n; ++p) {
        const REAL zero3[3] = {0, 0, 0};
        const REAL psd = cp ? cp->signed_distance[p] : K(0.0);
        const REAL *pn = cp ? cp->normal + 3 * p : zero3;
        const REAL *prv = cp ? cp->rigid_vel + 3 * p : zero3;
        REAL *ps = (tag[p] == 2 || tag[p] == 3 || tag[p] == 5) ? pstate + 3 * p : NULL;
        if (particle_update_one(tag[p], params + 24 * (size_t)p, h, dt, damping[p], fixed[p], vol0[p], pv + 3 * p,
                                pg + 9 * p, pdiv[p], psd, pn, prv, x + 3 * p, v + 3 * p, C + 9 * p, F + 9 * p, ps,
                                &mass[p], &enabled[p]) != 0)
            bad_model = 1;
    }
    if (bad_model) rc = -2;

    /* --- body integration (src_mpm/pipeline.rs:510-520, rigid_impulses.rs:66-140): the
     *     cuboids are massless coupled bodies, so each pose advances with its velocity
     *     about its centre (Velocity::integrate, src_rbd_shaders/dynamics/body.rs:253-268) --- */
    for (int i = 0; i < ncol; ++i) integrate_cuboid(&cols[i], (float)sim[3], sim);

    /* --- keep inspection data --- */
    ws_reserve(ws, nb, nsorted);
    ws->nb = nb;
    ws->nbase = nbase;
    ws->nsorted = nsorted;
    memcpy(ws->vid, bvid, sizeof(int32_t) * 3 * nb);
    memcpy(ws->np, npb, sizeof(uint32_t) * nb);
    memcpy(ws->npe, npe, sizeof(uint32_t) * nb);
    memcpy(ws->first, first, sizeof(uint32_t) * nb);
    memcpy(ws->buckets, cnt, sizeof(uint32_t) * NUM_BUCKETS * nb);
    memcpy(ws->nbh, nbh, sizeof(uint32_t) * 7 * nb);
    memcpy(ws->sorted, sorted, sizeof(uint32_t) * nsorted);
    memcpy(ws->mv, gmv, sizeof(REAL) * 64 * 3 * nb);
    memcpy(ws->m, gm, sizeof(REAL) * 64 * nb);
    memcpy(ws->vel, gvel, sizeof(REAL) * 64 * 3 * nb);
    ws->has_cpic = cp != NULL;
    if (cp) {
        memcpy(ws->ndist, ndist, sizeof(REAL) * 64 * nb);
        memcpy(ws->naff, naff, sizeof(uint32_t) * 64 * nb);
        memcpy(ws->nclosest, nclosest, sizeof(uint32_t) * 64 * nb);
        memcpy(ws->mv_i, gmv_i, sizeof(REAL) * 64 * 3 * nb);
        memcpy(ws->m_i, gm_i, sizeof(REAL) * 64 * nb);
        memcpy(ws->vel_i, gvel_i, sizeof(REAL) * 64 * 3 * nb);
    }

    free(ndist); free(naff); free(nclosest); free(gmv_i); free(gm_i); free(gvel_i);
    free(pv); free(pg); free(pdiv);
    free(gmv); free(gm); free(gvel);
    free(sorted); free(cur); free(npe); free(npb); free(first); free(cnt); free(nbh);
    free(bvid); hmap_free(&hm);
    free(pvid); free(ploc); free(pblk);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Building blocks for unit tests. */

int orc_svd3(const REAL *F, REAL *U, REAL *s, REAL *V)
{
    m3 u, vv;
    v3 sv;
    svd3_impl(mload(F), &u, &sv, &vv);
    mstore(U, u);
    mstore(V, vv);
    vstore(s, sv);
    return 0;
}

int orc_model_update(uint32_t tag, const float *params, REAL *pstate, REAL *F, const REAL *gradv,
                     REAL *tau)
{
    m3 Fm = mload(F), t;
    REAL st[3] = {pstate[0], pstate[1], pstate[2]};
    int rc = model_update(tag, params, st, &Fm, mload(gradv), &t);
    pstate[0] = st[0]; pstate[1] = st[1]; pstate[2] = st[2];
    mstore(F, Fm);
    mstore(tau, t);
    return rc;
}

int orc_particle_update(uint32_t tag, const float *params, float cell_width, float dt, REAL damping, uint32_t fixed,
                        REAL vol0, const REAL *gv, const REAL *gradv, REAL div, REAL psd, const REAL *normal,
                        const REAL *rigid_vel, REAL *x, REAL *v, REAL *C, REAL *F, REAL *pstate, REAL *mass,
                        uint32_t *enabled)
{
    const int has_state = tag == 2 || tag == 3 || tag == 5;
    return particle_update_one(tag, params, (REAL)cell_width, (REAL)dt, damping, fixed, vol0, gv, gradv, div, psd,
                               normal, rigid_vel, x, v, C, F, has_state ? pstate : NULL, mass, enabled);
}

int orc_node_cdf(const orc_cuboid *cols, int ncol, float cell_width, const REAL *pt, REAL *dist, uint32_t *aff,
                 uint32_t *closest)
{
    if (ncol > CPIC_MAX_COLLIDERS) return -3;
    node_cdf(cols, ncol, (REAL)cell_width, vload(pt), dist, aff, closest);
    return 0;
}

int orc_particle_cdf(const REAL *refp, float cell_width, const REAL *dist27, const uint32_t *aff27, uint32_t prev,
                     uint32_t *aff, REAL *normal, REAL *sd)
{
    const REAL *d[27];
    const uint32_t *a[27];
    for (int i = 0; i < 27; ++i) { d[i] = dist27 + i; a[i] = aff27 + i; }
    v3 n;
    particle_cdf(vload(refp), (REAL)cell_width, d, a, prev, aff, &n, sd);
    vstore(normal, n);
    return 0;
}

REAL orc_m4_solve(const REAL *m, const REAL *b, REAL *out)
{
    REAL s6[6], c6[6];
    REAL det = m4_minors(m, s6, c6);
    m4_solve(m, s6, c6, det, b, out);
    return det;
}

int orc_project_cuboid(const orc_cuboid *c, const REAL *pt, REAL *proj, int *inside)
{
    v3 p = project_shape(c, vload(pt), inside);
    vstore(proj, p);
    return 0;
}

int orc_grid_node_bc(const orc_cuboid *cols, int ncol, const float *sim, const REAL *node_pos, REAL *vel)
{
    v3 r = collide_node(cols, ncol, (REAL)sim[4], (REAL)sim[3], vload(node_pos), vload(vel), NULL);
    vstore(vel, r);
    return 0;
}

int orc_block_keys(int64_t n, const REAL *x, float cell_width, int32_t *vid, uint32_t *local)
{
    for (int64_t p = 0; p < n; ++p) particle_block(vload(x + 3 * p), (REAL)cell_width, vid + 3 * p, local + 3 * p);
    return 0;
}

/* ------------------------------------------------------------------------- */
/* CFL timestep bound (src_mpm_shaders/solver/timestep_bound.rs:61-111): per enabled
 * particle the smaller of the model's sound-speed bound (models/default.rs:268-330,
 * models/utils.rs:157-172: cfl h / max(|v|, sqrt((K + 4/3 G) / (rho0 / max(det F, 1e-6)))),
 * K = lambda + 2 mu / 3 and G = mu for the elastic and sand tags (utils.rs:28-36), K = bulk
 * gamma and G = 0 for the fluid (fluid.rs:88-106)) and the velocity bound
 * h / (|v| + (h^2/4) |A|_F / m * 6 sqrt(3) / h); the minimum over particles of
 * (u32)(dt * 1e12) (saturating, as Rust's `as`), u32::MAX when none is enabled. */
static uint32_t secs_to_int(REAL dt)
{
    REAL x = dt * (REAL)1.0e12f;
    if (!(x > (REAL)0)) return 0u;               /* NaN and negatives saturate to 0 */
    if (x >= (REAL)4294967295.0) return 0xFFFFFFFFu;
    return (uint32_t)x;
}

uint32_t orc_timestep_bound(int64_t n, const REAL *v, const REAL *C, const REAL *F, const REAL *mass,
                            const REAL *vol0, const uint32_t *enabled, const uint32_t *tag, const float *params,
                            float cell_width)
{
    const REAL h = (REAL)cell_width;
    uint32_t best = 0xFFFFFFFFu;
    for (int64_t p = 0; p < n; ++p) {
        if (!enabled[p]) continue;
        const float *w = params + 24 * p;
        REAL bulk, shear, cfl;
        switch (tag[p]) {
        case 0: case 1: bulk = (REAL)w[0] + K(2.0) * (REAL)w[1] / K(3.0); shear = (REAL)w[1]; cfl = (REAL)w[2]; break;
        case 2: case 3: bulk = (REAL)w[10] + K(2.0) * (REAL)w[11] / K(3.0); shear = (REAL)w[11]; cfl = (REAL)w[12]; break;
        case 4: bulk = (REAL)w[0] * (REAL)w[1]; shear = K(0.0); cfl = (REAL)w[3]; break;
        case 5: { /* snow: hardened linear backbone; word 0 holds the current det Fp */
            const REAL hard = snow_hardening((REAL)w[3], (REAL)w[0]);
            bulk = ((REAL)w[4] + K(2.0) * (REAL)w[5] / K(3.0)) * hard; shear = (REAL)w[5] * hard; cfl = (REAL)w[6];
            break;
        }
        default: bulk = shear = cfl = K(0.0); break;
        }
        const m3 Fm = mload(F + 9 * p);
        const v3 vel = vload(v + 3 * p);
        REAL dt;
        if (tag[p] <= 5u) {
            const REAL density0 = mass[p] / vol0[p];
            const REAL det = mdet(Fm);
            const REAL density = density0 / RFMAX(det, (REAL)1.0e-6f);
            const REAL c = RSQRT((bulk + shear * K(4.0) / K(3.0)) / density);
            const REAL speed = RFMAX(vlen(vel), c);
            dt = cfl * h / speed;
        } else {
            dt = K(0.0);
        }
        const m3 A = mload(C + 9 * p);
        const REAL na2 = vdot(A.c0, A.c0) + vdot(A.c1, A.c1) + vdot(A.c2, A.c2);
        const REAL d = (h * h) / K(4.0);
        const REAL norm_b = d * RSQRT(na2) / mass[p];
        const REAL apic_v = norm_b * K(6.0) * RSQRT(K(3.0)) / h;
        const REAL vv = vlen(vel) + apic_v;
        dt = RFMIN(dt, h / vv);
        const uint32_t cand = secs_to_int(dt);
        if (cand < best) best = cand;
    }
    return best;
}
