/*
 * smcrt_oracle.c — FP64 CPU restatement of signedMCRT's photon-packet loop.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the GPU engine is compared against and the
 * CPU baseline bench.py times; nothing in the product (signedmcrt_amd/) links or calls it.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load liboracle.so.
 *
 * It follows the reference's algorithm step by step, in the reference's working precision (FP64,
 * src/constants.f90:18) and with the reference's constants (surface eps 1e-8, crossing step
 * min|d|+2eps, DDA wall nudge 1e-8, normal step h 1e-6, bounce cap 1000).  Differences, all
 * deliberate and stated:
 *   - RNG: the reference draws from the compiler's `random_number` per thread
 *     (src/random_mod.f90:83-90), whose bit stream is compiler-specific.  Here every packet draws
 *     from its own Philox-4x32-10 stream keyed by (seed, global packet id), the same stream the
 *     GPU kernel uses, so the two can be compared packet by packet for the first events
 *     (common random numbers).  Each event takes one Philox block (see rng_block).
 *   - The fluence accumulator is FP64 (the reference adds FP32 `real(dcell,sp)` into an FP32
 *     `jmean`, src/inttau2.f90:272; src/iarray.f90:12).
 *   - `error stop` paths (src/inttau2.f90:156,337-343) increment diagnostic counters and end the
 *     packet instead of ending the process.
 *   - The reference's HG scatter draws one extra number for a disabled biased-scattering branch
 *     (p = 0, src/photon.f90:731-747); it is not drawn here.
 *
 * Parity pins (tests/test_oracle.py, tests/test_oracle_pins.py): SDF values against the
 * reference's own test vectors (test/SDF/test_SDF.f90); SmoothUnion, onion, twist, elongate and
 * revolution against closed forms and identities between SDFs (band limits, k/6 on the bisector,
 * value/slope/curvature continuity at the band edge; ||p|-R|-t; twist k = 0 and of a z-cylinder;
 * elongated sphere = capsule / rounded box; revolved sphere = torus, revolved box = rectangle
 * distance); normals against analytic gradients; Fresnel/reflect/refract against the flang-built
 * reference surfaces.f90 (oracle/_ref) and the textbook equations; the interface table and Fresnel
 * events against a clear n = 1.5 slab under a collimated beam (killed = R + (1-R) R,
 * transmitted (1-R)^2, Snell directions, TIR, the grazing-crossing kill); the circular and uniform
 * emitters against the reference's own circle test (test/photon/test_photon.f90:195-210) and
 * their uniform distributions; scatter against the HG distribution moments; the DDA against
 * brute-force chord clipping; the whole loop against the reference's end-to-end KATs
 * (test/end_to_end/test_scat.f90), the closed-form scatter-site moments and the infinite-medium
 * total path 1/mu_a.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/smcrt.h"

#define ORC_PI 3.14159265358979323846
#define ORC_TWOPI (2.0 * ORC_PI)

/* ---------------------------------------------------------------------------------------
 * Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as easy as 1, 2, 3",
 * SC'11).  Counter = (block, id_lo, id_hi, 0), key = (seed_lo, seed_hi).
 * ------------------------------------------------------------------------------------- */
typedef struct {
    uint32_t key0, key1, id_lo, id_hi, block;
    uint32_t buf[4];
    int idx;
} orc_rng;

static void philox_block(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c[1] ^ k0, n1 = lo1, n2 = hi0 ^ c[3] ^ k1, n3 = lo0;
        c[0] = n0; c[1] = n1; c[2] = n2; c[3] = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

static void rng_init(orc_rng* g, uint64_t seed, uint64_t id) {
    g->key0 = (uint32_t)seed;
    g->key1 = (uint32_t)(seed >> 32);
    g->id_lo = (uint32_t)id;
    g->id_hi = (uint32_t)(id >> 32);
    g->block = 0;
    g->idx = 4;
}

static double u_of(uint32_t x) { return (double)(((x >> 9) << 1) | 1u) * (1.0 / 16777216.0); }

/* The packet loop draws one Philox block per event, u[0..3], whatever it uses of it:
 *   emission       u[0], u[1] the emitter's draws, u[2] the first flight's optical depth;
 *   flight end     u[0] the albedo test (weight estimator: the roulette draw), u[1] cos theta,
 *                  u[2] the azimuth, u[3] the next flight's optical depth;
 *   Fresnel event  u[0] reflect or refract, u[1] the resampled optical depth.
 * (The reference draws from one compiler RNG stream per thread; its bitstream is not
 * reproducible, so the mapping of draws to the stream is this build's own.) */
static void rng_block(orc_rng* g, double u[4]) {
    uint32_t c[4] = {g->block, g->id_lo, g->id_hi, 0u};
    philox_block(c, g->key0, g->key1);
    g->block++;
    for (int i = 0; i < 4; ++i) u[i] = u_of(c[i]);
}

/* uniform in (0,1): an odd multiple of 2^-24, exactly representable in FP32 and FP64
 * (sequential draws of the stream: the raw-stream helper orc_uniforms) */
static double rng_u(orc_rng* g) {
    if (g->idx == 4) {
        uint32_t c[4] = {g->block, g->id_lo, g->id_hi, 0u};
        philox_block(c, g->key0, g->key1);
        memcpy(g->buf, c, sizeof c);
        g->block++;
        g->idx = 0;
    }
    return u_of(g->buf[g->idx++]);
}

/* ---------------------------------------------------------------------------------------
 * vec3 helpers (src/vector_class.f90)
 * ------------------------------------------------------------------------------------- */
typedef struct { double x, y, z; } v3;
static inline v3 mk(double x, double y, double z) { v3 r = {x, y, z}; return r; }
static inline v3 add(v3 a, v3 b) { return mk(a.x + b.x, a.y + b.y, a.z + b.z); }
static inline v3 sub(v3 a, v3 b) { return mk(a.x - b.x, a.y - b.y, a.z - b.z); }
static inline v3 scl(v3 a, double s) { return mk(a.x * s, a.y * s, a.z * s); }
static inline double dot(v3 a, v3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
static inline double len(v3 a) { return sqrt(dot(a, a)); }
static inline v3 vabs(v3 a) { return mk(fabs(a.x), fabs(a.y), fabs(a.z)); }
static inline v3 vmax0(v3 a) { return mk(fmax(a.x, 0.0), fmax(a.y, 0.0), fmax(a.z, 0.0)); }
static inline double clampd(double x, double lo, double hi) { return fmin(fmax(x, lo), hi); }
static inline v3 norm(v3 a) { return scl(a, 1.0 / len(a)); }

/* [p,1]·T with T given as rows 1..4 of columns 1..3 (src/vector_class.f90:292-304) */
static inline v3 xform(const double* t, v3 p) {
    return mk(p.x * t[0] + p.y * t[3] + p.z * t[6] + t[9],
              p.x * t[1] + p.y * t[4] + p.z * t[7] + t[10],
              p.x * t[2] + p.y * t[5] + p.z * t[8] + t[11]);
}

/* ---------------------------------------------------------------------------------------
 * SDF primitives (src/sdfs/sdfs.f90:494-735), operators and modifiers
 * (src/sdfs/sdfModifiers.f90:273-475)
 * ------------------------------------------------------------------------------------- */
static double prim_eval(int kind, const double* P, v3 p) {
    switch (kind) {
    case SMCRT_SPHERE: /* :494-508 */
        return sqrt(p.x * p.x + p.y * p.y + p.z * p.z) - P[0];
    case SMCRT_BOX: { /* :510-525 */
        v3 q = sub(vabs(p), mk(P[0], P[1], P[2]));
        return len(vmax0(q)) + fmin(fmax(q.x, fmax(q.y, q.z)), 0.0);
    }
    case SMCRT_TORUS: { /* :527-542 */
        v3 q = mk(len(mk(p.x, 0.0, p.z)) - P[0], p.y, 0.0);
        return len(q) - P[1];
    }
    case SMCRT_CYLINDER: { /* :544-581 */
        v3 a = mk(P[0], P[1], P[2]), b = mk(P[3], P[4], P[5]);
        v3 ba = sub(b, a), pa = sub(p, a);
        double baba = dot(ba, ba), paba = dot(pa, ba);
        double x = len(sub(scl(pa, baba), scl(ba, paba))) - P[6] * baba;
        double y = fabs(paba - baba * 0.5) - baba * 0.5;
        double x2 = x * x, y2 = y * y * baba, d;
        if (fmax(x, y) < 0.0) d = -fmin(x2, y2);
        else if (x > 0.0 && y > 0.0) d = x2 + y2;
        else if (x > 0.0) d = x2;
        else if (y > 0.0) d = y2;
        else d = 0.0;
        return copysign(sqrt(fabs(d)) / baba, d);
    }
    case SMCRT_TRIPRISM: { /* :583-597 */
        v3 q = vabs(p);
        return fmax(q.z - P[1], fmax(q.x * 0.866025 + p.y * 0.5, -p.y) - P[0] * 0.5);
    }
    case SMCRT_SEGMENT: /* :599-626, thickness 0.1 */
    case SMCRT_CAPSULE: { /* :628-648 */
        v3 a = mk(P[0], P[1], P[2]), b = mk(P[3], P[4], P[5]);
        v3 pa = sub(p, a), ba = sub(b, a);
        double h = clampd(dot(pa, ba) / dot(ba, ba), 0.0, 1.0);
        double r = (kind == SMCRT_SEGMENT) ? 0.1 : P[6];
        return len(sub(pa, scl(ba, h))) - r;
    }
    case SMCRT_CONE: { /* :650-686 */
        v3 a = mk(P[0], P[1], P[2]), b = mk(P[3], P[4], P[5]);
        double ra = P[6], rb = P[7];
        double rba = rb - ra;
        double baba = dot(sub(b, a), sub(b, a));
        double papa = dot(sub(p, a), sub(p, a));
        double paba = dot(sub(p, a), sub(b, a)) / baba;
        double x = sqrt(papa - baba * paba * paba);
        double cax = (paba < 0.5) ? fmax(0.0, x - ra) : fmax(0.0, x - rb);
        double cay = fabs(paba - 0.5) - 0.5;
        double k = rba * rba + baba;
        double f = clampd((rba * (x - ra) + paba * baba) / k, 0.0, 1.0);
        double cbx = x - ra - f * rba, cby = paba - f;
        double s = (cbx < 0.0 && cay < 0.0) ? -1.0 : 1.0;
        return s * sqrt(fmin(cax * cax + baba * cay * cay, cbx * cbx + baba * cby * cby));
    }
    case SMCRT_EGG: { /* :688-718 */
        double r1 = P[0], r2 = P[1], h = P[2];
        v3 pin = mk(fabs(p.x), p.y, p.z);
        double r = r1 - r2, hin = h + r;
        double l = (hin * hin - r * r) / (2.0 * r);
        if (pin.y <= 0.0) return len(pin) - r1;
        if ((pin.y - hin) * l > pin.x * hin)
            return len(sub(pin, mk(0.0, hin, 0.0))) - ((r1 + l) - len(mk(hin, l, 0.0)));
        return len(add(pin, mk(l, 0.0, 0.0))) - (r1 + l);
    }
    case SMCRT_PLANE: /* :720-735 */
        return dot(p, mk(P[0], P[1], P[2]));
    }
    return 0.0;
}

static double csg(int kind, double d1, double d2, double k) {
    switch (kind) {
    case SMCRT_UNION: return fmin(d1, d2);
    case SMCRT_SMOOTH_UNION: {
        double h = fmax(k - fabs(d1 - d2), 0.0) / k;
        return fmin(d1, d2) - h * h * h * k * (1.0 / 6.0);
    }
    case SMCRT_SUBTRACTION: return fmax(-d1, d2);
    case SMCRT_INTERSECTION: return fmax(d1, d2);
    }
    return d1;
}

static v3 mod_pos(int kind, const double* P, v3 p) {
    switch (kind) {
    case SMCRT_REVOLUTION: /* :290-305 */
        return mk(len(mk(p.x, p.z, 0.0)) - P[0], p.y, 0.0);
    case SMCRT_ELONGATE: /* :319-335 */
        return vmax0(sub(vabs(p), mk(P[0], P[1], P[2])));
    case SMCRT_TWIST: { /* :337-355 */
        double c = cos(P[0] * p.z), s = sin(P[0] * p.z);
        return mk(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
    }
    case SMCRT_BEND: { /* :357-375 */
        double c = cos(P[0] * p.x), s = sin(P[0] * p.x);
        return mk(c * p.x - s * p.y, s * p.x + c * p.y, p.z);
    }
    }
    return p; /* extrude, onion: unchanged */
}

static double mod_post(int kind, const double* P, v3 p, double d) {
    switch (kind) {
    case SMCRT_EXTRUDE: { /* :273-288 */
        double wx = d, wy = fabs(p.z) - P[0];
        double mx = fmax(wx, 0.0), my = fmax(wy, 0.0);
        return fmin(fmax(wx, wy), 0.0) + sqrt(mx * mx + my * my);
    }
    case SMCRT_ONION: /* :307-317 */
        return fabs(d) - P[0];
    case SMCRT_ELONGATE: {
        v3 q = sub(vabs(p), mk(P[0], P[1], P[2]));
        return d + fmin(fmax(q.x, fmax(q.y, q.z)), 0.0);
    }
    }
    return d;
}

/* closed polygon in the xy plane (SMCRT_POLYGON), brute force over every edge: distance to the
   nearest edge, sign by the crossing-number test of Inigo Quilez's published polygon distance
   (negative inside, either orientation).  The library evaluates the same shape through its
   uniform-grid accelerator with a local sign rule; this is the independent check. */
static double polygon_brute(const smcrt_scene_desc* s, const double* P, v3 p) {
    const int f = (int)P[0], c = (int)P[1];
    const double* v = s->poly_vertices + 2 * (size_t)f;
    double d = INFINITY, sg = 1.0;
    for (int i = 0, j = c - 1; i < c; j = i, i++) {
        const double ex = v[2 * j] - v[2 * i], ey = v[2 * j + 1] - v[2 * i + 1];
        const double wx = p.x - v[2 * i], wy = p.y - v[2 * i + 1];
        double t = (wx * ex + wy * ey) / (ex * ex + ey * ey);
        t = t < 0.0 ? 0.0 : (t > 1.0 ? 1.0 : t);
        const double bx = wx - ex * t, by = wy - ey * t;
        d = fmin(d, bx * bx + by * by);
        const int c1 = p.y >= v[2 * i + 1], c2 = p.y < v[2 * j + 1], c3 = ex * wy > ey * wx;
        if ((c1 && c2 && c3) || (!c1 && !c2 && !c3)) sg = -sg;
    }
    return sg * sqrt(d);
}

/* evaluate region r at p: the postfix program of smcrt.h */
static double region_eval(const smcrt_scene_desc* s, int r, v3 p) {
    double vs[SMCRT_MAX_VALUE_STACK + 1];
    v3 ps[SMCRT_MAX_POS_STACK + 1];
    int nv = 0, np = 0;
    ps[np++] = p;
    const smcrt_region* R = &s->regions[r];
    for (int i = 0; i < R->n_nodes; ++i) {
        const smcrt_node* n = &s->nodes[R->first_node + i];
        switch (n->op) {
        case SMCRT_OP_PRIM: {
            v3 q = xform(n->t, ps[np - 1]);
            vs[nv++] = n->kind == SMCRT_POLYGON ? polygon_brute(s, n->p, q) : prim_eval(n->kind, n->p, q);
            break;
        }
        case SMCRT_OP_CSG: { double d2 = vs[--nv]; double d1 = vs[--nv]; vs[nv++] = csg(n->kind, d1, d2, n->k); break; }
        case SMCRT_OP_MOD_BEGIN: { v3 q = mod_pos(n->kind, n->p, ps[np - 1]); ps[np++] = q; break; }
        case SMCRT_OP_MOD_END: { --np; vs[nv - 1] = mod_post(n->kind, n->p, ps[np - 1], vs[nv - 1]); break; }
        }
    }
    return vs[nv - 1];
}

/* `calcNormal` (src/sdfs/sdf_base.f90:165-189): tetrahedral finite differences, h = 1e-6 */
static v3 calc_normal(const smcrt_scene_desc* s, int r, v3 p) {
    const double h = 1e-6;
    v3 e0 = mk(1, -1, -1), e1 = mk(-1, -1, 1), e2 = mk(-1, 1, -1), e3 = mk(1, 1, 1);
    v3 n = add(add(scl(e0, region_eval(s, r, add(p, scl(e0, h)))),
                   scl(e1, region_eval(s, r, add(p, scl(e1, h))))),
               add(scl(e2, region_eval(s, r, add(p, scl(e2, h)))),
                   scl(e3, region_eval(s, r, add(p, scl(e3, h))))));
    return norm(n);
}

/* ---------------------------------------------------------------------------------------
 * Fresnel (src/surfaces.f90)
 * ------------------------------------------------------------------------------------- */
static double fresnel(v3 I, v3 N, double n1, double n2) { /* :86-120 */
    double costt = fabs(dot(I, N));
    double sintt = sqrt(1.0 - costt * costt);
    double sint2 = n1 / n2 * sintt;
    if (sint2 > 1.0) return 1.0;
    if (costt == 1.0) return 0.0;
    double cost2 = sqrt(1.0 - sint2 * sint2);
    double f1 = (n1 * costt - n2 * cost2) / (n1 * costt + n2 * cost2);
    double f2 = (n1 * cost2 - n2 * costt) / (n1 * cost2 + n2 * costt);
    return 0.5 * (f1 * f1 + f2 * f2);
}
static v3 reflect(v3 I, v3 N) { return sub(I, scl(N, 2.0 * dot(N, I))); } /* :42-55 */
static v3 refract(v3 I, v3 N, double eta) {                             /* :57-84 */
    double c1 = dot(N, I);
    if (c1 < 0.0) c1 = -c1;
    else N = scl(N, -1.0);
    double c2 = sqrt(1.0 - eta * eta * (1.0 - c1 * c1));
    return add(scl(I, eta), scl(N, eta * c1 - c2));
}

/* ---------------------------------------------------------------------------------------
 * scene constants
 * ------------------------------------------------------------------------------------- */
typedef struct {
    const smcrt_scene_desc* d;
    int nreg;
    double kappa[SMCRT_MAX_REGIONS], albedo[SMCRT_MAX_REGIONS], hgg[SMCRT_MAX_REGIONS],
        g2[SMCRT_MAX_REGIONS], n[SMCRT_MAX_REGIONS], mua[SMCRT_MAX_REGIONS], mus[SMCRT_MAX_REGIONS];
    int mat[SMCRT_MAX_REGIONS];
    double xmax, ymax, zmax;
    int nx, ny, nz;
    double circ_t[12];
    int circ_swap;
    v3 src_dir;
} orc_scene;

static void scene_setup(orc_scene* S, const smcrt_scene_desc* d) {
    S->d = d;
    S->nreg = d->n_regions;
    for (int r = 0; r < d->n_regions; ++r) {
        const smcrt_material* m = &d->materials[d->regions[r].material];
        S->mat[r] = d->regions[r].material;
        S->mus[r] = m->mus;
        S->mua[r] = m->mua;
        S->kappa[r] = m->mus + m->mua;                               /* opticalProperties.f90:114 */
        S->albedo[r] = (m->mua < 1e-9) ? 1.0 : m->mus / S->kappa[r]; /* :115-119 */
        S->hgg[r] = m->hgg;
        S->g2[r] = m->hgg * m->hgg;
        S->n[r] = m->n;
    }
    S->xmax = d->grid.xmax; S->ymax = d->grid.ymax; S->zmax = d->grid.zmax;
    S->nx = d->grid.nx; S->ny = d->grid.ny; S->nz = d->grid.nz;
    const smcrt_source* src = &d->source;
    S->src_dir = mk(src->dir[0], src->dir[1], src->dir[2]);
    if (src->kind == SMCRT_SRC_POINT) S->src_dir = mk(0, 0, -1);
    /* circular source (src/photon.f90:238-266): t = rotationAlign(a, b) . invert(translate(pos)) */
    v3 a = mk(1, 0, 0), b = norm(S->src_dir);
    S->circ_swap = 0;
    if (fabs(a.x) == fabs(b.x) && fabs(a.y) == fabs(b.y) && fabs(a.z) == fabs(b.z)) {
        a = mk(0, 0, 1);
        S->circ_swap = 1;
    }
    v3 v = mk(a.y * b.z - a.z * b.y, -a.x * b.z + a.z * b.x, a.x * b.y - a.y * b.x);
    double c = dot(a, b), k = 1.0 / (1.0 + c);
    /* skew matrix vx (column-major as in src/sdfs/sdfHelpers.f90:137-140), 3x3 block */
    double vx[3][3] = {{0, v.z, -v.y}, {-v.z, 0, v.x}, {v.y, -v.x, 0}}; /* vx[row][col] */
    double R[3][3];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s2 = 0;
            for (int m = 0; m < 3; ++m) s2 += vx[i][m] * vx[m][j];
            R[i][j] = (i == j ? 1.0 : 0.0) + vx[i][j] + s2 * k;
        }
    /* R (4x4, zero translation) times translate(-pos): rows 1..3 = R, row 4 = -pos */
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) S->circ_t[i * 3 + j] = R[i][j];
    for (int j = 0; j < 3; ++j) S->circ_t[9 + j] = -src->pos[j];
}

/* ---------------------------------------------------------------------------------------
 * packet state
 * ------------------------------------------------------------------------------------- */
enum { END_NONE = 0, END_ABSORB, END_ESCAPE, END_SURFACE, END_BOUNCE, END_ERROR };

typedef struct {
    v3 pos, dir;
    double weight; /* weight_scatter's packet weight */
    int layer, tflag, reason, bounces;
    orc_rng rng;
    double tau0; /* optical depth of the next flight (drawn with the preceding event) */
    /* adjoint accumulators */
    int grad;
    double C;                        /* weighted in-grid path tally so far */
    double L[SMCRT_MAX_MATERIALS], A[SMCRT_MAX_MATERIALS], B[SMCRT_MAX_MATERIALS];
    double K[SMCRT_MAX_MATERIALS];
    float* tr; /* trace slot of this packet (or NULL) */
    int tc, te;
} orc_packet;

typedef struct {
    double* jmean; /* may be NULL */
    const float* w; /* adjoint weights, may be NULL (w = 1) */
    uint64_t sweeps, region_evals, normals, steps, crossings, ifaces, fresnel, err_unreach, err_dcell;
} orc_thread;

static inline void tag_end(orc_packet* k, int why) {
    if (!k->tflag) { k->tflag = 1; k->reason = why; }
}

/* `update_voxels` (src/inttau2.f90:414-441), 1-based cells, -1 outside */
static void voxel_of(const orc_scene* S, v3 p, int* ci, int* cj, int* ck) {
    *ci = (int)floor(S->nx * p.x / (2.0 * S->xmax)) + 1;
    *cj = (int)floor(S->ny * p.y / (2.0 * S->ymax)) + 1;
    *ck = (int)floor(S->nz * p.z / (2.0 * S->zmax)) + 1;
    if (*ci > S->nx || *ci < 1) *ci = -1;
    if (*cj > S->ny || *cj < 1) *cj = -1;
    if (*ck > S->nz || *ck < 1) *ck = -1;
}

static inline double xface(const orc_scene* S, int i) { return (i - 1) * 2.0 * S->xmax / S->nx; }
static inline double yface(const orc_scene* S, int i) { return (i - 1) * 2.0 * S->ymax / S->ny; }
static inline double zface(const orc_scene* S, int i) { return (i - 1) * 2.0 * S->zmax / S->nz; }

static void tally_add(const orc_scene* S, orc_thread* T, orc_packet* k, int ci, int cj, int ck,
                      double dcell) {
    size_t idx = (size_t)(ci - 1) + (size_t)S->nx * ((size_t)(cj - 1) + (size_t)S->ny * (size_t)(ck - 1));
    T->crossings++;
    if (T->jmean) {
#pragma omp atomic
        T->jmean[idx] += dcell;
    }
    if (k->grad) {
        double w = T->w ? (double)T->w[idx] : 1.0;
        int m = S->mat[k->layer];
        k->A[m] += k->C * dcell + 0.5 * w * dcell * dcell;
        k->C += w * dcell;
    }
}

/* `update_grids` + `wall_dist` + `update_pos` (src/inttau2.f90:213-411): walk the voxels along
   a step of length dstep from pos, tallying each chord.  Returns the tallied length. */
static double update_grids(const orc_scene* S, orc_thread* T, orc_packet* k, v3 pos, v3 dir,
                           double dstep) {
    const double delta = 1e-8;
    v3 o = mk(pos.x + S->xmax, pos.y + S->ymax, pos.z + S->zmax);
    int ci, cj, ck;
    voxel_of(S, o, &ci, &cj, &ck);
    if (ci == -1 || cj == -1 || ck == -1) { tag_end(k, END_ESCAPE); return 0.0; }
    double d = 0.0;
    for (;;) {
        double dx = -999.0, dy = -999.0, dz = -999.0;
        if (dir.x > 0.0) dx = (xface(S, ci + 1) - o.x) / dir.x;
        else if (dir.x < 0.0) dx = (xface(S, ci) - o.x) / dir.x;
        else dx = 100000.0;
        if (dir.y > 0.0) dy = (yface(S, cj + 1) - o.y) / dir.y;
        else if (dir.y < 0.0) dy = (yface(S, cj) - o.y) / dir.y;
        else dy = 100000.0;
        if (dir.z > 0.0) dz = (zface(S, ck + 1) - o.z) / dir.z;
        else if (dir.z < 0.0) dz = (zface(S, ck) - o.z) / dir.z;
        else dz = 100000.0;
        double dcell = fmin(dx, fmin(dy, dz));
        if (dcell < 0.0) { /* src/inttau2.f90:337-343 `error stop` */
            T->err_dcell++;
            tag_end(k, END_ERROR);
            return d;
        }
        int lx = dcell == dx, ly = dcell == dy; /* otherwise the z wall */
        if (d + dcell > dstep) {
            dcell = dstep - d;
            d = dstep;
            tally_add(S, T, k, ci, cj, ck, dcell);
            break;
        }
        d += dcell;
        tally_add(S, T, k, ci, cj, ck, dcell);
        if (lx) {
            o.x = (dir.x > 0.0) ? xface(S, ci + 1) + delta : xface(S, ci) - delta;
            o.y += dir.y * dcell; o.z += dir.z * dcell;
        } else if (ly) {
            o.y = (dir.y > 0.0) ? yface(S, cj + 1) + delta : yface(S, cj) - delta;
            o.x += dir.x * dcell; o.z += dir.z * dcell;
        } else {
            o.z = (dir.z > 0.0) ? zface(S, ck + 1) + delta : zface(S, ck) - delta;
            o.x += dir.x * dcell; o.y += dir.y * dcell;
        }
        voxel_of(S, o, &ci, &cj, &ck);
        if (ci == -1 || cj == -1 || ck == -1) { tag_end(k, END_ESCAPE); break; }
    }
    return d;
}

/* one sphere-trace step of length d in the packet's current layer (the two branches of
   src/inttau2.f90:64-83 share this bookkeeping) */
static void march(const orc_scene* S, orc_thread* T, orc_packet* k, double d) {
    T->steps++;
    int was_tagged = k->tflag;
    double tallied = update_grids(S, T, k, k->pos, k->dir, d);
    if (k->grad && !was_tagged) {
        int m = S->mat[k->layer];
        if (k->tflag) {
            /* left the grid during this step: the outcome is fixed at the exit point, so the
               path density (and the score) stops there */
            k->L[m] += tallied;
        } else {
            k->L[m] += d;
            k->A[m] += k->C * (d - tallied); /* C is constant after the in-grid part */
        }
    }
    k->pos = add(k->pos, scl(k->dir, d));
}

static void sweep(const orc_scene* S, orc_thread* T, v3 p, double* ds) {
    for (int i = 0; i < S->nreg; ++i) ds[i] = region_eval(S->d, i, p);
    T->sweeps++;
}

static void trace_put(float* tr, int ne, int* cnt, v3 p, int ev);

/* `tauint2` (src/inttau2.f90:15-210) */
static void tauint2(const orc_scene* S, orc_thread* T, orc_packet* k, int max_bounces) {
    const double eps = 1e-8;
    double ds[SMCRT_MAX_REGIONS], dstmp[SMCRT_MAX_REGIONS];
    v3 oldpos = k->pos;
    double tau = k->tau0; /* flight optical depth (drawn at emission or the last interaction) */
    double taurun = 0.0;
    int nr = S->nreg;
    for (;;) {
        sweep(S, T, k->pos, ds);
        double d = INFINITY;
        for (int i = 0; i < nr; ++i) d = fmin(d, fabs(ds[i]));
        if (d < eps) { tag_end(k, END_SURFACE); break; }
        while (d > eps) {
            double kap = S->kappa[k->layer];
            double t = d * kap;
            if (taurun + t <= tau) {
                taurun += t;
                march(S, T, k, d);
            } else {
                d = (tau - taurun) / kap;
                taurun = tau;
                march(S, T, k, d);
                break;
            }
            sweep(S, T, k->pos, ds);
            d = INFINITY;
            double smin = INFINITY;
            for (int i = 0; i < nr; ++i) { d = fmin(d, fabs(ds[i])); smin = fmin(smin, ds[i]); }
            if (smin >= 0.0) { tag_end(k, END_ESCAPE); break; }
        }
        if (taurun >= tau || k->tflag) break;

        /* interface (src/inttau2.f90:104-185) */
        T->ifaces++;
        sweep(S, T, k->pos, dstmp);
        double dmin = INFINITY;
        for (int i = 0; i < nr; ++i) dmin = fmin(dmin, fabs(dstmp[i]));
        double dstep = dmin + 2.0 * eps;
        oldpos = k->pos;
        k->pos = add(k->pos, scl(k->dir, dstep));
        sweep(S, T, k->pos, ds);
        int new_layer = 0; /* minloc of an all-zero `new` is the first index */
        for (int i = 0; i < nr; ++i)
            if (ds[i] < 0.0) { new_layer = i; break; }
        int layer = k->layer;
        double n1 = S->n[layer], n2 = S->n[new_layer];
        if (n1 != n2) {
            T->fresnel++;
            int use;
            if (ds[layer] < 0.0 && ds[new_layer] < 0.0)
                use = (fabs(ds[new_layer]) < fabs(ds[layer])) ? new_layer : layer;
            else if (dstmp[layer] < 0.0 && dstmp[new_layer] < 0.0)
                use = (dstmp[new_layer] > dstmp[layer]) ? new_layer : layer;
            else if (ds[layer] > 0.0 && ds[new_layer] < 0.0)
                use = layer;
            else if (ds[layer] > 0.0 && ds[new_layer] > 0.0) {
                tag_end(k, END_ESCAPE);
                break;
            } else {
                T->err_unreach++; /* "This should not be reached!" (:156) */
                tag_end(k, END_ERROR);
                break;
            }
            T->normals++;
            v3 N = calc_normal(S->d, use, k->pos);
            /* reflect_refract (src/surfaces.f90:14-40): draw the branch number */
            double Ri = fresnel(k->dir, N, n1, n2);
            int rflag = 0;
            double u[4];
            rng_block(&k->rng, u); /* the Fresnel event's block */
            if (u[0] <= Ri) { k->dir = reflect(k->dir, N); rflag = 1; }
            else k->dir = refract(k->dir, N, n1 / n2);
            k->weight *= Ri; /* either branch (:167) */
            tau = -log(u[1]); /* resampled optical depth (:168) */
            taurun = 0.0;
            if (!rflag) {
                k->layer = new_layer;
                trace_put(k->tr, k->te, &k->tc, k->pos, 3);
            } else {
                k->pos = oldpos;
                trace_put(k->tr, k->te, &k->tc, k->pos, 4);
                k->bounces++;
                if (k->bounces > max_bounces) { tag_end(k, END_BOUNCE); break; }
            }
        } else {
            k->layer = new_layer;
        }
        if (k->tflag) break;
    }
    /* epilogue grid check (:201-209) */
    if (fabs(k->pos.x) > S->xmax || fabs(k->pos.y) > S->ymax || fabs(k->pos.z) > S->zmax)
        tag_end(k, END_ESCAPE);
}

/* HG / isotropic scatter (src/photon.f90:713-779); uc, uphi: the cos theta and azimuth draws */
static void scatter(orc_packet* k, double hgg, double g2, double uc, double uphi) {
    double cost;
    if (hgg == 0.0) {
        cost = 2.0 * uc - 1.0;
    } else {
        double temp = (1.0 - g2) / (1.0 - hgg + 2.0 * hgg * uc);
        cost = (1.0 + g2 - temp * temp) / (2.0 * hgg);
    }
    double sint = sqrt(1.0 - cost * cost);
    double phi = ORC_TWOPI * uphi;
    double cosp = cos(phi);
    double sinp = (phi < ORC_PI) ? sqrt(1.0 - cosp * cosp) : -sqrt(1.0 - cosp * cosp);
    v3 u = k->dir, n;
    if (1.0 - fabs(u.z) <= 1e-12) {
        n = mk(sint * cosp, sint * sinp, copysign(cost, u.z));
    } else {
        double temp = sqrt(1.0 - u.z * u.z);
        n.x = sint * (u.x * u.z * cosp - u.y * sinp) / temp + u.x * cost;
        n.y = sint * (u.y * u.z * cosp + u.x * sinp) / temp + u.y * cost;
        n.z = -sint * cosp * temp + u.z * cost;
    }
    k->dir = n;
}

/* emitters (src/photon.f90); the emission block also gives the first flight's optical depth */
static void emit(const orc_scene* S, orc_packet* k) {
    const smcrt_source* src = &S->d->source;
    double u[4];
    rng_block(&k->rng, u);
    k->tau0 = -log(u[2]);
    switch (src->kind) {
    case SMCRT_SRC_PENCIL: /* :464-505 */
        k->pos = mk(src->pos[0], src->pos[1], src->pos[2]);
        k->dir = mk(src->dir[0], src->dir[1], src->dir[2]);
        break;
    case SMCRT_SRC_UNIFORM: { /* :397-461 */
        double rx = u[0], ry = u[1]; /* the two surface coordinates */
        k->pos = mk(src->p1[0] + rx * src->p2[0] + ry * src->p3[0],
                    src->p1[1] + rx * src->p2[1] + ry * src->p3[1],
                    src->p1[2] + rx * src->p2[2] + ry * src->p3[2]);
        k->dir = mk(src->dir[0], src->dir[1], src->dir[2]);
        break;
    }
    case SMCRT_SRC_CIRCULAR: { /* :214-287 */
        double r = src->radius * sqrt(u[0]); /* radius */
        double th = u[1] * ORC_TWOPI;         /* angle */
        v3 p = S->circ_swap ? mk(r * cos(th), r * sin(th), 0.0) : mk(0.0, r * cos(th), r * sin(th));
        k->pos = xform(S->circ_t, p);
        k->dir = mk(src->dir[0], src->dir[1], src->dir[2]);
        break;
    }
    case SMCRT_SRC_POINT: { /* :290-338 */
        double phi = u[0] * ORC_TWOPI;  /* azimuth */
        double cost = 2.0 * u[1] - 1.0; /* cos theta */
        double sint = sqrt(1.0 - cost * cost);
        k->pos = mk(src->pos[0], src->pos[1], src->pos[2]);
        k->dir = mk(sint * cos(phi), sint * sin(phi), cost);
        break;
    }
    }
}

/* ---------------------------------------------------------------------------------------
 * exported API (ctypes)
 * ------------------------------------------------------------------------------------- */
typedef struct orc_opts {
    uint64_t n_packets, seed, first_packet;
    uint32_t flags;
    int32_t max_bounces;
    int32_t threads;        /* 0 = OpenMP default */
    int32_t trace_packets, trace_events;
    int32_t grad;           /* 1 = accumulate pMC scores */
    const float* weights;   /* adjoint weight grid or NULL */
} orc_opts;

typedef struct orc_out {
    double* jmean;          /* may be NULL */
    float* trace;           /* trace_packets*trace_events*4 or NULL */
    double* d_r_dmua, *d_r_dmus, *d_w_dmua, *d_w_dmus; /* n_materials each, when grad */
    double w_tally;
    smcrt_stats stats;
    double seconds;         /* wall time of the packet loop (omp_get_wtime) */
    uint64_t* det_counts;   /* every detector's bins, concatenated (NULL: not recorded) */
    double weight_deposit;  /* weight estimator: total weight deposited at interactions */
} orc_out;

/* ---------------------------------------------------------------------------------------
 * detectors: constructors (src/detectors/detectors.f90 init_circle_dect, init_annulus_dect,
 * init_camera; the circle's direction normalised by src/parse/parse_detectors.f90), hit tests
 * (check_hit_circle with intersectCircle/intersectPlane, src/geometryMod.f90:217-271;
 * check_hit_annulus; check_hit_camera) and binning (record_hit_1D_sub / record_hit_2D_sub,
 * src/detectors/detector_base.f90:106-162)
 * ------------------------------------------------------------------------------------- */
typedef struct {
    int kind, layer, nb;  /* layer 1-based; nb = nbins + 1 bins (per axis for the camera) */
    size_t off;
    v3 pos, dir, e1, e2, n;
    double radius, r1, r2, width, height, bw, bwx, bwy;
} orc_det;

static size_t det_setup(const smcrt_detector* in, orc_det* d) {
    memset(d, 0, sizeof *d);
    d->kind = in->kind;
    d->layer = in->layer;
    d->nb = in->nbins + 1;
    d->pos = mk(in->pos[0], in->pos[1], in->pos[2]);
    d->dir = mk(in->dir[0], in->dir[1], in->dir[2]);
    if (in->kind == SMCRT_DET_CIRCLE) {
        d->dir = norm(d->dir); /* dir%magnitude() */
        d->radius = in->radius1;
        d->bw = in->nbins == 0 ? 1.0 : in->maxval / (double)(in->nbins - 1);
        return (size_t)d->nb;
    }
    if (in->kind == SMCRT_DET_ANNULUS) {
        d->r1 = in->radius1;
        d->r2 = in->radius2;
        d->bw = in->nbins == 0 ? 1.0 : in->maxval / (double)in->nbins;
        return (size_t)d->nb;
    }
    v3 p2 = mk(in->p2[0], in->p2[1], in->p2[2]), p3 = mk(in->p3[0], in->p3[1], in->p3[2]);
    d->e1 = sub(p2, d->pos);
    d->e2 = sub(p3, d->pos);
    d->width = sqrt(dot(d->e1, d->e1));
    d->height = sqrt(dot(d->e2, d->e2));
    v3 c = mk(d->e2.y * d->e1.z - d->e2.z * d->e1.y, d->e2.z * d->e1.x - d->e2.x * d->e1.z,
              d->e2.x * d->e1.y - d->e2.y * d->e1.x); /* e2 .cross. e1 */
    d->n = norm(c);
    d->bwx = in->nbins == 0 ? 1.0 : in->maxval / (double)d->nb;
    d->bwy = d->bwx;
    return (size_t)d->nb * (size_t)d->nb;
}

/* nint(q) + 1 clamped to nb, as a 0-based bin; -1 where the reference's index would leave the
   array (q negative or NaN) */
static long bin1d(double q, int nb) {
    if (!(q >= 0.0)) return -1;
    long i = q < 9.0e18 ? (long)llround(q) + 1 : nb;
    return (i < nb ? i : nb) - 1;
}

/* int(q) + 1 clamped to nb, wrapped to nb below 1 (record_hit_2D_sub); 0 for NaN */
static long axis_bin(double q, int nb) {
    if (q != q) return 0;
    long i = q >= 9.0e18 ? nb : (q <= -9.0e18 ? 0 : (long)q + 1);
    if (i > nb) i = nb;
    return i < 1 ? nb : i;
}

/* record one hit point (pos, dir, value, layer 1-based) on detector d */
static void det_record(const orc_det* d, v3 pos, v3 dir, double value, int layer, uint64_t* counts) {
    if (layer != d->layer) return;
    long idx = -1; /* 0-based bin */
    if (d->kind == SMCRT_DET_CIRCLE) {
        double denom = dot(d->dir, dir);
        if (!(denom > 1e-6)) return;
        double t = dot(sub(d->pos, pos), d->dir) / denom;
        if (!(t >= 0.0)) return;
        v3 v = sub(add(pos, scl(dir, t)), d->pos);
        if (!(sqrt(dot(v, v)) <= d->radius)) return;
        if (t > 5e-3) return;
        idx = bin1d(value / d->bw, d->nb);
    } else if (d->kind == SMCRT_DET_ANNULUS) {
        v3 v = sub(pos, d->pos);
        double r = sqrt(dot(v, v));
        if (!(r >= d->r1 && r <= d->r2)) return;
        idx = bin1d(value / d->bw, d->nb);
    } else {
        double t = dot(sub(d->pos, pos), d->n) / dot(dir, d->n);
        if (!(t >= 0.0)) return;
        v3 v = sub(add(pos, scl(dir, t)), d->pos);
        double q1 = dot(v, d->e1) / d->width, q2 = dot(v, d->e2) / d->height;
        if (!(q1 < d->width && q1 > 0.0 && q2 < d->height && q2 > 0.0)) return;
        double x = pos.z + d->pos.x, y = pos.y + d->pos.y; /* as record_hit_2D_sub */
        long ix = axis_bin(x / d->bwx, d->nb), iy = axis_bin(y / d->bwy, d->nb);
        if (ix < 1 || iy < 1) return;
        idx = (ix - 1) + (long)d->nb * (iy - 1);
    }
    if (idx < 0) return;
#pragma omp atomic
    counts[d->off + (size_t)idx] += 1;
}

static void trace_put(float* tr, int ne, int* cnt, v3 p, int ev) {
    if (!tr || *cnt >= ne) return;
    float* e = tr + 4 * (*cnt);
    e[0] = (float)p.x; e[1] = (float)p.y; e[2] = (float)p.z; e[3] = (float)ev;
    (*cnt)++;
}

int orc_run(const smcrt_scene_desc* d, const orc_opts* o, orc_out* out) {
    orc_scene S;
    if (d->n_regions < 1 || d->n_regions > SMCRT_MAX_REGIONS) return -1;
    if (d->n_materials > SMCRT_MAX_MATERIALS) return -1;
    scene_setup(&S, d);
    int max_b = o->max_bounces > 0 ? o->max_bounces : 1000;
    int nm = d->n_materials;
    memset(&out->stats, 0, sizeof out->stats);
    out->w_tally = 0.0;
    if (o->grad) {
        for (int m = 0; m < nm; ++m)
            out->d_r_dmua[m] = out->d_r_dmus[m] = out->d_w_dmua[m] = out->d_w_dmus[m] = 0.0;
    }
    double t0 = 0.0;
#ifdef _OPENMP
    if (o->threads > 0) omp_set_num_threads(o->threads);
    t0 = omp_get_wtime();
#endif
    uint64_t a_abs = 0, a_r = 0, a_t = 0, a_ks = 0, a_kb = 0, a_sc = 0, a_err = 0;
    uint64_t e_sw = 0, e_re = 0, e_no = 0, e_st = 0, e_cr = 0, e_if = 0, e_fr = 0, e_un = 0, e_dc = 0;
    double wt = 0.0;
    int64_t N = (int64_t)o->n_packets;
    orc_det dets[SMCRT_MAX_DETECTORS];
    int ndet = out->det_counts ? d->n_detectors : 0;
    const int wmode = (o->flags & SMCRT_FLAG_WEIGHT) != 0;
    double wdep = 0.0;
    {
        size_t off = 0;
        for (int i = 0; i < ndet; ++i) {
            size_t nb = det_setup(&d->detectors[i], &dets[i]);
            dets[i].off = off;
            for (size_t b = 0; b < nb; ++b) out->det_counts[off + b] = 0;
            off += nb;
        }
    }
#pragma omp parallel reduction(+ : a_abs, a_r, a_t, a_ks, a_kb, a_sc, a_err, e_sw, e_re, e_no, e_st, e_cr, e_if, e_fr, e_un, e_dc, wt, wdep)
    {
        orc_thread T;
        memset(&T, 0, sizeof T);
        T.jmean = out->jmean;
        T.w = o->weights;
        double mom[2][4][3];
        memset(mom, 0, sizeof mom);
        double gra[4][SMCRT_MAX_MATERIALS];
        memset(gra, 0, sizeof gra);
        orc_packet k;
#pragma omp for schedule(dynamic, 256)
        for (int64_t j = 0; j < N; ++j) {
            memset(&k, 0, sizeof k);
            k.grad = o->grad;
            rng_init(&k.rng, o->seed, o->first_packet + (uint64_t)j);
            emit(&S, &k);
            float* tr = (o->trace_packets > j && out->trace) ? out->trace + (size_t)j * o->trace_events * 4 : NULL;
            k.tr = tr;
            k.te = o->trace_events;
            k.tc = 0;
            trace_put(tr, o->trace_events, &k.tc, k.pos, 1);
            /* layer at emission (src/kernelsMod.f90:275-280) */
            {
                double best = INFINITY;
                int bi = 0;
                for (int i = 0; i < S.nreg; ++i) {
                    double di = region_eval(d, i, k.pos);
                    if (di > 0.0) di = -999.0;
                    if (fabs(di) < best) { best = fabs(di); bi = i; }
                }
                k.layer = bi;
                T.sweeps++;
            }
            k.weight = 1.0;
            tauint2(&S, &T, &k, max_b);
            int nsc = 0;
            while (!k.tflag && wmode) {
                /* weight_scatter (src/kernelsMod.f90:83-111) */
                double u[4];
                rng_block(&k.rng, u); /* the interaction's block */
                double wa = k.weight * (1.0 - S.albedo[k.layer]);
                k.weight -= wa;
                int ci, cj, ck;
                voxel_of(&S, mk(k.pos.x + S.xmax, k.pos.y + S.ymax, k.pos.z + S.zmax), &ci, &cj, &ck);
                if (ci < 1 || cj < 1 || ck < 1) { tag_end(&k, END_ESCAPE); break; }
                if (out->jmean) {
                    size_t idx = (size_t)(ci - 1) + (size_t)S.nx * ((size_t)(cj - 1) + (size_t)S.ny * (size_t)(ck - 1));
#pragma omp atomic
                    out->jmean[idx] += wa;
                }
                wdep += wa;
                scatter(&k, S.hgg[k.layer], S.g2[k.layer], u[1], u[2]);
                nsc++;
                a_sc++;
                trace_put(tr, o->trace_events, &k.tc, k.pos, 2);
                if (k.weight < 0.01) { /* THRESHOLD, CHANCE (src/constants.f90:28-30) */
                    if (u[0] < 0.1) k.weight /= 0.1; /* roulette */
                    else { tag_end(&k, END_ABSORB); break; }
                }
                k.tau0 = -log(u[3]);
                tauint2(&S, &T, &k, max_b);
            }
            while (!k.tflag) {
                double u[4];
                rng_block(&k.rng, u); /* the flight end's block */
                if (u[0] < S.albedo[k.layer]) { /* albedo test */
                    if (k.grad) {
                        int m = S.mat[k.layer];
                        k.K[m] += 1.0;
                        k.B[m] += k.C;
                    }
                    scatter(&k, S.hgg[k.layer], S.g2[k.layer], u[1], u[2]);
                    k.tau0 = -log(u[3]);
                    nsc++;
                    a_sc++;
                    trace_put(tr, o->trace_events, &k.tc, k.pos, 2);
                    if ((o->flags & SMCRT_FLAG_MOMENTS) && nsc <= 4) {
                        mom[0][nsc - 1][0] += k.pos.x; mom[0][nsc - 1][1] += k.pos.y; mom[0][nsc - 1][2] += k.pos.z;
                        mom[1][nsc - 1][0] += k.pos.x * k.pos.x; mom[1][nsc - 1][1] += k.pos.y * k.pos.y;
                        mom[1][nsc - 1][2] += k.pos.z * k.pos.z;
                    }
                    if ((o->flags & SMCRT_FLAG_END_EARLY) && nsc >= 5) { tag_end(&k, END_ABSORB); break; }
                } else {
                    tag_end(&k, END_ABSORB);
                    break;
                }
                tauint2(&S, &T, &k, max_b);
            }
            int isr = 0;
            switch (k.reason) {
            case END_ABSORB: a_abs++; trace_put(tr, o->trace_events, &k.tc, k.pos, 5); break;
            case END_ESCAPE:
                if (dot(k.dir, S.src_dir) < 0.0) { a_r++; isr = 1; } else a_t++;
                trace_put(tr, o->trace_events, &k.tc, k.pos, 6);
                break;
            case END_SURFACE: a_ks++; trace_put(tr, o->trace_events, &k.tc, k.pos, 7); break;
            case END_BOUNCE: a_kb++; trace_put(tr, o->trace_events, &k.tc, k.pos, 7); break;
            default: a_err++; break;
            }
            /* detectors see every ended history (src/kernelsMod.f90:302-306, :141-145); the hit
               value is sqrt(x^2+y^2), or the packet weight for weight_scatter */
            if (k.reason != END_ERROR) {
                const double hv = wmode ? k.weight : sqrt(k.pos.x * k.pos.x + k.pos.y * k.pos.y);
                for (int i = 0; i < ndet; ++i) det_record(&dets[i], k.pos, k.dir, hv, k.layer + 1, out->det_counts);
            }
            if (k.grad) {
                wt += k.C;
                for (int m = 0; m < nm; ++m) {
                    double mus = d->materials[m].mus;
                    double sc_s = (mus > 0.0) ? k.K[m] / mus : 0.0;
                    if (isr) {
                        gra[0][m] += -k.L[m];
                        gra[1][m] += sc_s - k.L[m];
                    }
                    double lw = k.L[m] * k.C - k.A[m];
                    double bw = (mus > 0.0) ? (k.K[m] * k.C - k.B[m]) / mus : 0.0;
                    gra[2][m] += -lw;
                    gra[3][m] += bw - lw;
                }
            }
        }
        e_sw += T.sweeps; e_re += T.region_evals; e_no += T.normals; e_st += T.steps;
        e_cr += T.crossings; e_if += T.ifaces; e_fr += T.fresnel; e_un += T.err_unreach; e_dc += T.err_dcell;
#pragma omp critical
        {
            for (int a = 0; a < 2; ++a)
                for (int b = 0; b < 4; ++b)
                    for (int c = 0; c < 3; ++c) out->stats.moments[a][b][c] += mom[a][b][c];
            if (o->grad)
                for (int m = 0; m < nm; ++m) {
                    out->d_r_dmua[m] += gra[0][m];
                    out->d_r_dmus[m] += gra[1][m];
                    out->d_w_dmua[m] += gra[2][m];
                    out->d_w_dmus[m] += gra[3][m];
                }
        }
    }
#ifdef _OPENMP
    out->seconds = omp_get_wtime() - t0;
#else
    out->seconds = 0.0;
#endif
    smcrt_stats* st = &out->stats;
    st->launched = (uint64_t)N;
    st->absorbed = a_abs; st->escaped_r = a_r; st->escaped_t = a_t;
    st->killed_surface = a_ks; st->killed_bounce = a_kb; st->scatters = a_sc;
    st->ev_sweeps = e_sw; st->ev_region_evals = e_re; st->ev_normals = e_no; st->ev_steps = e_st;
    st->ev_crossings = e_cr; st->ev_interfaces = e_if; st->ev_fresnel = e_fr;
    st->err_unreachable = e_un + a_err;
    st->err_nan = e_dc;
    out->weight_deposit = wdep;
    if (o->grad && N > 0) {
        out->w_tally = wt / (double)N;
        for (int m = 0; m < nm; ++m) {
            out->d_r_dmua[m] /= (double)N; out->d_r_dmus[m] /= (double)N;
            out->d_w_dmua[m] /= (double)N; out->d_w_dmus[m] /= (double)N;
        }
    }
    if (out->jmean) {
        double s = 0.0;
        size_t nv = (size_t)S.nx * S.ny * S.nz;
        for (size_t i = 0; i < nv; ++i) s += out->jmean[i];
        st->tally_sum = s;
    }
    return 0;
}

/* --- unit-level entry points for the parity tests ------------------------------------- */

void orc_eval(const smcrt_scene_desc* d, int region, int64_t n, const double* pts, double* dist,
              double* normals) {
    for (int64_t i = 0; i < n; ++i) {
        v3 p = mk(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2]);
        if (dist) dist[i] = region_eval(d, region, p);
        if (normals) {
            v3 nn = calc_normal(d, region, p);
            normals[3 * i] = nn.x; normals[3 * i + 1] = nn.y; normals[3 * i + 2] = nn.z;
        }
    }
}

/* Fresnel coefficient and both candidate directions for each (I, N, n1, n2) row */
void orc_fresnel(int64_t n, const double* I, const double* N, const double* n1, const double* n2,
                 double* R, double* refl, double* refr) {
    for (int64_t i = 0; i < n; ++i) {
        v3 a = mk(I[3 * i], I[3 * i + 1], I[3 * i + 2]), b = mk(N[3 * i], N[3 * i + 1], N[3 * i + 2]);
        R[i] = fresnel(a, b, n1[i], n2[i]);
        v3 r1 = reflect(a, b), r2 = refract(a, b, n1[i] / n2[i]);
        refl[3 * i] = r1.x; refl[3 * i + 1] = r1.y; refl[3 * i + 2] = r1.z;
        refr[3 * i] = r2.x; refr[3 * i + 1] = r2.y; refr[3 * i + 2] = r2.z;
    }
}

/* scatter n directions dir_in with g, each packet with its own Philox stream (seed, id0+i) */
void orc_scatter(int64_t n, const double* dir_in, double g, uint64_t seed, uint64_t id0,
                 double* dir_out) {
    for (int64_t i = 0; i < n; ++i) {
        orc_packet k;
        memset(&k, 0, sizeof k);
        rng_init(&k.rng, seed, id0 + (uint64_t)i);
        k.dir = mk(dir_in[3 * i], dir_in[3 * i + 1], dir_in[3 * i + 2]);
        double u[4];
        rng_block(&k.rng, u);
        scatter(&k, g, g * g, u[1], u[2]);
        dir_out[3 * i] = k.dir.x; dir_out[3 * i + 1] = k.dir.y; dir_out[3 * i + 2] = k.dir.z;
    }
}

/* emit n packets (ids id0..id0+n-1): positions, directions */
void orc_emit(const smcrt_scene_desc* d, int64_t n, uint64_t seed, uint64_t id0, double* pos,
              double* dir) {
    orc_scene S;
    scene_setup(&S, d);
    for (int64_t i = 0; i < n; ++i) {
        orc_packet k;
        memset(&k, 0, sizeof k);
        rng_init(&k.rng, seed, id0 + (uint64_t)i);
        emit(&S, &k);
        pos[3 * i] = k.pos.x; pos[3 * i + 1] = k.pos.y; pos[3 * i + 2] = k.pos.z;
        dir[3 * i] = k.dir.x; dir[3 * i + 1] = k.dir.y; dir[3 * i + 2] = k.dir.z;
    }
}

/* one `update_grids` walk (raw tally into jmean), returns the tallied length and the
   grid-exit flag */
double orc_walk(const smcrt_scene_desc* d, const double* pos, const double* dir, double dstep,
                double* jmean, int* left_grid) {
    orc_scene S;
    scene_setup(&S, d);
    orc_thread T;
    memset(&T, 0, sizeof T);
    T.jmean = jmean;
    orc_packet k;
    memset(&k, 0, sizeof k);
    double r = update_grids(&S, &T, &k, mk(pos[0], pos[1], pos[2]), mk(dir[0], dir[1], dir[2]), dstep);
    *left_grid = k.tflag;
    return r;
}

/* raw Philox uniforms for stream (seed, id): the first n draws */
void orc_uniforms(uint64_t seed, uint64_t id, int64_t n, double* out) {
    orc_rng g;
    rng_init(&g, seed, id);
    for (int64_t i = 0; i < n; ++i) out[i] = rng_u(&g);
}

/* detector hit test and binning on given hit points (tests pin these against hand-worked
   geometry): n points with layer (1-based); counts has the detector's bins */
void orc_detector_hits(const smcrt_detector* det, int64_t n, const double* pos, const double* dir,
                       const int* layer, uint64_t* counts) {
    orc_det d;
    size_t nb = det_setup(det, &d);
    d.off = 0;
    for (size_t b = 0; b < nb; ++b) counts[b] = 0;
    for (int64_t i = 0; i < n; ++i)
        det_record(&d, mk(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]),
                   mk(dir[3 * i], dir[3 * i + 1], dir[3 * i + 2]),
                   sqrt(pos[3 * i] * pos[3 * i] + pos[3 * i + 1] * pos[3 * i + 1]), layer[i], counts);
}

/* `render_sub` (src/sdfs/sdf_base.f90:302-346): ns = nint(samples/2), wid = extent/ns, image
   (column-major) = min over every region at ((i-ns) wid, (j-ns) wid, (k-ns) wid), i 1-based */
void orc_render(const smcrt_scene_desc* d, const int32_t* samples, double* image) {
    double ns[3], wid[3];
    const double ext[3] = {d->grid.xmax, d->grid.ymax, d->grid.zmax};
    for (int a = 0; a < 3; ++a) {
        ns[a] = (double)llround(samples[a] / 2.0);
        wid[a] = ext[a] / ns[a];
    }
    const int64_t sx = samples[0], sy = samples[1], sz = samples[2];
#pragma omp parallel for schedule(static)
    for (int64_t k = 0; k < sz; ++k)
        for (int64_t j = 0; j < sy; ++j)
            for (int64_t i = 0; i < sx; ++i) {
                v3 p = mk((i + 1 - ns[0]) * wid[0], (j + 1 - ns[1]) * wid[1], (k + 1 - ns[2]) * wid[2]);
                double m = INFINITY;
                for (int r = 0; r < d->n_regions; ++r) m = fmin(m, region_eval(d, r, p));
                image[i + sx * (j + sy * k)] = m;
            }
}
