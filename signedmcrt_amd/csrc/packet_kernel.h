// packet_kernel.h — the persistent, wavefront-regenerating packet kernel.
//
// One lane owns one packet at a time.  Each loop iteration, every busy lane that is not in the
// middle of a voxel walk performs one SDF sweep (all regions evaluated at its position) and
// then the state-machine step that the reference runs after that sweep inside `tauint2`
// (src/inttau2.f90:15-210):
//
//   EMIT     sweep at the emission point: layer = nearest containing surface
//            (src/kernelsMod.f90:275-280) (tau was drawn with the emission), then as START
//   START    flight start (src/inttau2.f90:52-61): kill if within eps of a surface, else march
//   MARCHED  after a step (:85-97): escape if outside every region, march again while d > eps,
//            otherwise cross: step min|d| + 2 eps along dir (:104-116)
//   CROSSED  after the crossing step (:117-185): first-negative region rule, normal choice,
//            Fresnel reflect/refract at the SDF normal, tau resample, bounce cap
//
// A march step of length d consumes d*kappa of optical depth (:64-83) and walks the voxel grid
// with an integer DDA (Amanatides & Woo 1987), adding each chord to the FP64 fluence grid — the
// path-length estimator of `update_grids` (:213-292).  A walk advances at most `dda_cap` chords
// per iteration and resumes in the next one, so a few long walks do not stall the wavefront.
// Consecutive chords in the same voxel (a step ending and the next one starting in it) are
// merged; a block keeps the hottest voxel box in an FP64 LDS tile, and the other chords go through a
// per-wavefront LDS queue to FP64 atomics into its XCD's partial grid.  When the flight's optical
// depth runs out, the albedo test absorbs or scatters (HG, src/photon.f90:713-779) and a new
// flight starts.
// A lane whose packet ends takes the next packet id from its wavefront's chunk at the top of
// the next iteration (regeneration); a wavefront claims a chunk of ids from a global counter
// with one atomic.  Packet j draws from Philox stream (seed, first_packet + j), one block per
// event, so results do not depend on grid size, block size, chunking, the walk cap or the
// number of resident waves.
#pragma once
#include "smcrt_device.h"

namespace smcrt {

// ST_SCATTER: the flight ended inside the grid; the interaction (albedo test, scatter) runs in
// the next iteration, after the sweep at the interaction point, with the iteration's one event
// block (shared by emission, Fresnel and interaction lanes: one Philox site per iteration).
enum : int { ST_IDLE = 0, ST_EMIT = 1, ST_START = 2, ST_MARCHED = 3, ST_CROSSED = 4, ST_SCATTER = 5 };
enum : int { END_ABSORB = 0, END_R = 1, END_T = 2, END_SURFACE = 3, END_BOUNCE = 4, END_ERR = 5 };

// stats layout (unsigned long long): per set [NOUT], then global event counters [NEV]
enum : int {
    O_LAUNCHED = 0, O_ABSORBED, O_R, O_T, O_KSURF, O_KBOUNCE, O_SCATTERS, NOUT
};
enum : int {
    E_SWEEPS = 0, E_REGION, E_NORMALS, E_STEPS, E_CROSS, E_IFACE, E_FRESNEL, E_UNREACH, E_NAN, E_ATOMICS, NEV
};

// most region-indexed material entries staged in LDS (nsets * nregions)
#define SMCRT_LDS_MATS 2048
// most regions with a gradient descriptor in LDS (later regions take the waterfall)
#define SMCRT_LDS_GRAD 128

// Parameters used off the hot path (emission, outcome classification, tracing, moments, the
// adjoint), kept in device memory and read with scalar loads where they are used so that they
// do not occupy SGPRs across the packet loop.
struct KCold {
    // source
    int src_kind, circ_swap;
    float spos[3], sdir[3], p1[3], p2[3], p3[3], radius;
    float circ_t[12];
    float rdir[3];  // R/T classification direction
    // sets, counters, tracing
    const unsigned long long* set_prefix;  // nsets+1 cumulative budgets
    const unsigned long long* set_first;   // global packet id of each set's first packet
    unsigned long long* stats;             // nsets*NOUT + NEV
    double* moments;                       // 24 doubles
    float* trace;
    int trace_packets, trace_events;
    uint32_t flags;
    int max_bounces;
    // adjoint
    const float* weights;  // nullptr: w = 1
    double* grad;          // [4][nmat] + [1] weighted tally
    const float* mat_mus;  // mus per adjoint slot
    // detectors: every ended history is offered to each (0: none, and no grid-exit continuation)
    int n_det;
    const struct DevDet* det;
    unsigned long long* det_counts;  // every detector's bins, concatenated
};
typedef __attribute__((address_space(4))) const KCold CKCold;

// A detector prepared for the hit test (src/detectors/detectors.f90).  Geometry in FP64 as the
// reference keeps it; the test runs once per history.
struct DevDet {
    int kind, layer0, nb, pad;  // kind 1 circle, 2 annulus, 3 camera; 0-based layer; bins per axis
    unsigned long long off;     // first bin in det_counts
    double p[3], n[3];          // circle: position, unit direction; annulus: position; camera: p1, unit normal
    double e1[3], e2[3];        // camera edges p2-p1, p3-p1
    double a, b;                // circle: radius; annulus: r1, r2; camera: width, height
    double bw, bwy;             // bin widths
};
typedef __attribute__((address_space(4))) const DevDet CDevDet;

struct KParams {
    const DevNode* nodes;
    const DevRegion* regions;
    const float* accel;  // polygon accelerators (nullptr: none)
    const DevMat* mats;  // [nsets][nreg] region-indexed material constants
    int nreg, nmat;      // regions, adjoint materials
    // grid
    int nx, ny, nz;
    float xmax, ymax, zmax;
    float sx, sy, sz;  // voxels per cm
    double* jmean;     // set 0 grid (nullptr: no tally)
    unsigned long long grid_stride;
    // per-XCD partial grids: block b tallies into xgrid + (b % xcopies) * xstride (same layout as
    // jmean), summed into jmean after the launch; nullptr: straight into jmean.  Atomics from
    // different XCDs on one cache line are far slower than atomics within one XCD.
    double* xgrid;
    unsigned long long xstride;
    int xcopies;
    float eps;       // march / crossing epsilon
    float eps_kill;  // flight-start kill threshold after emission or scatter
    // run
    uint32_t seed_lo, seed_hi;
    unsigned long long total;              // work items (sum of set budgets)
    int nsets;
    unsigned long long* work;              // work counter
    int chunk, dda_cap;
    int refill_min;  // regeneration policy: refill once at least this many lanes are idle
    // LDS-privatised tally tile: voxels [t0, t0 + tn) per axis accumulate in FP64 LDS per block
    // and are flushed into the FP64 grid at the end (tn[0] == 0: no tile)
    int t0[3], tn[3];
    // Binned tally (kernels built with BIN, single-set forward launches): the grid's flat index
    // range is cut into bins of 2^SMCRT_BIN_SHIFT voxels; entries of the nhot hot bins
    // (binmap[bin] = slot, SMCRT_BIN_COLD otherwise) are staged per block in LDS, SMCRT_BIN_B to a
    // slot, and a full slot is appended with plain stores to the slot's current page of `pool`
    // (SMCRT_BIN_PAGE entries; block b takes pages from [b * ppb, (b + 1) * ppb)); the pages are
    // summed per bin after the launch.  Other entries, and any once the block's pages run out,
    // take the FP64 atomics.
    int nbins, nhot;
    const uint16_t* binmap;
    uint2* pool;              // pages of (grid index, FP32 length) entries
    uint32_t* page_bin;       // bin of each page
    uint32_t* page_fill;      // entries in each page
    uint32_t* block_pages;    // pages each block took
    unsigned ppb;             // pages per block
    const struct KCold* cold;  // rarely used parameters (constant memory)
};

// Binned tally geometry (see KParams): bin size (voxels), entries per staged slot, page size
#define SMCRT_BIN_SHIFT 14
#define SMCRT_BIN_B 8
#define SMCRT_BIN_PAGE 512
#define SMCRT_BIN_COLD 0xFFFFu
#define SMCRT_PAGE_NONE 0xFFFFFFFFu

// LDS bytes of a block's binned-tally stage for `waves` wavefronts: the bin map (u16 per bin),
// per hot bin a reserve and a write counter, its current page and that page's fill (u32 each),
// the block's page counter (16 bytes), per wavefront 128 u32 of flush scratch, then the staged
// entries
__host__ __device__ inline size_t bins_map_bytes(int nbins) { return ((size_t)nbins * 2 + 15) & ~(size_t)15; }
__host__ __device__ inline size_t bins_lds_bytes(int nbins, int nhot, int waves) {
    return nhot > 0 ? bins_map_bytes(nbins) + 16 * (size_t)nhot + 16 + 512 * (size_t)waves + 8 * (size_t)SMCRT_BIN_B * nhot
                    : 0;
}

struct Lane {
    float px, py, pz, dx, dy, dz;
    float opx, opy, opz;  // position before the crossing step
    float tau, taurun, dlay_old, dprev;
    float dfn_old;        // the pre-crossing sweep's first negative region (fn_old) and its distance
    int fn_old;
    float weight;         // packet weight (the weight_scatter estimator; 1 otherwise)
    float rx, ry, rz;     // rounding error of the march moves so far (see move_axis)
    int layer, st, bounces, nsc, set, tcount;
    bool iface;           // the current flight started at an interface
    bool ghost;           // left the grid mid-flight: moves on, untallied, to the flight's end
                          // or the next surface (only with detectors: the final position)
    unsigned iters;
    unsigned long long pid;
    Rng rng;
};

// voxel walk in progress
struct Walk {
    int ix, iy, iz, idx;
    float tmx, tmy, tmz, t, d;
    bool active, flight_over, inside;
    float aint;  // adjoint: integral of C over the step
};

// p += inc with the rounding error carried in r (Knuth's TwoSum).  A march step is the distance
// to the nearest surface; a packet leaving a surface at a grazing angle takes steps of its height
// h above it and rises h*|cos| per step.  Below half an ulp of the coordinate that rise rounds
// away in FP32, the height never grows and the packet crawls along the surface in steps of h
// (1e5 steps for 0.1 cm at h = 1e-6).  Carried, the rises accumulate and the height grows
// geometrically, as in the reference's FP64 arithmetic.
SMCRT_DEV void move_axis(float& p, float& r, float inc) {
    const float b = inc + r;
    const float s = p + b;
    const float bb = s - p;
    r = (p - (s - bb)) + (b - bb);
    p = s;
}

SMCRT_DEV void move(Lane& L, float d) {
    move_axis(L.px, L.rx, d * L.dx);
    move_axis(L.py, L.ry, d * L.dy);
    move_axis(L.pz, L.rz, d * L.dz);
}

// the packet's next event block of four uniforms
SMCRT_DEV void draw_block(Lane& L, const KParams& P, float u[4]) { rng_block(L.rng, L.pid, P.seed_lo, P.seed_hi, u); }

SMCRT_DEV uint64_t wave_ballot(bool b) { return __ballot(b); }

SMCRT_DEV void trace_event(CKCold* C, Lane& L, int kind) {
    if (C->trace && L.pid < (unsigned long long)C->trace_packets && L.tcount < C->trace_events) {
        float* e = C->trace + ((size_t)L.pid * C->trace_events + L.tcount) * 4;
        e[0] = L.px; e[1] = L.py; e[2] = L.pz; e[3] = (float)kind;
        L.tcount++;
    }
}

// Whether the emitter's position depends on its draws (uniform, circular): those draw the
// emission block at emission, before the sweep; the others (pencil, point) start at a fixed
// position and draw it with the iteration's event block after the sweep.
template <int KSRC>
SMCRT_DEV bool emit_draws_position(CKCold* C) {
    const int k = KSRC >= 0 ? KSRC : C->src_kind;
    return k == 2 || k == 3;
}

// emitters (src/photon.f90): pencil :464-505, uniform :397-461, circular :214-287, point :290-338.
// The emission block u: u0, u1 the emitter's draws, u2 the first flight's optical depth.
template <int KSRC>
SMCRT_DEV void emit(CKCold* C, Lane& L, const float u[4]) {
    L.tau = f_neglog(u[2]);
    switch (KSRC >= 0 ? KSRC : C->src_kind) {
    case 2: {  // uniform: draws rx, ry
        float rx = u[0], ry = u[1];
        L.px = C->p1[0] + rx * C->p2[0] + ry * C->p3[0];
        L.py = C->p1[1] + rx * C->p2[1] + ry * C->p3[1];
        L.pz = C->p1[2] + rx * C->p2[2] + ry * C->p3[2];
        L.dx = C->sdir[0]; L.dy = C->sdir[1]; L.dz = C->sdir[2];
        break;
    }
    case 3: {  // circular: draws radius, angle; disc rotated onto dir, centred at -position
        float r = C->radius * f_sqrt(u[0]);
        float th = u[1];
        float c = f_cos_rev(th), s = f_sin_rev(th);
        float qx, qy, qz;
        if (C->circ_swap) { qx = r * c; qy = r * s; qz = 0.f; }
        else { qx = 0.f; qy = r * c; qz = r * s; }
        const __attribute__((address_space(4))) float* t = C->circ_t;
        L.px = qx * t[0] + qy * t[3] + qz * t[6] + t[9];
        L.py = qx * t[1] + qy * t[4] + qz * t[7] + t[10];
        L.pz = qx * t[2] + qy * t[5] + qz * t[8] + t[11];
        L.dx = C->sdir[0]; L.dy = C->sdir[1]; L.dz = C->sdir[2];
        break;
    }
    case 4: {  // point: draws azimuth, cos theta
        float ph = u[0];
        float ct = 2.f * u[1] - 1.f;
        float stt = f_sqrt(fmaxf(1.f - ct * ct, 0.f));
        L.px = C->spos[0]; L.py = C->spos[1]; L.pz = C->spos[2];
        L.dx = stt * f_cos_rev(ph); L.dy = stt * f_sin_rev(ph); L.dz = ct;
        break;
    }
    default: {  // pencil
        L.px = C->spos[0]; L.py = C->spos[1]; L.pz = C->spos[2];
        L.dx = C->sdir[0]; L.dy = C->sdir[1]; L.dz = C->sdir[2];
        break;
    }
    }
}

// HG / isotropic scatter (src/photon.f90:713-779); uc, ph: the cos theta and azimuth draws.
// The near-axis branch keeps the reference's Fortran `sign(cost, nzp)` = |cos theta|*sign(nzp).
SMCRT_DEV void scatter(Lane& L, float hgg, float g2, float uc, float ph) {
    float cost;
    if (hgg == 0.f) {
        cost = 2.f * uc - 1.f;
    } else {
        float temp = f_div(1.f - g2, 1.f - hgg + 2.f * hgg * uc);
        cost = f_div(1.f + g2 - temp * temp, 2.f * hgg);
        cost = fminf(fmaxf(cost, -1.f), 1.f);
    }
    float sint = f_sqrt(fmaxf(1.f - cost * cost, 0.f));
    float cosp = f_cos_rev(ph);
    float sinp = f_sqrt(fmaxf(1.f - cosp * cosp, 0.f));
    if (ph >= 0.5f) sinp = -sinp;  // phi >= pi
    float ux = L.dx, uy = L.dy, uz = L.dz;
    float nx, ny, nz;
    if (1.f - fabsf(uz) <= 1e-12f) {
        nx = sint * cosp;
        ny = sint * sinp;
        nz = uz < 0.f ? -fabsf(cost) : fabsf(cost);
    } else {
        float temp = f_sqrt(1.f - uz * uz);
        float rt = f_rcp(temp);
        nx = sint * (ux * uz * cosp - uy * sinp) * rt + ux * cost;
        ny = sint * (uy * uz * cosp + ux * sinp) * rt + uy * cost;
        nz = -sint * cosp * temp + uz * cost;
    }
    float inv = f_rsqrt(nx * nx + ny * ny + nz * nz);
    L.dx = nx * inv; L.dy = ny * inv; L.dz = nz * inv;
}

// Fresnel (src/surfaces.f90:86-120)
SMCRT_DEV float fresnel(float ix, float iy, float iz, float nx, float ny, float nz, float n1, float n2) {
    float costt = fminf(fabsf(ix * nx + iy * ny + iz * nz), 1.f);
    float sintt = f_sqrt(fmaxf(1.f - costt * costt, 0.f));
    float sint2 = f_div(n1, n2) * sintt;
    if (sint2 > 1.f) return 1.f;
    if (costt == 1.f) return 0.f;
    float cost2 = f_sqrt(fmaxf(1.f - sint2 * sint2, 0.f));
    float f1 = f_div(n1 * costt - n2 * cost2, n1 * costt + n2 * cost2);
    float f2 = f_div(n1 * cost2 - n2 * costt, n1 * cost2 + n2 * costt);
    return 0.5f * (f1 * f1 + f2 * f2);
}

// Evaluate, per lane, the region r_lane at p (only lanes with `active` set).  Waterfall over the
// distinct region ids in the wavefront so each evaluation runs the wave-uniform interpreter.
template <class T, uint32_t KM>
SMCRT_DEV T eval_lane_region(CDevNode* nodes, CDevRegion* regions, bool active, int r_lane, const V3<T>& p,
                             const float* __restrict__ accel) {
    T res = g_const<T>(0.f);
    uint64_t pending = wave_ballot(active);
    while (pending) {
        int leader = __builtin_ctzll(pending);
        const int r = __builtin_amdgcn_readlane(r_lane, leader);
        const bool mine = active && r_lane == r;
        // the region's bounds are wave-uniform: read them (scalar) before the divergent branch
        const int first = __builtin_amdgcn_readfirstlane(regions[r].first);
        const int count = __builtin_amdgcn_readfirstlane(regions[r].count);
        if (mine) res = region_eval<T, KM>(nodes, first, count, p, accel);
        pending &= ~wave_ballot(mine);
    }
    return res;
}

// One primitive node (index i, wave-uniform) at p: transform (unless the identity), then the
// primitive.  readfirstlane keeps the node reads scalar wherever the call sits.
template <class T, uint32_t KM>
SMCRT_DEV T prim_node(CDevNode* nodes, int i, const V3<T>& p, const float* __restrict__ accel) {
    CDevNode* n = nodes + __builtin_amdgcn_readfirstlane(i);
    float pp[8], qq[8], tt[12];
#pragma unroll
    for (int j = 0; j < 8; ++j) pp[j] = n->p[j];
#pragma unroll
    for (int j = 0; j < 8; ++j) qq[j] = n->q[j];
    V3<T> q = p;
    if (n->ident == 0) {
#pragma unroll
        for (int j = 0; j < 12; ++j) tt[j] = n->t[j];
        q = xform<T>(tt, p);
    } else if (n->ident == 2) {  // translation only
        q = V3<T>{p.x + n->t[9], p.y + n->t[10], p.z + n->t[11]};
    }
    return prim_eval<T, KM>(n->kind, pp, qq, q, accel);
}

SMCRT_DEV V3<float> dual_grad(const Dual& d) { return V3<float>{d.x, d.y, d.z}; }

// Closed-form gradient of a sphere (radius unused) or a box (half sizes hx, hy, hz) at q in the
// primitive's frame: the derivative the Dual evaluation of the same expressions gives, branch for
// branch (|x| has slope +1 at 0, max keeps its first argument on ties, sqrt has slope 0 at 0).
SMCRT_DEV V3<float> closed_form_gradient(bool sphere, float hx, float hy, float hz, const V3<float>& q) {
    if (sphere) {
        const float s = f_sqrt(q.x * q.x + q.y * q.y + q.z * q.z);
        const float h = s > 0.f ? f_rcp(s) : 0.f;
        return V3<float>{q.x * h, q.y * h, q.z * h};
    }
    const float sx = q.x < 0.f ? -1.f : 1.f, sy = q.y < 0.f ? -1.f : 1.f, sz = q.z < 0.f ? -1.f : 1.f;
    const float dx = fabsf(q.x) - hx, dy = fabsf(q.y) - hy, dz = fabsf(q.z) - hz;
    const float mx = dx > 0.f ? dx : 0.f, my = dy > 0.f ? dy : 0.f, mz = dz > 0.f ? dz : 0.f;
    const float s = f_sqrt(mx * mx + my * my + mz * mz);
    const float h = s > 0.f ? f_rcp(s) : 0.f;
    V3<float> gq{sx * mx * h, sy * my * h, sz * mz * h};
    // min(max(dx, max(dy, dz)), 0): the largest component when negative
    const bool zy = dz > dy;
    const float inner = zy ? dz : dy;
    const bool ix = inner > dx;
    if ((ix ? inner : dx) < 0.f) {
        if (!ix) gq.x += sx;
        else if (zy) gq.z += sz;
        else gq.y += sy;
    }
    return gq;
}

// Closed-form gradient of a single sphere or box node at p (the derivative the Dual evaluation
// of the same expressions gives, branch for branch: |x| has slope +1 at 0, max keeps its first
// argument on ties, sqrt has slope 0 at 0), pulled back through the node's affine transform.
// Returns false for other kinds (the caller differentiates them with Dual numbers).
template <uint32_t KM>
SMCRT_DEV bool fast_prim_gradient(CDevNode* nodes, int i, const V3<float>& p, V3<float>& g) {
    CDevNode* n = nodes + __builtin_amdgcn_readfirstlane(i);
    const int kind = n->kind;
    const bool sphere = KM_HAS(KM, KM_PRIM(SMCRT_SPHERE)) && kind == SMCRT_SPHERE;
    const bool box = KM_HAS(KM, KM_PRIM(SMCRT_BOX)) && kind == SMCRT_BOX;
    if (n->op != SMCRT_OP_PRIM || !(sphere || box)) return false;
    float tt[12];
    V3<float> q = p;
    // the gradient is pulled back through the 3x3 part only: none for the identity and for a
    // translation
    const bool ident = n->ident != 0;
    if (n->ident == 0) {
#pragma unroll
        for (int j = 0; j < 12; ++j) tt[j] = n->t[j];
        q = xform<float>(tt, p);
    } else if (n->ident == 2) {
        q = V3<float>{p.x + n->t[9], p.y + n->t[10], p.z + n->t[11]};
    }
    const V3<float> gq = closed_form_gradient(sphere, n->p[0], n->p[1], n->p[2], q);
    if (ident) {
        g = gq;
    } else {  // q_j = sum_i p_i t[3i + j] + t[9 + j]
        g = V3<float>{gq.x * tt[0] + gq.y * tt[1] + gq.z * tt[2], gq.x * tt[3] + gq.y * tt[4] + gq.z * tt[5],
                      gq.x * tt[6] + gq.y * tt[7] + gq.z * tt[8]};
    }
    return true;
}

// Exact gradient of region r at p.  A flat fold of primitives by union or SmoothUnion (a `model`
// of primitives, DevRegion.flat) is differentiated lazily, forward-mode through the fold:
// running value a, gradient g; each child c_k enters as a <- op(a, c_k), g <- alpha g + beta
// grad(c_k) with alpha, beta the partials of op, and grad(c_k) (a Dual evaluation of one
// primitive) is computed only where beta != 0.  The walk starts at the last child that wins
// outright (alpha = 0 there), found in a first float pass, so children that cannot shape the
// surface at p cost one float evaluation.  Any other program runs the Dual interpreter once.
//
// The program's node indices are wave-uniform; readfirstlane keeps them (and so every node read)
// scalar inside the waterfall's divergent branches.
template <uint32_t KM>
SMCRT_DEV V3<float> region_gradient(CDevNode* nodes, int first, int count, int flat, const V3<float>& p,
                                    const float* __restrict__ accel) {
    const V3<Dual> pd{Dual{p.x, 1.f, 0.f, 0.f}, Dual{p.y, 0.f, 1.f, 0.f}, Dual{p.z, 0.f, 0.f, 1.f}};
    // flat folds exist only in programs with a union or SmoothUnion
    if (count == 1) {
        V3<float> g;
        if (fast_prim_gradient<KM>(nodes, first, p, g)) return g;
    }
    if constexpr (!KM_HAS(KM, KM_CSG(1) | KM_CSG(2))) return dual_grad(region_eval<Dual, KM>(nodes, first, count, pd, accel));
    if (!flat) return dual_grad(region_eval<Dual, KM>(nodes, first, count, pd, accel));
    const int nch = (count + 1) / 2;
    const bool smooth = KM_HAS(KM, KM_CSG(2)) && flat == 2;
    const int ki = __builtin_amdgcn_readfirstlane(first + 2);  // the fold's first operator node
    const float k = smooth ? nodes[ki].k : 0.f;
    float kq[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) kq[j] = smooth ? nodes[ki].q[j] : 0.f;
    // pass 1: index of the last child that wins outright
    float a = prim_node<float, KM>(nodes, first, p, accel);
    int last = 0;
    for (int j = 1; j < nch; ++j) {
        const float c = prim_node<float, KM>(nodes, first + 2 * j - 1, p, accel);
        const bool wins = smooth ? (a - c >= k) : (c < a);
        last = wins ? j : last;
        a = smooth ? csg_eval<float, KM>(SMCRT_SMOOTH_UNION, a, c, k, kq) : fminf(a, c);
    }
    // pass 2: forward-mode from that child on
    int li = last == 0 ? first : first + 2 * last - 1;
    V3<float> g{0.f, 0.f, 0.f};
    a = 0.f;
    {
        // the starting child's gradient (a waterfall over the lanes' distinct start children)
        uint64_t pending = wave_ballot(true);
        while (pending) {
            const int lead = __builtin_ctzll(pending);
            const int ls = __builtin_amdgcn_readlane(li, lead);
            const bool mine = li == ls;
            if (mine) {
                const Dual d = prim_node<Dual, KM>(nodes, ls, pd, accel);
                a = d.v;
                g = dual_grad(d);
            }
            pending &= ~wave_ballot(mine);
        }
    }
    for (int j = 1; j < nch; ++j) {
        if (j <= last) continue;
        const int ni = first + 2 * j - 1;
        const float c = prim_node<float, KM>(nodes, ni, p, accel);
        float alpha, beta;
        if (smooth) {
            const float h = fmaxf(k - fabsf(a - c), 0.f) * kq[0], h2 = 0.5f * h * h;
            alpha = c < a ? h2 : 1.f - h2;
            beta = c < a ? 1.f - h2 : h2;
        } else {
            alpha = c < a ? 0.f : 1.f;
            beta = c < a ? 1.f : 0.f;
        }
        V3<float> gc{0.f, 0.f, 0.f};
        if (beta != 0.f) gc = dual_grad(prim_node<Dual, KM>(nodes, ni, pd, accel));
        g = V3<float>{alpha * g.x + beta * gc.x, alpha * g.y + beta * gc.y, alpha * g.z + beta * gc.z};
        a = smooth ? csg_eval<float, KM>(SMCRT_SMOOTH_UNION, a, c, k, kq) : fminf(a, c);
    }
    return g;
}

// A region's closed-form gradient, when it is one sphere or box node without rotation: kind 1
// sphere, 2 box, 0 none; the node's translation (zero for the identity) and half sizes.  Staged
// in LDS per block so the per-lane reads wait on LDS, not on the wave's outstanding atomics.
struct alignas(16) GradDesc {
    int kind;
    float tx, ty, tz, hx, hy, hz, pad;
};

__host__ __device__ inline int grad_desc_count(int nreg) { return nreg < SMCRT_LDS_GRAD ? nreg : SMCRT_LDS_GRAD; }

template <uint32_t KM>
SMCRT_DEV GradDesc grad_desc(CDevNode* nodes, CDevRegion* regions, int r) {
    GradDesc g{0, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    const DevRegion R = ((const DevRegion*)regions)[r];
    if (R.count != 1) return g;
    const DevNode* n = (const DevNode*)nodes + R.first;
    const bool sphere = KM_HAS(KM, KM_PRIM(SMCRT_SPHERE)) && n->kind == SMCRT_SPHERE;
    const bool box = KM_HAS(KM, KM_PRIM(SMCRT_BOX)) && n->kind == SMCRT_BOX;
    if (!(sphere || box) || n->op != SMCRT_OP_PRIM || n->ident == 0) return g;
    g.kind = sphere ? 1 : 2;
    if (n->ident == 2) { g.tx = n->t[9]; g.ty = n->t[10]; g.tz = n->t[11]; }
    g.hx = n->p[0]; g.hy = n->p[1]; g.hz = n->p[2];
    return g;
}

// The gradient of each lane's region.  Regions with a descriptor (one sphere or box node without
// rotation) are done for every lane at once; the others in a waterfall over the distinct regions
// of the remaining lanes.
template <uint32_t KM>
SMCRT_DEV V3<float> lane_region_gradient(CDevNode* nodes, CDevRegion* regions, const GradDesc* gd, int n_gd,
                                         bool active, int r_lane, const V3<float>& p, const float* __restrict__ accel) {
    V3<float> res{0.f, 0.f, 0.f};
    bool done = false;
    if constexpr (KM_HAS(KM, KM_PRIM(SMCRT_SPHERE)) || KM_HAS(KM, KM_PRIM(SMCRT_BOX))) {
        if (active && r_lane < n_gd) {
            const GradDesc g = gd[r_lane];
            if (g.kind != 0) {
                const V3<float> q{p.x + g.tx, p.y + g.ty, p.z + g.tz};
                res = closed_form_gradient(g.kind == 1, g.hx, g.hy, g.hz, q);
                done = true;
            }
        }
    }
    uint64_t pending = wave_ballot(active && !done);
    while (pending) {
        const int leader = __builtin_ctzll(pending);
        const int r = __builtin_amdgcn_readlane(r_lane, leader);
        const bool mine = active && !done && r_lane == r;
        const int first = __builtin_amdgcn_readfirstlane(regions[r].first);
        const int count = __builtin_amdgcn_readfirstlane(regions[r].count);
        const int flat = __builtin_amdgcn_readfirstlane(regions[r].flat);
        if (mine) res = region_gradient<KM>(nodes, first, count, flat, p, accel);
        pending &= ~wave_ballot(mine);
    }
    return res;
}

// LDS layout for the adjoint: per-lane scores float [4][nmat][blockDim], then the block's
// gradient sums double [4][nmat] (LDS FP64 atomics when a packet ends)
struct AdjLds {
    float* sc;
    double* g;
    int stride;  // blockDim
};

SMCRT_DEV float& adj_sc(const AdjLds& A, int which, int m, int nmat) {
    return A.sc[(which * nmat + m) * A.stride + threadIdx.x];
}
SMCRT_DEV void adj_g_add(const AdjLds& A, int which, int m, int nmat, double v) {
    atomicAdd(A.g + which * nmat + m, v);
}

// Offer an ended history to every detector: hit point = (final position, direction,
// value = sqrt(x^2 + y^2), layer), src/kernelsMod.f90:302-306; tests and binning of
// src/detectors/detectors.f90 and src/detectors/detector_base.f90:106-162.
// 1-D bin (0-based): min(nint(value / width) + 1, bins), nint rounding half away from zero;
// -1 (not counted) where the reference's index would leave the array (negative or NaN)
SMCRT_DEV long long bin1d(double value, double width, int bins) {
    const double q = value / width;
    if (!(q >= 0.0)) return -1;
    const long long i = q < 9.0e18 ? llround(q) + 1 : (long long)bins;
    return (i < bins ? i : bins) - 1;
}

// camera axis bin (1-based): min(int(q) + 1, bins), wrapped to `bins` below 1; 0 for NaN
SMCRT_DEV long long cam_bin(double q, int bins) {
    if (q != q) return 0;
    long long i = q >= 9.0e18 ? (long long)bins : (q <= -9.0e18 ? 0 : (long long)q + 1);
    i = i < bins ? i : bins;
    return i < 1 ? bins : i;
}

SMCRT_DEV void record_detectors(CKCold* C, const Lane& L) {
    const double px = L.px, py = L.py, pz = L.pz, dx = L.dx, dy = L.dy, dz = L.dz;
    // hit value: sqrt(x^2 + y^2) (pathlength_scatter, src/kernelsMod.f90:303) or the packet
    // weight (weight_scatter, :142)
    const double value = (C->flags & 16u) ? (double)L.weight : sqrt(px * px + py * py);
    for (int i = 0; i < C->n_det; ++i) {
        CDevDet* D = (CDevDet*)C->det + i;
        if (L.layer != D->layer0) continue;
        long long bin = -1;
        if (D->kind == 1) {  // circle: intersectCircle then t <= 5e-3
            const double den = D->n[0] * dx + D->n[1] * dy + D->n[2] * dz;
            if (den > 1e-6) {
                const double t = ((D->p[0] - px) * D->n[0] + (D->p[1] - py) * D->n[1] + (D->p[2] - pz) * D->n[2]) / den;
                if (t >= 0.0) {
                    const double vx = px + dx * t - D->p[0], vy = py + dy * t - D->p[1], vz = pz + dz * t - D->p[2];
                    if (sqrt(vx * vx + vy * vy + vz * vz) <= D->a && !(t > 5e-3))
                        bin = bin1d(value, D->bw, D->nb);
                }
            }
        } else if (D->kind == 2) {  // annulus
            const double vx = px - D->p[0], vy = py - D->p[1], vz = pz - D->p[2];
            const double r = sqrt(vx * vx + vy * vy + vz * vz);
            if (r >= D->a && r <= D->b) bin = bin1d(value, D->bw, D->nb);
        } else {  // camera
            const double t = ((D->p[0] - px) * D->n[0] + (D->p[1] - py) * D->n[1] + (D->p[2] - pz) * D->n[2]) /
                             (dx * D->n[0] + dy * D->n[1] + dz * D->n[2]);
            if (t >= 0.0) {
                const double vx = px + t * dx - D->p[0], vy = py + t * dy - D->p[1], vz = pz + t * dz - D->p[2];
                const double q1 = (vx * D->e1[0] + vy * D->e1[1] + vz * D->e1[2]) / D->a;
                const double q2 = (vx * D->e2[0] + vy * D->e2[1] + vz * D->e2[2]) / D->b;
                if (q1 < D->a && q1 > 0.0 && q2 < D->b && q2 > 0.0) {
                    // record_hit_2D_sub bins (final z + p1.x, final y + p1.y)
                    const long long ix = cam_bin((pz + D->p[0]) / D->bw, D->nb);
                    const long long iy = cam_bin((py + D->p[1]) / D->bwy, D->nb);
                    if (ix > 0 && iy > 0) bin = (ix - 1) + (long long)D->nb * (iy - 1);
                }
            }
        }
        if (bin >= 0) atomicAdd(C->det_counts + D->off + bin, 1ull);
    }
}

SMCRT_DEV int classify_escape(CKCold* C, const Lane& L) {
    return (L.dx * C->rdir[0] + L.dy * C->rdir[1] + L.dz * C->rdir[2] < 0.f) ? END_R : END_T;
}

// Set up the voxel walk of a step of length d from the lane's position.  Returns false when the
// start lies outside the grid (the packet ends there, src/inttau2.f90:256-259).
// Global tally atomics, staged per wavefront.  A global atomic instruction costs the wavefront
// about the same whatever its number of active lanes, and a walk step hands only a few lanes'
// chords to the grid (the rest merge or land in the LDS tile), so lanes append (voxel, length)
// pairs to the wavefront's LDS queue and every 64 entries one atomic instruction with all 64
// lanes active drains them.  push/drain run in wave-converged code only (`n` is wave-uniform);
// between calls the queue holds < 64 entries, so 128 entries per wavefront suffice.
constexpr int TALLY_QUEUE = 128;

// A block's binned-tally stage in LDS (KParams: nbins, nhot, ...), at `c`: per hot bin (slot)
// cnt[slot] reserves places in the slot's buffer of SMCRT_BIN_B entries and wr[slot] counts the
// entries written.  Every writer counts its entry after writing it (LDS executes a wavefront's
// instructions in order), so the lane whose count completes the buffer knows it is full: it
// flushes it and resets wr, then cnt; no lane waits on another wavefront.  Entries that find the
// buffer full, or whose bin is cold, take an atomic.  pg[slot], fill[slot]: the slot's current
// page and its entries (only the slot's flusher touches them, and flushes of one slot follow each
// other through the reset).  Layout after the bin map:
//   cnt[nhot], wr[nhot], pg[nhot], fill[nhot], used (the block's page counter) + 3 pad,
//   scratch[128] per wavefront, buf[nhot][SMCRT_BIN_B]
template <bool BIN>
struct TallyQueue {
    uint2* e;     // this wavefront's entries: (grid element index, FP32 length bits)
    int n;        // entries held (wave-uniform)
    uint32_t* c;  // BIN: the stage's counters
};

SMCRT_DEV void lds_order() { __asm__ volatile("" ::: "memory"); }

// Hand lanes [0, n) of the queue to the binned stage (converged code: every lane present).
// Returns the entries that will not take an atomic (staged, less those the block had no page for).
SMCRT_DEV uint32_t bin_stage_drain(TallyQueue<true>& Q, const KParams& P, double* __restrict__ base, int lane, int n) {
    const int nh = P.nhot;
    uint32_t* const cnt = Q.c;
    uint32_t* const wr = cnt + nh;
    uint32_t* const pgs = cnt + 2 * nh;
    uint32_t* const fills = cnt + 3 * nh;
    uint32_t* const used = cnt + 4 * nh;
    uint32_t* const scr = used + 4 + (threadIdx.x / 64) * 128;
    uint2* const buf = reinterpret_cast<uint2*>(used + 4 + (blockDim.x / 64) * 128);
    const uint16_t* const map = reinterpret_cast<const uint16_t*>(reinterpret_cast<unsigned char*>(cnt) - bins_map_bytes(P.nbins));
    bool flusher = false, staged = false;
    uint32_t info = 0u;  // slot | bin << 16 of a full buffer
    if (lane < n) {
        const uint2 e = Q.e[lane];
        const uint32_t bin = e.x >> SMCRT_BIN_SHIFT;
        const uint32_t slot = bin < (uint32_t)P.nbins ? (uint32_t)map[bin] : SMCRT_BIN_COLD;
        bool direct = true;
        if (slot != SMCRT_BIN_COLD) {
            const uint32_t s = __hip_atomic_fetch_add(cnt + slot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
            if (s < SMCRT_BIN_B) {
                buf[slot * SMCRT_BIN_B + s] = e;
                lds_order();
                const uint32_t w = __hip_atomic_fetch_add(wr + slot, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                flusher = w == SMCRT_BIN_B - 1;
                info = slot | (bin << 16);
                direct = false;
                staged = true;
            }
        }
        if (direct) unsafeAtomicAdd(base + e.x, (double)__uint_as_float(e.y));
    }
    uint32_t kept = (uint32_t)__builtin_popcountll(__ballot(staged));
    const uint64_t fl = __ballot(flusher);
    if (fl == 0ull) return kept;
    // the full buffers, listed in the wavefront's scratch, eight at a time (lane = 8 * buffer +
    // entry): the first lane of a buffer's group advances its slot's page (a new one from the
    // block's range when it is full) and passes the place on through the scratch; the group
    // writes the buffer there with one 64-byte store
    const uint32_t nf = (uint32_t)__builtin_popcountll(fl);
    if (flusher) scr[__builtin_amdgcn_mbcnt_hi((uint32_t)(fl >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)fl, 0u))] = info;
    lds_order();
    const uint32_t pg0 = blockIdx.x * P.ppb;
    for (uint32_t g0 = 0; g0 < nf; g0 += 64 / SMCRT_BIN_B) {
        const uint32_t g = g0 + (uint32_t)lane / SMCRT_BIN_B, j = (uint32_t)lane % SMCRT_BIN_B;
        const bool mine = g < nf;
        const uint32_t inf = mine ? scr[g] : 0u;
        const uint32_t kf = inf & 0xFFFFu;
        if (mine && j == 0) {
            uint32_t pg = pgs[kf], off = fills[kf];
            if (pg == SMCRT_PAGE_NONE || off == SMCRT_BIN_PAGE) {
                if (pg != SMCRT_PAGE_NONE) P.page_fill[pg] = SMCRT_BIN_PAGE;  // full
                const uint32_t np = __hip_atomic_fetch_add(used, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
                pg = np < P.ppb ? pg0 + np : SMCRT_PAGE_NONE;  // none left: the entries take atomics
                off = 0u;
                if (pg != SMCRT_PAGE_NONE) P.page_bin[pg] = inf >> 16;
            }
            pgs[kf] = pg;
            fills[kf] = pg != SMCRT_PAGE_NONE ? off + SMCRT_BIN_B : 0u;
            scr[64 + g % 64] = pg;
        }
        lds_order();
        const uint32_t pg = mine ? scr[64 + g % 64] : SMCRT_PAGE_NONE;
        const uint32_t off = mine && pg != SMCRT_PAGE_NONE ? fills[kf] - SMCRT_BIN_B : 0u;
        const uint2 v = mine ? buf[kf * SMCRT_BIN_B + j] : make_uint2(0u, 0u);
        if (mine) {
            if (pg != SMCRT_PAGE_NONE) P.pool[(size_t)pg * SMCRT_BIN_PAGE + off + j] = v;
            else unsafeAtomicAdd(base + v.x, (double)__uint_as_float(v.y));
        }
        kept -= (uint32_t)__builtin_popcountll(__ballot(mine && pg == SMCRT_PAGE_NONE));
        lds_order();  // the buffer is read (in order before these writes): release it
        if (mine && j == 0) {
            wr[kf] = 0u;
            lds_order();
            cnt[kf] = 0u;
        }
    }
    return kept;
}

// Returns the drained entries that did not take an atomic (binned stage).
template <bool BIN>
SMCRT_DEV uint32_t tally_drain(TallyQueue<BIN>& Q, const KParams& P, double* __restrict__ base, int lane, bool all) {
    uint32_t kept = 0u;
    while (Q.n >= 64 || (all && Q.n > 0)) {
        const int n = Q.n < 64 ? Q.n : 64;
        if constexpr (BIN) {
            kept += bin_stage_drain(Q, P, base, lane, n);
        } else if (lane < n) {
            const uint2 e = Q.e[lane];
            unsafeAtomicAdd(base + e.x, (double)__uint_as_float(e.y));
        }
        const int rem = Q.n - n;
        if (lane < rem) {
            const uint2 e = Q.e[n + lane];
            Q.e[lane] = e;
        }
        Q.n = rem;
    }
    return kept;
}

// Returns the number of entries appended (wave-uniform), less those a drain staged (the count of
// FP64 tally atomics, up to entries still queued).
template <bool BIN>
SMCRT_DEV unsigned tally_push(TallyQueue<BIN>& Q, const KParams& P, bool want, uint32_t idx, float v,
                              double* __restrict__ base, int lane) {
    const uint64_t m = __ballot(want);
    if (m == 0ull) return 0u;
    if (want) {
        const unsigned rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
        Q.e[Q.n + (int)rank] = make_uint2(idx, __float_as_uint(v));
    }
    const unsigned cnt = (unsigned)__builtin_popcountll(m);
    Q.n += (int)cnt;
    if (Q.n >= 64) return cnt - tally_drain(Q, P, base, lane, false);
    return cnt;
}

SMCRT_DEV bool walk_begin(const KParams& P, const Lane& L, Walk& W, float d) {
    const float gx = (L.px + P.xmax) * P.sx, gy = (L.py + P.ymax) * P.sy, gz = (L.pz + P.zmax) * P.sz;
    W.ix = (int)floorf(gx);
    W.iy = (int)floorf(gy);
    W.iz = (int)floorf(gz);
    W.t = 0.f;
    W.d = d;
    W.aint = 0.f;
    if ((unsigned)W.ix >= (unsigned)P.nx || (unsigned)W.iy >= (unsigned)P.ny || (unsigned)W.iz >= (unsigned)P.nz)
        return false;
    const float INF = __builtin_inff();
    const float adx = fabsf(L.dx) * P.sx, ady = fabsf(L.dy) * P.sy, adz = fabsf(L.dz) * P.sz;
    W.tmx = adx > 0.f ? (L.dx > 0.f ? (float)(W.ix + 1) - gx : gx - (float)W.ix) * f_rcp(adx) : INF;
    W.tmy = ady > 0.f ? (L.dy > 0.f ? (float)(W.iy + 1) - gy : gy - (float)W.iy) * f_rcp(ady) : INF;
    W.tmz = adz > 0.f ? (L.dz > 0.f ? (float)(W.iz + 1) - gz : gz - (float)W.iz) * f_rcp(adz) : INF;
    W.idx = W.ix + P.nx * (W.iy + P.ny * W.iz);
    return true;
}

// How the sweep evaluates the scene's regions.  The library's kernels interpret the SDF program
// (region_eval); a kernel compiled at run time for one scene passes a policy whose region<KM>(r,
// p, accel, h) is that region's program as straight-line code (constants as immediates, no node
// loads or opcode dispatch, independent regions interleaved by the compiler; h: the lane's
// gathered polygon, from poly_hit<KM>(p, accel) when kGather).
struct InterpretedSweep {
    static constexpr bool kCompiled = false;
    static constexpr int kRegions = 0;
    static constexpr bool kDiagnostics = true;  // event trace, scatter moments, end-early (tests)
    static constexpr bool kDetectors = true;    // detector records and the ghost continuation
    static constexpr bool kWeight = true;       // the weight_scatter estimator (SMCRT_FLAG_WEIGHT)
    static constexpr int kSource = -1;          // source kind (-1: read at run time)
    static constexpr bool kSets = true;         // launches may hold several sets (batches)
    static constexpr bool kGather = false;      // gathered polygon evaluation (PolyHit)
    // a compiled policy overwrites the scene's fixed launch parameters (grid, epsilons, region
    // count) with constants
    static SMCRT_DEV void bake(KParams&) {}
    template <uint32_t KM>
    static SMCRT_DEV float region(int, const V3<float>&, const float* __restrict__, const PolyHit&) { return 0.f; }
};

// Adds the outcome counts of the lanes active here (divergent code: the lanes whose set changes) to
// their old set's counters (stats[set * NOUT + i]) and clears them.  For each distinct set among
// them, each counter is summed over its lanes in scalar registers (readlane) and added by one
// atomic: atomics per lane on the same few counters serialise in L2, and the lanes cross each set
// boundary of a batch at nearly the same time.
SMCRT_DEV void flush_outcomes(unsigned long long* stats, int set, unsigned (&out)[NOUT], int lane) {
    uint64_t act = __builtin_amdgcn_read_exec();
    while (act) {
        const int lead = __builtin_ctzll(act);
        const int s0 = __builtin_amdgcn_readlane(set, lead);
        const uint64_t grp = wave_ballot(set == s0) & act;
        act &= ~grp;
#pragma unroll
        for (int i = 0; i < NOUT; ++i) {
            unsigned long long sum = 0;
            for (uint64_t g = grp; g; g &= g - 1)
                sum += (unsigned)__builtin_amdgcn_readlane((int)out[i], __builtin_ctzll(g));
            if (lane == lead && sum) atomicAdd(stats + (size_t)s0 * NOUT + i, sum);
        }
    }
#pragma unroll
    for (int i = 0; i < NOUT; ++i) out[i] = 0u;
}

// The lanes' outcome counts at the end of the launch: for each distinct set among the wavefront's
// lanes, each counter summed over that set's lanes (shuffles) and added by one lane.  Call
// converged.
SMCRT_DEV void flush_outcomes_wave(unsigned long long* stats, int set, const unsigned (&out)[NOUT], int lane) {
    uint64_t todo = ~0ull;
    while (todo) {
        const int lead = __builtin_ctzll(todo);
        const int s0 = __shfl(set, lead);
        const bool mine = set == s0;
        todo &= ~wave_ballot(mine);
        for (int i = 0; i < NOUT; ++i) {
            unsigned long long v = mine ? (unsigned long long)out[i] : 0ull;
            for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o);
            if (lane == lead && v) atomicAdd(stats + (size_t)s0 * NOUT + i, v);
        }
    }
}

// WAVES: resident waves per SIMD the registers are allocated for (2: at most 256 VGPRs, 3: at most
// 168).  SW: the sweep policy.  BIN: the binned tally's stage (forward single-set launches of
// the kernels compiled at run time; KParams).
template <bool ADJ, uint32_t KM, int WAVES = 2, class SW = InterpretedSweep, bool BIN = false>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(WAVES))) packet_kernel(KParams P) {
    extern __shared__ __attribute__((aligned(16))) unsigned char smem[];
    constexpr bool kDiag = SW::kDiagnostics, kDet = SW::kDetectors, kWeight = SW::kWeight;
    SW::bake(P);
    CDevNode* const nodes = (CDevNode*)P.nodes;
    CKCold* const C = (CKCold*)P.cold;
    CDevRegion* const regions = (CDevRegion*)P.regions;

    // region-indexed material tables of every set, staged in LDS
    DevMat* const s_mat = reinterpret_cast<DevMat*>(smem);
    const int n_mats = P.nsets * P.nreg;
    for (int i = threadIdx.x; i < n_mats; i += blockDim.x) s_mat[i] = P.mats[i];
    // the tally tile follows the materials (FP64 sums: the LDS adds round like FP32 register adds,
    // and an FP32 sum of a block's ~1e4-1e5 chords into a hot voxel drifts by ~1e-5 relative)
    // the regions' gradient descriptors follow the materials
    GradDesc* const s_gd = reinterpret_cast<GradDesc*>(smem + sizeof(DevMat) * n_mats);
    const int n_gd = grad_desc_count(P.nreg);
    for (int r = threadIdx.x; r < n_gd; r += blockDim.x) s_gd[r] = grad_desc<KM>(nodes, regions, r);
    const size_t lds_head = sizeof(DevMat) * n_mats + sizeof(GradDesc) * n_gd;
    double* const s_tile = reinterpret_cast<double*>(smem + lds_head);
    const int tile_vox = P.tn[0] * P.tn[1] * P.tn[2];
    for (int i = threadIdx.x; i < tile_vox; i += blockDim.x) s_tile[i] = 0.0;
    // this block's tally target (its XCD's partial grid; workgroups go to XCDs round-robin)
    double* const tbase = P.xgrid ? P.xgrid + (size_t)(blockIdx.x % (unsigned)P.xcopies) * P.xstride : P.jmean;
    // each wavefront's tally queue comes last (after the tile and the adjoint's scores)
    TallyQueue<BIN> Q;
    Q.n = 0;
    Q.c = nullptr;
    {
        size_t qoff = lds_head + sizeof(double) * ((tile_vox + 1) & ~1);
        if (ADJ) qoff += sizeof(float) * 4 * P.nmat * blockDim.x + sizeof(double) * 4 * P.nmat;
        Q.e = reinterpret_cast<uint2*>(smem + qoff) + (threadIdx.x / 64) * TALLY_QUEUE;
        // the binned-tally stage follows the queues: the bin map, then the counters (zeroed)
        if constexpr (BIN) {
            unsigned char* sb = smem + qoff + sizeof(uint2) * TALLY_QUEUE * (blockDim.x / 64);
            uint16_t* map = reinterpret_cast<uint16_t*>(sb);
            for (int i = threadIdx.x; i < P.nbins; i += blockDim.x) map[i] = P.binmap[i];
            Q.c = reinterpret_cast<uint32_t*>(sb + bins_map_bytes(P.nbins));
            const int nh = P.nhot;
            for (int i = threadIdx.x; i < 4 * nh + 4; i += blockDim.x) Q.c[i] = (i >= 2 * nh && i < 3 * nh) ? SMCRT_PAGE_NONE : 0u;
        }
    }
    AdjLds A;
    if (ADJ) {
        A.stride = blockDim.x;
        unsigned char* base = smem + lds_head + sizeof(double) * ((tile_vox + 1) & ~1);
        A.sc = reinterpret_cast<float*>(base);
        A.g = reinterpret_cast<double*>(base + sizeof(float) * 4 * P.nmat * blockDim.x);
        for (int i = threadIdx.x; i < 4 * P.nmat; i += blockDim.x) A.g[i] = 0.0;
    }
    __syncthreads();
    float adjC = 0.f;
    double adjW = 0.0;

    Lane L;
    L.st = ST_IDLE;
    auto trace_ev = [&](int kind) {
        if constexpr (kDiag) trace_event(C, L, kind);
    };
    L.set = 0;
    L.layer = 0;
    L.tcount = 0;
    L.iface = false;
    L.dprev = 0.f;
    L.iters = 0;
    Walk W;
    W.active = false;
    int pend_idx = -1;  // merged chord waiting for its atomic (grid index)
    int pend_li = -1;   // its tile index, -1 outside the tile
    float pend_len = 0.f;
    // Outcome counters: per lane where a lane's set can differ from its neighbours' (batched
    // launches), wave-uniform ballot counts in single-set compiled kernels.
    constexpr bool kWaveOut = SW::kCompiled && !SW::kSets;
    unsigned out[NOUT];
#pragma unroll
    for (int i = 0; i < NOUT; ++i) out[i] = 0;
    unsigned ev[NEV];  // event counters: wave-uniform (ballot counts at converged points)
#pragma unroll
    for (int i = 0; i < NEV; ++i) ev[i] = 0;

    unsigned long long base = 0;
    unsigned avail = 0;
    bool exhausted = false;
    // the sets' queue positions and first packet ids, read-only here: scalar loads where the index
    // is wave-uniform
    typedef __attribute__((address_space(4))) const unsigned long long CU64;
    CU64* const sprefix = (CU64*)C->set_prefix;
    CU64* const sfirst = (CU64*)C->set_first;
    // one set: packet id = first id + queue position (read once, scalar)
    const unsigned long long first0 = sfirst[0] - sprefix[0];
    // several sets: the wavefront's current set cs holds queue positions [cs_lo, cs_hi) from packet
    // id cs_first on; it moves (a scalar binary search) when a claimed chunk starts past it, and
    // only lanes of a chunk that straddles a set boundary search on their own (vector loads, which
    // wait for the wavefront's tally atomics in flight: once per refill they cost batches a quarter
    // of their time)
    int cs = 0;
    unsigned long long cs_lo = 0, cs_hi = 0, cs_first = 0;
    const float eps = P.eps;
    const float eps_kill = P.eps_kill;
    const unsigned chunk = (unsigned)P.chunk;
    const int cap = P.dda_cap;
    const int lane_id = __lane_id();

    for (;;) {
        // ---- regeneration: idle lanes take the next packet ids of the wavefront's chunk ----
        if (!exhausted) {
            uint64_t need = wave_ballot(L.st == ST_IDLE);
            // regeneration policy: wait until enough lanes are idle (or the whole wave is)
            if (need != ~0ull && __builtin_popcountll(need) < P.refill_min) need = 0;
            while (need) {
                if (avail == 0) {
                    unsigned long long b = 0;
                    if (lane_id == 0) b = atomicAdd(P.work, (unsigned long long)chunk);
                    b = __shfl(b, 0);
                    b = ((unsigned long long)(uint32_t)__builtin_amdgcn_readfirstlane((int)(b >> 32)) << 32) |
                        (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)b);  // wave-uniform
                    if (b >= P.total) { exhausted = true; break; }
                    if (P.nsets > 1 && b >= cs_hi) {
                        int lo = 0, hi = P.nsets - 1;
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (sprefix[mid] <= b) lo = mid; else hi = mid - 1;
                        }
                        cs = lo;
                        cs_lo = sprefix[lo];
                        cs_hi = sprefix[lo + 1];
                        cs_first = sfirst[lo];
                    }
                    base = b;
                    unsigned long long rem = P.total - b;
                    avail = rem < chunk ? (unsigned)rem : chunk;
                }
                unsigned rank = __builtin_amdgcn_mbcnt_hi((uint32_t)(need >> 32),
                                                          __builtin_amdgcn_mbcnt_lo((uint32_t)need, 0u));
                unsigned cnt = (unsigned)__builtin_popcountll(need);
                if (L.st == ST_IDLE && rank < avail) {
                    unsigned long long w = base + rank;
                    int s = 0;
                    unsigned long long pid = first0 + w;
                    if (P.nsets > 1) {
                        if (w >= cs_lo && w < cs_hi) {
                            s = cs;
                            pid = cs_first + (w - cs_lo);
                        } else {
                            int lo = 0, hi = P.nsets - 1;
                            while (lo < hi) {
                                const int mid = (lo + hi + 1) >> 1;
                                if (sprefix[mid] <= w) lo = mid; else hi = mid - 1;
                            }
                            s = lo;
                            pid = sfirst[s] + (w - sprefix[s]);
                        }
                    }
                    // per-lane counters belong to the lane's set: flushed when it changes
                    if (s != L.set) flush_outcomes(C->stats, L.set, out, lane_id);
                    L.set = s;
                    L.pid = pid;
                    L.st = ST_EMIT;
                }
                unsigned taken = cnt < avail ? cnt : avail;
                base += taken;
                avail -= taken;
                need = wave_ballot(L.st == ST_IDLE);
            }
        }
        if (wave_ballot(L.st != ST_IDLE) == 0) {
            if (exhausted) break;
            continue;
        }

        const DevMat* mt = s_mat + L.set * P.nreg;
        const bool grid = P.jmean != nullptr;  // tallying (wave-uniform)
        const uint32_t gbase = (uint32_t)L.set * (uint32_t)P.grid_stride;

        // ---- emission ----
        const bool emitting = L.st == ST_EMIT;
        if constexpr (kWaveOut) out[O_LAUNCHED] += __builtin_popcountll(wave_ballot(emitting));
        const bool emit_now = emit_draws_position<SW::kSource>(C);  // wave-uniform
        if (emitting) {
            rng_init(L.rng);
            if (emit_now) {
                float u[4];
                draw_block(L, P, u);
                emit<SW::kSource>(C, L, u);
            } else {  // fixed start: the draws follow with the iteration's event block
                L.px = C->spos[0]; L.py = C->spos[1]; L.pz = C->spos[2];
            }
            L.rx = 0.f; L.ry = 0.f; L.rz = 0.f;
            L.bounces = 0;
            L.nsc = 0;
            L.tcount = 0;
            L.iters = 0;
            L.ghost = false;
            if constexpr (kWeight) L.weight = 1.f;
            if (ADJ) {
                adjC = 0.f;
                for (int w = 0; w < 4; ++w)
                    for (int m = 0; m < P.nmat; ++m) adj_sc(A, w, m, P.nmat) = 0.f;
            }
            trace_ev(1);
        }

        // ---- the sweep: every region at the lane's position (lanes not mid-walk) ----
        const bool sweeping = L.st != ST_IDLE && !W.active;
        float dmin = __builtin_inff(), smin = __builtin_inff(), dfirst = 0.f, dlay = 0.f, d0 = 0.f;
        float bestv = __builtin_inff();
        int fn = -1, best = 0;
        bool runaway = false;
        if (sweeping) {
            V3<float> p{L.px, L.py, L.pz};
            auto visit = [&](int r, float d) {
                dmin = fminf(dmin, fabsf(d));
                smin = fminf(smin, d);
                if (r == 0) d0 = d;
                if (fn < 0 && d < 0.f) { fn = r; dfirst = d; }
                if (r == L.layer) dlay = d;
                float c = d > 0.f ? 999.f : -d;  // positive distances count as -999
                if (c < bestv) { bestv = c; best = r; }
            };
            if constexpr (SW::kCompiled) {
                // the scene's program compiled into the kernel: straight-line code per region
                // (scenes of several polygons: each lane's own polygon evaluated first, gathered)
                PolyHit h{-1, 0.f};
                if constexpr (SW::kGather) h = SW::template poly_hit<KM>(p, P.accel);
#pragma unroll
                for (int r = 0; r < SW::kRegions; ++r) visit(r, SW::template region<KM>(r, p, P.accel, h));
            } else {
                for (int r = 0; r < P.nreg; ++r)
                    visit(r, region_eval<float, KM>(nodes, regions[r].first, regions[r].count, p, P.accel));
            }
            if (fn < 0) { fn = 0; dfirst = d0; }  // minloc of an all-zero array is index 1
            runaway = ++L.iters > (1u << 24);     // runaway guard: never expected
        }
        ev[E_SWEEPS] += __builtin_popcountll(wave_ballot(sweeping));

        if (sweeping && L.st == ST_EMIT) {
            L.layer = best;  // nearest containing surface (src/kernelsMod.f90:275-280)
            // (the first flight's optical depth was drawn with the emission block)
            L.taurun = 0.f;
            L.iface = false;
            L.st = ST_START;
        }

        int end_kind = runaway ? END_ERR : -1;  // >= 0: the packet ends this iteration
        ev[E_NAN] += __builtin_popcountll(wave_ballot(runaway));
        bool want_dstmp = false, want_normal = false, do_march = false, cross = false, fres = false;
        bool unreach = false;  // the reference's "This should not be reached!" branch
        int use_region = 0;
        const int st0 = (sweeping && !runaway) ? L.st : ST_IDLE;
        if (st0 == ST_START) {
            // FP32 distances are quantised (multiples of 2^-24 near |p| = 1), so d == eps ties
            // that never occur in the reference's FP64 are common here: a tie counts as inside
            // eps, which kills a reflected packet (pos = oldpos, d0 <= eps) as the reference does.
            // After an interface the threshold is the march epsilon (the crossing-step ratio
            // that kills grazing crossings and reflected packets); after emission or a scatter
            // it is the reference's own 1e-8 (src/inttau2.f90:58-61).
            if (L.iface ? (dmin <= eps) : (dmin < eps_kill)) end_kind = END_SURFACE;
            else do_march = true;
        } else if (st0 == ST_MARCHED) {
            // src/inttau2.f90:85-97
            if (smin >= 0.f) end_kind = classify_escape(C, L);
            else if (dmin > eps) do_march = true;
            else if (kDet && L.ghost) end_kind = classify_escape(C, L);  // outside the grid: ends at the surface
            else if (dmin <= L.dprev || dmin <= eps_kill) cross = true;  // converged onto a surface
            else do_march = true;  // within eps but moving away from the surface: keep marching
        } else if (st0 == ST_CROSSED) {
            // the sweep ran at the crossed position (src/inttau2.f90:117-162)
            const int nl = fn;
            const float n1 = mt[L.layer].n, n2 = mt[nl].n;
            if (n1 != n2) {
                fres = true;
                const float dl = dlay, dn = dfirst;
                if (dl < 0.f && dn < 0.f) {
                    use_region = fabsf(dn) < fabsf(dl) ? nl : L.layer;
                    want_normal = true;
                } else if (L.dlay_old < 0.f) {
                    if (nl == L.layer) {
                        use_region = L.layer;  // maxloc of two equal values -> the first
                        want_normal = true;
                    } else {
                        want_dstmp = true;  // needs region nl at the pre-crossing position
                    }
                } else if (dl > 0.f && dn < 0.f) {
                    use_region = L.layer;
                    want_normal = true;
                } else if (dl > 0.f && dn > 0.f) {
                    end_kind = classify_escape(C, L);
                } else {
                    end_kind = END_ERR;  // "This should not be reached!" (:156)
                    unreach = true;
                }
            } else {
                L.layer = nl;  // index-matched: switch region, keep tau (:182-184)
                L.iface = true;
                L.st = ST_START;
            }
        }
        ev[E_FRESNEL] += __builtin_popcountll(wave_ballot(fres));
        ev[E_UNREACH] += __builtin_popcountll(wave_ballot(unreach));

        if (cross) {
            // interface: step a little past the surface (src/inttau2.f90:104-116)
            L.opx = L.px; L.opy = L.py; L.opz = L.pz;
            L.dlay_old = dlay;
            L.fn_old = fn;
            L.dfn_old = dfirst;
            const float s = dmin + 2.f * eps;
            L.px += s * L.dx; L.py += s * L.dy; L.pz += s * L.dz;
            L.st = ST_CROSSED;
        }
        ev[E_IFACE] += __builtin_popcountll(wave_ballot(cross));

        // dstmp(new): region nl at the pre-crossing position.  The sweep there found its first
        // negative region fn_old: regions before it were >= 0 (only the sign of dstmp matters
        // for them), and fn_old's own distance was kept; other regions are evaluated again
        // (waterfall over distinct regions).
        if (wave_ballot(want_dstmp)) {
            float dt = fn < L.fn_old ? 0.f : L.dfn_old;
            const bool need = want_dstmp && fn > L.fn_old;
            if (const uint64_t mn = wave_ballot(need)) {
                ev[E_REGION] += __builtin_popcountll(mn);
                const float de = eval_lane_region<float, KM>(nodes, regions, need, fn, V3<float>{L.opx, L.opy, L.opz}, P.accel);
                if (need) dt = de;
            }
            if (want_dstmp) {
                const float dl = dlay, dn = dfirst;
                if (dt < 0.f) {
                    use_region = dt > L.dlay_old ? fn : L.layer;
                    want_normal = true;
                } else if (dl > 0.f && dn < 0.f) {
                    use_region = L.layer;
                    want_normal = true;
                } else if (dl > 0.f && dn > 0.f) {
                    end_kind = classify_escape(C, L);
                } else {
                    end_kind = END_ERR;  // "This should not be reached!" (:156)
                    unreach = true;
                }
            }
        }

        ev[E_UNREACH] += __builtin_popcountll(wave_ballot(unreach && want_dstmp));  // set by the dstmp branch
        // exact gradient of the normal's region (one forward-mode pass; flat folds lazily)
        V3<float> gxyz{0.f, 0.f, 0.f};
        if (wave_ballot(want_normal))
            gxyz = lane_region_gradient<KM>(nodes, regions, s_gd, n_gd, want_normal, use_region, V3<float>{L.px, L.py, L.pz},
                                            P.accel);

        // ---- the iteration's event block: emission (fixed-start emitters), Fresnel, interaction ----
        const bool emit_late = emitting && !emit_now && st0 == ST_START;
        const bool interact = st0 == ST_SCATTER;
        const bool blk = emit_late || want_normal || interact;
        bool scat = false;   // this lane's packet scattered
        bool wpush = false;  // weight_scatter deposit for the global tally
        uint32_t widx = 0u;
        float wlen = 0.f;
        if (wave_ballot(blk)) {
            float u[4] = {0.f, 0.f, 0.f, 0.f};
            if (blk) draw_block(L, P, u);
            if (emit_late) emit<SW::kSource>(C, L, u);
            // normal + Fresnel (src/inttau2.f90:163-181, src/surfaces.f90:14-84)
            if (const uint64_t m = wave_ballot(want_normal)) {
                ev[E_NORMALS] += __builtin_popcountll(m);
                if (want_normal) {
                    float inv = f_rsqrt(gxyz.x * gxyz.x + gxyz.y * gxyz.y + gxyz.z * gxyz.z);
                    float nx = gxyz.x * inv, ny = gxyz.y * inv, nz = gxyz.z * inv;
                    const int nl = fn;
                    const float n1 = mt[L.layer].n, n2 = mt[nl].n;
                    float R = fresnel(L.dx, L.dy, L.dz, nx, ny, nz, n1, n2);
                    bool rflag = u[0] <= R;  // reflect or refract
                    if constexpr (kWeight) L.weight *= R;  // on either branch (src/inttau2.f90:167)
                    if (rflag) {
                        float c = 2.f * (nx * L.dx + ny * L.dy + nz * L.dz);
                        L.dx -= c * nx; L.dy -= c * ny; L.dz -= c * nz;
                    } else {
                        float eta = f_div(n1, n2);
                        float c1 = nx * L.dx + ny * L.dy + nz * L.dz;
                        if (c1 < 0.f) c1 = -c1;
                        else { nx = -nx; ny = -ny; nz = -nz; }
                        float c2 = f_sqrt(fmaxf(1.f - eta * eta * (1.f - c1 * c1), 0.f));
                        float k = eta * c1 - c2;
                        L.dx = eta * L.dx + k * nx; L.dy = eta * L.dy + k * ny; L.dz = eta * L.dz + k * nz;
                    }
                    L.tau = f_neglog(u[1]);  // resampled optical depth (:168)
                    L.taurun = 0.f;
                    L.iface = true;
                    if (!rflag) {
                        L.layer = nl;
                        trace_ev(3);
                    } else {
                        L.px = L.opx; L.py = L.opy; L.pz = L.opz;  // step back inside (:174)
                        trace_ev(4);
                        if (++L.bounces > C->max_bounces) end_kind = END_BOUNCE;
                    }
                    if (end_kind < 0) L.st = ST_START;
                }
            }
            // interaction at the end of a flight inside the grid
            bool absorbed = false;  // ended here (the sweep at this point served nothing)
            if (interact) {
                const DevMat m = mt[L.layer];
                if (kWeight && (C->flags & 16u)) {
                    // weight_scatter (src/kernelsMod.f90:83-111): deposit the absorbed fraction in
                    // the interaction voxel (leaving the grid there ends the packet), scatter, then
                    // Russian roulette below THRESHOLD = 0.01 with CHANCE = 0.1 (src/constants.f90)
                    const float wa = L.weight * (1.f - m.albedo);
                    L.weight -= wa;
                    const int vx = (int)floorf((L.px + P.xmax) * P.sx), vy = (int)floorf((L.py + P.ymax) * P.sy),
                              vz = (int)floorf((L.pz + P.zmax) * P.sz);
                    if ((unsigned)vx >= (unsigned)P.nx || (unsigned)vy >= (unsigned)P.ny || (unsigned)vz >= (unsigned)P.nz) {
                        end_kind = classify_escape(C, L);
                    } else {
                        if (grid) {
                            wpush = true;
                            widx = gbase + (uint32_t)(vx + P.nx * (vy + P.ny * vz));
                            wlen = wa;
                        }
                        L.nsc++;
                        scat = true;
                        scatter(L, m.hgg, m.g2, u[1], u[2]);
                        trace_ev(2);
                        if (L.weight < 0.01f) {
                            if (u[0] < 0.1f) L.weight *= 10.f;  // roulette
                            else end_kind = END_ABSORB;
                        }
                    }
                } else if (u[0] < m.albedo) {
                    // albedo test (src/kernelsMod.f90:287-297)
                    if (ADJ && m.mat >= 0) {
                        adj_sc(A, 2, m.mat, P.nmat) += 1.f;
                        adj_sc(A, 3, m.mat, P.nmat) += adjC;
                    }
                    L.nsc++;
                    scat = true;
                    if (kDiag && (C->flags & 1u) && L.nsc <= 4) {
                        double* mo = C->moments + (L.nsc - 1) * 3;
                        atomicAdd(mo + 0, (double)L.px); atomicAdd(mo + 1, (double)L.py); atomicAdd(mo + 2, (double)L.pz);
                        atomicAdd(mo + 12, (double)L.px * L.px); atomicAdd(mo + 13, (double)L.py * L.py);
                        atomicAdd(mo + 14, (double)L.pz * L.pz);
                    }
                    scatter(L, m.hgg, m.g2, u[1], u[2]);
                    trace_ev(2);
                    if (kDiag && (C->flags & 2u) && L.nsc >= 5) end_kind = END_ABSORB;
                } else {
                    end_kind = END_ABSORB;
                }
                absorbed = end_kind >= 0;
                if (!absorbed) {
                    // the next flight starts here: its optical depth, then the flight-start test
                    // against the sweep already made at this point (src/inttau2.f90:58-61)
                    L.tau = f_neglog(u[3]);
                    L.taurun = 0.f;
                    L.iface = false;
                    L.st = ST_START;
                    if (dmin < eps_kill) end_kind = END_SURFACE;
                    else do_march = true;
                }
            }
            ev[E_SWEEPS] -= __builtin_popcountll(wave_ballot(absorbed));
        }
        if constexpr (kWaveOut) out[O_SCATTERS] += __builtin_popcountll(wave_ballot(scat));
        if (kWeight && P.jmean) ev[E_ATOMICS] += tally_push(Q, P, wpush, widx, wlen, tbase, lane_id);

        // ---- march step (src/inttau2.f90:64-83): optical depth, then the voxel walk ----
        ev[E_STEPS] += __builtin_popcountll(wave_ballot(do_march));
        if (do_march) {
            L.dprev = dmin;
            const float kap = mt[L.layer].kappa;
            const float t = dmin * kap;
            W.flight_over = !(L.taurun + t <= L.tau);
            float d = dmin;
            if (W.flight_over) d = f_div(L.tau - L.taurun, kap);
            else L.taurun += t;
            if (kDet && L.ghost) {  // outside the grid: move without a walk (nothing to tally)
                move(L, d);
                if (W.flight_over) end_kind = classify_escape(C, L);
                else L.st = ST_MARCHED;
            } else {
                W.inside = walk_begin(P, L, W, d);
                W.active = true;
            }
        }

        // ---- voxel walk: at most `cap` chords per lane this iteration ----
        // (the loop is wave-uniform so the tally queue can be filled between lane steps)
        bool walk_done = W.active && !W.inside;  // started outside the grid
        bool walking = W.active && W.inside;
        // per-axis step of the crossing distance (the direction is fixed over the walk)
        const float dtx = f_rcp(fabsf(L.dx) * P.sx), dty = f_rcp(fabsf(L.dy) * P.sy), dtz = f_rcp(fabsf(L.dz) * P.sz);
        const int sgx = L.dx > 0.f ? 1 : -1, sgy = L.dy > 0.f ? 1 : -1, sgz = L.dz > 0.f ? 1 : -1;
        // flat-index step per axis (x: sgx)
        const int ystep = L.dy > 0.f ? P.nx : -P.nx, zstep = L.dz > 0.f ? P.nx * P.ny : -(P.nx * P.ny);
        // one chord per lane and trip, written as selects rather than nested branches (lanes of a
        // wavefront differ in which axis they step and whether they merge, flush or leave)
        for (int k = 0; k < cap; ++k) {
            if (wave_ballot(walking) == 0ull) break;
            bool gpush = false;  // this lane hands its merged chord to the global tally
            bool crossed = false;
            uint32_t gidx = 0u;
            float glen = 0.f;
            if (walking) {
                const float tn = fminf(W.tmx, fminf(W.tmy, W.tmz));
                const float len = fminf(tn, W.d) - W.t;
                crossed = len > 0.f;
                if (crossed) {
                    if (ADJ) {
                        const float w = C->weights ? C->weights[W.idx] : 1.f;
                        W.aint += adjC * len + 0.5f * w * len * len;
                        adjC += w * len;
                    }
                }
                if (grid) {
                    // a chord in the pending voxel merges; one in another voxel hands the pending
                    // chord to the LDS tile or the global queue and becomes pending itself
                    const bool flush = crossed && W.idx != pend_idx;
                    if (flush && pend_li >= 0) atomicAdd(s_tile + pend_li, (double)pend_len);
                    gpush = flush && pend_li < 0 && pend_idx >= 0;
                    gidx = gbase + (uint32_t)pend_idx;
                    glen = pend_len;
                    const unsigned lx = (unsigned)(W.ix - P.t0[0]), ly = (unsigned)(W.iy - P.t0[1]),
                                   lz = (unsigned)(W.iz - P.t0[2]);
                    // (bitwise ands: no short-circuit branches; 24-bit products: the tile holds at
                    // most 8192 voxels)
                    const bool in_tile = (lx < (unsigned)P.tn[0]) & (ly < (unsigned)P.tn[1]) & (lz < (unsigned)P.tn[2]);
                    const int li = in_tile ? (int)(lx + __umul24((unsigned)P.tn[0], ly + __umul24((unsigned)P.tn[1], lz))) : -1;
                    pend_len = flush ? len : (crossed ? pend_len + len : pend_len);
                    pend_li = flush ? li : pend_li;
                    pend_idx = flush ? W.idx : pend_idx;
                }
                // next voxel boundary: the axis with the nearest crossing (x, then y, on ties); the
                // candidate indices of all three axes and selects, no branches
                const bool ax = (W.tmx <= W.tmy) & (W.tmx <= W.tmz);
                const bool ay = !ax & (W.tmy <= W.tmz);
                const bool az = !ax & !ay;
                const int jx = W.ix + (ax ? sgx : 0), jy = W.iy + (ay ? sgy : 0), jz = W.iz + (az ? sgz : 0);
                const bool out = ((unsigned)jx >= (unsigned)P.nx) | ((unsigned)jy >= (unsigned)P.ny) |
                                 ((unsigned)jz >= (unsigned)P.nz);
                const bool ends = tn >= W.d;     // the step ends inside this voxel
                const bool leaves = !ends & out;  // into the grid's outside
                const bool steps = !ends & !out;
                W.t = ends ? W.t : tn;
                W.inside = !leaves;
                W.ix = steps ? jx : W.ix;
                W.iy = steps ? jy : W.iy;
                W.iz = steps ? jz : W.iz;
                W.idx += steps ? (ax ? sgx : (ay ? ystep : zstep)) : 0;
                W.tmx = (steps & ax) ? W.tmx + dtx : W.tmx;
                W.tmy = (steps & ay) ? W.tmy + dty : W.tmy;
                W.tmz = (steps & az) ? W.tmz + dtz : W.tmz;
                walk_done = !steps;
                walking = steps;
            }
            ev[E_CROSS] += __builtin_popcountll(wave_ballot(crossed));
            if (P.jmean) ev[E_ATOMICS] += tally_push(Q, P, gpush, gidx, glen, tbase, lane_id);
        }

        // ---- step finished: move, then end of flight (interaction next iteration) or next march step ----
        if (walk_done) {
            W.active = false;
            const float d = W.d;
            if (ADJ && mt[L.layer].mat >= 0) {  // mat: adjoint slot of the material, -1 untracked
                const int m = mt[L.layer].mat;
                if (W.inside) {
                    adj_sc(A, 0, m, P.nmat) += d;
                    adj_sc(A, 1, m, P.nmat) += W.aint;
                } else {  // left the grid: the score stops at the exit point
                    adj_sc(A, 0, m, P.nmat) += W.t;
                    adj_sc(A, 1, m, P.nmat) += W.aint;
                }
            }
            move(L, d);
            if (!W.inside) {
                // left the grid (src/inttau2.f90:256-259,282-285): the outcome is fixed.  With
                // detectors the packet moves on, untallied, to the end of its flight or the next
                // surface, as the reference's loop does (:64-100), for its final position.
                if (kDet && C->n_det > 0 && !W.flight_over) {
                    L.ghost = true;
                    L.st = ST_MARCHED;
                } else {
                    end_kind = classify_escape(C, L);
                }
            } else {
                L.st = W.flight_over ? ST_SCATTER : ST_MARCHED;
            }
        }

        // non-finite guard
        const bool nonfinite = end_kind < 0 && L.st != ST_IDLE && !W.active &&
                               !(fabsf(L.px) < 1e30f && fabsf(L.py) < 1e30f && fabsf(L.pz) < 1e30f);
        ev[E_NAN] += __builtin_popcountll(wave_ballot(nonfinite));
        if (nonfinite) end_kind = END_ERR;
        if constexpr (kWaveOut) {
            out[O_ABSORBED] += __builtin_popcountll(wave_ballot(end_kind == END_ABSORB));
            out[O_R] += __builtin_popcountll(wave_ballot(end_kind == END_R));
            out[O_T] += __builtin_popcountll(wave_ballot(end_kind == END_T));
            out[O_KSURF] += __builtin_popcountll(wave_ballot(end_kind == END_SURFACE));
            out[O_KBOUNCE] += __builtin_popcountll(wave_ballot(end_kind == END_BOUNCE));
        }
        bool epush = false;  // the packet's last merged chord for the global tally
        uint32_t eidx = 0u;
        float elen = 0.f;
        if (end_kind >= 0) {
            // flush the merged chord before the lane changes packet/set
            if (pend_li >= 0) atomicAdd(s_tile + pend_li, (double)pend_len);
            epush = pend_li < 0 && pend_idx >= 0;
            eidx = gbase + (uint32_t)pend_idx;
            elen = pend_len;
            pend_idx = -1;
            pend_li = -1;
            if (kDet && C->n_det > 0 && end_kind != END_ERR) record_detectors(C, L);
            // per-lane outcome counters, once per packet (its launch and scatters too: counting
            // in the event loop, per lane, costs the sets kernels a third of their time)
            if constexpr (!kWaveOut) {
                out[O_LAUNCHED] += 1u;
                out[O_SCATTERS] += (unsigned)L.nsc;
                out[O_ABSORBED] += end_kind == END_ABSORB ? 1u : 0u;
                out[O_R] += end_kind == END_R ? 1u : 0u;
                out[O_T] += end_kind == END_T ? 1u : 0u;
                out[O_KSURF] += end_kind == END_SURFACE ? 1u : 0u;
                out[O_KBOUNCE] += end_kind == END_BOUNCE ? 1u : 0u;
            }
            if constexpr (kDiag) {
                switch (end_kind) {
                case END_ABSORB: trace_ev(5); break;
                case END_R: case END_T: trace_ev(6); break;
                case END_SURFACE: case END_BOUNCE: trace_ev(7); break;
                default: break;
                }
            }
            if (ADJ) {
                const bool isr = end_kind == END_R;
                adjW += (double)adjC;
                for (int m = 0; m < P.nmat; ++m) {
                    float Lm = adj_sc(A, 0, m, P.nmat), Am = adj_sc(A, 1, m, P.nmat);
                    float Km = adj_sc(A, 2, m, P.nmat), Bm = adj_sc(A, 3, m, P.nmat);
                    if (Lm == 0.f && Am == 0.f && Km == 0.f && Bm == 0.f) continue;  // material not visited
                    float mus = C->mat_mus[m];
                    float scs = mus > 0.f ? Km / mus : 0.f;
                    if (isr) {
                        adj_g_add(A, 0, m, P.nmat, (double)(-Lm));
                        adj_g_add(A, 1, m, P.nmat, (double)(scs - Lm));
                    }
                    float lw = Lm * adjC - Am;
                    float bw = mus > 0.f ? (Km * adjC - Bm) / mus : 0.f;
                    adj_g_add(A, 2, m, P.nmat, (double)(-lw));
                    adj_g_add(A, 3, m, P.nmat, (double)(bw - lw));
                }
            }
            L.st = ST_IDLE;
        }
        if (P.jmean) ev[E_ATOMICS] += tally_push(Q, P, epush, eidx, elen, tbase, lane_id);
    }
    if (P.jmean) ev[E_ATOMICS] -= tally_drain(Q, P, tbase, lane_id, true);

    // ---- flush the tally tile (single-set launches only) into the FP64 grid ----
    if (tile_vox > 0 || BIN) __syncthreads();
    if (tile_vox > 0) {
        for (int i = threadIdx.x; i < tile_vox; i += blockDim.x) {
            const double v = s_tile[i];
            if (v != 0.0) {
                const int lx = i % P.tn[0], r = i / P.tn[0];
                const int ly = r % P.tn[1], lz = r / P.tn[1];
                const size_t g = (size_t)(P.t0[0] + lx) + (size_t)P.nx * ((size_t)(P.t0[1] + ly) + (size_t)P.ny * (P.t0[2] + lz));
                unsafeAtomicAdd(tbase + g, v);
            }
        }
    }
    // ---- binned tally: entries left in the slots' buffers take atomics; the pages' fills ----
    if constexpr (BIN) {
        const int nh = P.nhot;
        const uint32_t* cnt = Q.c;
        const uint32_t* pgs = Q.c + 2 * nh;
        const uint32_t* fills = Q.c + 3 * nh;
        const uint32_t used = Q.c[4 * nh];
        const uint2* buf = reinterpret_cast<const uint2*>(Q.c + 4 * nh + 4 + (blockDim.x / 64) * 128);
        for (int k = threadIdx.x; k < nh; k += blockDim.x) {
            const uint32_t c = cnt[k] < SMCRT_BIN_B ? cnt[k] : SMCRT_BIN_B;
            for (uint32_t i = 0; i < c; ++i) {
                const uint2 v = buf[k * SMCRT_BIN_B + i];
                unsafeAtomicAdd(tbase + v.x, (double)__uint_as_float(v.y));
            }
            if (pgs[k] != SMCRT_PAGE_NONE) P.page_fill[pgs[k]] = fills[k];
        }
        if (threadIdx.x == 0) P.block_pages[blockIdx.x] = used < P.ppb ? used : P.ppb;
    }

    // ---- flush counters ----
    if constexpr (kWaveOut) {
#pragma unroll
        for (int i = 0; i < NOUT; ++i)
            if (out[i] && lane_id == 0) atomicAdd(C->stats + i, (unsigned long long)out[i]);
    } else {
        flush_outcomes_wave(C->stats, L.set, out, lane_id);
    }
    unsigned long long* evg = C->stats + (size_t)P.nsets * NOUT;
#pragma unroll
    for (int i = 0; i < NEV; ++i)
        if (lane_id == 0 && ev[i]) atomicAdd(evg + i, (unsigned long long)ev[i]);
    if (ADJ) {
        __syncthreads();  // every wave has left the packet loop: the block's sums are final
        for (int i = threadIdx.x; i < 4 * P.nmat; i += blockDim.x)
            if (A.g[i] != 0.0) atomicAdd(C->grad + i, A.g[i]);
        double v = adjW;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id == 0) atomicAdd(C->grad + 4 * P.nmat, v);
    }
}

}  // namespace smcrt
