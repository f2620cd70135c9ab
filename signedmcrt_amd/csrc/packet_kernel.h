// packet_kernel.h — the persistent, wavefront-regenerating packet kernel.
//
// One lane owns one packet at a time.  Every loop iteration each busy lane performs exactly one
// SDF sweep (all regions evaluated at its position) and then the state-machine step that the
// reference runs after that sweep inside `tauint2` (src/inttau2.f90:15-210):
//
//   EMIT     sweep at the emission point: layer = nearest containing surface
//            (src/kernelsMod.f90:275-280), draw tau, then as START
//   START    flight start (src/inttau2.f90:52-61): kill if within eps of a surface, else march
//   MARCHED  after a step (:85-97): escape if outside every region, march again while d > eps,
//            otherwise cross: step min|d| + 2 eps along dir (:104-116)
//   CROSSED  after the crossing step (:117-185): first-negative region rule, normal choice,
//            Fresnel reflect/refract at the SDF normal, tau resample, bounce cap
//
// A march step of length d consumes d*kappa of optical depth (:64-83) and walks the voxel grid
// with an integer DDA (Amanatides & Woo 1987) adding each chord to the FP64 fluence grid — the
// path-length estimator of `update_grids` (:213-292).  When the flight's optical depth runs out
// the albedo test absorbs or scatters (HG, src/photon.f90:713-779) and a new flight starts.
// A lane whose packet ends takes the next packet id from its wavefront's chunk at the top of
// the next iteration (regeneration); a wavefront claims a new chunk of ids from a global
// counter with one atomic.  Packet j draws from Philox stream (seed, first_packet + j), so
// results do not depend on grid size, block size or chunking.
#pragma once
#include "smcrt_device.h"

namespace smcrt {

enum : int { ST_IDLE = 0, ST_EMIT = 1, ST_START = 2, ST_MARCHED = 3, ST_CROSSED = 4 };
enum : int { END_ABSORB = 0, END_R = 1, END_T = 2, END_SURFACE = 3, END_BOUNCE = 4, END_ERR = 5 };

// stats layout (unsigned long long): per set [NOUT], then global event counters [NEV]
enum : int {
    O_LAUNCHED = 0, O_ABSORBED, O_R, O_T, O_KSURF, O_KBOUNCE, O_SCATTERS, NOUT
};
enum : int {
    E_SWEEPS = 0, E_REGION, E_NORMALS, E_STEPS, E_CROSS, E_IFACE, E_FRESNEL, E_UNREACH, E_NAN, NEV
};

struct KParams {
    const DevNode* nodes;
    const DevRegion* regions;
    const DevMat* mats;  // [nsets][nreg] region-indexed material constants
    int nreg, nmat;      // regions, adjoint materials
    // grid
    int nx, ny, nz;
    float xmax, ymax, zmax;
    float sx, sy, sz;  // voxels per cm
    double* jmean;     // set 0 grid (nullptr: no tally)
    unsigned long long grid_stride;
    // source
    int src_kind, circ_swap;
    float spos[3], sdir[3], p1[3], p2[3], p3[3], radius;
    float circ_t[12];
    float rdir[3];  // R/T classification direction
    float eps;       // march / crossing epsilon
    float eps_kill;  // flight-start kill threshold after emission or scatter
    // run
    uint32_t seed_lo, seed_hi;
    unsigned long long total;          // work items (sum of set budgets)
    int nsets;
    const unsigned long long* set_prefix;  // nsets+1 cumulative budgets
    const unsigned long long* set_first;   // global packet id of each set's first packet
    unsigned long long* work;              // work counter
    unsigned long long* stats;             // nsets*NOUT + NEV
    double* moments;                       // 24 doubles
    float* trace;
    int trace_packets, trace_events;
    uint32_t flags;
    int max_bounces, chunk;
    // adjoint
    const float* weights;  // nullptr: w = 1
    double* grad;          // [4][nmat] + [1] weighted tally
    const int* region_mat; // region -> adjoint material index
    const float* mat_mus;  // mus per adjoint material
};

struct Lane {
    float px, py, pz, dx, dy, dz;
    float opx, opy, opz;  // position before the crossing step
    float tau, taurun, dlay_old;
    int layer, st, bounces, nsc, set, tcount;
    float dprev;     // distance at the previous march step
    bool iface;      // the current flight started at an interface
    unsigned iters;
    unsigned long long pid;
    Rng rng;
};

SMCRT_DEV float rnd(Lane& L, const KParams& P) { return rng_u(L.rng, P.seed_lo, P.seed_hi); }

SMCRT_DEV uint64_t wave_ballot(bool b) { return __ballot(b); }

SMCRT_DEV void trace_event(const KParams& P, Lane& L, int kind) {
    if (L.pid < (unsigned long long)P.trace_packets && L.tcount < P.trace_events) {
        float* e = P.trace + ((size_t)L.pid * P.trace_events + L.tcount) * 4;
        e[0] = L.px; e[1] = L.py; e[2] = L.pz; e[3] = (float)kind;
        L.tcount++;
    }
}

// emitters (src/photon.f90): pencil :464-505, uniform :397-461, circular :214-287, point :290-338
SMCRT_DEV void emit(const KParams& P, Lane& L) {
    switch (P.src_kind) {
    case 2: {  // uniform: draws rx, ry
        float rx = rnd(L, P), ry = rnd(L, P);
        L.px = P.p1[0] + rx * P.p2[0] + ry * P.p3[0];
        L.py = P.p1[1] + rx * P.p2[1] + ry * P.p3[1];
        L.pz = P.p1[2] + rx * P.p2[2] + ry * P.p3[2];
        L.dx = P.sdir[0]; L.dy = P.sdir[1]; L.dz = P.sdir[2];
        break;
    }
    case 3: {  // circular: draws radius, angle; disc rotated onto dir, centred at -position
        float r = P.radius * f_sqrt(rnd(L, P));
        float th = rnd(L, P);
        float c = f_cos_rev(th), s = f_sin_rev(th);
        float qx, qy, qz;
        if (P.circ_swap) { qx = r * c; qy = r * s; qz = 0.f; }
        else { qx = 0.f; qy = r * c; qz = r * s; }
        const float* t = P.circ_t;
        L.px = qx * t[0] + qy * t[3] + qz * t[6] + t[9];
        L.py = qx * t[1] + qy * t[4] + qz * t[7] + t[10];
        L.pz = qx * t[2] + qy * t[5] + qz * t[8] + t[11];
        L.dx = P.sdir[0]; L.dy = P.sdir[1]; L.dz = P.sdir[2];
        break;
    }
    case 4: {  // point: draws azimuth, cos theta
        float ph = rnd(L, P);
        float ct = 2.f * rnd(L, P) - 1.f;
        float stt = f_sqrt(fmaxf(1.f - ct * ct, 0.f));
        L.px = P.spos[0]; L.py = P.spos[1]; L.pz = P.spos[2];
        L.dx = stt * f_cos_rev(ph); L.dy = stt * f_sin_rev(ph); L.dz = ct;
        break;
    }
    default: {  // pencil
        L.px = P.spos[0]; L.py = P.spos[1]; L.pz = P.spos[2];
        L.dx = P.sdir[0]; L.dy = P.sdir[1]; L.dz = P.sdir[2];
        break;
    }
    }
}

// HG / isotropic scatter (src/photon.f90:713-779).  Draws cos theta, then azimuth.  The
// near-axis branch keeps the reference's Fortran `sign(cost, nzp)` = |cos theta|*sign(nzp).
SMCRT_DEV void scatter(const KParams& P, Lane& L, float hgg, float g2) {
    float cost;
    if (hgg == 0.f) {
        cost = 2.f * rnd(L, P) - 1.f;
    } else {
        float temp = (1.f - g2) / (1.f - hgg + 2.f * hgg * rnd(L, P));
        cost = (1.f + g2 - temp * temp) / (2.f * hgg);
        cost = fminf(fmaxf(cost, -1.f), 1.f);
    }
    float sint = f_sqrt(fmaxf(1.f - cost * cost, 0.f));
    float ph = rnd(L, P);
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
    float sint2 = n1 / n2 * sintt;
    if (sint2 > 1.f) return 1.f;
    if (costt == 1.f) return 0.f;
    float cost2 = f_sqrt(fmaxf(1.f - sint2 * sint2, 0.f));
    float f1 = (n1 * costt - n2 * cost2) / (n1 * costt + n2 * cost2);
    float f2 = (n1 * cost2 - n2 * costt) / (n1 * cost2 + n2 * costt);
    return 0.5f * (f1 * f1 + f2 * f2);
}

// Evaluate, per lane, the region r_lane at p (only lanes with active set).  Waterfall over the
// distinct region ids in the wavefront so each evaluation runs the wave-uniform interpreter.
template <class T>
SMCRT_DEV T eval_lane_region(const KParams& P, bool active, int r_lane, const V3<T>& p) {
    T res = g_const<T>(0.f);
    uint64_t pending = wave_ballot(active);
    while (pending) {
        int leader = __builtin_ctzll(pending);
        int r = __builtin_amdgcn_readlane(r_lane, leader);
        bool mine = active && r_lane == r;
        if (mine) {
            const DevRegion reg = P.regions[r];
            res = region_eval<T>(P.nodes, reg.first, reg.count, p);
        }
        pending &= ~wave_ballot(mine);
    }
    return res;
}

template <bool ADJ>
struct AdjState {
    float C;      // weighted in-grid path so far
    double wsum;  // sum of C over ended packets
};

// LDS layout for the adjoint: scores float [4][nmat][blockDim], gradients double [4][nmat][blockDim]
struct AdjLds {
    float* sc;
    double* g;
    int stride;  // blockDim
};

SMCRT_DEV float& adj_sc(const AdjLds& A, int which, int m, int nmat) {
    return A.sc[(which * nmat + m) * A.stride + threadIdx.x];
}
SMCRT_DEV double& adj_g(const AdjLds& A, int which, int m, int nmat) {
    return A.g[(which * nmat + m) * A.stride + threadIdx.x];
}

// Walk the voxels along a step of length d from the lane's position and tally the chords.
// Returns false if the step leaves the grid (the packet then ends, src/inttau2.f90:256-285);
// *in_grid receives the tallied length.  For the adjoint, accumulates C and the step integral
// of C into *cacc / *aint.
template <bool ADJ>
SMCRT_DEV bool dda(const KParams& P, double* __restrict__ grid, const Lane& L, float d,
                   float* in_grid, unsigned& crossings, float* cacc, float* aint) {
    float gx = (L.px + P.xmax) * P.sx, gy = (L.py + P.ymax) * P.sy, gz = (L.pz + P.zmax) * P.sz;
    int ix = (int)floorf(gx), iy = (int)floorf(gy), iz = (int)floorf(gz);
    *in_grid = 0.f;
    if ((unsigned)ix >= (unsigned)P.nx || (unsigned)iy >= (unsigned)P.ny || (unsigned)iz >= (unsigned)P.nz)
        return false;
    const float INF = __builtin_inff();
    float adx = fabsf(L.dx) * P.sx, ady = fabsf(L.dy) * P.sy, adz = fabsf(L.dz) * P.sz;
    float tdx = adx > 0.f ? f_rcp(adx) : INF;
    float tdy = ady > 0.f ? f_rcp(ady) : INF;
    float tdz = adz > 0.f ? f_rcp(adz) : INF;
    float tmx = adx > 0.f ? (L.dx > 0.f ? (float)(ix + 1) - gx : gx - (float)ix) * tdx : INF;
    float tmy = ady > 0.f ? (L.dy > 0.f ? (float)(iy + 1) - gy : gy - (float)iy) * tdy : INF;
    float tmz = adz > 0.f ? (L.dz > 0.f ? (float)(iz + 1) - gz : gz - (float)iz) * tdz : INF;
    const int sxs = L.dx > 0.f ? 1 : -1, sys = L.dy > 0.f ? 1 : -1, szs = L.dz > 0.f ? 1 : -1;
    const int nxny = P.nx * P.ny;
    int idx = ix + P.nx * (iy + P.ny * iz);
    float t = 0.f;
    bool inside = true;
    for (;;) {
        float tn = fminf(tmx, fminf(tmy, tmz));
        float end = fminf(tn, d);
        float len = end - t;
        if (len > 0.f) {
            if (grid) unsafeAtomicAdd(grid + idx, (double)len);
            crossings++;
            if (ADJ) {
                float w = P.weights ? P.weights[idx] : 1.f;
                *aint += *cacc * len + 0.5f * w * len * len;
                *cacc += w * len;
            }
        }
        if (tn >= d) break;
        t = tn;
        if (tmx <= tmy && tmx <= tmz) {
            ix += sxs;
            if ((unsigned)ix >= (unsigned)P.nx) { inside = false; break; }
            idx += sxs;
            tmx += tdx;
        } else if (tmy <= tmz) {
            iy += sys;
            if ((unsigned)iy >= (unsigned)P.ny) { inside = false; break; }
            idx += sys * P.nx;
            tmy += tdy;
        } else {
            iz += szs;
            if ((unsigned)iz >= (unsigned)P.nz) { inside = false; break; }
            idx += szs * nxny;
            tmz += tdz;
        }
    }
    *in_grid = inside ? d : t;
    return inside;
}

template <bool ADJ>
__global__ void __launch_bounds__(256) packet_kernel(KParams P) {
    extern __shared__ unsigned char smem[];
    AdjLds A;
    if (ADJ) {
        A.stride = blockDim.x;
        A.sc = reinterpret_cast<float*>(smem);
        A.g = reinterpret_cast<double*>(smem + sizeof(float) * 4 * P.nmat * blockDim.x);
        for (int w = 0; w < 4; ++w)
            for (int m = 0; m < P.nmat; ++m) adj_g(A, w, m, P.nmat) = 0.0;
    }
    float adjC = 0.f;
    double adjW = 0.0;

    Lane L;
    L.st = ST_IDLE;
    L.set = 0;
    L.layer = 0;
    L.tcount = 0;
    L.iface = false;
    L.dprev = 0.f;
    L.iters = 0;
    unsigned out[NOUT];
#pragma unroll
    for (int i = 0; i < NOUT; ++i) out[i] = 0;
    unsigned ev[NEV];
#pragma unroll
    for (int i = 0; i < NEV; ++i) ev[i] = 0;
    int cur_set = 0;

    unsigned long long base = 0;
    unsigned avail = 0;
    bool exhausted = false;
    const float eps = P.eps;          // march / crossing epsilon (FP32-safe)
    const float eps_kill = P.eps_kill; // flight-start kill threshold after emission or scatter
    const unsigned chunk = (unsigned)P.chunk;
    const int lane_id = __lane_id();

    for (;;) {
        // ---- regeneration: idle lanes take the next packet ids of the wavefront's chunk ----
        if (!exhausted) {
            uint64_t need = wave_ballot(L.st == ST_IDLE);
            while (need) {
                if (avail == 0) {
                    unsigned long long b = 0;
                    if (lane_id == 0) b = atomicAdd(P.work, (unsigned long long)chunk);
                    b = __shfl(b, 0);
                    if (b >= P.total) { exhausted = true; break; }
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
                    if (P.nsets > 1) {
                        int lo = 0, hi = P.nsets - 1;
                        while (lo < hi) {
                            int mid = (lo + hi + 1) >> 1;
                            if (P.set_prefix[mid] <= w) lo = mid; else hi = mid - 1;
                        }
                        s = lo;
                    }
                    if (s != cur_set) {
                        for (int i = 0; i < NOUT; ++i)
                            if (out[i]) atomicAdd(P.stats + (size_t)cur_set * NOUT + i, (unsigned long long)out[i]);
                        for (int i = 0; i < NOUT; ++i) out[i] = 0;
                        cur_set = s;
                    }
                    L.set = s;
                    L.pid = P.set_first[s] + (w - P.set_prefix[s]);
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

        const bool busy = L.st != ST_IDLE;
        const DevMat* __restrict__ mt = P.mats + (size_t)L.set * P.nreg;
        double* __restrict__ grid = P.jmean ? P.jmean + (size_t)L.set * P.grid_stride : nullptr;

        // ---- emission ----
        if (L.st == ST_EMIT) {
            rng_init(L.rng, L.pid);
            emit(P, L);
            L.bounces = 0;
            L.nsc = 0;
            L.tcount = 0;
            L.iters = 0;
            out[O_LAUNCHED]++;
            if (ADJ) {
                adjC = 0.f;
                for (int w = 0; w < 4; ++w)
                    for (int m = 0; m < P.nmat; ++m) adj_sc(A, w, m, P.nmat) = 0.f;
            }
            trace_event(P, L, 1);
        }

        // ---- the sweep: every region at the lane's position ----
        float dmin = __builtin_inff(), smin = __builtin_inff(), dfirst = 0.f, dlay = 0.f, d0 = 0.f;
        float bestv = __builtin_inff();
        int fn = -1, best = 0;
        if (busy) {
            V3<float> p{L.px, L.py, L.pz};
            for (int r = 0; r < P.nreg; ++r) {
                const DevRegion reg = P.regions[r];
                float d = region_eval<float>(P.nodes, reg.first, reg.count, p);
                dmin = fminf(dmin, fabsf(d));
                smin = fminf(smin, d);
                if (r == 0) d0 = d;
                if (fn < 0 && d < 0.f) { fn = r; dfirst = d; }
                if (r == L.layer) dlay = d;
                float c = d > 0.f ? 999.f : -d;  // positive distances count as -999
                if (c < bestv) { bestv = c; best = r; }
            }
            ev[E_SWEEPS]++;
            if (fn < 0) { fn = 0; dfirst = d0; }  // minloc of an all-zero array is index 1
            if (++L.iters > (1u << 24)) {         // runaway guard: never expected
                ev[E_NAN]++;
                L.st = ST_IDLE;
            }
        }

        if (L.st == ST_EMIT) {
            L.layer = best;  // nearest containing surface (src/kernelsMod.f90:275-280)
            L.tau = f_neglog(rnd(L, P));
            L.taurun = 0.f;
            L.iface = false;
            L.st = ST_START;
        }

        int end_kind = -1;  // >= 0: the packet ends this iteration
        bool want_dstmp = false, want_normal = false, do_march = false, cross = false;
        int use_region = 0;
        const int st0 = L.st;
        if (st0 == ST_START) {
            // src/inttau2.f90:52-63.  After an interface the threshold is the march epsilon (the
            // crossing-step ratio that kills grazing crossings and reflected packets); after
            // emission or a scatter it is the reference's own 1e-8.
            // FP32 distances are quantised (multiples of 2^-24 near |p| = 1), so d == eps ties
            // that never occur in the reference's FP64 are common here: a tie counts as inside
            // eps, which kills a reflected packet (pos = oldpos, d0 <= eps) as the reference does.
            if (L.iface ? (dmin <= eps) : (dmin < eps_kill)) end_kind = END_SURFACE;
            else do_march = true;
        } else if (st0 == ST_MARCHED) {
            // src/inttau2.f90:85-97
            if (smin >= 0.f) end_kind = (L.dx * P.rdir[0] + L.dy * P.rdir[1] + L.dz * P.rdir[2] < 0.f) ? END_R : END_T;
            else if (dmin > eps) do_march = true;
            else if (dmin <= L.dprev || dmin <= eps_kill) cross = true;  // converged onto a surface
            else do_march = true;  // within eps but moving away from the surface: keep marching
        } else if (st0 == ST_CROSSED) {
            // the sweep ran at the crossed position (src/inttau2.f90:117-162)
            const int nl = fn;
            const float n1 = mt[L.layer].n, n2 = mt[nl].n;
            if (n1 != n2) {
                ev[E_FRESNEL]++;
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
                    end_kind = (L.dx * P.rdir[0] + L.dy * P.rdir[1] + L.dz * P.rdir[2] < 0.f) ? END_R : END_T;
                } else {
                    ev[E_UNREACH]++;
                    end_kind = END_ERR;
                }
            } else {
                L.layer = nl;  // index-matched: switch region, keep tau (:182-184)
                L.iface = true;
                L.st = ST_START;
            }
        }

        if (cross) {
            // interface: step a little past the surface (src/inttau2.f90:104-116)
            ev[E_IFACE]++;
            L.opx = L.px; L.opy = L.py; L.opz = L.pz;
            L.dlay_old = dlay;
            const float s = dmin + 2.f * eps;
            L.px += s * L.dx; L.py += s * L.dy; L.pz += s * L.dz;
            L.st = ST_CROSSED;
        }

        // dstmp(new): region nl at the pre-crossing position (waterfall over distinct regions)
        if (wave_ballot(want_dstmp)) {
            V3<float> op{L.opx, L.opy, L.opz};
            float dt = eval_lane_region<float>(P, want_dstmp, fn, op);
            if (want_dstmp) {
                ev[E_REGION]++;
                const float dl = dlay, dn = dfirst;
                if (dt < 0.f) {
                    use_region = dt > L.dlay_old ? fn : L.layer;
                    want_normal = true;
                } else if (dl > 0.f && dn < 0.f) {
                    use_region = L.layer;
                    want_normal = true;
                } else if (dl > 0.f && dn > 0.f) {
                    end_kind = (L.dx * P.rdir[0] + L.dy * P.rdir[1] + L.dz * P.rdir[2] < 0.f) ? END_R : END_T;
                } else {
                    ev[E_UNREACH]++;
                    end_kind = END_ERR;
                }
            }
        }

        // normal + Fresnel (src/inttau2.f90:163-181, src/surfaces.f90:14-84)
        if (wave_ballot(want_normal)) {
            V3<Dual> pd{Dual{L.px, 1.f, 0.f, 0.f}, Dual{L.py, 0.f, 1.f, 0.f}, Dual{L.pz, 0.f, 0.f, 1.f}};
            Dual g = eval_lane_region<Dual>(P, want_normal, use_region, pd);
            if (want_normal) {
                ev[E_NORMALS]++;
                float inv = f_rsqrt(g.x * g.x + g.y * g.y + g.z * g.z);
                float nx = g.x * inv, ny = g.y * inv, nz = g.z * inv;
                const int nl = fn;
                const float n1 = mt[L.layer].n, n2 = mt[nl].n;
                float R = fresnel(L.dx, L.dy, L.dz, nx, ny, nz, n1, n2);
                bool rflag = rnd(L, P) <= R;  // draw: reflect or refract
                if (rflag) {
                    float c = 2.f * (nx * L.dx + ny * L.dy + nz * L.dz);
                    L.dx -= c * nx; L.dy -= c * ny; L.dz -= c * nz;
                } else {
                    float eta = n1 / n2;
                    float c1 = nx * L.dx + ny * L.dy + nz * L.dz;
                    if (c1 < 0.f) c1 = -c1;
                    else { nx = -nx; ny = -ny; nz = -nz; }
                    float c2 = f_sqrt(fmaxf(1.f - eta * eta * (1.f - c1 * c1), 0.f));
                    float k = eta * c1 - c2;
                    L.dx = eta * L.dx + k * nx; L.dy = eta * L.dy + k * ny; L.dz = eta * L.dz + k * nz;
                }
                L.tau = f_neglog(rnd(L, P));  // draw: resampled optical depth (:168)
                L.taurun = 0.f;
                L.iface = true;
                if (!rflag) {
                    L.layer = nl;
                    trace_event(P, L, 3);
                } else {
                    L.px = L.opx; L.py = L.opy; L.pz = L.opz;  // step back inside (:174)
                    trace_event(P, L, 4);
                    if (++L.bounces > P.max_bounces) end_kind = END_BOUNCE;
                }
                if (end_kind < 0) L.st = ST_START;
            }
        }

        // ---- march step (src/inttau2.f90:64-83) ----
        if (do_march) {
            ev[E_STEPS]++;
            L.dprev = dmin;
            const float kap = mt[L.layer].kappa;
            const float t = dmin * kap;
            const bool flight_over = !(L.taurun + t <= L.tau);
            float d = dmin;
            if (flight_over) d = (L.tau - L.taurun) / kap;
            else L.taurun += t;
            float in_grid = 0.f, aint = 0.f;
            const bool inside = dda<ADJ>(P, grid, L, d, &in_grid, ev[E_CROSS], &adjC, &aint);
            if (ADJ) {
                const int m = P.region_mat[L.layer];
                if (inside) {
                    adj_sc(A, 0, m, P.nmat) += d;
                    adj_sc(A, 1, m, P.nmat) += aint;
                } else {  // left the grid: the score stops at the exit point
                    adj_sc(A, 0, m, P.nmat) += in_grid;
                    adj_sc(A, 1, m, P.nmat) += aint;
                }
            }
            L.px += d * L.dx; L.py += d * L.dy; L.pz += d * L.dz;
            if (!inside) {
                end_kind = (L.dx * P.rdir[0] + L.dy * P.rdir[1] + L.dz * P.rdir[2] < 0.f) ? END_R : END_T;
            } else if (flight_over) {
                // albedo test (src/kernelsMod.f90:287-297)
                const DevMat m = mt[L.layer];
                if (rnd(L, P) < m.albedo) {
                    if (ADJ) {
                        const int am = P.region_mat[L.layer];
                        adj_sc(A, 2, am, P.nmat) += 1.f;
                        adj_sc(A, 3, am, P.nmat) += adjC;
                    }
                    L.nsc++;
                    out[O_SCATTERS]++;
                    if ((P.flags & 1u) && L.nsc <= 4) {
                        double* mo = P.moments + (L.nsc - 1) * 3;
                        atomicAdd(mo + 0, (double)L.px); atomicAdd(mo + 1, (double)L.py); atomicAdd(mo + 2, (double)L.pz);
                        atomicAdd(mo + 12, (double)L.px * L.px); atomicAdd(mo + 13, (double)L.py * L.py);
                        atomicAdd(mo + 14, (double)L.pz * L.pz);
                    }
                    scatter(P, L, m.hgg, m.g2);
                    trace_event(P, L, 2);
                    if ((P.flags & 2u) && L.nsc >= 5) {
                        end_kind = END_ABSORB;
                    } else {
                        L.tau = f_neglog(rnd(L, P));  // draw: next flight's optical depth
                        L.taurun = 0.f;
                        L.iface = false;
                        L.st = ST_START;
                    }
                } else {
                    end_kind = END_ABSORB;
                }
            } else {
                L.st = ST_MARCHED;
            }
        }

        // non-finite guard
        if (end_kind < 0 && L.st != ST_IDLE && !(fabsf(L.px) < 1e30f && fabsf(L.py) < 1e30f && fabsf(L.pz) < 1e30f)) {
            ev[E_NAN]++;
            end_kind = END_ERR;
        }
        if (end_kind >= 0) {
            switch (end_kind) {
            case END_ABSORB: out[O_ABSORBED]++; trace_event(P, L, 5); break;
            case END_R: out[O_R]++; trace_event(P, L, 6); break;
            case END_T: out[O_T]++; trace_event(P, L, 6); break;
            case END_SURFACE: out[O_KSURF]++; trace_event(P, L, 7); break;
            case END_BOUNCE: out[O_KBOUNCE]++; trace_event(P, L, 7); break;
            default: break;
            }
            if (ADJ) {
                const bool isr = end_kind == END_R;
                adjW += (double)adjC;
                for (int m = 0; m < P.nmat; ++m) {
                    float Lm = adj_sc(A, 0, m, P.nmat), Am = adj_sc(A, 1, m, P.nmat);
                    float Km = adj_sc(A, 2, m, P.nmat), Bm = adj_sc(A, 3, m, P.nmat);
                    float mus = P.mat_mus[m];
                    float scs = mus > 0.f ? Km / mus : 0.f;
                    if (isr) {
                        adj_g(A, 0, m, P.nmat) += (double)(-Lm);
                        adj_g(A, 1, m, P.nmat) += (double)(scs - Lm);
                    }
                    float lw = Lm * adjC - Am;
                    float bw = mus > 0.f ? (Km * adjC - Bm) / mus : 0.f;
                    adj_g(A, 2, m, P.nmat) += (double)(-lw);
                    adj_g(A, 3, m, P.nmat) += (double)(bw - lw);
                }
            }
            L.st = ST_IDLE;
        }
    }

    // ---- flush counters ----
    for (int i = 0; i < NOUT; ++i)
        if (out[i]) atomicAdd(P.stats + (size_t)cur_set * NOUT + i, (unsigned long long)out[i]);
    unsigned long long* evg = P.stats + (size_t)P.nsets * NOUT;
    for (int i = 0; i < NEV; ++i) {
        unsigned long long v = ev[i];
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id == 0 && v) atomicAdd(evg + i, v);
    }
    if (ADJ) {
        for (int w = 0; w < 4; ++w)
            for (int m = 0; m < P.nmat; ++m) {
                double v = adj_g(A, w, m, P.nmat);
                for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
                if (lane_id == 0) atomicAdd(P.grad + w * P.nmat + m, v);
            }
        double v = adjW;
        for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off);
        if (lane_id == 0) atomicAdd(P.grad + 4 * P.nmat, v);
    }
}

}  // namespace smcrt
