// smcrt_device.h — device-side building blocks of the packet kernel (gfx950).
//
//  - float helpers on hardware instructions (v_sqrt_f32, v_rcp_f32, v_log_f32, v_sin/cos_f32 in
//    revolutions) — the transport runs in FP32;
//  - Dual: forward-mode derivative (value + gradient wrt position) so the SDF normal is the
//    exact gradient of the SDF program instead of the reference's FP64 tetrahedral finite
//    difference (src/sdfs/sdf_base.f90:165-189), whose h = 1e-6 is below FP32 resolution;
//  - the SDF program interpreter, written once as a template over float and Dual;
//  - Philox-4x32-10 keyed by (seed, packet id).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define SMCRT_DEV __device__ __forceinline__

// ------------------------------------------------------------------------------------------
// device scene layout
// ------------------------------------------------------------------------------------------
// One SDF program node, 32 dwords so a wave-uniform index loads it with scalar loads.
struct alignas(16) DevNode {
    int op;       // SMCRT_OP_*
    int kind;     // primitive / CSG / modifier kind
    int ident;    // PRIM transform: 1 the identity (skipped), 2 a translation only (3 adds), 0 general
    float k;      // CSG smoothing
    float p[8];
    float t[12];
    float q[8];   // constants derived from p and k on the host (see derive_node in smcrt_api.hip)
};
static_assert(sizeof(DevNode) == 128, "DevNode must stay 32 dwords");

struct DevRegion {
    int first, count, material;
    int flat;  // 1: flat fold of primitives by union, 2: by SmoothUnion (lazy gradients), 0: other
};

// Per-region material constants (one table per property set); `mat` is the region's material
// index (the adjoint's parameter index).
struct alignas(16) DevMat {
    float kappa, albedo, hgg, g2;
    float n, mus, mua;
    int mat;
};

// The scene program is wave-uniform: reading it through the constant address space lets the
// compiler fetch it with scalar loads into SGPRs (the generic pointer would force vector loads,
// since the kernel also writes global memory and the reads could otherwise be clobbered).
typedef __attribute__((address_space(4))) const DevNode CDevNode;
typedef __attribute__((address_space(4))) const DevRegion CDevRegion;

// ------------------------------------------------------------------------------------------
// float helpers
// ------------------------------------------------------------------------------------------
SMCRT_DEV float f_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }
SMCRT_DEV float f_rcp(float x) { return __builtin_amdgcn_rcpf(x); }
SMCRT_DEV float f_div(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
SMCRT_DEV float f_rsqrt(float x) { return __builtin_amdgcn_rsqf(x); }
// -ln(u) via v_log_f32 (log2)
SMCRT_DEV float f_neglog(float u) { return -0.69314718055994531f * __builtin_amdgcn_logf(u); }
// sin/cos of 2*pi*r for r in revolutions (v_sin_f32 / v_cos_f32 take revolutions)
SMCRT_DEV float f_sin_rev(float r) { return __builtin_amdgcn_sinf(r); }
SMCRT_DEV float f_cos_rev(float r) { return __builtin_amdgcn_cosf(r); }
SMCRT_DEV float f_sin(float x) { return __builtin_amdgcn_sinf(x * 0.15915494309189535f); }
SMCRT_DEV float f_cos(float x) { return __builtin_amdgcn_cosf(x * 0.15915494309189535f); }

SMCRT_DEV float val(float x) { return x; }

// ------------------------------------------------------------------------------------------
// Dual numbers: value and d/dx, d/dy, d/dz
// ------------------------------------------------------------------------------------------
struct Dual {
    float v, x, y, z;
};
SMCRT_DEV float val(const Dual& a) { return a.v; }
SMCRT_DEV Dual dconst(float c) { return Dual{c, 0.f, 0.f, 0.f}; }
SMCRT_DEV Dual operator+(const Dual& a, const Dual& b) { return Dual{a.v + b.v, a.x + b.x, a.y + b.y, a.z + b.z}; }
SMCRT_DEV Dual operator-(const Dual& a, const Dual& b) { return Dual{a.v - b.v, a.x - b.x, a.y - b.y, a.z - b.z}; }
SMCRT_DEV Dual operator-(const Dual& a) { return Dual{-a.v, -a.x, -a.y, -a.z}; }
SMCRT_DEV Dual operator+(const Dual& a, float b) { return Dual{a.v + b, a.x, a.y, a.z}; }
SMCRT_DEV Dual operator-(const Dual& a, float b) { return Dual{a.v - b, a.x, a.y, a.z}; }
SMCRT_DEV Dual operator-(float b, const Dual& a) { return Dual{b - a.v, -a.x, -a.y, -a.z}; }
SMCRT_DEV Dual operator*(const Dual& a, float s) { return Dual{a.v * s, a.x * s, a.y * s, a.z * s}; }
SMCRT_DEV Dual operator*(float s, const Dual& a) { return a * s; }
SMCRT_DEV Dual operator*(const Dual& a, const Dual& b) {
    return Dual{a.v * b.v, a.x * b.v + a.v * b.x, a.y * b.v + a.v * b.y, a.z * b.v + a.v * b.z};
}
SMCRT_DEV Dual operator/(const Dual& a, const Dual& b) {
    float r = f_rcp(b.v);
    float q = a.v * r;
    return Dual{q, (a.x - q * b.x) * r, (a.y - q * b.y) * r, (a.z - q * b.z) * r};
}
SMCRT_DEV Dual operator/(const Dual& a, float b) { return a * f_rcp(b); }
SMCRT_DEV Dual d_sqrt(const Dual& a) {
    float s = f_sqrt(a.v);
    float h = s > 0.f ? 0.5f * f_rcp(s) : 0.f;
    return Dual{s, a.x * h, a.y * h, a.z * h};
}
SMCRT_DEV Dual d_abs(const Dual& a) { return a.v < 0.f ? -a : a; }
SMCRT_DEV Dual d_min(const Dual& a, const Dual& b) { return b.v < a.v ? b : a; }
SMCRT_DEV Dual d_max(const Dual& a, const Dual& b) { return b.v > a.v ? b : a; }
SMCRT_DEV Dual d_sin(const Dual& a) { float s = f_sin(a.v), c = f_cos(a.v); return Dual{s, c * a.x, c * a.y, c * a.z}; }
SMCRT_DEV Dual d_cos(const Dual& a) { float s = f_sin(a.v), c = f_cos(a.v); return Dual{c, -s * a.x, -s * a.y, -s * a.z}; }

// D1: value and one directional derivative.  Three passes (seed x, y, z) give the gradient at a
// third of Dual's register footprint; the packet kernel uses it for interface normals.
struct D1 {
    float v, d;
};
SMCRT_DEV float val(const D1& a) { return a.v; }
SMCRT_DEV D1 operator+(const D1& a, const D1& b) { return D1{a.v + b.v, a.d + b.d}; }
SMCRT_DEV D1 operator-(const D1& a, const D1& b) { return D1{a.v - b.v, a.d - b.d}; }
SMCRT_DEV D1 operator-(const D1& a) { return D1{-a.v, -a.d}; }
SMCRT_DEV D1 operator+(const D1& a, float b) { return D1{a.v + b, a.d}; }
SMCRT_DEV D1 operator-(const D1& a, float b) { return D1{a.v - b, a.d}; }
SMCRT_DEV D1 operator-(float b, const D1& a) { return D1{b - a.v, -a.d}; }
SMCRT_DEV D1 operator*(const D1& a, float s) { return D1{a.v * s, a.d * s}; }
SMCRT_DEV D1 operator*(float s, const D1& a) { return D1{a.v * s, a.d * s}; }
SMCRT_DEV D1 operator*(const D1& a, const D1& b) { return D1{a.v * b.v, a.d * b.v + a.v * b.d}; }
SMCRT_DEV D1 operator/(const D1& a, const D1& b) {
    float r = f_rcp(b.v), q = a.v * r;
    return D1{q, (a.d - q * b.d) * r};
}
SMCRT_DEV D1 operator/(const D1& a, float b) { return a * f_rcp(b); }

// generic wrappers so one template body serves float, Dual and D1
SMCRT_DEV D1 g_sqrt(const D1& a) {
    float s = f_sqrt(a.v);
    return D1{s, s > 0.f ? a.d * (0.5f * f_rcp(s)) : 0.f};
}
SMCRT_DEV D1 g_abs(const D1& a) { return a.v < 0.f ? -a : a; }
SMCRT_DEV D1 g_min(const D1& a, const D1& b) { return b.v < a.v ? b : a; }
SMCRT_DEV D1 g_max(const D1& a, const D1& b) { return b.v > a.v ? b : a; }
SMCRT_DEV D1 g_max0(const D1& a) { return a.v > 0.f ? a : D1{0.f, 0.f}; }
SMCRT_DEV D1 g_min0(const D1& a) { return a.v < 0.f ? a : D1{0.f, 0.f}; }
SMCRT_DEV D1 g_clamp01(const D1& a) { return a.v < 0.f ? D1{0.f, 0.f} : (a.v > 1.f ? D1{1.f, 0.f} : a); }
SMCRT_DEV D1 g_div(const D1& a, const D1& b) { return a / b; }
SMCRT_DEV D1 g_sin(const D1& a) { return D1{f_sin(a.v), f_cos(a.v) * a.d}; }
SMCRT_DEV D1 g_cos(const D1& a) { return D1{f_cos(a.v), -f_sin(a.v) * a.d}; }
SMCRT_DEV D1 g_neg_if(const D1& a, bool neg) { return neg ? -a : a; }
SMCRT_DEV float g_sqrt(float a) { return f_sqrt(a); }
SMCRT_DEV Dual g_sqrt(const Dual& a) { return d_sqrt(a); }
SMCRT_DEV float g_abs(float a) { return fabsf(a); }
SMCRT_DEV Dual g_abs(const Dual& a) { return d_abs(a); }
SMCRT_DEV float g_min(float a, float b) { return fminf(a, b); }
SMCRT_DEV Dual g_min(const Dual& a, const Dual& b) { return d_min(a, b); }
SMCRT_DEV float g_max(float a, float b) { return fmaxf(a, b); }
SMCRT_DEV Dual g_max(const Dual& a, const Dual& b) { return d_max(a, b); }
SMCRT_DEV float g_max0(float a) { return fmaxf(a, 0.f); }
SMCRT_DEV Dual g_max0(const Dual& a) { return a.v > 0.f ? a : dconst(0.f); }
SMCRT_DEV float g_min0(float a) { return fminf(a, 0.f); }
SMCRT_DEV Dual g_min0(const Dual& a) { return a.v < 0.f ? a : dconst(0.f); }
SMCRT_DEV float g_clamp01(float a) { return fminf(fmaxf(a, 0.f), 1.f); }
SMCRT_DEV Dual g_clamp01(const Dual& a) { return a.v < 0.f ? dconst(0.f) : (a.v > 1.f ? dconst(1.f) : a); }
SMCRT_DEV float g_div(float a, float b) { return f_div(a, b); }
SMCRT_DEV Dual g_div(const Dual& a, const Dual& b) { return a / b; }
SMCRT_DEV float g_sin(float a) { return f_sin(a); }
SMCRT_DEV Dual g_sin(const Dual& a) { return d_sin(a); }
SMCRT_DEV float g_cos(float a) { return f_cos(a); }
SMCRT_DEV Dual g_cos(const Dual& a) { return d_cos(a); }
template <class T> SMCRT_DEV T g_const(float c);
template <> SMCRT_DEV float g_const<float>(float c) { return c; }
template <> SMCRT_DEV Dual g_const<Dual>(float c) { return dconst(c); }
template <> SMCRT_DEV D1 g_const<D1>(float c) { return D1{c, 0.f}; }
// sign(s) * a with s's sign as a float multiplier
SMCRT_DEV float g_neg_if(float a, bool neg) { return neg ? -a : a; }
SMCRT_DEV Dual g_neg_if(const Dual& a, bool neg) { return neg ? -a : a; }

template <class T>
struct V3 {
    T x, y, z;
};

// plain float (the sweep's evaluations) as opposed to Dual numbers (gradients)
template <class T> struct is_plain_float { static constexpr bool value = false; };
template <> struct is_plain_float<float> { static constexpr bool value = true; };
// keeps a rarely taken wave-uniform branch a branch (the compiler would otherwise compute its
// body for every wavefront and select)
SMCRT_DEV void rare_path() { __asm__ volatile(""); }
template <class T> SMCRT_DEV T vdot(const V3<T>& a, const V3<T>& b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
template <class T> SMCRT_DEV V3<T> vsub(const V3<T>& a, const V3<T>& b) { return V3<T>{a.x - b.x, a.y - b.y, a.z - b.z}; }
template <class T> SMCRT_DEV V3<T> vsubc(const V3<T>& a, float bx, float by, float bz) { return V3<T>{a.x - bx, a.y - by, a.z - bz}; }
template <class T> SMCRT_DEV V3<T> vabs(const V3<T>& a) { return V3<T>{g_abs(a.x), g_abs(a.y), g_abs(a.z)}; }
template <class T> SMCRT_DEV T vlen(const V3<T>& a) { return g_sqrt(vdot(a, a)); }
template <class T> SMCRT_DEV V3<T> vmax0(const V3<T>& a) { return V3<T>{g_max0(a.x), g_max0(a.y), g_max0(a.z)}; }

// ------------------------------------------------------------------------------------------
// Closed polygon in the xy plane through its uniform-grid accelerator (built on the host by
// smcrt_scene_create).  Accelerator layout, floats from `base` (a multiple of 4: every record
// below is read as aligned vectors):
//   header[16]: nx, ny, gx0, gy0, 1/cell, cell, gx1, gy1, margin, edges, cells, lists, n_edges
//   edges[8 each]: vx, vy, ex, ey | start-vertex pseudo-normal (2), end-vertex pseudo-normal (2)
//   cells[4 each]: far cell:  0, 0, signed centre distance, 0
//                  near cell: rest (int bits), pairs (int bits, >= 1), first pair of candidate
//                  edge offsets (int bits)
//   lists: the near cells' further pairs of candidate edge offsets (absolute float offsets of
//          edge records, int bits); an odd list is padded with its first edge
// A near cell lists every edge that can be nearest anywhere in the cell (distance from the
// centre <= D_c + 2r, r the half diagonal), so its distance is exact; its sign comes from the
// nearest feature (edge side, or the vertex pseudo-normal when the nearest point is a vertex).
// A far cell (D_c > 2r, no edge inside) returns sign * (D_c - |q - c|); outside the grid the
// distance to the grid rectangle plus the margin — both lower bounds, which keep sphere
// tracing safe and are only used away from the surface.  Latency, not arithmetic, bounds this
// evaluation (a few L1/L2 hits per polygon per sweep): the cell record carries the first pair
// of offsets, a pair's two 16-byte edge loads are in flight together, and the nearest edge's
// pseudo-normals are read once, after the loop.
// ------------------------------------------------------------------------------------------
SMCRT_DEV float4 ld_f4(const float* __restrict__ A, int i) { return *reinterpret_cast<const float4*>(A + i); }

template <class T>
SMCRT_DEV T polygon_eval(const float* __restrict__ A, int base, const V3<T>& p) {
    const float* H = A + base;
    const float gx0 = H[2], gy0 = H[3], inv = H[4], cell = H[5], gx1 = H[6], gy1 = H[7];
    const float qx = val(p.x), qy = val(p.y);
    if (!(qx >= gx0 && qx < gx1 && qy >= gy0 && qy < gy1)) {
        T dx = g_max0(g_max(gx0 - p.x, p.x - gx1));
        T dy = g_max0(g_max(gy0 - p.y, p.y - gy1));
        return g_sqrt(dx * dx + dy * dy) + H[8];
    }
    const int nx = (int)H[0], ny = (int)H[1];
    int ix = (int)((qx - gx0) * inv), iy = (int)((qy - gy0) * inv);
    ix = ix < nx - 1 ? ix : nx - 1;
    iy = iy < ny - 1 ? iy : ny - 1;
    const float4 c = ld_f4(A, (int)H[10] + 4 * (ix + nx * iy));
    const int np = __float_as_int(c.y);
    if (np == 0) {
        const float sD = c.z;
        T ddx = p.x - (gx0 + ((float)ix + 0.5f) * cell), ddy = p.y - (gy0 + ((float)iy + 0.5f) * cell);
        T r = fabsf(sD) - g_sqrt(ddx * ddx + ddy * ddy);
        return g_neg_if(r, sD < 0.f);
    }
    const int rest = __float_as_int(c.x);
    int o0 = __float_as_int(c.z), o1 = __float_as_int(c.w);
    T best = g_const<T>(3.0e38f);
    float bestv = 3.0e38f;
    int be = o0;  // the nearest edge's record (a valid record even if no distance compares)
    auto edge = [&](const float4& e, int off) {
        T wx = p.x - e.x, wy = p.y - e.y;
        T t = g_clamp01((wx * e.z + wy * e.w) * f_rcp(e.z * e.z + e.w * e.w));
        T bx = wx - t * e.z, by = wy - t * e.w;
        T d2 = bx * bx + by * by;
        const float d2v = val(d2);
        if (d2v < bestv) {
            bestv = d2v;
            best = d2;
            be = off;
        }
    };
    for (int k = 0;;) {
        const float4 e0 = ld_f4(A, o0), e1 = ld_f4(A, o1);
        edge(e0, o0);
        edge(e1, o1);
        if (++k >= np) break;
        const int2 v = *reinterpret_cast<const int2*>(A + rest + 2 * (k - 1));
        o0 = v.x;
        o1 = v.y;
    }
    // the sign from the nearest feature of the nearest edge
    const float4 e = ld_f4(A, be), pn = ld_f4(A, be + 4);
    const float wxv = qx - e.x, wyv = qy - e.y;
    const float tv = (wxv * e.z + wyv * e.w) * f_rcp(e.z * e.z + e.w * e.w);
    bool inside;
    if (tv <= 0.f) inside = wxv * pn.x + wyv * pn.y < 0.f;
    else if (tv >= 1.f) inside = (wxv - e.z) * pn.z + (wyv - e.w) * pn.w < 0.f;
    else inside = e.z * wyv - e.w * wxv > 0.f;  // left of a counter-clockwise edge
    return g_neg_if(g_sqrt(best), inside);
}

// Gathered polygon evaluation for kernels compiled for one scene.  A point lies inside few
// polygons' grid rectangles (one, where rectangles do not overlap); every other polygon's value is
// the rectangle distance plus the margin, arithmetic on constants.  So the compiled sweep first
// finds, per lane, the first polygon (in program order) whose rectangle holds its position and
// evaluates that one through its accelerator (PolyHit), all lanes at once; a polygon's region
// code then takes the lane's gathered value, the rectangle formula, or (a lane inside a second
// rectangle, a wave-level branch no lane takes where rectangles are disjoint) a full evaluation.
// The values are those of polygon_eval.  Q (the node's, host-filled): gx0, gy0, gx1, gy1, margin.
struct PolyHit {
    int k;    // the lane's gathered polygon (-1: inside none)
    float d;  // its value
};
SMCRT_DEV bool polygon_rect_inside(const float* __restrict__ Q, const V3<float>& p) {
    return (p.x >= Q[0]) & (p.x < Q[2]) & (p.y >= Q[1]) & (p.y < Q[3]);
}
SMCRT_DEV float polygon_outside(const float* __restrict__ Q, const V3<float>& p) {
    const float dx = fmaxf(fmaxf(Q[0] - p.x, p.x - Q[2]), 0.f);
    const float dy = fmaxf(fmaxf(Q[1] - p.y, p.y - Q[3]), 0.f);
    return f_sqrt(dx * dx + dy * dy) + Q[4];
}
// Scenes whose polygons share one position frame and have pairwise disjoint rectangles (checked on
// the host): a lane inside polygon j's rectangle gathered polygon j.
SMCRT_DEV float poly_value_disjoint(int j, const float* __restrict__ Q, const V3<float>& p, const PolyHit& h) {
    return h.k == j ? h.d : polygon_outside(Q, p);
}
SMCRT_DEV float poly_value(int j, int base, const float* __restrict__ Q, const V3<float>& p,
                           const float* __restrict__ A, const PolyHit& h) {
    if (h.k == j) return h.d;
    const bool in = polygon_rect_inside(Q, p);
    float v = polygon_outside(Q, p);
    if (__builtin_amdgcn_ballot_w64(in)) {
        rare_path();
        if (in) v = polygon_eval<float>(A, base, p);
    }
    return v;
}

// ------------------------------------------------------------------------------------------
// Kind masks.  The packet kernel is compiled for a few sets of primitive / operator / modifier
// kinds (the host launches the smallest variant covering the scene): the interpreter's code is a
// large part of the kernel, and a kernel whose loop outgrows the instruction cache runs far
// slower, so scenes that use two primitives get a kernel that contains two.
// ------------------------------------------------------------------------------------------
#define KM_PRIM(k) (1u << (k))         // primitive kinds 1..11
#define KM_CSG(k) (1u << (15 + (k)))   // CSG kinds 1..4 -> bits 16..19
#define KM_MOD(k) (1u << (23 + (k)))   // modifier kinds 1..6 -> bits 24..29
#define KM_ANY_CSG (KM_CSG(1) | KM_CSG(2) | KM_CSG(3) | KM_CSG(4))
#define KM_ANY_MOD (KM_MOD(1) | KM_MOD(2) | KM_MOD(3) | KM_MOD(4) | KM_MOD(5) | KM_MOD(6))
#define KM_FULL (0xFFEu | KM_ANY_CSG | KM_ANY_MOD)
#define KM_HAS(KM, bit) (((KM) & (bit)) != 0u)

// ------------------------------------------------------------------------------------------
// SDF primitives (src/sdfs/sdfs.f90:494-735)
// ------------------------------------------------------------------------------------------
template <class T, uint32_t KM = KM_FULL>
SMCRT_DEV T prim_eval(int kind, const float* __restrict__ P, const float* __restrict__ Q, const V3<T>& p,
                      const float* __restrict__ accel) {
    // a kind outside the mask is never dispatched (the host picks a covering variant): its case
    // compiles to a break
    switch (kind) {
    case 1: if constexpr (!KM_HAS(KM, KM_PRIM(1))) break; {  // sphere
        return g_sqrt(vdot(p, p)) - P[0];
    }
    case 2: if constexpr (!KM_HAS(KM, KM_PRIM(2))) break; {  // box
        V3<T> q = vsubc(vabs(p), P[0], P[1], P[2]);
        if constexpr (is_plain_float<T>::value) {
            // length(max(q, 0)) + min(max(q), 0) is max(q) unless two or more components are
            // positive (near an edge or corner): the square root only where a lane needs it
            float d = fmaxf(q.x, fmaxf(q.y, q.z));
            const bool edge = __builtin_amdgcn_fmed3f(q.x, q.y, q.z) > 0.f;  // two or more positive
            if (__builtin_amdgcn_ballot_w64(edge)) {
                rare_path();
                const float mx = fmaxf(q.x, 0.f), my = fmaxf(q.y, 0.f), mz = fmaxf(q.z, 0.f);
                const float e = f_sqrt(mx * mx + my * my + mz * mz);
                d = edge ? e : d;
            }
            return d;
        }
        return vlen(vmax0(q)) + g_min0(g_max(q.x, g_max(q.y, q.z)));
    }
    case 3: if constexpr (!KM_HAS(KM, KM_PRIM(3))) break; {  // torus
        T a = g_sqrt(p.x * p.x + p.z * p.z) - P[0];
        return g_sqrt(a * a + p.y * p.y) - P[1];
    }
    case 4: if constexpr (!KM_HAS(KM, KM_PRIM(4))) break; {  // cylinder(a, b, r)
        // Q: b - a, |b - a|^2, 1 / |b - a|^2, r |b - a|^2, |b - a|^2 / 2, axis (1-3 when b - a
        // lies along x, y or z, else 0)
        const float bax = Q[0], bay = Q[1], baz = Q[2], baba = Q[3], hb = Q[6];
        V3<T> pa = vsubc(p, P[0], P[1], P[2]);
        if (Q[7] != 0.f) {
            // b - a along one axis (length L = |b - a|): the same distance with the projections
            // taken directly — X = (distance from the axis) - r, Y = |along - L/2| - L/2 are the
            // general formula's x / |b - a|^2 and y / |b - a|
            const int ax = (int)Q[7];
            const T u = ax == 1 ? pa.y : pa.x, v = ax == 3 ? pa.y : pa.z, w = ax == 1 ? pa.x : (ax == 2 ? pa.y : pa.z);
            const float L = ax == 1 ? bax : (ax == 2 ? bay : baz);  // signed length along the axis
            const T X = g_sqrt(u * u + v * v) - P[6];
            const T Y = g_abs(w - 0.5f * L) - 0.5f * fabsf(L);
            if constexpr (is_plain_float<T>::value) {
                // max(X, Y) unless both are positive (beyond the rim of a cap): the square root
                // only where a lane needs it
                float d = fmaxf(X, Y);
                const bool rim = (X > 0.f) & (Y > 0.f);
                if (__builtin_amdgcn_ballot_w64(rim)) {
                    rare_path();
                    const float e = f_sqrt(X * X + Y * Y);
                    d = rim ? e : d;
                }
                return d;
            }
            const float xv = val(X), yv = val(Y);
            if (fmaxf(xv, yv) < 0.f) return g_max(X, Y);
            const T mx = g_max0(X), my = g_max0(Y);
            return g_sqrt(mx * mx + my * my);
        }
        T paba = pa.x * bax + pa.y * bay + pa.z * baz;
        V3<T> w{pa.x * baba - paba * bax, pa.y * baba - paba * bay, pa.z * baba - paba * baz};
        T x = vlen(w) - Q[5];
        T y = g_abs(paba - hb) - hb;
        T x2 = x * x;
        T y2 = (y * y) * baba;
        float xv = val(x), yv = val(y);
        T d;
        if (fmaxf(xv, yv) < 0.f) d = -g_min(x2, y2);
        else if (xv > 0.f && yv > 0.f) d = x2 + y2;
        else if (xv > 0.f) d = x2;
        else if (yv > 0.f) d = y2;
        else d = g_const<T>(0.f);
        T r = g_sqrt(g_abs(d)) * Q[4];
        return g_neg_if(r, val(d) < 0.f);
    }
    case 5: if constexpr (!KM_HAS(KM, KM_PRIM(5))) break; {  // triprism(h1, h2)
        V3<T> q = vabs(p);
        return g_max(q.z - P[1], g_max(q.x * 0.866025f + p.y * 0.5f, -p.y) - P[0] * 0.5f);
    }
    case 6:    // segment (tube thickness 0.1)
    case 7: if constexpr (!KM_HAS(KM, KM_PRIM(6) | KM_PRIM(7))) break; {  // capsule(a, b, r)
        const float bax = Q[0], bay = Q[1], baz = Q[2];  // Q: b - a, |b - a|^2, 1 / |b - a|^2
        V3<T> pa = vsubc(p, P[0], P[1], P[2]);
        T h = g_clamp01((pa.x * bax + pa.y * bay + pa.z * baz) * Q[4]);
        V3<T> w{pa.x - h * bax, pa.y - h * bay, pa.z - h * baz};
        return vlen(w) - (kind == 6 ? 0.1f : P[6]);
    }
    case 8: if constexpr (!KM_HAS(KM, KM_PRIM(8))) break; {  // capped cone(a, b, ra, rb)
        float ra = P[6], rb = P[7], rba = rb - ra;
        float bax = P[3] - P[0], bay = P[4] - P[1], baz = P[5] - P[2];
        float baba = bax * bax + bay * bay + baz * baz;
        V3<T> pa = vsubc(p, P[0], P[1], P[2]);
        T papa = vdot(pa, pa);
        T paba = (pa.x * bax + pa.y * bay + pa.z * baz) * (1.0f / baba);
        T x = g_sqrt(g_max0(papa - baba * paba * paba));
        T cax = val(paba) < 0.5f ? g_max0(x - ra) : g_max0(x - rb);
        T cay = g_abs(paba - 0.5f) - 0.5f;
        float k = rba * rba + baba;
        T f = g_clamp01((rba * (x - ra) + paba * baba) * (1.0f / k));
        T cbx = x - ra - f * rba;
        T cby = paba - f;
        bool neg = val(cbx) < 0.f && val(cay) < 0.f;
        T r = g_sqrt(g_min(cax * cax + baba * (cay * cay), cbx * cbx + baba * (cby * cby)));
        return g_neg_if(r, neg);
    }
    case 9: if constexpr (!KM_HAS(KM, KM_PRIM(9))) break; {  // egg(r1, r2, h)
        float r1 = P[0], r2 = P[1], h = P[2];
        float r = r1 - r2, hin = h + r;
        float l = (hin * hin - r * r) / (2.f * r);
        T px = g_abs(p.x);
        if (val(p.y) <= 0.f) return g_sqrt(px * px + p.y * p.y + p.z * p.z) - r1;
        if ((val(p.y) - hin) * l > val(px) * hin) {
            T dy = p.y - hin;
            return g_sqrt(px * px + dy * dy + p.z * p.z) - ((r1 + l) - sqrtf(hin * hin + l * l));
        }
        T dx = px + l;
        return g_sqrt(dx * dx + p.y * p.y + p.z * p.z) - (r1 + l);
    }
    case 10: if constexpr (!KM_HAS(KM, KM_PRIM(10))) break; {  // plane(a)
        return p.x * P[0] + p.y * P[1] + p.z * P[2];
    }
    case 11:  // polygon (P[0]: accelerator offset)
        if constexpr (KM_HAS(KM, KM_PRIM(11))) return polygon_eval<T>(accel, (int)P[0], p);
        break;
    }
    return g_const<T>(0.f);
}

// CSG operators (src/sdfs/sdfModifiers.f90:412-475)
// Q: 1 / k, k / 6 (SmoothUnion)
template <class T, uint32_t KM = KM_FULL>
SMCRT_DEV T csg_eval(int kind, const T& d1, const T& d2, float k, const float* __restrict__ Q) {
    switch (kind) {
    case 1: if constexpr (KM_HAS(KM, KM_CSG(1))) return g_min(d1, d2);
        break;
    case 2: if constexpr (KM_HAS(KM, KM_CSG(2))) {
        T h = g_max0(k - g_abs(d1 - d2)) * Q[0];
        return g_min(d1, d2) - h * h * h * Q[1];
    }
        break;
    case 3: if constexpr (KM_HAS(KM, KM_CSG(3))) return g_max(-d1, d2);
        break;
    case 4: if constexpr (KM_HAS(KM, KM_CSG(4))) return g_max(d1, d2);
        break;
    }
    return d1;
}

// modifier position maps (src/sdfs/sdfModifiers.f90:290-375)
template <class T, uint32_t KM = KM_FULL>
SMCRT_DEV V3<T> mod_pos(int kind, const float* __restrict__ P, const V3<T>& p) {
    switch (kind) {
    case 2: if constexpr (!KM_HAS(KM, KM_MOD(2))) break; {  // revolution
        T a = g_sqrt(p.x * p.x + p.z * p.z) - P[0];
        return V3<T>{a, p.y, g_const<T>(0.f)};
    }
    case 4:  // elongate
        if constexpr (KM_HAS(KM, KM_MOD(4))) return vmax0(vsubc(vabs(p), P[0], P[1], P[2]));
        break;
    case 5: if constexpr (!KM_HAS(KM, KM_MOD(5))) break; {  // twist
        T a = p.z * P[0];
        T c = g_cos(a), s = g_sin(a);
        return V3<T>{c * p.x - s * p.y, s * p.x + c * p.y, p.z};
    }
    case 6: if constexpr (!KM_HAS(KM, KM_MOD(6))) break; {  // bend
        T a = p.x * P[0];
        T c = g_cos(a), s = g_sin(a);
        return V3<T>{c * p.x - s * p.y, s * p.x + c * p.y, p.z};
    }
    }
    return p;  // extrude, onion
}

// modifier post-processing (src/sdfs/sdfModifiers.f90:273-335)
template <class T, uint32_t KM = KM_FULL>
SMCRT_DEV T mod_post(int kind, const float* __restrict__ P, const V3<T>& p, const T& d) {
    switch (kind) {
    case 1: if constexpr (!KM_HAS(KM, KM_MOD(1))) break; {  // extrude
        T wy = g_abs(p.z) - P[0];
        T mx = g_max0(d), my = g_max0(wy);
        return g_min0(g_max(d, wy)) + g_sqrt(mx * mx + my * my);
    }
    case 3:  // onion
        if constexpr (KM_HAS(KM, KM_MOD(3))) return g_abs(d) - P[0];
        break;
    case 4: if constexpr (!KM_HAS(KM, KM_MOD(4))) break; {  // elongate
        V3<T> q = vsubc(vabs(p), P[0], P[1], P[2]);
        return d + g_min0(g_max(q.x, g_max(q.y, q.z)));
    }
    }
    return d;
}

template <class T>
SMCRT_DEV V3<T> xform(const float* __restrict__ t, const V3<T>& p) {
    return V3<T>{p.x * t[0] + p.y * t[3] + p.z * t[6] + t[9],
                 p.x * t[1] + p.y * t[4] + p.z * t[7] + t[10],
                 p.x * t[2] + p.y * t[5] + p.z * t[8] + t[11]};
}

// Evaluate one region's postfix program at p.  Node range and node contents are wave-uniform,
// so the opcode dispatch is a scalar branch and parameters come from scalar loads.  The value
// and position stacks are shift registers (no dynamically indexed register arrays).
template <class T, uint32_t KM = KM_FULL>
SMCRT_DEV T region_eval(CDevNode* nodes, int first, int count, const V3<T>& p0, const float* __restrict__ accel) {
    T v0 = g_const<T>(0.f), v1 = v0, v2 = v0, v3 = v0;
    V3<T> q0 = p0, q1 = p0, q2 = p0;  // q0 = top of the position stack
    for (int i = 0; i < count; ++i) {
        CDevNode* n = nodes + first + i;
        const int op = n->op;
        const int kind = n->kind;
        float pp[8], qq[8], tt[12];
#pragma unroll
        for (int j = 0; j < 8; ++j) pp[j] = n->p[j];
#pragma unroll
        for (int j = 0; j < 8; ++j) qq[j] = n->q[j];
        if (op == 1) {  // PRIM
            V3<T> q = q0;
            if (n->ident == 0) {
#pragma unroll
                for (int j = 0; j < 12; ++j) tt[j] = n->t[j];
                q = xform<T>(tt, q0);
            } else if (n->ident == 2) {  // translation only
                q = V3<T>{q0.x + n->t[9], q0.y + n->t[10], q0.z + n->t[11]};
            }
            T d = prim_eval<T, KM>(kind, pp, qq, q, accel);
            v3 = v2; v2 = v1; v1 = v0; v0 = d;
        } else if (op == 2) {  // CSG: d1 = v1, d2 = v0
            if constexpr (KM_HAS(KM, KM_ANY_CSG)) {
                T d = csg_eval<T, KM>(kind, v1, v0, n->k, qq);
                v0 = d; v1 = v2; v2 = v3;
            }
        } else if (op == 3) {  // MOD_BEGIN
            if constexpr (KM_HAS(KM, KM_ANY_MOD)) {
                V3<T> q = mod_pos<T, KM>(kind, pp, q0);
                q2 = q1; q1 = q0; q0 = q;
            }
        } else {  // MOD_END
            if constexpr (KM_HAS(KM, KM_ANY_MOD)) {
                q0 = q1; q1 = q2;
                v0 = mod_post<T, KM>(kind, pp, q0, v0);
            }
        }
    }
    return v0;
}

// ------------------------------------------------------------------------------------------
// Philox-4x32-10 (Salmon et al., SC'11).  Counter (block, id_lo, id_hi, 0), key (seed_lo, seed_hi).
// ------------------------------------------------------------------------------------------
SMCRT_DEV void philox10(uint32_t& c0, uint32_t& c1, uint32_t& c2, uint32_t& c3, uint32_t k0, uint32_t k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c1 = (uint32_t)p1;
        c3 = (uint32_t)p0;
        c0 = n0;
        c2 = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// uniform in (0,1): odd multiple of 2^-24, bit-identical to the oracle's draw
SMCRT_DEV float u_of(uint32_t x) { return (float)(((x >> 9) << 1) | 1u) * (1.0f / 16777216.0f); }

// A packet's stream: one Philox block per event, counter (block, id_lo, id_hi, 0), key (seed).
// The events and the use of their four uniforms (the oracle's mapping): emission u0, u1 for the
// emitter, u2 the first flight's optical depth; flight end u0 albedo test (weight estimator:
// roulette), u1 cos theta, u2 azimuth, u3 the next optical depth; Fresnel u0 reflect or
// refract, u1 the resampled optical depth.  One Philox call per event keeps one inlined copy
// per event site, and no buffered draws in registers.
struct Rng {
    uint32_t block;
};

SMCRT_DEV void rng_init(Rng& g) { g.block = 0u; }

SMCRT_DEV void rng_block(Rng& g, uint64_t id, uint32_t k0, uint32_t k1, float u[4]) {
    uint32_t c0 = g.block, c1 = (uint32_t)id, c2 = (uint32_t)(id >> 32), c3 = 0u;
    philox10(c0, c1, c2, c3, k0, k1);
    g.block++;
    u[0] = u_of(c0);
    u[1] = u_of(c1);
    u[2] = u_of(c2);
    u[3] = u_of(c3);
}
