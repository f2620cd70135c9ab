// smcrt_api.hip — the C ABI of include/smcrt.h: scene flattening/validation, device set-up, the
// packet-loop launches (single scene, adjoint, batched + HIP graph), SDF evaluation and the
// host-side output helpers.
#include <hip/hip_runtime.h>
#include <hip/hiprtc.h>
#include <sys/stat.h>
#include <unistd.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <chrono>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/smcrt.h"
#include "packet_kernel.h"
#include "jit_sources.inc"  // the device headers as strings (generated at build time)

using namespace smcrt;

// ------------------------------------------------------------------------------------------
// errors
// ------------------------------------------------------------------------------------------
static thread_local std::string g_err;

static int fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

// error reporting for the library's other translation units (smcrt_config.cpp)
__attribute__((visibility("hidden"))) int smcrt_internal_fail(int code, const char* msg) {
    g_err = msg;
    return code;
}

#define HIPCHECK(x)                                                                          \
    do {                                                                                     \
        hipError_t e_ = (x);                                                                 \
        if (e_ != hipSuccess) return fail(SMCRT_EDEVICE, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

// ------------------------------------------------------------------------------------------
// small device kernels
// ------------------------------------------------------------------------------------------
__global__ void k_fill_u64(unsigned long long* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0ull;
}
__global__ void k_fill_f64(double* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0.0;
}
__global__ void k_sum_f64(const double* __restrict__ p, size_t n, double* out) {
    __shared__ double part[256];
    double s = 0.0;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) s += p[i];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int off = blockDim.x / 2; off > 0; off >>= 1) {
        if ((int)threadIdx.x < off) part[threadIdx.x] += part[threadIdx.x + off];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(out, part[0]);
}

__global__ void k_fill_u32(uint32_t* p, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) p[i] = 0u;
}

// ---- binned tally: after the packet kernel, the pages of every bin are summed in LDS ----
// Block b of the packet kernel took pages [b * ppb, b * ppb + block_pages[b]), each holding
// page_fill entries of one bin.  One workgroup here per packet-kernel block counts its pages per
// bin in LDS (at most 4096 bins) and adds the counts to the bins' totals:
constexpr int BIN_MAX = 4096;
__global__ void k_bin_count(const uint32_t* __restrict__ page_bin, const uint32_t* __restrict__ block_pages,
                            unsigned ppb, int nbins, uint32_t* __restrict__ bin_pages) {
    __shared__ uint32_t h[BIN_MAX];
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const size_t p0 = (size_t)blockIdx.x * ppb;
    const uint32_t n = block_pages[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) atomicAdd(h + page_bin[p0 + j], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (h[i]) atomicAdd(bin_pages + i, h[i]);
}
// ... then lists the pages bin by bin (bin_off: each bin's first list position): the block's run
// of each bin is reserved with one atomic, its pages placed with LDS atomics
__global__ void k_bin_list(const uint32_t* __restrict__ page_bin, const uint32_t* __restrict__ block_pages,
                           unsigned ppb, int nbins, const uint32_t* __restrict__ bin_off, uint32_t* __restrict__ bin_cur,
                           uint32_t* __restrict__ list) {
    __shared__ uint32_t h[BIN_MAX];  // counts, then the block's next list position per bin
    for (int i = threadIdx.x; i < nbins; i += blockDim.x) h[i] = 0u;
    __syncthreads();
    const size_t p0 = (size_t)blockIdx.x * ppb;
    const uint32_t n = block_pages[blockIdx.x];
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) atomicAdd(h + page_bin[p0 + j], 1u);
    __syncthreads();
    for (int i = threadIdx.x; i < nbins; i += blockDim.x)
        if (h[i]) h[i] = bin_off[i] + atomicAdd(bin_cur + i, h[i]);
    __syncthreads();
    for (uint32_t j = threadIdx.x; j < n; j += blockDim.x) list[atomicAdd(h + page_bin[p0 + j], 1u)] = (uint32_t)(p0 + j);
}
// One work item: list positions [l0, l1) of one bin's pages; `shared`: the bin has other items
struct BinWork {
    uint32_t bin, l0, l1, shared;
};
// Sum a work item's entries in an FP64 LDS tile of the bin's voxels, then add the tile to the
// grid (plain read-modify-write when the item holds the whole bin, atomics otherwise).  The
// item's page ids and fills are read into LDS first; each wavefront then streams whole pages,
// two at a time (sixteen independent 8-byte loads in flight per lane).
constexpr int BIN_VOX = 1 << SMCRT_BIN_SHIFT;
constexpr uint32_t BIN_PAGES_PER_ITEM = 2048;  // reduction work item: 1M entries
__global__ void __launch_bounds__(1024) k_bin_accum(const BinWork* __restrict__ work, const uint32_t* __restrict__ list,
                                                    const uint2* __restrict__ pool, const uint32_t* __restrict__ page_fill,
                                                    double* __restrict__ grid, size_t nvox) {
    __shared__ double t[BIN_VOX];
    __shared__ uint32_t pid[BIN_PAGES_PER_ITEM], pfill[BIN_PAGES_PER_ITEM];
    const BinWork w = work[blockIdx.x];
    const uint32_t np = w.l1 - w.l0;  // <= BIN_PAGES_PER_ITEM
    for (int i = threadIdx.x; i < BIN_VOX; i += blockDim.x) t[i] = 0.0;
    for (uint32_t i = threadIdx.x; i < np; i += blockDim.x) {
        const uint32_t p = list[w.l0 + i];
        pid[i] = p;
        pfill[i] = page_fill[p];
    }
    __syncthreads();
    // each wavefront streams two whole pages at a time: page id and fill read once per page,
    // sixteen entry loads in flight per lane
    const uint32_t lane = threadIdx.x & 63u, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
    for (uint32_t k0 = 2u * wid; k0 < np; k0 += 2u * nw) {
        uint2 e[16];
        uint32_t f[2];
#pragma unroll
        for (int q = 0; q < 2; ++q) {
            const uint32_t k = k0 + (uint32_t)q;
            f[q] = k < np ? pfill[k] : 0u;
            const uint2* src = pool + (size_t)(k < np ? pid[k] : 0u) * SMCRT_BIN_PAGE;
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const uint32_t i = lane + 64u * (uint32_t)u;
                e[8 * q + u] = i < f[q] ? src[i] : make_uint2(0u, 0u);
            }
        }
#pragma unroll
        for (int q = 0; q < 2; ++q)
#pragma unroll
            for (int u = 0; u < 8; ++u)
                if (lane + 64u * (uint32_t)u < f[q])
                    atomicAdd(t + (e[8 * q + u].x & (BIN_VOX - 1)), (double)__uint_as_float(e[8 * q + u].y));
    }
    __syncthreads();
    const size_t g0 = (size_t)w.bin << SMCRT_BIN_SHIFT;
    for (int i = threadIdx.x; i < BIN_VOX; i += blockDim.x) {
        const double v = t[i];
        if (v != 0.0 && g0 + i < nvox) {
            if (w.shared) unsafeAtomicAdd(grid + g0 + i, v);
            else grid[g0 + i] += v;
        }
    }
}

// dst[i] += sum over the partial grids c of src[c * n + i] (per-XCD tallies, fixed order)
__global__ void k_reduce_copies(double* __restrict__ dst, const double* __restrict__ src, size_t n, int copies) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        double v = 0.0;
        for (int c = 0; c < copies; ++c) v += src[(size_t)c * n + i];
        if (v != 0.0) dst[i] += v;
    }
}

// Partial grids for a tally of n doubles: how many (at least 2, else 0 = tally straight into the
// grid).  Copies per XCD from the flags' SMCRT_FLAG_GRID_COPIES field, else `per_xcd`; within 8
// GiB for one per XCD and 40 GiB for more, and within half of `avail` (the device's free memory
// plus what the copies already hold); whole multiples of the XCD count where more than one fits
// (block b tallies into copy b % copies, and blocks go to the XCDs round-robin).
static int xgrid_copies(int n_xcc, size_t n, uint32_t flags, int per_xcd, size_t avail) {
    if ((flags & SMCRT_FLAG_SHARED_GRID) || n_xcc < 2 || n == 0) return 0;
    const uint32_t f = (flags & SMCRT_FLAG_GRID_COPIES_MASK) >> SMCRT_FLAG_GRID_COPIES_SHIFT;
    const size_t per = f ? (size_t)1 << (f - 1) : (size_t)(per_xcd > 1 ? per_xcd : 1);
    size_t cap = per > 1 ? (size_t)40 << 30 : (size_t)8 << 30;  // bytes
    if (per > 1 && cap > avail / 2) cap = std::max(avail / 2, (size_t)8 << 30);
    size_t c = std::min(cap / (n * sizeof(double)), (size_t)n_xcc * per);
    if (c > (size_t)n_xcc) c -= c % (size_t)n_xcc;
    return c >= 2 ? (int)c : 0;
}

// Device memory a set of partial grids may use: free memory plus the `held` bytes their buffer
// already holds.
static size_t xgrid_avail(size_t held) {
    size_t fr = 0, tot = 0;
    if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
    return fr + held;
}

// SDF values / exact gradients of one region at n points (one point per thread; the region's
// program is uniform across the grid).
__global__ void k_evaluate(const DevNode* __restrict__ gnodes, const DevRegion* __restrict__ gregions, int region,
                           long long n, const float* __restrict__ pts, float* __restrict__ dist,
                           float* __restrict__ nrm, const float* __restrict__ accel) {
    CDevNode* nodes = (CDevNode*)gnodes;
    CDevRegion* R = (CDevRegion*)gregions + region;
    long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    float x = pts[3 * i], y = pts[3 * i + 1], z = pts[3 * i + 2];
    if (dist) dist[i] = region_eval<float>(nodes, R->first, R->count, V3<float>{x, y, z}, accel);
    if (nrm) {
        // the packet kernel's gradient (lazy through flat folds)
        const V3<float> g = region_gradient<KM_FULL>(nodes, R->first, R->count, R->flat, V3<float>{x, y, z}, accel);
        float inv = 1.0f / sqrtf(g.x * g.x + g.y * g.y + g.z * g.z);
        nrm[3 * i] = g.x * inv;
        nrm[3 * i + 1] = g.y * inv;
        nrm[3 * i + 2] = g.z * inv;
    }
}

// render_sub: minimum over every region at one voxel centre per thread (regions are
// wave-uniform: scalar program loads as in the packet kernel)
__global__ void k_render(const DevNode* __restrict__ gnodes, const DevRegion* __restrict__ gregions, int nreg,
                         int sx, int sy, int sz, float nsx, float nsy, float nsz, float wx, float wy, float wz,
                         float* __restrict__ image, const float* __restrict__ accel) {
    CDevNode* nodes = (CDevNode*)gnodes;
    CDevRegion* regions = (CDevRegion*)gregions;
    const long long n = (long long)sx * sy * sz;
    for (long long v = blockIdx.x * (long long)blockDim.x + threadIdx.x; v < n; v += (long long)gridDim.x * blockDim.x) {
        const int i = (int)(v % sx), j = (int)((v / sx) % sy), k = (int)(v / ((long long)sx * sy));
        const V3<float> p{((float)(i + 1) - nsx) * wx, ((float)(j + 1) - nsy) * wy, ((float)(k + 1) - nsz) * wz};
        float d = __builtin_inff();
        for (int r = 0; r < nreg; ++r) d = fminf(d, region_eval<float>(nodes, regions[r].first, regions[r].count, p, accel));
        image[v] = d;
    }
}

static inline size_t fill_blocks(size_t n) {
    size_t b = (n + 255) / 256;
    return b < 4096 ? (b ? b : 1) : 4096;
}

// ------------------------------------------------------------------------------------------
// scene
// ------------------------------------------------------------------------------------------
struct DevBuf {
    void* p = nullptr;
    size_t bytes = 0;
    int ensure(size_t want) {
        if (want <= bytes) return SMCRT_OK;
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
        if (hipMalloc(&p, want) != hipSuccess) return fail(SMCRT_ENOMEM, "hipMalloc(%zu) failed", want);
        bytes = want;
        return SMCRT_OK;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        bytes = 0;
    }
    template <class T> T* as() const { return reinterpret_cast<T*>(p); }
};

// Packet kernels compiled at run time for one scene's SDF program (see jit_build).
struct JitKernels {
    bool tried = false, ok = false, cached = false;
    hipModule_t mod = nullptr;
    // [waves - 2]: forward, adjoint, forward with the binned tally's stage, forward over several sets
    hipFunction_t fwd[3] = {nullptr, nullptr, nullptr}, adj[3] = {nullptr, nullptr, nullptr}, bin[3] = {nullptr, nullptr, nullptr},
                  sets[3] = {nullptr, nullptr, nullptr};
    std::string error;
    double seconds = 0.0;
};

struct smcrt_scene {
    std::vector<DevNode> nodes;
    std::vector<DevRegion> regions;
    std::vector<int> region_mat;
    std::vector<smcrt_material> mats;
    smcrt_grid grid;
    smcrt_source src;
    float eps, eps_kill;
    float circ_t[12];
    int circ_swap;
    float rdir[3];
    // polygon accelerators (floats, see polygon_eval in smcrt_device.h)
    std::vector<float> accel;
    // primitive, operator and modifier kinds the program uses (KM_* bits of smcrt_device.h)
    uint32_t kind_mask = 0;
    JitKernels jit;
    std::string kernel_name;  // smcrt_scene_kernel's answer
    // detectors (prepared for the device) and the offset of each one's bins
    std::vector<smcrt_detector> dets;
    std::vector<DevDet> ddets;
    size_t det_bins = 0;
    // LDS tally tile (voxel box) for register occupancies of 2, 3 and 4 blocks per CU, chosen from a
    // pilot run on first use
    bool tile_ready = false;
    int tile_t0[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}}, tile_tn[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
    // binned tally plan, from the same pilot run, per occupancy of 2, 3 and 4 blocks per CU: the
    // hot bins' slots (SMCRT_BIN_COLD for the others), their number, the tile that fits beside
    // their LDS stage, the share of the pilot's path length in the hot bins and in the direct
    // tally's tile; tally entries per packet (merged chords, no tile) in the pilot
    int nbins = 0;
    std::vector<uint16_t> binmap[3];
    int nhot[3] = {0, 0, 0};
    int btile_t0[3][3] = {}, btile_tn[3][3] = {};
    double hot_frac[3] = {0.0, 0.0, 0.0}, tile_capture[3] = {0.0, 0.0, 0.0};
    double pilot_entries = 0.0, pilot_entries_per_sweep = 0.0;
    int binmap_up = -1;  // which occupancy's map d_binmap holds
    bool last_binned = false;  // the last forward launch ran the binned tally
    std::vector<uint32_t> h_bin_pages, h_bin_off;
    std::vector<BinWork> h_bin_work;
    // device
    int device = -1;
    int num_cu = 0;
    int n_xcc = 1;  // XCDs of the device (one FP64 partial grid each)
    bool uploaded = false;
    bool mats_dirty = true;
    hipStream_t stream = nullptr;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    DevBuf d_nodes, d_regions, d_mats, d_region_mat, d_mat_mus, d_work, d_stats, d_moments,
        d_grad, d_setinfo, d_jmean, d_weights, d_trace, d_pts, d_out, d_sum, d_cold, d_det, d_det_counts, d_accel,
        d_adj_mats, d_adj_mus, d_xgrid, d_binmap, d_pool, d_page_bin, d_page_fill, d_block_pages, d_bin_pages,
        d_bin_off, d_bin_cur, d_bin_list, d_bin_work;
    // host copies of the region-indexed material table and per-material mu_s the allocation-free
    // entry point uploads from (kept alive across its asynchronous copies; rewritten only when
    // the materials change)
    std::vector<DevMat> na_mats;
    std::vector<float> na_mus;
};

// Per-node constants the device would otherwise recompute (in VALU) at every evaluation, from
// the FP64 parameters: segment/capsule/cylinder b - a, |b - a|^2 and its reciprocal, r |b - a|^2
// and |b - a|^2 / 2; SmoothUnion 1/k and k/6.
static void derive_node(DevNode& dn, const smcrt_node& n) {
    for (int j = 0; j < 8; ++j) dn.q[j] = 0.f;
    if (n.op == SMCRT_OP_PRIM && (n.kind == SMCRT_SEGMENT || n.kind == SMCRT_CAPSULE || n.kind == SMCRT_CYLINDER)) {
        const double bx = n.p[3] - n.p[0], by = n.p[4] - n.p[1], bz = n.p[5] - n.p[2];
        const double baba = bx * bx + by * by + bz * bz;
        dn.q[0] = (float)bx;
        dn.q[1] = (float)by;
        dn.q[2] = (float)bz;
        dn.q[3] = (float)baba;
        dn.q[4] = (float)(1.0 / baba);
        dn.q[5] = (float)(n.p[6] * baba);
        dn.q[6] = (float)(0.5 * baba);
        // b - a along one coordinate axis (the other two components exactly 0): the cylinder's
        // axis-aligned evaluation
        if (n.kind == SMCRT_CYLINDER && baba > 0.0) {
            const int nz = (bx != 0.0) + (by != 0.0) + (bz != 0.0);
            if (nz == 1) dn.q[7] = bx != 0.0 ? 1.f : (by != 0.0 ? 2.f : 3.f);
        }
    } else if (n.op == SMCRT_OP_CSG && n.kind == SMCRT_SMOOTH_UNION) {
        dn.q[0] = (float)(1.0 / n.k);
        dn.q[1] = (float)(n.k / 6.0);
    }
}

// ---- polygon accelerator (host build) ------------------------------------------------------
static double seg_dist(double px, double py, double ax, double ay, double ex, double ey) {
    const double ee = ex * ex + ey * ey;
    double t = ((px - ax) * ex + (py - ay) * ey) / ee;
    t = t < 0 ? 0 : (t > 1 ? 1 : t);
    const double bx = px - ax - t * ex, by = py - ay - t * ey;
    return std::sqrt(bx * bx + by * by);
}

// even-odd crossing test of a horizontal ray from (px, py) to +x
static bool point_in_polygon(const std::vector<double>& X, const std::vector<double>& Y, double px, double py) {
    bool in = false;
    const size_t n = X.size();
    for (size_t i = 0, j = n - 1; i < n; j = i++) {
        if ((Y[i] > py) != (Y[j] > py) && px < (X[j] - X[i]) * (py - Y[i]) / (Y[j] - Y[i]) + X[i]) in = !in;
    }
    return in;
}

// Build the accelerator of one polygon (n vertices at xy) and append it to A; returns its base
// (a multiple of 4 floats, so every record the device reads as a 16-byte vector is aligned).
// Grid cells are about one mean edge length (8..512 cells across the polygon).
static int build_polygon_accel(const double* xy, int n, std::vector<float>& A) {
    std::vector<double> X(n), Y(n);
    for (int i = 0; i < n; ++i) { X[i] = xy[2 * i]; Y[i] = xy[2 * i + 1]; }
    double area2 = 0;
    for (int i = 0; i < n; ++i) area2 += X[i] * Y[(i + 1) % n] - X[(i + 1) % n] * Y[i];
    if (area2 < 0) { std::reverse(X.begin(), X.end()); std::reverse(Y.begin(), Y.end()); }
    std::vector<double> EX(n), EY(n), NX(n), NY(n);
    double xmin = X[0], xmax = X[0], ymin = Y[0], ymax = Y[0], perim = 0;
    for (int i = 0; i < n; ++i) {
        EX[i] = X[(i + 1) % n] - X[i];
        EY[i] = Y[(i + 1) % n] - Y[i];
        const double l = std::sqrt(EX[i] * EX[i] + EY[i] * EY[i]);
        perim += l;
        NX[i] = EY[i] / l;  // outward normal of a counter-clockwise edge
        NY[i] = -EX[i] / l;
        xmin = std::fmin(xmin, X[i]); xmax = std::fmax(xmax, X[i]);
        ymin = std::fmin(ymin, Y[i]); ymax = std::fmax(ymax, Y[i]);
    }
    const double W = std::fmax(xmax - xmin, ymax - ymin);
    const double cell = std::fmax(W / 512.0, std::fmin(W / 8.0, perim / n));
    const double m = 2.0 * cell;
    const double gx0 = xmin - m, gy0 = ymin - m;
    const int nx = (int)std::ceil((xmax + m - gx0) / cell), ny = (int)std::ceil((ymax + m - gy0) / cell);
    const double r = cell * std::sqrt(0.5);
    const double slack = 1e-5 * std::fmax(W, 1.0);
    A.resize((A.size() + 3) & ~(size_t)3, 0.f);
    const int base = (int)A.size();
    const int eoff = base + 16, coff = eoff + 8 * n, loff = coff + 4 * nx * ny;
    A.resize(loff, 0.f);
    auto put_int = [](float* f, int v) { memcpy(f, &v, sizeof v); };
    float* H = A.data() + base;
    H[0] = (float)nx; H[1] = (float)ny; H[2] = (float)gx0; H[3] = (float)gy0;
    H[4] = (float)(1.0 / cell); H[5] = (float)cell;
    H[6] = (float)(gx0 + nx * cell); H[7] = (float)(gy0 + ny * cell);
    H[8] = (float)m; H[9] = (float)eoff; H[10] = (float)coff; H[11] = (float)loff; H[12] = (float)n;
    for (int i = 0; i < n; ++i) {
        float* E = A.data() + eoff + 8 * i;
        const int ip = (i + n - 1) % n, in = (i + 1) % n;
        double n0x = NX[ip] + NX[i], n0y = NY[ip] + NY[i];  // vertex pseudo-normals (bisectors)
        double n1x = NX[i] + NX[in], n1y = NY[i] + NY[in];
        double l0 = std::sqrt(n0x * n0x + n0y * n0y), l1 = std::sqrt(n1x * n1x + n1y * n1y);
        if (l0 < 1e-12) { n0x = NX[i]; n0y = NY[i]; l0 = 1; }
        if (l1 < 1e-12) { n1x = NX[i]; n1y = NY[i]; l1 = 1; }
        E[0] = (float)X[i]; E[1] = (float)Y[i]; E[2] = (float)EX[i]; E[3] = (float)EY[i];
        E[4] = (float)(n0x / l0); E[5] = (float)(n0y / l0); E[6] = (float)(n1x / l1); E[7] = (float)(n1y / l1);
    }
    std::vector<double> dist(n);
    std::vector<float> lists;
    std::vector<int> cand;
    for (int iy = 0; iy < ny; ++iy)
        for (int ix = 0; ix < nx; ++ix) {
            const double cx = gx0 + (ix + 0.5) * cell, cy = gy0 + (iy + 0.5) * cell;
            double D = 1e300;
            for (int i = 0; i < n; ++i) {
                dist[i] = seg_dist(cx, cy, X[i], Y[i], EX[i], EY[i]);
                D = std::fmin(D, dist[i]);
            }
            const bool in = point_in_polygon(X, Y, cx, cy);
            float* Cc = A.data() + coff + 4 * (ix + nx * iy);
            Cc[2] = (float)(in ? -D : D);
            if (D > 2.0 * r + slack) {  // far: no edge within the cell
                put_int(Cc + 0, 0);
                put_int(Cc + 1, 0);
                continue;
            }
            cand.clear();
            for (int i = 0; i < n; ++i)
                if (dist[i] <= D + 2.0 * r + slack) cand.push_back(eoff + 8 * i);
            // pairs of edge offsets, an odd list padded with its first edge (a repeated edge never
            // replaces the nearest: the device keeps the first of equal distances); the first pair
            // sits in the cell record, the rest in the lists
            if (cand.size() % 2) cand.push_back(cand[0]);
            const int np = (int)cand.size() / 2;
            put_int(Cc + 0, np > 1 ? loff + (int)lists.size() : 0);
            put_int(Cc + 1, np);
            put_int(Cc + 2, cand[0]);
            put_int(Cc + 3, cand[1]);
            for (size_t k = 2; k < cand.size(); ++k) {
                float f;
                memcpy(&f, &cand[k], sizeof f);
                lists.push_back(f);
            }
        }
    A.insert(A.end(), lists.begin(), lists.end());
    // every offset the device follows must stay inside this polygon's records
    const size_t end = A.size();
    auto get_int = [&](size_t i) { int v; memcpy(&v, A.data() + i, sizeof v); return v; };
    auto edge_ok = [&](int e) { return e >= eoff && e + 8 <= coff && (e - eoff) % 8 == 0; };
    for (int c = 0; c < nx * ny; ++c) {
        const size_t ci = (size_t)coff + 4 * (size_t)c;
        const int rest = get_int(ci), np = get_int(ci + 1);
        if (np == 0) continue;
        if (np < 0 || !edge_ok(get_int(ci + 2)) || !edge_ok(get_int(ci + 3))) return -1;
        if (np == 1) continue;
        if (rest < loff || (rest - loff) % 2 != 0 || (size_t)rest + 2 * (size_t)(np - 1) > end) return -1;
        for (int k = 0; k < 2 * (np - 1); ++k)
            if (!edge_ok(get_int((size_t)rest + k))) return -1;
    }
    return base;
}

static bool finite_all(const double* v, int n);

// A detector as the reference constructs it (src/detectors/detectors.f90: init_circle_dect,
// init_annulus_dect, init_camera): bin counts and widths, the circle's normalised direction
// (src/parse/parse_detectors.f90: dir = dir%magnitude()), the camera's edges and unit normal.
static int prepare_detector(const smcrt_detector& in, int idx, DevDet* out, size_t* bins) {
    DevDet d;
    memset(&d, 0, sizeof d);
    if (in.kind < SMCRT_DET_CIRCLE || in.kind > SMCRT_DET_CAMERA)
        return fail(SMCRT_EINVAL, "detector %d: Invalid detector type. Valid types are [circle, annulus, camera]", idx);
    if (in.nbins < 0) return fail(SMCRT_EINVAL, "detector %d: nbins must be >= 0", idx);
    if (!finite_all(in.pos, 3) || !finite_all(in.dir, 3) || !finite_all(in.p2, 3) || !finite_all(in.p3, 3) ||
        !std::isfinite(in.radius1) || !std::isfinite(in.radius2) || !std::isfinite(in.maxval))
        return fail(SMCRT_EINVAL, "detector %d: non-finite parameter", idx);
    d.kind = in.kind;
    d.layer0 = in.layer - 1;
    d.nb = in.nbins + 1;
    for (int j = 0; j < 3; ++j) d.p[j] = in.pos[j];
    if (in.kind == SMCRT_DET_CIRCLE) {
        const double l = std::sqrt(in.dir[0] * in.dir[0] + in.dir[1] * in.dir[1] + in.dir[2] * in.dir[2]);
        for (int j = 0; j < 3; ++j) d.n[j] = in.dir[j] / l;
        d.a = in.radius1;
        d.bw = in.nbins == 0 ? 1.0 : in.maxval / (double)(in.nbins - 1);
        *bins = (size_t)d.nb;
    } else if (in.kind == SMCRT_DET_ANNULUS) {
        if (!(in.radius2 > in.radius1))
            return fail(SMCRT_EINVAL, "detector %d: Radii are invalid (Expected radius2 (%g) > radius 1 (%g))", idx,
                        in.radius2, in.radius1);
        d.a = in.radius1;
        d.b = in.radius2;
        d.bw = in.nbins == 0 ? 1.0 : in.maxval / (double)in.nbins;
        *bins = (size_t)d.nb;
    } else {
        double e1[3], e2[3];
        for (int j = 0; j < 3; ++j) { e1[j] = in.p2[j] - in.pos[j]; e2[j] = in.p3[j] - in.pos[j]; }
        // n = e2 x e1, normalised
        double n[3] = {e2[1] * e1[2] - e2[2] * e1[1], e2[2] * e1[0] - e2[0] * e1[2], e2[0] * e1[1] - e2[1] * e1[0]};
        const double nl = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
        if (!(nl > 0)) return fail(SMCRT_EINVAL, "detector %d: camera corners are collinear", idx);
        for (int j = 0; j < 3; ++j) { d.e1[j] = e1[j]; d.e2[j] = e2[j]; d.n[j] = n[j] / nl; }
        d.a = std::sqrt(e1[0] * e1[0] + e1[1] * e1[1] + e1[2] * e1[2]);
        d.b = std::sqrt(e2[0] * e2[0] + e2[1] * e2[1] + e2[2] * e2[2]);
        d.bw = in.nbins == 0 ? 1.0 : in.maxval / (double)d.nb;
        d.bwy = d.bw;
        *bins = (size_t)d.nb * (size_t)d.nb;
    }
    *out = d;
    return SMCRT_OK;
}

static int n_values_pushed(int op) { return op == SMCRT_OP_PRIM ? 1 : (op == SMCRT_OP_CSG ? -1 : 0); }

static bool finite_all(const double* v, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(v[i])) return false;
    return true;
}

static void region_mats(const smcrt_scene* s, const smcrt_material* mats, DevMat* out) {
    for (size_t r = 0; r < s->regions.size(); ++r) {
        const smcrt_material& m = mats[s->region_mat[r]];
        DevMat d;
        d.kappa = (float)(m.mus + m.mua);
        d.albedo = (m.mua < 1e-9) ? 1.0f : (float)(m.mus / (m.mus + m.mua));
        d.hgg = (float)m.hgg;
        d.g2 = (float)(m.hgg * m.hgg);
        d.n = (float)m.n;
        d.mus = (float)m.mus;
        d.mua = (float)m.mua;
        d.mat = s->region_mat[r];
        out[r] = d;
    }
}

static int validate_materials(const smcrt_material* m, int n) {
    for (int i = 0; i < n; ++i) {
        const smcrt_material& x = m[i];
        if (!(std::isfinite(x.mus) && std::isfinite(x.mua) && std::isfinite(x.hgg) && std::isfinite(x.n)))
            return fail(SMCRT_EINVAL, "material %d has a non-finite property", i);
        if (x.mus < 0 || x.mua < 0) return fail(SMCRT_EINVAL, "material %d: mus and mua must be >= 0", i);
        if (!(x.hgg > -1.0 && x.hgg < 1.0)) return fail(SMCRT_EINVAL, "material %d: hgg must lie in (-1, 1)", i);
        if (!(x.n > 0)) return fail(SMCRT_EINVAL, "material %d: refractive index must be > 0", i);
    }
    return SMCRT_OK;
}

extern "C" int smcrt_abi_version(void) { return SMCRT_ABI_VERSION; }

extern "C" int smcrt_last_error(char* buf, size_t len) {
    if (!buf || len == 0) return SMCRT_EINVAL;
    snprintf(buf, len, "%s", g_err.c_str());
    return SMCRT_OK;
}

extern "C" int64_t smcrt_sizeof(const char* name) {
    if (!name) return -1;
    struct { const char* n; int64_t s; } t[] = {
        {"node", sizeof(smcrt_node)}, {"region", sizeof(smcrt_region)},
        {"material", sizeof(smcrt_material)}, {"grid", sizeof(smcrt_grid)},
        {"source", sizeof(smcrt_source)}, {"scene_desc", sizeof(smcrt_scene_desc)},
        {"run_opts", sizeof(smcrt_run_opts)}, {"stats", sizeof(smcrt_stats)},
        {"tallies", sizeof(smcrt_tallies)}, {"grad_out", sizeof(smcrt_grad_out)},
        {"batch_set", sizeof(smcrt_batch_set)}, {"detector", sizeof(smcrt_detector)},
        {"config", sizeof(smcrt_config)}};
    for (auto& e : t)
        if (!strcmp(e.n, name)) return e.s;
    return -1;
}

extern "C" int smcrt_scene_create(const smcrt_scene_desc* d, smcrt_scene** out) {
    if (!d || !out) return fail(SMCRT_EINVAL, "null descriptor or output handle");
    *out = nullptr;
    if (d->abi_version != SMCRT_ABI_VERSION)
        return fail(SMCRT_EINVAL, "abi_version %u, library speaks %d", d->abi_version, SMCRT_ABI_VERSION);
    if (d->n_regions < 1 || d->n_regions > SMCRT_MAX_REGIONS)
        return fail(SMCRT_ERANGE, "n_regions %d outside [1, %d]", d->n_regions, SMCRT_MAX_REGIONS);
    if (d->n_nodes < 1 || d->n_nodes > SMCRT_MAX_NODES)
        return fail(SMCRT_ERANGE, "n_nodes %d outside [1, %d]", d->n_nodes, SMCRT_MAX_NODES);
    if (d->n_materials < 1 || d->n_materials > SMCRT_MAX_MATERIALS)
        return fail(SMCRT_ERANGE, "n_materials %d outside [1, %d]", d->n_materials, SMCRT_MAX_MATERIALS);
    if (!d->nodes || !d->regions || !d->materials) return fail(SMCRT_EINVAL, "null nodes/regions/materials");
    int rc = validate_materials(d->materials, d->n_materials);
    if (rc) return rc;
    const smcrt_grid& g = d->grid;
    if (g.nx < 1 || g.ny < 1 || g.nz < 1) return fail(SMCRT_EINVAL, "grid counts must be >= 1");
    // the kernel addresses grid elements with 32-bit indices
    if ((uint64_t)g.nx * (uint64_t)g.ny * (uint64_t)g.nz >= (1ull << 32))
        return fail(SMCRT_ERANGE, "grid of %d x %d x %d voxels: at most 2^32 - 1 voxels", g.nx, g.ny, g.nz);
    if ((double)g.nx * g.ny * g.nz > 2147483647.0) return fail(SMCRT_ERANGE, "grid has more than 2^31-1 voxels");
    if (!(g.xmax > 0 && g.ymax > 0 && g.zmax > 0)) return fail(SMCRT_EINVAL, "grid half sizes must be > 0");

    smcrt_scene* s = new smcrt_scene();
    s->grid = g;
    s->src = d->source;
    s->mats.assign(d->materials, d->materials + d->n_materials);
    for (int r = 0; r < d->n_regions; ++r) {
        const smcrt_region& R = d->regions[r];
        if (R.n_nodes < 1 || R.first_node < 0 || R.first_node + R.n_nodes > d->n_nodes) {
            delete s;
            return fail(SMCRT_EINVAL, "region %d node range [%d, %d) outside the program", r, R.first_node,
                        R.first_node + R.n_nodes);
        }
        if (R.material < 0 || R.material >= d->n_materials) {
            delete s;
            return fail(SMCRT_EINVAL, "region %d material %d out of range", r, R.material);
        }
        int nv = 0, np = 1;
        DevRegion dr{(int)s->nodes.size(), R.n_nodes, R.material, 0};  // flat set below
        for (int i = 0; i < R.n_nodes; ++i) {
            const smcrt_node& n = d->nodes[R.first_node + i];
            bool ok_kind = false;
            switch (n.op) {
            case SMCRT_OP_PRIM: ok_kind = n.kind >= SMCRT_SPHERE && n.kind <= SMCRT_POLYGON; break;
            case SMCRT_OP_CSG: ok_kind = n.kind >= SMCRT_UNION && n.kind <= SMCRT_INTERSECTION; break;
            case SMCRT_OP_MOD_BEGIN:
            case SMCRT_OP_MOD_END: ok_kind = n.kind >= SMCRT_EXTRUDE && n.kind <= SMCRT_BEND; break;
            default: break;
            }
            if (!ok_kind) {
                delete s;
                return fail(SMCRT_EINVAL, "region %d node %d: bad op/kind (%d, %d)", r, i, n.op, n.kind);
            }
            s->kind_mask |= n.op == SMCRT_OP_PRIM ? KM_PRIM(n.kind) : (n.op == SMCRT_OP_CSG ? KM_CSG(n.kind) : KM_MOD(n.kind));
            if (!finite_all(n.p, 8) || !finite_all(n.t, 12) || !std::isfinite(n.k)) {
                delete s;
                return fail(SMCRT_EINVAL, "region %d node %d: non-finite parameter", r, i);
            }
            if (n.op == SMCRT_OP_CSG && n.kind == SMCRT_SMOOTH_UNION && !(n.k > 0)) {
                delete s;
                return fail(SMCRT_EINVAL, "region %d node %d: SmoothUnion needs k > 0", r, i);
            }
            if (n.op == SMCRT_OP_CSG && nv < 2) {
                delete s;
                return fail(SMCRT_EINVAL, "region %d node %d: operator needs two distances", r, i);
            }
            if (n.op == SMCRT_OP_MOD_END && (np < 2 || nv < 1)) {
                delete s;
                return fail(SMCRT_EINVAL, "region %d node %d: unmatched modifier end", r, i);
            }
            nv += n_values_pushed(n.op);
            if (n.op == SMCRT_OP_MOD_BEGIN) ++np;
            if (n.op == SMCRT_OP_MOD_END) --np;
            if (nv > SMCRT_MAX_VALUE_STACK || np > SMCRT_MAX_POS_STACK) {
                delete s;
                return fail(SMCRT_ERANGE, "region %d: program deeper than the evaluation stacks (%d/%d)", r,
                            SMCRT_MAX_VALUE_STACK, SMCRT_MAX_POS_STACK);
            }
            DevNode dn;
            memset(&dn, 0, sizeof dn);
            dn.op = n.op;
            dn.kind = n.kind;
            dn.k = (float)n.k;
            for (int j = 0; j < 8; ++j) dn.p[j] = (float)n.p[j];
            derive_node(dn, n);
            if (n.op == SMCRT_OP_PRIM && n.kind == SMCRT_POLYGON) {
                const double f = n.p[0], c = n.p[1];
                if (!(f >= 0 && c >= 3 && f == std::floor(f) && c == std::floor(c) && f + c <= d->n_poly_vertices) ||
                    !d->poly_vertices) {
                    delete s;
                    return fail(SMCRT_EINVAL, "region %d node %d: polygon vertices [%g, %g) outside the %d given", r, i,
                                f, f + c, d->n_poly_vertices);
                }
                const double* xy = d->poly_vertices + 2 * (size_t)f;
                for (int v = 0; v < (int)c; ++v) {
                    const double ex = xy[2 * ((v + 1) % (int)c)] - xy[2 * v], ey = xy[2 * ((v + 1) % (int)c) + 1] - xy[2 * v + 1];
                    if (!std::isfinite(xy[2 * v]) || !std::isfinite(xy[2 * v + 1]) || !(ex * ex + ey * ey > 0)) {
                        delete s;
                        return fail(SMCRT_EINVAL, "region %d node %d: polygon vertex %d non-finite or repeated", r, i, v);
                    }
                }
                if (s->accel.size() + 20 + 8 * c > (1u << 24)) {
                    delete s;
                    return fail(SMCRT_ERANGE, "polygon accelerators exceed 2^24 floats");
                }
                const int pb = build_polygon_accel(xy, (int)c, s->accel);
                if (pb < 0) {
                    delete s;
                    return fail(SMCRT_EINVAL, "region %d node %d: polygon accelerator failed its index check", r, i);
                }
                dn.p[0] = (float)pb;
                // the accelerator's grid rectangle and margin (gx0, gy0, gx1, gy1, margin): constants
                // of the gathered evaluation in the scene's compiled kernel
                const float* H = s->accel.data() + pb;
                dn.q[0] = H[2];
                dn.q[1] = H[3];
                dn.q[2] = H[6];
                dn.q[3] = H[7];
                dn.q[4] = H[8];
                if (s->accel.size() > (1u << 24)) {
                    delete s;
                    return fail(SMCRT_ERANGE, "polygon accelerators exceed 2^24 floats");
                }
            }
            bool ident = true, linear_ident = true;
            static const double I12[12] = {1, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0};
            for (int j = 0; j < 12; ++j) {
                dn.t[j] = (float)n.t[j];
                if (n.t[j] != I12[j]) {
                    ident = false;
                    if (j < 9) linear_ident = false;
                }
            }
            // 1: identity; 2: translation only (q = p + t, the same rounding as the full
            // product for finite p: the unit and zero products are exact); 0: general
            dn.ident = ident ? 1 : (linear_ident ? 2 : 0);
            s->nodes.push_back(dn);
        }
        if (nv != 1 || np != 1) {
            delete s;
            return fail(SMCRT_EINVAL, "region %d: program leaves %d distances and %d positions (want 1, 1)", r, nv, np);
        }
        // flat fold of primitives (P P CSG P CSG ...) by one union or SmoothUnion operator
        if (R.n_nodes >= 3 && R.n_nodes % 2 == 1) {
            const smcrt_node* N = d->nodes + R.first_node;
            const int kind = N[2].kind;
            bool flat = (kind == SMCRT_UNION || kind == SMCRT_SMOOTH_UNION) && N[0].op == SMCRT_OP_PRIM;
            for (int i2 = 1; flat && i2 < R.n_nodes; ++i2) {
                const bool want_csg = (i2 % 2) == 0;
                flat = want_csg ? (N[i2].op == SMCRT_OP_CSG && N[i2].kind == kind && N[i2].k == N[2].k)
                                : N[i2].op == SMCRT_OP_PRIM;
            }
            if (flat) dr.flat = kind == SMCRT_UNION ? 1 : 2;
        }
        s->regions.push_back(dr);
        s->region_mat.push_back(R.material);
    }

    // source
    const smcrt_source& src = d->source;
    if (src.kind < SMCRT_SRC_PENCIL || src.kind > SMCRT_SRC_POINT) {
        delete s;
        return fail(SMCRT_EINVAL, "No such source! (kind %d)", src.kind);
    }
    double dn = std::sqrt(src.dir[0] * src.dir[0] + src.dir[1] * src.dir[1] + src.dir[2] * src.dir[2]);
    if (src.kind != SMCRT_SRC_POINT && std::fabs(dn - 1.0) > 1e-6) {
        delete s;
        return fail(SMCRT_EINVAL, "source direction must be a unit vector (|dir| = %g)", dn);
    }
    if (src.kind == SMCRT_SRC_CIRCULAR && !(src.radius >= 0)) {
        delete s;
        return fail(SMCRT_EINVAL, "circular source radius must be >= 0");
    }
    // circular: t = rotationAlign(a, dir) . translate(-position) (src/photon.f90:238-266)
    {
        double b[3] = {src.dir[0], src.dir[1], src.dir[2]};
        if (dn > 0) for (double& v : b) v /= dn;
        double a[3] = {1, 0, 0};
        s->circ_swap = 0;
        if (std::fabs(b[0]) == 1.0 && b[1] == 0.0 && b[2] == 0.0) {
            a[0] = 0; a[2] = 1;
            s->circ_swap = 1;
        }
        double v[3] = {a[1] * b[2] - a[2] * b[1], -a[0] * b[2] + a[2] * b[0], a[0] * b[1] - a[1] * b[0]};
        double c = a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
        double k = 1.0 / (1.0 + c);
        double vx[3][3] = {{0, v[2], -v[1]}, {-v[2], 0, v[0]}, {v[1], -v[0], 0}};  // [row][col]
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s2 = 0;
                for (int m = 0; m < 3; ++m) s2 += vx[i][m] * vx[m][j];
                s->circ_t[i * 3 + j] = (float)((i == j ? 1.0 : 0.0) + vx[i][j] + s2 * k);
            }
        for (int j = 0; j < 3; ++j) s->circ_t[9 + j] = (float)(-src.pos[j]);
    }
    if (src.kind == SMCRT_SRC_POINT) { s->rdir[0] = 0; s->rdir[1] = 0; s->rdir[2] = -1; }
    else for (int j = 0; j < 3; ++j) s->rdir[j] = (float)(src.dir[j] / dn);

    // detectors
    if (d->n_detectors < 0 || d->n_detectors > SMCRT_MAX_DETECTORS) {
        delete s;
        return fail(SMCRT_ERANGE, "n_detectors %d outside [0, %d]", d->n_detectors, SMCRT_MAX_DETECTORS);
    }
    if (d->n_detectors > 0 && !d->detectors) {
        delete s;
        return fail(SMCRT_EINVAL, "null detectors");
    }
    for (int i = 0; i < d->n_detectors; ++i) {
        DevDet dd;
        size_t nb = 0;
        int rc2 = prepare_detector(d->detectors[i], i, &dd, &nb);
        if (rc2) {
            delete s;
            return rc2;
        }
        dd.off = s->det_bins;
        s->det_bins += nb;
        s->dets.push_back(d->detectors[i]);
        s->ddets.push_back(dd);
    }

    // epsilons: the march/crossing epsilon is scaled to FP32 resolution at the scene's size
    // (2^-17 of the largest grid half-size, about 64 ulp); the kill threshold after emission or
    // a scatter keeps the reference's 1e-8 (src/inttau2.f90:44)
    double scale = std::fmax(g.xmax, std::fmax(g.ymax, g.zmax));
    s->eps = (float)std::fmax(1e-8, scale * (1.0 / 131072.0));
    s->eps_kill = 1e-8f;
    *out = s;
    return SMCRT_OK;
}

extern "C" int smcrt_scene_destroy(smcrt_scene* s) {
    if (!s) return SMCRT_OK;
    if (s->device >= 0) {
        (void)hipSetDevice(s->device);
        DevBuf* bufs[] = {&s->d_nodes, &s->d_regions, &s->d_mats, &s->d_region_mat, &s->d_mat_mus, &s->d_work,
                          &s->d_stats, &s->d_moments, &s->d_grad, &s->d_setinfo, &s->d_jmean, &s->d_weights, &s->d_xgrid,
                          &s->d_trace, &s->d_pts, &s->d_out, &s->d_sum, &s->d_cold, &s->d_det,
                          &s->d_det_counts, &s->d_accel, &s->d_adj_mats, &s->d_adj_mus, &s->d_binmap, &s->d_pool,
                          &s->d_page_bin, &s->d_page_fill, &s->d_block_pages, &s->d_bin_pages, &s->d_bin_off,
                          &s->d_bin_cur, &s->d_bin_list, &s->d_bin_work};
        for (DevBuf* b : bufs) b->release();
        if (s->ev0) (void)hipEventDestroy(s->ev0);
        if (s->ev1) (void)hipEventDestroy(s->ev1);
        if (s->stream) (void)hipStreamDestroy(s->stream);
        if (s->jit.mod) (void)hipModuleUnload(s->jit.mod);
    }
    delete s;
    return SMCRT_OK;
}

static int ensure_device(smcrt_scene* s);

extern "C" int smcrt_render(smcrt_scene* s, const int32_t samples[3], float* image) {
    if (!s || !samples || !image) return fail(SMCRT_EINVAL, "null scene, samples or image");
    for (int a = 0; a < 3; ++a)
        if (samples[a] < 1) return fail(SMCRT_EINVAL, "render samples must be >= 1 (axis %d: %d)", a, samples[a]);
    const size_t nv = (size_t)samples[0] * samples[1] * samples[2];
    if (nv > ((size_t)1 << 31)) return fail(SMCRT_ERANGE, "render grid has more than 2^31 voxels");
    int rc = ensure_device(s);
    if (rc) return rc;
    if ((rc = s->d_out.ensure(nv * sizeof(float)))) return rc;
    // ns = nint(samples / 2) (Fortran nint: half away from zero), wid = extent / ns
    const double ext[3] = {s->grid.xmax, s->grid.ymax, s->grid.zmax};
    float ns[3], wid[3];
    for (int a = 0; a < 3; ++a) {
        ns[a] = (float)std::floor(samples[a] / 2.0 + 0.5);
        wid[a] = (float)(ext[a] / (double)ns[a]);
    }
    const unsigned blocks = (unsigned)std::min<size_t>((nv + 255) / 256, 65536);
    hipLaunchKernelGGL(k_render, dim3(blocks), dim3(256), 0, s->stream, s->d_nodes.as<DevNode>(),
                       s->d_regions.as<DevRegion>(), (int)s->regions.size(), samples[0], samples[1], samples[2],
                       ns[0], ns[1], ns[2], wid[0], wid[1], wid[2], s->d_out.as<float>(),
                       s->accel.empty() ? nullptr : s->d_accel.as<float>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    HIPCHECK(hipMemcpy(image, s->d_out.p, nv * sizeof(float), hipMemcpyDeviceToHost));
    return SMCRT_OK;
}

extern "C" int64_t smcrt_detector_bins(const smcrt_scene* s, int32_t i) {
    if (!s || i < 0 || i >= (int32_t)s->ddets.size()) return 0;
    const size_t end = (size_t)i + 1 < s->ddets.size() ? s->ddets[i + 1].off : s->det_bins;
    return (int64_t)(end - s->ddets[i].off);
}

extern "C" int smcrt_scene_set_materials(smcrt_scene* s, const smcrt_material* m, int32_t n) {
    if (!s || !m) return fail(SMCRT_EINVAL, "null scene or materials");
    if (n != (int32_t)s->mats.size()) return fail(SMCRT_EINVAL, "expected %zu materials, got %d", s->mats.size(), n);
    int rc = validate_materials(m, n);
    if (rc) return rc;
    s->mats.assign(m, m + n);
    s->mats_dirty = true;
    return SMCRT_OK;
}

static int ensure_device(smcrt_scene* s) {
    if (s->device < 0) {
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return fail(SMCRT_EDEVICE, "no HIP device available");
        int dev = 0;
        HIPCHECK(hipGetDevice(&dev));
        s->device = dev;
        HIPCHECK(hipDeviceGetAttribute(&s->num_cu, hipDeviceAttributeMultiprocessorCount, dev));
        if (hipDeviceGetAttribute(&s->n_xcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || s->n_xcc < 1)
            s->n_xcc = 1;
        HIPCHECK(hipStreamCreate(&s->stream));
        HIPCHECK(hipEventCreate(&s->ev0));
        HIPCHECK(hipEventCreate(&s->ev1));
    }
    HIPCHECK(hipSetDevice(s->device));
    if (!s->uploaded) {
        int rc;
        if ((rc = s->d_nodes.ensure(s->nodes.size() * sizeof(DevNode)))) return rc;
        if ((rc = s->d_regions.ensure(s->regions.size() * sizeof(DevRegion)))) return rc;
        if ((rc = s->d_region_mat.ensure(s->region_mat.size() * sizeof(int)))) return rc;
        if ((rc = s->d_work.ensure(8 * sizeof(unsigned long long)))) return rc;
        if ((rc = s->d_stats.ensure((NOUT + NEV) * sizeof(unsigned long long)))) return rc;
        if ((rc = s->d_moments.ensure(24 * sizeof(double)))) return rc;
        if ((rc = s->d_setinfo.ensure(4 * sizeof(unsigned long long)))) return rc;
        if ((rc = s->d_sum.ensure(sizeof(double)))) return rc;
        HIPCHECK(hipMemcpy(s->d_nodes.p, s->nodes.data(), s->nodes.size() * sizeof(DevNode), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(s->d_regions.p, s->regions.data(), s->regions.size() * sizeof(DevRegion), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(s->d_region_mat.p, s->region_mat.data(), s->region_mat.size() * sizeof(int), hipMemcpyHostToDevice));
        if (!s->accel.empty()) {
            if ((rc = s->d_accel.ensure(s->accel.size() * sizeof(float)))) return rc;
            HIPCHECK(hipMemcpy(s->d_accel.p, s->accel.data(), s->accel.size() * sizeof(float), hipMemcpyHostToDevice));
        }
        if (!s->ddets.empty()) {
            if ((rc = s->d_det.ensure(s->ddets.size() * sizeof(DevDet)))) return rc;
            if ((rc = s->d_det_counts.ensure(s->det_bins * sizeof(unsigned long long)))) return rc;
            HIPCHECK(hipMemcpy(s->d_det.p, s->ddets.data(), s->ddets.size() * sizeof(DevDet), hipMemcpyHostToDevice));
        }
        s->uploaded = true;
    }
    if (s->mats_dirty) {
        int rc;
        std::vector<DevMat> dm(s->regions.size());
        region_mats(s, s->mats.data(), dm.data());
        std::vector<float> mus(s->mats.size());
        for (size_t i = 0; i < s->mats.size(); ++i) mus[i] = (float)s->mats[i].mus;
        if ((rc = s->d_mats.ensure(dm.size() * sizeof(DevMat)))) return rc;
        if ((rc = s->d_mat_mus.ensure(mus.size() * sizeof(float)))) return rc;
        HIPCHECK(hipMemcpy(s->d_mats.p, dm.data(), dm.size() * sizeof(DevMat), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(s->d_mat_mus.p, mus.data(), mus.size() * sizeof(float), hipMemcpyHostToDevice));
        s->mats_dirty = false;
    }
    return SMCRT_OK;
}

static void fill_common(const smcrt_scene* s, KParams& P, KCold& Cc) {
    memset(&P, 0, sizeof P);
    memset(&Cc, 0, sizeof Cc);
    P.nodes = s->d_nodes.as<DevNode>();
    P.regions = s->d_regions.as<DevRegion>();
    P.accel = s->accel.empty() ? nullptr : s->d_accel.as<float>();
    P.mats = s->d_mats.as<DevMat>();
    P.nreg = (int)s->regions.size();
    P.nmat = (int)s->mats.size();
    P.nx = s->grid.nx; P.ny = s->grid.ny; P.nz = s->grid.nz;
    P.xmax = (float)s->grid.xmax; P.ymax = (float)s->grid.ymax; P.zmax = (float)s->grid.zmax;
    P.sx = (float)(s->grid.nx / (2.0 * s->grid.xmax));
    P.sy = (float)(s->grid.ny / (2.0 * s->grid.ymax));
    P.sz = (float)(s->grid.nz / (2.0 * s->grid.zmax));
    Cc.src_kind = s->src.kind;
    Cc.circ_swap = s->circ_swap;
    for (int j = 0; j < 3; ++j) {
        Cc.spos[j] = (float)s->src.pos[j];
        Cc.sdir[j] = (float)s->src.dir[j];
        Cc.p1[j] = (float)s->src.p1[j];
        Cc.p2[j] = (float)s->src.p2[j];
        Cc.p3[j] = (float)s->src.p3[j];
        Cc.rdir[j] = s->rdir[j];
    }
    Cc.radius = (float)s->src.radius;
    memcpy(Cc.circ_t, s->circ_t, sizeof Cc.circ_t);
    P.eps = s->eps;
    P.eps_kill = s->eps_kill;
    Cc.mat_mus = s->d_mat_mus.as<float>();
}

struct LaunchShape {
    int block, blocks, chunk, walk_cap, refill_min, waves;
    bool compiled;  // the scene's run-time compiled kernels (JitKernels) instead of the library's
    bool bin;       // the compiled forward kernel with the binned tally's stage
    bool sets;      // the compiled forward kernel over several sets (batches)
    size_t lds;
};

// The packet kernel is compiled for a few kind sets; a scene runs the smallest one that covers
// every kind its program uses (a kernel without the unused interpreter cases is about 40% less
// code, and the kernel is instruction-cache bound).  SMCRT_KERNEL=full in the environment forces
// the general variant.
typedef void (*PacketKernelFn)(KParams);
static constexpr uint32_t KM_BASIC = KM_PRIM(SMCRT_SPHERE) | KM_PRIM(SMCRT_BOX) | KM_CSG(SMCRT_UNION);
static constexpr uint32_t KM_LAYERS = KM_BASIC | KM_PRIM(SMCRT_CYLINDER) | KM_CSG(SMCRT_SMOOTH_UNION);
static constexpr uint32_t KM_EXTRUDED = KM_PRIM(SMCRT_BOX) | KM_PRIM(SMCRT_POLYGON) | KM_MOD(SMCRT_EXTRUDE);
struct KernelVariant {
    uint32_t mask;
    const char* name;
    PacketKernelFn fwd, adj;
    PacketKernelFn fwd3, adj3;  // built for 3 resident waves per SIMD (nullptr: none)
};
#define SMCRT_VARIANT(KM, NAME)                                                                 \
    {                                                                                           \
        KM, NAME, packet_kernel<false, KM>, packet_kernel<true, KM>, packet_kernel<false, KM, 3>, \
            packet_kernel<true, KM, 3>                                                          \
    }
static const KernelVariant kVariants[] = {
    SMCRT_VARIANT(KM_BASIC, "basic"),
    SMCRT_VARIANT(KM_LAYERS, "layers"),
    SMCRT_VARIANT(KM_EXTRUDED, "extruded"),
    {KM_FULL, "full", packet_kernel<false, KM_FULL>, packet_kernel<true, KM_FULL>, nullptr, nullptr},
};
#undef SMCRT_VARIANT

static const KernelVariant& kernel_variant(const smcrt_scene* s) {
    const size_t nv = sizeof kVariants / sizeof kVariants[0];
    const char* force = getenv("SMCRT_KERNEL");
    if (force && !strcmp(force, "full")) return kVariants[nv - 1];
    for (size_t i = 0; i < nv; ++i)
        if ((s->kind_mask & ~kVariants[i].mask) == 0u) return kVariants[i];
    return kVariants[nv - 1];
}

// The kernel a launch runs: the scene's variant, adjoint or forward, and the 3-wave build when
// asked for and available.
static PacketKernelFn packet_kernel_for(const smcrt_scene* s, bool adj, int waves) {
    const KernelVariant& v = kernel_variant(s);
    const PacketKernelFn k3 = adj ? v.adj3 : v.fwd3;
    return (waves == 3 && k3) ? k3 : (adj ? v.adj : v.fwd);
}

// Resident waves per SIMD a launch of the library's kernels runs with: 3 (the default; also for
// 4, which only the run-time compiled kernels are built for) where a 3-wave build exists, else 2.
static int launch_waves(const smcrt_scene* s, const smcrt_run_opts* o, bool adj) {
    const int w = (o && o->waves > 0) ? o->waves : 3;
    const KernelVariant& v = kernel_variant(s);
    return (w >= 3 && (adj ? v.adj3 : v.fwd3)) ? 3 : 2;
}

// The same for the run-time compiled kernels, built for 2, 3 and 4 waves (default 3).
static int compiled_waves(const smcrt_run_opts* o) {
    const int w = (o && o->waves > 0) ? o->waves : 3;
    return w < 2 ? 2 : (w > 4 ? 4 : w);
}

// ---- packet kernels compiled at run time for one scene -------------------------------------
// The library's kernels interpret the SDF program; for large launches the scene's program is
// turned into straight-line device code (one `case` per region: every node's parameters as
// exact hexadecimal float constants, the interpreter's value and position stacks as named
// temporaries), compiled with hipRTC together with the packet kernel (the device headers are
// embedded in the library), and launched from the loaded module.  Same device functions, same
// arithmetic: only the node loads and the opcode dispatch disappear.  The code object is cached
// on disk by a hash of its source.  Any failure leaves the library's kernels in use.

static void jit_float(std::string& o, float v) {
    char b[64];
    if (std::isnan(v)) snprintf(b, sizeof b, "__builtin_nanf(\"\")");
    else if (std::isinf(v)) snprintf(b, sizeof b, v > 0.f ? "__builtin_inff()" : "(-__builtin_inff())");
    else snprintf(b, sizeof b, "%af", (double)v);
    o += b;
}

static void jit_array(std::string& o, const std::string& name, const float* v, int n) {
    o += "            const float " + name + "[" + std::to_string(n) + "] = {";
    for (int i = 0; i < n; ++i) {
        if (i) o += ", ";
        jit_float(o, v[i]);
    }
    o += "};\n";
}

// Instantiation name expressions, per kind and waves (2..4).  Kind 0: forward, 1: adjoint, 2:
// forward with the binned tally's stage, 3: forward over several sets (batches)
static constexpr int kJitKinds = 4;
static std::string jit_name(uint32_t km, int kind, int waves) {
    return std::string("smcrt::packet_kernel<") + (kind == 1 ? "true" : "false") + ", " + std::to_string(km) + "u, " +
           std::to_string(waves) + (kind == 3 ? ", smcrt::SceneSweepSets, " : ", smcrt::SceneSweep, ") +
           (kind == 2 ? "true" : "false") + ">";
}

// The statements computing node i's input position inside region R (the modifiers open at node
// i, then its own transform), with `pre` prefixed to the temporaries' names; returns the
// position's name.
static std::string jit_pos_chain(std::string& o, const smcrt_scene* s, const DevRegion& R, int i, const std::string& pre) {
    std::vector<int> open;
    for (int k = 0; k < i; ++k) {
        const DevNode& n = s->nodes[R.first + k];
        if (n.op == SMCRT_OP_MOD_BEGIN) open.push_back(k);
        else if (n.op == SMCRT_OP_MOD_END && !open.empty()) open.pop_back();
    }
    std::string q = "p";
    for (int k : open) {
        const DevNode& n = s->nodes[R.first + k];
        const std::string id = pre + std::to_string(k);
        jit_array(o, "P" + id, n.p, 8);
        o += "            const V3<float> x" + id + " = mod_pos<float, KM>(" + std::to_string(n.kind) + ", P" + id + ", " +
             q + ");\n";
        q = "x" + id;
    }
    const DevNode& n = s->nodes[R.first + i];
    const std::string id = pre + std::to_string(i);
    if (n.ident == 0) {
        jit_array(o, "T" + id, n.t, 12);
        o += "            const V3<float> x" + id + " = xform<float>(T" + id + ", " + q + ");\n";
        q = "x" + id;
    } else if (n.ident == 2) {  // translation only
        o += "            const V3<float> x" + id + "{" + q + ".x + ";
        jit_float(o, n.t[9]);
        o += ", " + q + ".y + ";
        jit_float(o, n.t[10]);
        o += ", " + q + ".z + ";
        jit_float(o, n.t[11]);
        o += "};\n";
        q = "x" + id;
    }
    return q;
}

static std::string jit_source(const smcrt_scene* s, uint32_t km) {
    // polygon primitives, numbered in program order (the gathered evaluation's indices)
    std::vector<std::pair<int, int>> polys;  // (region, node)
    for (size_t r = 0; r < s->regions.size(); ++r)
        for (int i = 0; i < s->regions[r].count; ++i) {
            const DevNode& n = s->nodes[s->regions[r].first + i];
            if (n.op == SMCRT_OP_PRIM && n.kind == SMCRT_POLYGON) polys.push_back({(int)r, i});
        }
    const bool gather = polys.size() >= 2;
    // pairwise disjoint grid rectangles in one shared position frame (the same modifiers and
    // transform before every polygon): a lane lies inside at most one rectangle
    bool disjoint = gather;
    if (gather) {
        std::string frame0;
        for (size_t j = 0; j < polys.size() && disjoint; ++j) {
            std::string f;
            (void)jit_pos_chain(f, s, s->regions[polys[j].first], polys[j].second, "");
            if (j == 0) frame0 = f;
            else disjoint = f == frame0;
        }
        for (size_t a = 0; a < polys.size() && disjoint; ++a)
            for (size_t b = a + 1; b < polys.size() && disjoint; ++b) {
                const float* qa = s->nodes[s->regions[polys[a].first].first + polys[a].second].q;
                const float* qb = s->nodes[s->regions[polys[b].first].first + polys[b].second].q;
                // [gx0, gx1) x [gy0, gy1) rectangles
                disjoint = qa[2] <= qb[0] || qb[2] <= qa[0] || qa[3] <= qb[1] || qb[3] <= qa[1];
            }
    }
    std::string o = "#include \"smcrt.h\"\n#include \"packet_kernel.h\"\nnamespace smcrt {\n"
                    "struct SceneSweep {\n"
                    "    static constexpr bool kCompiled = true;\n"
                    "    static constexpr int kRegions = " + std::to_string(s->regions.size()) + ";\n"
                    "    static constexpr bool kDiagnostics = false;\n"
                    "    static constexpr bool kDetectors = " + std::string(s->ddets.empty() ? "false" : "true") + ";\n"
                    "    static constexpr bool kWeight = false;\n"
                    "    static constexpr int kSource = " + std::to_string(s->src.kind) + ";\n"
                    "    static constexpr bool kSets = false;\n"
                    "    static constexpr bool kGather = " + std::string(gather ? "true" : "false") + ";\n";
    if (gather) {
        // poly_hit: the first polygon whose grid rectangle holds the lane's position, evaluated
        o += "    template <uint32_t KM>\n"
             "    static SMCRT_DEV PolyHit poly_hit(const V3<float>& p, const float* __restrict__ accel) {\n"
             "        int k = -1, base = 0;\n"
             "        V3<float> xs = p;\n";
        for (size_t j = 0; j < polys.size(); ++j) {
            const DevRegion& R = s->regions[polys[j].first];
            const DevNode& n = s->nodes[R.first + polys[j].second];
            o += "        {\n";
            const std::string q = jit_pos_chain(o, s, R, polys[j].second, "");
            jit_array(o, "Q", n.q, 8);
            o += "            const bool t = (k < 0) & polygon_rect_inside(Q, " + q + ");\n"
                 "            k = t ? " + std::to_string(j) + " : k;\n"
                 "            base = t ? " + std::to_string((int)n.p[0]) + " : base;\n"
                 "            xs = t ? " + q + " : xs;\n"
                 "        }\n";
        }
        o += "        PolyHit h{k, 0.f};\n"
             "        if (k >= 0) h.d = polygon_eval<float>(accel, base, xs);\n"
             "        return h;\n"
             "    }\n";
    }
    o += "    template <uint32_t KM>\n"
         "    static SMCRT_DEV float region(int r, const V3<float>& p, const float* __restrict__ accel, const PolyHit& h) {\n"
         "        switch (r) {\n";
    size_t pj = 0;  // next polygon index
    for (size_t r = 0; r < s->regions.size(); ++r) {
        const DevRegion& R = s->regions[r];
        o += "        case " + std::to_string(r) + ": {\n";
        std::vector<std::string> vs, ps{"p"};
        for (int i = 0; i < R.count; ++i) {
            const DevNode& n = s->nodes[R.first + i];
            const std::string id = std::to_string(i), kind = std::to_string(n.kind);
            if (n.op == SMCRT_OP_PRIM) {
                jit_array(o, "P" + id, n.p, 8);
                jit_array(o, "Q" + id, n.q, 8);
                std::string q = ps.back();
                if (n.ident == 0) {
                    jit_array(o, "T" + id, n.t, 12);
                    o += "            const V3<float> x" + id + " = xform<float>(T" + id + ", " + q + ");\n";
                    q = "x" + id;
                } else if (n.ident == 2) {  // translation only
                    o += "            const V3<float> x" + id + "{" + q + ".x + ";
                    jit_float(o, n.t[9]);
                    o += ", " + q + ".y + ";
                    jit_float(o, n.t[10]);
                    o += ", " + q + ".z + ";
                    jit_float(o, n.t[11]);
                    o += "};\n";
                    q = "x" + id;
                }
                if (gather && n.kind == SMCRT_POLYGON && disjoint)
                    o += "            const float v" + id + " = poly_value_disjoint(" + std::to_string(pj++) + ", Q" + id +
                         ", " + q + ", h);\n";
                else if (gather && n.kind == SMCRT_POLYGON)
                    o += "            const float v" + id + " = poly_value(" + std::to_string(pj++) + ", " +
                         std::to_string((int)n.p[0]) + ", Q" + id + ", " + q + ", accel, h);\n";
                else
                    o += "            const float v" + id + " = prim_eval<float, KM>(" + kind + ", P" + id + ", Q" + id + ", " +
                         q + ", accel);\n";
                vs.push_back("v" + id);
            } else if (n.op == SMCRT_OP_CSG) {
                jit_array(o, "Q" + id, n.q, 8);
                const std::string b = vs.back();
                vs.pop_back();
                const std::string a = vs.back();
                vs.pop_back();
                o += "            const float v" + id + " = csg_eval<float, KM>(" + kind + ", " + a + ", " + b + ", ";
                jit_float(o, n.k);
                o += ", Q" + id + ");\n";
                vs.push_back("v" + id);
            } else if (n.op == SMCRT_OP_MOD_BEGIN) {
                jit_array(o, "P" + id, n.p, 8);
                o += "            const V3<float> x" + id + " = mod_pos<float, KM>(" + kind + ", P" + id + ", " + ps.back() +
                     ");\n";
                ps.push_back("x" + id);
            } else {  // MOD_END: the modifier's post-step at the enclosing position
                jit_array(o, "P" + id, n.p, 8);
                ps.pop_back();
                const std::string v = vs.back();
                vs.pop_back();
                o += "            const float v" + id + " = mod_post<float, KM>(" + kind + ", P" + id + ", " + ps.back() +
                     ", " + v + ");\n";
                vs.push_back("v" + id);
            }
        }
        o += "            return " + (vs.empty() ? std::string("0.f") : vs.back()) + ";\n        }\n";
    }
    o += "        }\n        return 0.f;\n    }\n";
    std::string bake;  // the scene's fixed launch parameters, exactly as fill_common sets them
    {
        KParams P;
        KCold Cc;
        fill_common(s, P, Cc);
        std::string& o = bake;
        o += "        P.nreg = " + std::to_string(P.nreg) + ";\n";
        o += "        P.nx = " + std::to_string(P.nx) + "; P.ny = " + std::to_string(P.ny) + "; P.nz = " + std::to_string(P.nz) +
             ";\n";
        const std::pair<const char*, float> fl[] = {{"xmax", P.xmax}, {"ymax", P.ymax}, {"zmax", P.zmax}, {"sx", P.sx},
                                                    {"sy", P.sy},     {"sz", P.sz},     {"eps", P.eps},   {"eps_kill", P.eps_kill}};
        for (const auto& f : fl) {
            o += std::string("        P.") + f.first + " = ";
            jit_float(o, f.second);
            o += ";\n";
        }
    }
    o += "    static SMCRT_DEV void bake(KParams& P) {\n"
         "        P.nsets = 1;  // single-set launches\n" + bake + "    }\n};\n";
    // the same program for batched launches: any number of sets of this scene (set lookup, per-lane
    // outcome counters)
    o += "struct SceneSweepSets : SceneSweep {\n"
         "    static constexpr bool kSets = true;\n"
         "    static SMCRT_DEV void bake(KParams& P) {\n" + bake + "    }\n};\n";
    for (int kind = 0; kind < kJitKinds; ++kind)
        for (int w = 2; w <= 4; ++w) o += "template __global__ void " + jit_name(km, kind, w).substr(7) + "(KParams);\n";
    o += "}  // namespace smcrt\n";
    return o;
}

static uint64_t fnv1a(const void* p, size_t n, uint64_t h = 1469598103934665603ull) {
    const unsigned char* c = (const unsigned char*)p;
    for (size_t i = 0; i < n; ++i) h = (h ^ c[i]) * 1099511628211ull;
    return h;
}

static std::string jit_cache_dir() {
    const char* d = getenv("SMCRT_CACHE_DIR");
    std::string dir;
    if (d && *d) dir = d;
    else {
        const char* home = getenv("HOME");
        dir = std::string(home && *home ? home : "/tmp") + "/.cache/smcrt";
    }
    dir += "/jit";
    // mkdir -p
    for (size_t i = 1; i <= dir.size(); ++i)
        if (i == dir.size() || dir[i] == '/') (void)mkdir(dir.substr(0, i).c_str(), 0755);
    return dir;
}

// cache file: u32 count, then per name (u32 length, bytes), then u64 code size and the code
// kernels in a compiled module: every kind, each built for 2, 3 and 4 waves
static constexpr uint32_t kJitKernels = 3 * kJitKinds;

static bool jit_cache_read(const std::string& path, std::vector<std::string>& names, std::vector<char>& code) {
    FILE* f = fopen(path.c_str(), "rb");
    if (!f) return false;
    bool ok = true;
    uint32_t n = 0;
    ok = ok && fread(&n, 4, 1, f) == 1 && n == kJitKernels;
    names.assign(ok ? n : 0, std::string());
    for (uint32_t i = 0; ok && i < n; ++i) {
        uint32_t len = 0;
        ok = fread(&len, 4, 1, f) == 1 && len < 4096;
        if (ok) {
            names[i].resize(len);
            ok = fread(&names[i][0], 1, len, f) == len;
        }
    }
    uint64_t cs = 0;
    ok = ok && fread(&cs, 8, 1, f) == 1 && cs > 0 && cs < (1ull << 30);
    if (ok) {
        code.resize(cs);
        ok = fread(code.data(), 1, cs, f) == cs;
    }
    fclose(f);
    return ok;
}

static void jit_cache_write(const std::string& path, const std::vector<std::string>& names, const std::vector<char>& code) {
    const std::string tmp = path + ".tmp" + std::to_string((unsigned long)getpid());
    FILE* f = fopen(tmp.c_str(), "wb");
    if (!f) return;
    bool ok = true;
    const uint32_t n = (uint32_t)names.size();
    ok = ok && fwrite(&n, 4, 1, f) == 1;
    for (const std::string& x : names) {
        const uint32_t len = (uint32_t)x.size();
        ok = ok && fwrite(&len, 4, 1, f) == 1 && fwrite(x.data(), 1, len, f) == len;
    }
    const uint64_t cs = code.size();
    ok = ok && fwrite(&cs, 8, 1, f) == 1 && fwrite(code.data(), 1, cs, f) == cs;
    ok = (fclose(f) == 0) && ok;
    if (ok) (void)rename(tmp.c_str(), path.c_str());
    else (void)remove(tmp.c_str());
}

// Compile (or load from the cache) the scene's kernels; true when they are usable.
static bool jit_build(smcrt_scene* s) {
    JitKernels& J = s->jit;
    if (J.tried) return J.ok;
    J.tried = true;
    const auto t0 = std::chrono::steady_clock::now();
    const uint32_t km = s->kind_mask;
    const std::string src = jit_source(s, km);
    static const char* const opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-munsafe-fp-atomics",
                                       "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"};
    const int nopts = (int)(sizeof opts / sizeof opts[0]);
    uint64_t h = fnv1a(src.data(), src.size());
    for (int i = 0; i < kJitHeaders; ++i) h = fnv1a(kJitHeaderSources[i], strlen(kJitHeaderSources[i]), h);
    for (int i = 0; i < nopts; ++i) h = fnv1a(opts[i], strlen(opts[i]), h);
    // the toolchain and the device the code object is for: a cache written by another hipRTC,
    // HIP runtime or GPU is not reused
    {
        int ver[3] = {0, 0, 0};
        (void)hiprtcVersion(&ver[0], &ver[1]);
        (void)hipRuntimeGetVersion(&ver[2]);
        h = fnv1a(ver, sizeof ver, h);
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, s->device) == hipSuccess)
            h = fnv1a(prop.gcnArchName, strnlen(prop.gcnArchName, sizeof prop.gcnArchName), h);
    }
    char hex[32];
    snprintf(hex, sizeof hex, "%016llx", (unsigned long long)h);
    const std::string path = jit_cache_dir() + "/packet_kernel_" + hex + ".co";
    // SMCRT_JIT_DUMP=1: also write the generated source beside the code object (inspection)
    if (const char* dump = getenv("SMCRT_JIT_DUMP"); dump && !strcmp(dump, "1")) {
        if (FILE* f = fopen((jit_cache_dir() + "/packet_kernel_" + hex + ".hip").c_str(), "w")) {
            fwrite(src.data(), 1, src.size(), f);
            fclose(f);
        }
    }
    std::vector<std::string> names;
    std::vector<char> code;
    if (jit_cache_read(path, names, code)) {
        J.cached = true;
    } else {
        // hipRTC has no C library headers: stddef.h / stdint.h / hip_runtime.h come from its own
        // runtime header, so the headers the sources include are given as stubs
        std::vector<const char*> hs, hn;
        for (int i = 0; i < kJitHeaders; ++i) {
            hs.push_back(kJitHeaderSources[i]);
            hn.push_back(kJitHeaderNames[i]);
        }
        static const char* const stdint_stub =
            "#pragma once\n"
            "typedef __hip_internal::int8_t int8_t;\ntypedef __hip_internal::int16_t int16_t;\n"
            "typedef __hip_internal::int32_t int32_t;\ntypedef __hip_internal::int64_t int64_t;\n"
            "typedef __hip_internal::uint8_t uint8_t;\ntypedef __hip_internal::uint16_t uint16_t;\n"
            "typedef __hip_internal::uint32_t uint32_t;\ntypedef __hip_internal::uint64_t uint64_t;\n";
        hs.push_back("#pragma once\n");
        hn.push_back("stddef.h");
        hs.push_back(stdint_stub);
        hn.push_back("stdint.h");
        hs.push_back("#pragma once\n");
        hn.push_back("hip/hip_runtime.h");
        hiprtcProgram prog = nullptr;
        if (hiprtcCreateProgram(&prog, src.c_str(), "scene_packet_kernel.hip", (int)hs.size(), hs.data(), hn.data()) !=
            HIPRTC_SUCCESS) {
            J.error = "hiprtcCreateProgram failed";
            return false;
        }
        std::vector<std::string> exprs;
        for (int kind = 0; kind < kJitKinds; ++kind)
            for (int w = 2; w <= 4; ++w) {
                exprs.push_back(jit_name(km, kind, w));
                (void)hiprtcAddNameExpression(prog, exprs.back().c_str());
            }
        const hiprtcResult rc = hiprtcCompileProgram(prog, nopts, opts);
        if (rc != HIPRTC_SUCCESS) {
            size_t ls = 0;
            (void)hiprtcGetProgramLogSize(prog, &ls);
            std::string log(ls + 1, '\0');
            (void)hiprtcGetProgramLog(prog, &log[0]);
            J.error = std::string("hipRTC: ") + hiprtcGetErrorString(rc) + ": " + log.substr(0, 2000);
            (void)hiprtcDestroyProgram(&prog);
            return false;
        }
        for (const std::string& e : exprs) {
            const char* lowered = nullptr;
            if (hiprtcGetLoweredName(prog, e.c_str(), &lowered) != HIPRTC_SUCCESS || !lowered) {
                J.error = "hipRTC: no lowered name for " + e;
                (void)hiprtcDestroyProgram(&prog);
                return false;
            }
            names.push_back(lowered);
        }
        size_t cs = 0;
        if (hiprtcGetCodeSize(prog, &cs) != HIPRTC_SUCCESS || cs == 0) {
            J.error = "hipRTC: empty code object";
            (void)hiprtcDestroyProgram(&prog);
            return false;
        }
        code.resize(cs);
        (void)hiprtcGetCode(prog, code.data());
        (void)hiprtcDestroyProgram(&prog);
        jit_cache_write(path, names, code);
    }
    if (hipSetDevice(s->device) != hipSuccess || hipModuleLoadData(&J.mod, code.data()) != hipSuccess) {
        J.error = "hipModuleLoadData failed";
        J.mod = nullptr;
        return false;
    }
    hipFunction_t* fn[kJitKernels] = {&J.fwd[0],  &J.fwd[1],  &J.fwd[2],  &J.adj[0], &J.adj[1], &J.adj[2],
                                      &J.bin[0],  &J.bin[1],  &J.bin[2],  &J.sets[0], &J.sets[1], &J.sets[2]};
    if (names.size() != kJitKernels) {
        J.error = "unexpected kernel count in the compiled module";
        (void)hipModuleUnload(J.mod);
        J.mod = nullptr;
        return false;
    }
    for (uint32_t i = 0; i < kJitKernels; ++i)
        if (hipModuleGetFunction(fn[i], J.mod, names[i].c_str()) != hipSuccess) {
            J.error = "hipModuleGetFunction failed for " + names[i];
            (void)hipModuleUnload(J.mod);
            J.mod = nullptr;
            return false;
        }
    J.seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    J.ok = true;
    return true;
}

// Run-time compiled kernels for this launch?  SMCRT_JIT=0 never, =1 always; by default for
// launches of at least kJitMinPackets = 2^19 packets (a compile costs a few seconds once per
// program; the disk cache makes later processes load it directly; a 1e6-packet slab launch runs
// 2.2x faster compiled).  Launches asking for the test diagnostics (event
// trace, scatter moments, end-early) or the weight_scatter estimator, which the compiled kernels
// leave out, use the library's kernels.
static constexpr unsigned long long kJitMinPackets = 1ull << 19;
static bool jit_env(unsigned long long packets, bool* decided) {
    const char* e = getenv("SMCRT_JIT");
    *decided = e && (!strcmp(e, "0") || !strcmp(e, "1"));
    if (*decided) return !strcmp(e, "1");
    return packets >= kJitMinPackets;
}
static bool jit_wanted(const smcrt_run_opts* o) {
    if (o && (o->trace_packets > 0 || (o->flags & (SMCRT_FLAG_MOMENTS | SMCRT_FLAG_END_EARLY | SMCRT_FLAG_WEIGHT))))
        return false;
    bool decided;
    const bool w = jit_env(o ? o->n_packets : 0, &decided);
    return decided ? w : (o && w);
}
// The same for one geometry group of a batch (its sets' packets together).
static bool batch_jit_wanted(const smcrt_run_opts* shape, unsigned long long packets) {
    if (shape && (shape->flags & SMCRT_FLAG_WEIGHT)) return false;
    bool decided;
    return jit_env(packets, &decided);
}

static hipFunction_t jit_function(const smcrt_scene* s, bool adj, int waves, bool bin = false, bool sets = false) {
    if (adj) return s->jit.adj[waves - 2];
    return bin ? s->jit.bin[waves - 2] : (sets ? s->jit.sets[waves - 2] : s->jit.fwd[waves - 2]);
}

static hipError_t kernel_occupancy(const smcrt_scene* s, bool adj, int waves, bool compiled, int block, size_t lds,
                                   int* active, bool bin = false, bool sets = false) {
    if (compiled)
        return hipModuleOccupancyMaxActiveBlocksPerMultiprocessor(active, jit_function(s, adj, waves, bin, sets), block,
                                                                  lds);
    return hipOccupancyMaxActiveBlocksPerMultiprocessor(active, packet_kernel_for(s, adj, waves), block, lds);
}

// 256-lane blocks per CU the kernel's registers allow, clamped to 2..4: the index + 2 of the tally
// tile sized for that occupancy (the 2-wave build of a small scene program can fit 4 waves).
static int register_blocks(const smcrt_scene* s, bool adj, int waves, bool compiled, bool bin = false) {
    int active = 0;
    if (kernel_occupancy(s, adj, waves, compiled, 256, 0, &active, bin) != hipSuccess || active < 2) return 2;
    return active > 4 ? 4 : active;
}

static hipError_t launch_packet_kernel(const smcrt_scene* s, bool adj, const LaunchShape& sh, hipStream_t stream,
                                       KParams& P) {
    void* args[] = {&P};
    if (sh.compiled)
        return hipModuleLaunchKernel(jit_function(s, adj, sh.waves, sh.bin, sh.sets), (unsigned)sh.blocks, 1, 1,
                                     (unsigned)sh.block, 1, 1, (unsigned)sh.lds, stream, args, nullptr);
    return hipLaunchKernel((const void*)packet_kernel_for(s, adj, sh.waves), dim3(sh.blocks), dim3(sh.block), args,
                           sh.lds, stream);
}

// Launch shape: block size, resident blocks, chunk, walk cap and dynamic LDS (the region-indexed
// material tables of n_sets sets, plus the adjoint's per-lane scores and gradients).
static LaunchShape pick_shape(const smcrt_scene* s, const smcrt_run_opts* o, bool adj, int n_sets, int tile_vox = 0,
                              int adj_slots = 0, bool compiled = false, int nhot = 0, bool sets = false) {
    LaunchShape L;
    L.compiled = compiled;
    L.bin = compiled && !adj && nhot > 0;
    L.sets = compiled && !adj && !L.bin && sets;
    L.block = (o && o->block_size > 0) ? o->block_size : 256;
    int bpc = (o && o->blocks_per_cu > 0) ? o->blocks_per_cu : 4;
    L.chunk = (o && o->chunk > 0) ? o->chunk : 64;
    L.walk_cap = (o && o->walk_cap > 0) ? o->walk_cap : 4;
    L.refill_min = (o && o->refill_min > 0) ? (o->refill_min > 64 ? 64 : o->refill_min) : 1;
    L.waves = compiled ? compiled_waves(o) : launch_waves(s, o, adj);
    // LDS: material tables, gradient descriptors, the tally tile, then the adjoint's and the queues
    const size_t mats = (size_t)n_sets * s->regions.size() * sizeof(DevMat) +
                        sizeof(GradDesc) * grad_desc_count((int)s->regions.size()) + sizeof(double) * ((tile_vox + 1) & ~1);
    L.block = (L.block / 64) * 64;
    if (L.block < 64) L.block = 64;
    if (L.block > 256) L.block = 256;
    // per lane: the wavefront's tally queue (TALLY_QUEUE 8-byte entries per 64 lanes)
    const size_t queue_lane = sizeof(uint2) * TALLY_QUEUE / 64;
    // the binned tally's stage follows the queues
    L.lds = mats + queue_lane * L.block + (L.bin ? bins_lds_bytes(s->nbins, nhot, L.block / 64) : 0);
    if (adj) {
        // per lane: 4 FP32 scores per adjoint material; per block: 4 FP64 sums per material
        const size_t per_lane = (size_t)16 * (size_t)adj_slots + queue_lane, per_block = (size_t)32 * (size_t)adj_slots;
        while (L.block > 64 && mats + per_lane * L.block + per_block > 65536) L.block /= 2;
        L.lds = mats + per_lane * L.block + per_block;
    }
    if (!(o && o->blocks_per_cu > 0)) {
        // persistent grid: as many blocks as can be resident at once
        int active = 0;
        hipError_t e = kernel_occupancy(s, adj, L.waves, compiled, L.block, L.lds, &active, L.bin, L.sets);
        if (e == hipSuccess && active > 0) bpc = active;
    }
    L.blocks = s->num_cu * bpc;
    return L;
}

// Greedy voxel box of at most `budget` voxels holding as much of `g`'s mass as possible: start
// at the heaviest voxel and repeatedly extend by the face slab with the highest mass per voxel.
static void choose_tile(const double* g, int nx, int ny, int nz, int budget, int t0[3], int tn[3]) {
    const size_t nv = (size_t)nx * ny * nz;
    size_t best = 0;
    for (size_t i = 1; i < nv; ++i)
        if (g[i] > g[best]) best = i;
    int lo[3] = {(int)(best % nx), (int)((best / nx) % ny), (int)(best / ((size_t)nx * ny))};
    int hi[3] = {lo[0] + 1, lo[1] + 1, lo[2] + 1};
    const int n[3] = {nx, ny, nz};
    auto at = [&](int i, int j, int k) { return g[(size_t)i + (size_t)nx * ((size_t)j + (size_t)ny * k)]; };
    for (;;) {
        const long vol = (long)(hi[0] - lo[0]) * (hi[1] - lo[1]) * (hi[2] - lo[2]);
        double best_d = -1.0;
        int best_f = -1;
        for (int f = 0; f < 6; ++f) {
            const int ax = f / 2, up = f % 2;
            if (up ? hi[ax] >= n[ax] : lo[ax] <= 0) continue;
            const long area = vol / (hi[ax] - lo[ax]);
            if (vol + area > budget) continue;
            const int c = up ? hi[ax] : lo[ax] - 1;
            int b0[3] = {lo[0], lo[1], lo[2]}, b1[3] = {hi[0], hi[1], hi[2]};
            b0[ax] = c;
            b1[ax] = c + 1;
            double m = 0.0;
            for (int k = b0[2]; k < b1[2]; ++k)
                for (int j = b0[1]; j < b1[1]; ++j)
                    for (int i = b0[0]; i < b1[0]; ++i) m += at(i, j, k);
            const double dens = m / (double)area;
            if (dens > best_d) { best_d = dens; best_f = f; }
        }
        if (best_f < 0) break;
        if (best_f % 2) hi[best_f / 2]++;
        else lo[best_f / 2]--;
    }
    for (int a = 0; a < 3; ++a) { t0[a] = lo[a]; tn[a] = hi[a] - lo[a]; }
}

struct AdjTables {
    const DevMat* mats;    // region-indexed table whose `mat` is the adjoint slot (-1: untracked)
    const float* mat_mus;  // mus per slot
    int nslots;
};
static int launch_run(smcrt_scene* s, const smcrt_run_opts* o, double* d_jmean, bool adj, const float* d_w,
                      hipStream_t stream, float* kernel_ms, const AdjTables* at = nullptr);

// Pick the LDS tally tile and the binned tally's hot bins from a pilot run of the scene (once per
// scene).
static int ensure_tile(smcrt_scene* s, const smcrt_run_opts* o, hipStream_t stream) {
    if (s->tile_ready) return SMCRT_OK;
    s->tile_ready = true;  // the pilot itself runs without a tile or bins
    const size_t nv = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
    double* pilot = nullptr;
    if (hipMalloc(&pilot, nv * sizeof(double)) != hipSuccess) return SMCRT_OK;  // no tile, not an error
    k_fill_f64<<<(unsigned)fill_blocks(nv), 256, 0, stream>>>(pilot, nv);
    smcrt_run_opts po = *o;
    po.n_packets = 100000;
    po.first_packet = (1ull << 62);
    po.flags = SMCRT_FLAG_DIRECT_TALLY;
    po.trace_packets = 0;
    po.trace_events = 0;
    float ms = 0.f;
    int rc = launch_run(s, &po, pilot, false, nullptr, stream, &ms);
    std::vector<double> h(nv);
    unsigned long long ev[NOUT + NEV];
    if (!rc && hipMemcpy(h.data(), pilot, nv * sizeof(double), hipMemcpyDeviceToHost) == hipSuccess &&
        hipMemcpy(ev, s->d_stats.p, sizeof ev, hipMemcpyDeviceToHost) == hipSuccess) {
        double total = 0.0;
        for (size_t i = 0; i < nv; ++i) total += h[i];
        auto box_mass = [&](const int t0[3], const int tn[3]) {
            double m = 0.0;
            for (int k = t0[2]; k < t0[2] + tn[2]; ++k)
                for (int j = t0[1]; j < t0[1] + tn[1]; ++j)
                    for (int i = t0[0]; i < t0[0] + tn[0]; ++i)
                        m += h[(size_t)i + (size_t)s->grid.nx * ((size_t)j + (size_t)s->grid.ny * k)];
            return m;
        };
        // bins of 2^SMCRT_BIN_SHIFT voxels by the pilot's path length (heaviest first)
        const size_t nb = (nv + BIN_VOX - 1) >> SMCRT_BIN_SHIFT;
        std::vector<double> bmass;
        std::vector<int> order;
        if (nb <= 4096 && nv < (1ull << 32) && total > 0.0) {
            bmass.assign(nb, 0.0);
            for (size_t i = 0; i < nv; ++i) bmass[i >> SMCRT_BIN_SHIFT] += h[i];
            for (size_t b = 0; b < nb; ++b)
                if (bmass[b] > 0.0) order.push_back((int)b);
            std::sort(order.begin(), order.end(), [&](int a, int b) { return bmass[a] > bmass[b]; });
            s->nbins = (int)nb;
            s->pilot_entries = (double)ev[NOUT + E_ATOMICS] / (double)po.n_packets;
            s->pilot_entries_per_sweep = (double)ev[NOUT + E_ATOMICS] / std::max(1.0, (double)ev[NOUT + E_SWEEPS]);
        }
        // per register occupancy of 2, 3 or 4 256-lane blocks per CU (a launch takes the one its
        // kernel's registers allow): the LDS left per block, a tile of at most 64 KiB of FP64 sums;
        // for the binned tally, hot bins in at most 60 % of that LDS (90 % where the direct
        // tally's tile holds under 10 % of the path length) and a tile in the rest
        int lds_cu = 0;
        if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, s->device) != hipSuccess ||
            lds_cu <= 0)
            lds_cu = 160 * 1024;
        for (int bpc = 2; bpc <= 4; ++bpc) {
            const int wi = bpc - 2;
            const long mats = (long)s->regions.size() * (long)sizeof(DevMat) +
                              (long)sizeof(GradDesc) * grad_desc_count((int)s->regions.size());
            const long queue = (long)(sizeof(uint2) * TALLY_QUEUE) * (256 / 64);
            const long per_block = (long)lds_cu / bpc - mats - queue - 256;
            const int budget = (int)std::max(0L, std::min(8192L, per_block / (long)sizeof(double)));
            if (budget > 0) {
                choose_tile(h.data(), s->grid.nx, s->grid.ny, s->grid.nz, budget, s->tile_t0[wi], s->tile_tn[wi]);
                s->tile_capture[wi] = total > 0.0 ? box_mass(s->tile_t0[wi], s->tile_tn[wi]) / total : 0.0;
            }
            if (s->nbins > 0 && per_block > 0) {
                // where the tile holds little (< 10 % of the path length) the bins take 90 %
                const double share = s->tile_capture[wi] < 0.1 ? 0.9 : 0.6;
                const long slot = 16 + 8 * SMCRT_BIN_B;
                const long kmax = ((long)(share * per_block) - (long)bins_lds_bytes(s->nbins, 1, 4) + slot) / slot;
                const int k = (int)std::max(0L, std::min({kmax, (long)order.size(), (long)SMCRT_BIN_COLD - 1}));
                s->binmap[wi].assign(nb, (uint16_t)SMCRT_BIN_COLD);
                double hm = 0.0;
                for (int j = 0; j < k; ++j) {
                    s->binmap[wi][order[j]] = (uint16_t)j;
                    hm += bmass[order[j]];
                }
                s->nhot[wi] = k;
                s->hot_frac[wi] = hm / total;
                const long left = per_block - (long)bins_lds_bytes(s->nbins, k, 4);
                const int bbudget = (int)std::max(0L, std::min(8192L, left / (long)sizeof(double)));
                if (k > 0 && bbudget > 0)
                    choose_tile(h.data(), s->grid.nx, s->grid.ny, s->grid.nz, bbudget, s->btile_t0[wi], s->btile_tn[wi]);
            }
        }
    }
    (void)hipFree(pilot);
    return rc;
}

// Binned tally: pages for a launch (the pilot's tally entries per packet in the hot bins x 1.25,
// plus one partly filled page per hot bin and block), within SMCRT_TALLY_POOL_MB or half the free
// device memory (at most 64 GiB); the bin map of the launch's occupancy; the kernel's pointers.
// Entries beyond a block's pages take the FP64 atomics, so a short pool only costs speed.
static int bin_pool(smcrt_scene* s, const smcrt_run_opts* o, const LaunchShape& sh, int wi, KParams& P,
                    hipStream_t stream) {
    const int nh = s->nhot[wi], nb = s->nbins;
    const double want = s->pilot_entries * s->hot_frac[wi] * (double)o->n_packets * 1.25;
    size_t pages = (size_t)(want / SMCRT_BIN_PAGE) + (size_t)sh.blocks * (size_t)(nh + 1);
    const size_t per_page = (size_t)SMCRT_BIN_PAGE * sizeof(uint2) + 3 * sizeof(uint32_t);  // pool, bin, fill, list
    size_t cap = 0;
    if (const char* e = getenv("SMCRT_TALLY_POOL_MB"); e && *e) {
        cap = (size_t)strtoull(e, nullptr, 10) << 20;
    } else {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) fr = 0;
        const size_t held = s->d_pool.bytes + s->d_page_bin.bytes + s->d_page_fill.bytes + s->d_bin_list.bytes;
        cap = std::min((size_t)64 << 30, (fr + held) / 2);
    }
    pages = std::min({pages, cap / per_page, (size_t)0xFFFFFFFFu});
    const size_t ppb = pages / (size_t)sh.blocks;
    pages = ppb * (size_t)sh.blocks;
    int rc = 0;
    if ((rc = s->d_pool.ensure(std::max<size_t>(pages, 1) * SMCRT_BIN_PAGE * sizeof(uint2))) ||
        (rc = s->d_page_bin.ensure(std::max<size_t>(pages, 1) * sizeof(uint32_t))) ||
        (rc = s->d_page_fill.ensure(std::max<size_t>(pages, 1) * sizeof(uint32_t))) ||
        (rc = s->d_bin_list.ensure(std::max<size_t>(pages, 1) * sizeof(uint32_t))) ||
        (rc = s->d_block_pages.ensure((size_t)sh.blocks * sizeof(uint32_t))) ||
        (rc = s->d_bin_pages.ensure((size_t)nb * sizeof(uint32_t))) ||
        (rc = s->d_bin_off.ensure((size_t)nb * sizeof(uint32_t))) ||
        (rc = s->d_bin_cur.ensure((size_t)nb * sizeof(uint32_t))) ||
        (rc = s->d_bin_work.ensure(((size_t)nb + pages / BIN_PAGES_PER_ITEM + 1) * sizeof(BinWork))) ||
        (rc = s->d_binmap.ensure((size_t)nb * sizeof(uint16_t))))
        return rc;
    if (s->binmap_up != wi) {
        HIPCHECK(hipMemcpyAsync(s->d_binmap.p, s->binmap[wi].data(), (size_t)nb * sizeof(uint16_t), hipMemcpyHostToDevice,
                                stream));
        s->binmap_up = wi;
    }
    P.nbins = nb;
    P.nhot = nh;
    P.binmap = s->d_binmap.as<uint16_t>();
    P.pool = s->d_pool.as<uint2>();
    P.page_bin = s->d_page_bin.as<uint32_t>();
    P.page_fill = s->d_page_fill.as<uint32_t>();
    P.block_pages = s->d_block_pages.as<uint32_t>();
    P.ppb = (unsigned)ppb;
    return SMCRT_OK;
}

// Binned tally after the launch: pages per bin (read back), the pages listed bin by bin, then one
// LDS-tile workgroup per item of up to 1M entries of a bin adds the sums into the grid.
static int bin_reduce(smcrt_scene* s, const LaunchShape& sh, const KParams& P, double* d_jmean, size_t nvox,
                      hipStream_t stream) {
    const int nb = s->nbins;
    uint32_t* const bp = s->d_bin_pages.as<uint32_t>();
    k_fill_u32<<<(unsigned)fill_blocks((size_t)nb), 256, 0, stream>>>(bp, (size_t)nb);
    k_bin_count<<<(unsigned)sh.blocks, 256, 0, stream>>>(P.page_bin, P.block_pages, P.ppb, nb, bp);
    HIPCHECK(hipGetLastError());
    s->h_bin_pages.assign(nb, 0u);
    HIPCHECK(hipMemcpyAsync(s->h_bin_pages.data(), bp, (size_t)nb * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    s->h_bin_off.assign(nb, 0u);
    s->h_bin_work.clear();
    uint32_t tot = 0;
    for (int b = 0; b < nb; ++b) {
        const uint32_t c = s->h_bin_pages[b];
        s->h_bin_off[b] = tot;
        for (uint32_t j = 0; j < c; j += BIN_PAGES_PER_ITEM)
            s->h_bin_work.push_back(
                BinWork{(uint32_t)b, tot + j, tot + std::min(j + BIN_PAGES_PER_ITEM, c), c > BIN_PAGES_PER_ITEM ? 1u : 0u});
        tot += c;
    }
    if (s->h_bin_work.empty()) return SMCRT_OK;
    if (s->h_bin_work.size() * sizeof(BinWork) > s->d_bin_work.bytes) return fail(SMCRT_EDEVICE, "bin work list overflow");
    HIPCHECK(hipMemcpyAsync(s->d_bin_off.p, s->h_bin_off.data(), (size_t)nb * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    HIPCHECK(hipMemcpyAsync(s->d_bin_work.p, s->h_bin_work.data(), s->h_bin_work.size() * sizeof(BinWork),
                            hipMemcpyHostToDevice, stream));
    k_fill_u32<<<(unsigned)fill_blocks((size_t)nb), 256, 0, stream>>>(s->d_bin_cur.as<uint32_t>(), (size_t)nb);
    k_bin_list<<<(unsigned)sh.blocks, 256, 0, stream>>>(P.page_bin, P.block_pages, P.ppb, nb, s->d_bin_off.as<uint32_t>(),
                                                        s->d_bin_cur.as<uint32_t>(), s->d_bin_list.as<uint32_t>());
    k_bin_accum<<<(unsigned)s->h_bin_work.size(), 1024, 0, stream>>>(s->d_bin_work.as<BinWork>(), s->d_bin_list.as<uint32_t>(),
                                                                     P.pool, P.page_fill, d_jmean, nvox);
    HIPCHECK(hipGetLastError());
    return SMCRT_OK;
}

// Launch the packet kernel for one scene (one set).  d_jmean may be null (no tally).
static int launch_run(smcrt_scene* s, const smcrt_run_opts* o, double* d_jmean, bool adj, const float* d_w,
                      hipStream_t stream, float* kernel_ms, const AdjTables* at) {
    if (o->n_packets == 0) {
        *kernel_ms = 0.f;
        return SMCRT_OK;
    }
    if (o->trace_packets < 0 || o->trace_events < 0) return fail(SMCRT_EINVAL, "negative trace sizes");
    const bool use_tile = !adj && d_jmean && !(o->flags & SMCRT_FLAG_NO_LDS_TILE);
    const bool may_bin = !adj && d_jmean && !(o->flags & SMCRT_FLAG_DIRECT_TALLY);
    if (use_tile || may_bin) {
        int rc = ensure_tile(s, o, stream);
        if (rc) return rc;
    }
    KParams P;
    KCold Cc;
    fill_common(s, P, Cc);
    if (adj && at) {
        P.mats = at->mats;
        P.nmat = at->nslots;
        Cc.mat_mus = at->mat_mus;
    }
    P.jmean = d_jmean;
    P.grid_stride = 0;
    P.xgrid = nullptr;
    P.xstride = 0;
    P.xcopies = 0;
    const size_t nvox = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
    const int copies = d_jmean ? xgrid_copies(s->n_xcc, nvox, o->flags, 1, xgrid_avail(s->d_xgrid.bytes)) : 0;
    if (copies) {
        int rc = s->d_xgrid.ensure((size_t)copies * nvox * sizeof(double));
        if (rc) return rc;
        P.xgrid = s->d_xgrid.as<double>();
        P.xstride = nvox;
        P.xcopies = copies;
        k_fill_f64<<<(unsigned)fill_blocks((size_t)copies * nvox), 256, 0, stream>>>(P.xgrid, (size_t)copies * nvox);
    }
    P.seed_lo = (uint32_t)o->seed;
    P.seed_hi = (uint32_t)(o->seed >> 32);
    P.total = o->n_packets;
    P.nsets = 1;
    unsigned long long info[4] = {0ull, o->n_packets, o->first_packet, 0ull};
    HIPCHECK(hipMemcpyAsync(s->d_setinfo.p, info, sizeof info, hipMemcpyHostToDevice, stream));
    Cc.set_prefix = s->d_setinfo.as<unsigned long long>();
    Cc.set_first = s->d_setinfo.as<unsigned long long>() + 2;
    P.work = s->d_work.as<unsigned long long>();
    Cc.stats = s->d_stats.as<unsigned long long>();
    Cc.moments = s->d_moments.as<double>();
    if (o->trace_packets > 0 && o->trace_events > 0) {
        size_t tb = (size_t)o->trace_packets * o->trace_events * 4 * sizeof(float);
        int rc = s->d_trace.ensure(tb);
        if (rc) return rc;
        HIPCHECK(hipMemsetAsync(s->d_trace.p, 0, tb, stream));
        Cc.trace = s->d_trace.as<float>();
        Cc.trace_packets = o->trace_packets;
        Cc.trace_events = o->trace_events;
    }
    Cc.flags = o->flags;
    Cc.max_bounces = o->max_bounces > 0 ? o->max_bounces : 1000;
    // the scene's run-time compiled kernels (built once per scene; on failure the library's)
    const bool compiled = jit_wanted(o) && jit_build(s);
    int tile_vox = 0;
    bool binned = false;
    int wi = 0;
    const int waves = compiled ? compiled_waves(o) : launch_waves(s, o, adj);
    if (use_tile || may_bin) wi = register_blocks(s, adj, waves, compiled) - 2;
    // binned tally (the compiled kernels' stage): asked for, or by default where the tile holds
    // little of the path length and the tally is heavy next to the sweeps (more than 1.5 global
    // entries per sweep in the pilot: the atomics then bound the launch); the plan of the binned
    // kernel's own occupancy
    if (may_bin && compiled && s->nbins > 0 && s->nhot[wi] > 0 &&
        ((o->flags & SMCRT_FLAG_BINNED_TALLY) || (s->tile_capture[wi] < 0.35 && s->pilot_entries_per_sweep > 1.5))) {
        wi = register_blocks(s, adj, waves, compiled, true) - 2;
        binned = s->nhot[wi] > 0;
    }
    if (use_tile) {
        for (int a = 0; a < 3; ++a) {
            P.t0[a] = binned ? s->btile_t0[wi][a] : s->tile_t0[wi][a];
            P.tn[a] = binned ? s->btile_tn[wi][a] : s->tile_tn[wi][a];
        }
        tile_vox = P.tn[0] * P.tn[1] * P.tn[2];
    }
    LaunchShape sh = pick_shape(s, o, adj, 1, tile_vox, P.nmat, compiled, binned ? s->nhot[wi] : 0);
    if (!adj && d_jmean) s->last_binned = binned;
    if (binned) {
        int rc = bin_pool(s, o, sh, wi, P, stream);
        if (rc) return rc;
    }
    P.chunk = sh.chunk;
    P.dda_cap = sh.walk_cap;
    P.refill_min = sh.refill_min;
    Cc.weights = d_w;
    Cc.grad = s->d_grad.as<double>();
    // detectors (not in adjoint launches: their final-position continuation changes nothing the
    // adjoint scores)
    if (!adj && !s->ddets.empty()) {
        Cc.n_det = (int)s->ddets.size();
        Cc.det = s->d_det.as<DevDet>();
        Cc.det_counts = s->d_det_counts.as<unsigned long long>();
        k_fill_u64<<<(unsigned)fill_blocks(s->det_bins), 256, 0, stream>>>(Cc.det_counts, s->det_bins);
    }
    {
        int rc = s->d_cold.ensure(sizeof(KCold));
        if (rc) return rc;
        HIPCHECK(hipMemcpyAsync(s->d_cold.p, &Cc, sizeof Cc, hipMemcpyHostToDevice, stream));
        P.cold = s->d_cold.as<KCold>();
    }
    k_fill_u64<<<1, 64, 0, stream>>>(P.work, 8);
    k_fill_u64<<<1, 64, 0, stream>>>(Cc.stats, NOUT + NEV);
    k_fill_f64<<<1, 64, 0, stream>>>(Cc.moments, 24);
    if (adj) k_fill_f64<<<1, 256, 0, stream>>>(Cc.grad, 4 * (size_t)P.nmat + 1);
    HIPCHECK(hipEventRecord(s->ev0, stream));
    HIPCHECK(launch_packet_kernel(s, adj, sh, stream, P));
    HIPCHECK(hipGetLastError());
    if (!binned) HIPCHECK(hipEventRecord(s->ev1, stream));
    if (copies) k_reduce_copies<<<(unsigned)fill_blocks(nvox), 256, 0, stream>>>(d_jmean, P.xgrid, nvox, copies);
    if (binned) {
        int rc = bin_reduce(s, sh, P, d_jmean, nvox, stream);
        if (rc) return rc;
        HIPCHECK(hipEventRecord(s->ev1, stream));
    }
    HIPCHECK(hipStreamSynchronize(stream));
    HIPCHECK(hipEventElapsedTime(kernel_ms, s->ev0, s->ev1));
    return SMCRT_OK;
}

static int read_stats(smcrt_scene* s, smcrt_stats* st, float kernel_ms) {
    unsigned long long h[NOUT + NEV];
    double mom[24];
    HIPCHECK(hipMemcpy(h, s->d_stats.p, sizeof h, hipMemcpyDeviceToHost));
    HIPCHECK(hipMemcpy(mom, s->d_moments.p, sizeof mom, hipMemcpyDeviceToHost));
    memset(st, 0, sizeof *st);
    st->launched = h[O_LAUNCHED];
    st->absorbed = h[O_ABSORBED];
    st->escaped_r = h[O_R];
    st->escaped_t = h[O_T];
    st->killed_surface = h[O_KSURF];
    st->killed_bounce = h[O_KBOUNCE];
    st->scatters = h[O_SCATTERS];
    const unsigned long long* e = h + NOUT;
    st->ev_sweeps = e[E_SWEEPS];
    st->ev_region_evals = e[E_REGION];
    st->ev_normals = e[E_NORMALS];
    st->ev_steps = e[E_STEPS];
    st->ev_crossings = e[E_CROSS];
    st->ev_interfaces = e[E_IFACE];
    st->ev_fresnel = e[E_FRESNEL];
    st->err_unreachable = e[E_UNREACH];
    st->err_nan = e[E_NAN];
    st->ev_atomics = e[E_ATOMICS];
    memcpy(st->moments, mom, sizeof mom);
    st->kernel_ms = kernel_ms;
    return SMCRT_OK;
}

static int device_sum(smcrt_scene* s, const double* d, size_t n, hipStream_t stream, double* out) {
    k_fill_f64<<<1, 64, 0, stream>>>(s->d_sum.as<double>(), 1);
    k_sum_f64<<<(unsigned)fill_blocks(n), 256, 0, stream>>>(d, n, s->d_sum.as<double>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipMemcpyAsync(out, s->d_sum.p, sizeof(double), hipMemcpyDeviceToHost, stream));
    HIPCHECK(hipStreamSynchronize(stream));
    return SMCRT_OK;
}

extern "C" int smcrt_run_device(smcrt_scene* s, const smcrt_run_opts* o, double* d_jmean, smcrt_stats* st,
                                void* stream) {
    if (!s || !o) return fail(SMCRT_EINVAL, "null scene or options");
    int rc = ensure_device(s);
    if (rc) return rc;
    hipStream_t strm = stream ? (hipStream_t)stream : s->stream;
    float ms = 0.f;
    if ((o->flags & SMCRT_FLAG_NO_TALLY)) d_jmean = nullptr;
    if ((rc = launch_run(s, o, d_jmean, false, nullptr, strm, &ms))) return rc;
    if (st) {
        if ((rc = read_stats(s, st, ms))) return rc;
        if (d_jmean) {
            size_t nv = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
            if ((rc = device_sum(s, d_jmean, nv, strm, &st->tally_sum))) return rc;
        }
    }
    return SMCRT_OK;
}

extern "C" int smcrt_run(smcrt_scene* s, const smcrt_run_opts* o, smcrt_tallies* out) {
    if (!s || !o || !out) return fail(SMCRT_EINVAL, "null scene, options or output");
    int rc = ensure_device(s);
    if (rc) return rc;
    const size_t nv = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
    const bool tally = !(o->flags & SMCRT_FLAG_NO_TALLY) && (out->jmean || out->jmean_f32);
    double* dj = nullptr;
    if (tally) {
        if ((rc = s->d_jmean.ensure(nv * sizeof(double)))) return rc;
        dj = s->d_jmean.as<double>();
        k_fill_f64<<<(unsigned)fill_blocks(nv), 256, 0, s->stream>>>(dj, nv);
        HIPCHECK(hipGetLastError());
    }
    float ms = 0.f;
    if ((rc = launch_run(s, o, dj, false, nullptr, s->stream, &ms))) return rc;
    if ((rc = read_stats(s, &out->stats, ms))) return rc;
    if (tally) {
        std::vector<double> tmp;
        double* host = out->jmean;
        if (!host) {
            tmp.resize(nv);
            host = tmp.data();
        }
        HIPCHECK(hipMemcpy(host, dj, nv * sizeof(double), hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (size_t i = 0; i < nv; ++i) sum += host[i];
        out->stats.tally_sum = sum;
        if (out->jmean_f32)
            for (size_t i = 0; i < nv; ++i) out->jmean_f32[i] = (float)host[i];
    }
    if (out->detector_counts && s->det_bins) {
        HIPCHECK(hipMemcpy(out->detector_counts, s->d_det_counts.p, s->det_bins * sizeof(unsigned long long),
                           hipMemcpyDeviceToHost));
    }
    if (out->trace && o->trace_packets > 0 && o->trace_events > 0) {
        HIPCHECK(hipMemcpy(out->trace, s->d_trace.p, (size_t)o->trace_packets * o->trace_events * 4 * sizeof(float),
                           hipMemcpyDeviceToHost));
    }
    return SMCRT_OK;
}

extern "C" int smcrt_evaluate(smcrt_scene* s, int32_t region, int64_t n, const float* pts, float* dist,
                              float* normals) {
    if (!s || !pts || n < 0) return fail(SMCRT_EINVAL, "bad arguments");
    if (region < 0 || region >= (int32_t)s->regions.size()) return fail(SMCRT_EINVAL, "region %d out of range", region);
    if (n == 0) return SMCRT_OK;
    int rc = ensure_device(s);
    if (rc) return rc;
    if ((rc = s->d_pts.ensure((size_t)n * 3 * sizeof(float)))) return rc;
    if ((rc = s->d_out.ensure((size_t)n * 4 * sizeof(float)))) return rc;
    HIPCHECK(hipMemcpy(s->d_pts.p, pts, (size_t)n * 3 * sizeof(float), hipMemcpyHostToDevice));
    float* dd = dist ? s->d_out.as<float>() : nullptr;
    float* dn = normals ? s->d_out.as<float>() + n : nullptr;
    hipLaunchKernelGGL(k_evaluate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s->stream, s->d_nodes.as<DevNode>(),
                       s->d_regions.as<DevRegion>(), (int)region, (long long)n, s->d_pts.as<float>(), dd, dn,
                       s->accel.empty() ? nullptr : s->d_accel.as<float>());
    HIPCHECK(hipGetLastError());
    HIPCHECK(hipStreamSynchronize(s->stream));
    if (dist) HIPCHECK(hipMemcpy(dist, dd, (size_t)n * sizeof(float), hipMemcpyDeviceToHost));
    if (normals) HIPCHECK(hipMemcpy(normals, dn, (size_t)n * 3 * sizeof(float), hipMemcpyDeviceToHost));
    return SMCRT_OK;
}

extern "C" int smcrt_grad(smcrt_scene* s, const smcrt_run_opts* o, const float* weights, double* jmean,
                          smcrt_grad_out* out) {
    if (!s || !o || !out) return fail(SMCRT_EINVAL, "null scene, options or output");
    const size_t nm = s->mats.size();
    // adjoint slots: the materials selected by the mask (all when it is empty), at most 16
    const uint64_t mask = ((uint64_t)o->grad_mask_hi << 32) | o->grad_mask_lo;
    std::vector<int> slot(nm, -1), slot_mat;
    for (size_t m = 0; m < nm; ++m)
        if (!mask || ((mask >> m) & 1u)) {
            slot[m] = (int)slot_mat.size();
            slot_mat.push_back((int)m);
        }
    if (mask >> nm) return fail(SMCRT_EINVAL, "grad mask selects materials beyond the scene's %zu", nm);
    if (slot_mat.size() > 16)
        return fail(SMCRT_ERANGE, "the adjoint tracks at most 16 materials (%zu selected; use grad_mask)", slot_mat.size());
    int rc = ensure_device(s);
    if (rc) return rc;
    const size_t nv = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
    const size_t ns = slot_mat.size();
    if ((rc = s->d_grad.ensure((4 * ns + 1) * sizeof(double)))) return rc;
    AdjTables at;
    {
        std::vector<DevMat> dm(s->regions.size());
        region_mats(s, s->mats.data(), dm.data());
        for (DevMat& x : dm) x.mat = slot[x.mat];
        std::vector<float> mus(ns > 0 ? ns : 1, 0.f);
        for (size_t k = 0; k < ns; ++k) mus[k] = (float)s->mats[slot_mat[k]].mus;
        if ((rc = s->d_adj_mats.ensure(dm.size() * sizeof(DevMat)))) return rc;
        if ((rc = s->d_adj_mus.ensure(mus.size() * sizeof(float)))) return rc;
        HIPCHECK(hipMemcpy(s->d_adj_mats.p, dm.data(), dm.size() * sizeof(DevMat), hipMemcpyHostToDevice));
        HIPCHECK(hipMemcpy(s->d_adj_mus.p, mus.data(), mus.size() * sizeof(float), hipMemcpyHostToDevice));
        at.mats = s->d_adj_mats.as<DevMat>();
        at.mat_mus = s->d_adj_mus.as<float>();
        at.nslots = (int)ns;
    }
    const float* dw = nullptr;
    if (weights) {
        if ((rc = s->d_weights.ensure(nv * sizeof(float)))) return rc;
        HIPCHECK(hipMemcpy(s->d_weights.p, weights, nv * sizeof(float), hipMemcpyHostToDevice));
        dw = s->d_weights.as<float>();
    }
    double* dj = nullptr;
    if (jmean) {
        if ((rc = s->d_jmean.ensure(nv * sizeof(double)))) return rc;
        dj = s->d_jmean.as<double>();
        k_fill_f64<<<(unsigned)fill_blocks(nv), 256, 0, s->stream>>>(dj, nv);
    }
    float ms = 0.f;
    if ((rc = launch_run(s, o, dj, true, dw, s->stream, &ms, &at))) return rc;
    if ((rc = read_stats(s, &out->stats, ms))) return rc;
    std::vector<double> g(4 * ns + 1);
    HIPCHECK(hipMemcpy(g.data(), s->d_grad.p, g.size() * sizeof(double), hipMemcpyDeviceToHost));
    const double N = (double)(o->n_packets ? o->n_packets : 1);
    double* dst[4] = {out->d_r_dmua, out->d_r_dmus, out->d_w_dmua, out->d_w_dmus};
    for (int w = 0; w < 4; ++w)
        if (dst[w])
            for (size_t m = 0; m < nm; ++m) dst[w][m] = slot[m] >= 0 ? g[w * ns + slot[m]] / N : 0.0;
    out->w_tally = g[4 * ns] / N;
    out->r_d = (double)out->stats.escaped_r / N;
    if (jmean) {
        HIPCHECK(hipMemcpy(jmean, dj, nv * sizeof(double), hipMemcpyDeviceToHost));
        double sum = 0.0;
        for (size_t i = 0; i < nv; ++i) sum += jmean[i];
        out->stats.tally_sum = sum;
    }
    return SMCRT_OK;
}

// ------------------------------------------------------------------------------------------
// batched launch
// ------------------------------------------------------------------------------------------
struct BatchGroup {
    smcrt_scene* scene;
    KCold Cc;
    std::vector<int> sets;  // batch set indices in this group
    KParams P;
    LaunchShape sh;
    size_t grid_off;        // first double of this group's grids
    size_t stats_off;       // first counter of this group's stats
};

struct smcrt_batch {
    std::vector<smcrt_scene*> scenes;
    std::vector<BatchGroup> groups;
    std::vector<int> set_group, set_slot;
    std::vector<size_t> set_grid_off;
    std::vector<size_t> set_nv;
    int device = -1;
    hipStream_t stream = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    DevBuf d_mats, d_grids, d_stats, d_info, d_work, d_cold, d_xgrid;
    size_t total_doubles = 0, total_stats = 0;
    int copies = 0;  // per-XCD partial grids (0: none)
    hipEvent_t ev0 = nullptr, ev1 = nullptr;  // around the last replay
    float replay_ms = 0.f;
};

extern "C" int smcrt_batch_create(smcrt_scene* const* scenes, int32_t n_scenes, const smcrt_batch_set* sets,
                                  int32_t n_sets, uint64_t seed, const smcrt_run_opts* shape,
                                  smcrt_batch** out) {
    if (!scenes || !sets || !out || n_scenes < 1 || n_sets < 1) return fail(SMCRT_EINVAL, "bad batch arguments");
    *out = nullptr;
    for (int i = 0; i < n_scenes; ++i) {
        if (!scenes[i]) return fail(SMCRT_EINVAL, "scene %d is null", i);
        int rc = ensure_device(scenes[i]);
        if (rc) return rc;
    }
    smcrt_batch* b = new smcrt_batch();
    b->scenes.assign(scenes, scenes + n_scenes);
    b->device = scenes[0]->device;
    for (int i = 0; i < n_sets; ++i) {
        const smcrt_batch_set& st = sets[i];
        if (st.scene < 0 || st.scene >= n_scenes) {
            delete b;
            return fail(SMCRT_EINVAL, "set %d: scene %d out of range", i, st.scene);
        }
        if (!st.materials) {
            delete b;
            return fail(SMCRT_EINVAL, "set %d: null material table", i);
        }
        int rc = validate_materials(st.materials, (int)scenes[st.scene]->mats.size());
        if (rc) { delete b; return rc; }
    }
    // groups: one per scene with sets
    b->set_group.assign(n_sets, -1);
    b->set_slot.assign(n_sets, -1);
    b->set_grid_off.assign(n_sets, 0);
    b->set_nv.assign(n_sets, 0);
    for (int sc = 0; sc < n_scenes; ++sc) {
        BatchGroup g;
        g.scene = scenes[sc];
        for (int i = 0; i < n_sets; ++i)
            if (sets[i].scene == sc) g.sets.push_back(i);
        if (!g.sets.empty()) b->groups.push_back(g);
    }
    // sizes
    size_t nmats = 0, ndoubles = 0, nstats = 0, ninfo = 0;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        BatchGroup& g = b->groups[gi];
        const smcrt_scene* s = g.scene;
        const size_t nv = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
        g.grid_off = ndoubles;
        g.stats_off = nstats;
        for (size_t k = 0; k < g.sets.size(); ++k) {
            int i = g.sets[k];
            b->set_group[i] = (int)gi;
            b->set_slot[i] = (int)k;
            b->set_grid_off[i] = ndoubles;
            b->set_nv[i] = nv;
            ndoubles += nv;
        }
        nmats += g.sets.size() * s->regions.size();
        nstats += g.sets.size() * NOUT + NEV;
        ninfo += 2 * g.sets.size() + 1;
    }
    if (ndoubles >= (1ull << 32)) {
        smcrt_batch_destroy(b);
        return fail(SMCRT_ERANGE, "a batch holds %zu grid voxels; at most 2^32 - 1 (32-bit tally indices)", ndoubles);
    }
    b->total_doubles = ndoubles;
    b->total_stats = nstats;
    // 8 partial grids per XCD by default: a batch's chords have no LDS tile, so its hot voxels
    // contend (256-set workload: 5.1 s per replay with one per XCD, 3.9 s with eight)
    HIPCHECK(hipSetDevice(b->device));
    b->copies = xgrid_copies(scenes[0]->n_xcc, ndoubles, shape ? shape->flags : 0u, 8, xgrid_avail(b->d_xgrid.bytes));
    int rc;
    if ((rc = b->d_mats.ensure(nmats * sizeof(DevMat))) || (rc = b->d_grids.ensure(ndoubles * sizeof(double))) ||
        (rc = b->d_stats.ensure(nstats * sizeof(unsigned long long))) ||
        (rc = b->d_info.ensure(ninfo * sizeof(unsigned long long))) ||
        (rc = b->d_work.ensure(b->groups.size() * sizeof(unsigned long long))) ||
        (b->copies && (rc = b->d_xgrid.ensure((size_t)b->copies * ndoubles * sizeof(double))))) {
        smcrt_batch_destroy(b);
        return rc;
    }
    std::vector<DevMat> hm(nmats);
    std::vector<unsigned long long> info(ninfo);
    size_t mo = 0, io = 0;
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        BatchGroup& g = b->groups[gi];
        smcrt_scene* s = g.scene;
        const size_t nr = s->regions.size();
        const size_t mat_off = mo, info_off = io;
        unsigned long long acc = 0;
        for (size_t k = 0; k < g.sets.size(); ++k) {
            const smcrt_batch_set& st = sets[g.sets[k]];
            region_mats(s, st.materials, hm.data() + mo);
            mo += nr;
            info[io + k] = acc;
            info[io + g.sets.size() + 1 + k] = st.first_packet;
            acc += st.n_packets;
        }
        info[io + g.sets.size()] = acc;
        io += 2 * g.sets.size() + 1;
        KParams& P = g.P;
        fill_common(s, P, g.Cc);
        P.mats = b->d_mats.as<DevMat>() + mat_off;
        P.jmean = b->d_grids.as<double>() + g.grid_off;
        P.grid_stride = (unsigned long long)s->grid.nx * s->grid.ny * s->grid.nz;
        P.xgrid = b->copies ? b->d_xgrid.as<double>() + g.grid_off : nullptr;
        P.xstride = b->copies ? ndoubles : 0;
        P.xcopies = b->copies;
        P.seed_lo = (uint32_t)seed;
        P.seed_hi = (uint32_t)(seed >> 32);
        P.total = acc;
        P.nsets = (int)g.sets.size();
        g.Cc.set_prefix = b->d_info.as<unsigned long long>() + info_off;
        g.Cc.set_first = b->d_info.as<unsigned long long>() + info_off + g.sets.size() + 1;
        P.work = b->d_work.as<unsigned long long>() + gi;
        g.Cc.stats = b->d_stats.as<unsigned long long>() + g.stats_off;
        g.Cc.moments = s->d_moments.as<double>();
        g.Cc.max_bounces = 1000;
        if (g.sets.size() * s->regions.size() > SMCRT_LDS_MATS) {
            smcrt_batch_destroy(b);
            return fail(SMCRT_ERANGE, "a geometry group holds %zu sets x %zu regions; at most %d material entries fit in LDS",
                        g.sets.size(), s->regions.size(), SMCRT_LDS_MATS);
        }
        // large groups run the scene's compiled kernel for several sets (same transport, the
        // program as straight-line code)
        const bool compiled = batch_jit_wanted(shape, acc) && jit_build(s);
        // Batches refill a wavefront once 8 of its lanes are idle, from chunks as large as give
        // every wavefront about 32 claims (64 to 2048 queue positions), unless the shape says
        // otherwise: a claim's returning atomic waits for the wavefront's tally atomics in flight,
        // which batches (no LDS tile) keep many of.  256-set workload: 3276 ms per replay with the
        // single-launch defaults (64, every idle lane), 2592 with 256 and 8, 2392 with 2048 and 8.
        smcrt_run_opts bo{};
        if (shape) bo = *shape;
        if (bo.chunk <= 0) {
            const unsigned long long per_wave = acc / ((unsigned long long)std::max(s->num_cu, 1) * 16ull * 32ull);
            bo.chunk = 64;
            while (bo.chunk < 2048 && 2ull * (unsigned long long)bo.chunk <= per_wave) bo.chunk *= 2;
        }
        if (bo.refill_min <= 0) bo.refill_min = 8;
        g.sh = pick_shape(s, &bo, false, (int)g.sets.size(), 0, 0, compiled, 0, compiled);
        P.chunk = g.sh.chunk;
        P.dda_cap = g.sh.walk_cap;
        P.refill_min = g.sh.refill_min;
    }
    if ((rc = b->d_cold.ensure(b->groups.size() * sizeof(KCold)))) {
        smcrt_batch_destroy(b);
        return rc;
    }
    for (size_t gi = 0; gi < b->groups.size(); ++gi) {
        HIPCHECK(hipMemcpy(b->d_cold.as<KCold>() + gi, &b->groups[gi].Cc, sizeof(KCold), hipMemcpyHostToDevice));
        b->groups[gi].P.cold = b->d_cold.as<KCold>() + gi;
    }
    HIPCHECK(hipMemcpy(b->d_mats.p, hm.data(), nmats * sizeof(DevMat), hipMemcpyHostToDevice));
    HIPCHECK(hipMemcpy(b->d_info.p, info.data(), ninfo * sizeof(unsigned long long), hipMemcpyHostToDevice));
    HIPCHECK(hipStreamCreate(&b->stream));
    // capture: clear grids, counters and work queues with fill kernels (no memset nodes), then
    // one packet-kernel node per geometry group
    HIPCHECK(hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    k_fill_f64<<<(unsigned)fill_blocks(ndoubles), 256, 0, b->stream>>>(b->d_grids.as<double>(), ndoubles);
    if (b->copies)
        k_fill_f64<<<(unsigned)fill_blocks((size_t)b->copies * ndoubles), 256, 0, b->stream>>>(
            b->d_xgrid.as<double>(), (size_t)b->copies * ndoubles);
    k_fill_u64<<<(unsigned)fill_blocks(nstats), 256, 0, b->stream>>>(b->d_stats.as<unsigned long long>(), nstats);
    k_fill_u64<<<1, 64, 0, b->stream>>>(b->d_work.as<unsigned long long>(), b->groups.size());
    for (BatchGroup& g : b->groups) {
        hipError_t le = launch_packet_kernel(g.scene, false, g.sh, b->stream, g.P);
        if (le != hipSuccess) {
            hipGraph_t partial = nullptr;
            (void)hipStreamEndCapture(b->stream, &partial);
            if (partial) (void)hipGraphDestroy(partial);
            smcrt_batch_destroy(b);
            return fail(SMCRT_EDEVICE, "packet kernel launch during capture failed: %s", hipGetErrorString(le));
        }
    }
    if (b->copies)
        k_reduce_copies<<<(unsigned)fill_blocks(ndoubles), 256, 0, b->stream>>>(b->d_grids.as<double>(),
                                                                                b->d_xgrid.as<double>(), ndoubles,
                                                                                b->copies);
    hipError_t ce = hipStreamEndCapture(b->stream, &b->graph);
    if (ce != hipSuccess) {
        smcrt_batch_destroy(b);
        return fail(SMCRT_EDEVICE, "graph capture failed: %s", hipGetErrorString(ce));
    }
    HIPCHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    *out = b;
    return SMCRT_OK;
}

extern "C" int smcrt_batch_run(smcrt_batch* b) {
    if (!b) return fail(SMCRT_EINVAL, "null batch");
    HIPCHECK(hipSetDevice(b->device));
    if (!b->ev0) HIPCHECK(hipEventCreate(&b->ev0));
    if (!b->ev1) HIPCHECK(hipEventCreate(&b->ev1));
    HIPCHECK(hipEventRecord(b->ev0, b->stream));
    HIPCHECK(hipGraphLaunch(b->exec, b->stream));
    HIPCHECK(hipEventRecord(b->ev1, b->stream));
    HIPCHECK(hipStreamSynchronize(b->stream));
    HIPCHECK(hipEventElapsedTime(&b->replay_ms, b->ev0, b->ev1));
    return SMCRT_OK;
}

extern "C" int smcrt_batch_read(smcrt_batch* b, int32_t set, double* jmean, smcrt_stats* st) {
    if (!b) return fail(SMCRT_EINVAL, "null batch");
    if (set < 0 || set >= (int32_t)b->set_group.size()) return fail(SMCRT_EINVAL, "set %d out of range", set);
    HIPCHECK(hipSetDevice(b->device));
    const BatchGroup& g = b->groups[b->set_group[set]];
    if (jmean)
        HIPCHECK(hipMemcpy(jmean, b->d_grids.as<double>() + b->set_grid_off[set], b->set_nv[set] * sizeof(double),
                           hipMemcpyDeviceToHost));
    if (st) {
        unsigned long long h[NOUT], e[NEV];
        HIPCHECK(hipMemcpy(h, b->d_stats.as<unsigned long long>() + g.stats_off + (size_t)b->set_slot[set] * NOUT,
                           sizeof h, hipMemcpyDeviceToHost));
        // the group's event counters follow its sets' outcome counters
        HIPCHECK(hipMemcpy(e, b->d_stats.as<unsigned long long>() + g.stats_off + g.sets.size() * NOUT, sizeof e,
                           hipMemcpyDeviceToHost));
        memset(st, 0, sizeof *st);
        st->ev_sweeps = e[E_SWEEPS];
        st->ev_region_evals = e[E_REGION];
        st->ev_normals = e[E_NORMALS];
        st->ev_steps = e[E_STEPS];
        st->ev_crossings = e[E_CROSS];
        st->ev_interfaces = e[E_IFACE];
        st->ev_fresnel = e[E_FRESNEL];
        st->err_unreachable = e[E_UNREACH];
        st->err_nan = e[E_NAN];
        st->ev_atomics = e[E_ATOMICS];
        st->kernel_ms = b->replay_ms;
        st->launched = h[O_LAUNCHED];
        st->absorbed = h[O_ABSORBED];
        st->escaped_r = h[O_R];
        st->escaped_t = h[O_T];
        st->killed_surface = h[O_KSURF];
        st->killed_bounce = h[O_KBOUNCE];
        st->scatters = h[O_SCATTERS];
        if (jmean) {
            double sum = 0.0;
            for (size_t i = 0; i < b->set_nv[set]; ++i) sum += jmean[i];
            st->tally_sum = sum;
        }
    }
    return SMCRT_OK;
}

extern "C" int smcrt_batch_destroy(smcrt_batch* b) {
    if (!b) return SMCRT_OK;
    if (b->device >= 0) (void)hipSetDevice(b->device);
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    if (b->ev0) (void)hipEventDestroy(b->ev0);
    if (b->ev1) (void)hipEventDestroy(b->ev1);
    b->d_mats.release();
    b->d_grids.release();
    b->d_stats.release();
    b->d_info.release();
    b->d_work.release();
    b->d_cold.release();
    b->d_xgrid.release();
    delete b;
    return SMCRT_OK;
}

// ------------------------------------------------------------------------------------------
// allocation-free packet loop (smcrt_run_device_noalloc)
// ------------------------------------------------------------------------------------------
// Every buffer the packet loop needs beyond the caller's grid lives in one file-scope device array
// (each GPU holds its own copy, at its own address, from the library's load until the process
// ends): the work counter, outcome/event counters, moments, set table, cold parameters, the scene
// program, materials and polygon accelerator, and as many per-XCD partial grids as fit.  In host
// code `scratch` names a host-side stand-in; its device address comes from hipGetSymbolAddress,
// once per device (kept in g_scratch).  Calls that overlap on one GPU share the array and race.
static __device__ __align__(256) unsigned char scratch[SMCRT_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static unsigned char* g_scratch[256];  // device address of `scratch` per device (nullptr: not yet)
static int g_scratch_cu[256], g_scratch_xcc[256];

static int scratch_for_device(int dev, unsigned char** base, int* num_cu, int* n_xcc) {
    if (dev < 0 || dev >= 256) return fail(SMCRT_EDEVICE, "device ordinal %d out of range", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(SMCRT_EDEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
        int cu = 0, xcc = 1;
        HIPCHECK(hipDeviceGetAttribute(&cu, hipDeviceAttributeMultiprocessorCount, dev));
        if (hipDeviceGetAttribute(&xcc, hipDeviceAttributeNumberOfXccs, dev) != hipSuccess || xcc < 1) xcc = 1;
        g_scratch_cu[dev] = cu;
        g_scratch_xcc[dev] = xcc;
        g_scratch[dev] = (unsigned char*)ptr;
    }
    *base = g_scratch[dev];
    *num_cu = g_scratch_cu[dev];
    *n_xcc = g_scratch_xcc[dev];
    return SMCRT_OK;
}

// The launch's cold parameters and set table, written into the scratch by a kernel from its
// by-value argument (no host-to-device copy), and the counters it clears.
struct NoallocInit {
    KCold cold;
    unsigned long long setinfo[4];
};
__global__ void k_noalloc_init(NoallocInit in, KCold* cold, unsigned long long* setinfo, unsigned long long* ctl,
                               int n_ctl) {
    const int t = threadIdx.x;
    if (t == 0) *cold = in.cold;
    if (t < 4) setinfo[t] = in.setinfo[t];
    for (int i = t; i < n_ctl; i += blockDim.x) ctl[i] = 0ull;
}

static size_t align256(size_t x) { return (x + 255) & ~(size_t)255; }

extern "C" int smcrt_run_device_noalloc(smcrt_scene* s, const smcrt_run_opts* o, double* d_jmean, smcrt_stats* st,
                                        void* stream) {
    if (!s || !o) return fail(SMCRT_EINVAL, "null scene or options");
    if (o->trace_packets != 0 || o->trace_events != 0)
        return fail(SMCRT_EINVAL, "event tracing needs a trace buffer: use smcrt_run");
    int dev = 0;
    HIPCHECK(hipGetDevice(&dev));
    unsigned char* base = nullptr;
    int num_cu = 0, n_xcc = 1;
    int rc = scratch_for_device(dev, &base, &num_cu, &n_xcc);
    if (rc) return rc;
    if ((o->flags & SMCRT_FLAG_NO_TALLY)) d_jmean = nullptr;
    if (o->n_packets == 0) {
        if (st) memset(st, 0, sizeof *st);
        return SMCRT_OK;
    }
    hipStream_t strm = (hipStream_t)stream;

    // the region-indexed material table and per-material mu_s (host copies the uploads read)
    {
        std::vector<DevMat> dm(s->regions.size());
        region_mats(s, s->mats.data(), dm.data());
        if (s->na_mats.size() != dm.size() || memcmp(s->na_mats.data(), dm.data(), dm.size() * sizeof(DevMat)))
            s->na_mats = dm;
        std::vector<float> mus(s->mats.size());
        for (size_t i = 0; i < s->mats.size(); ++i) mus[i] = (float)s->mats[i].mus;
        if (s->na_mus != mus) s->na_mus = mus;
    }

    // scratch layout
    const size_t n_ctl = 8 + (NOUT + NEV) + 24 + 1;  // work, stats, moments, tally sum
    size_t off = 0;
    const size_t o_ctl = off;       off = align256(off + n_ctl * sizeof(unsigned long long));
    const size_t o_set = off;       off = align256(off + 4 * sizeof(unsigned long long));
    const size_t o_cold = off;      off = align256(off + sizeof(KCold));
    const size_t o_nodes = off;     off = align256(off + s->nodes.size() * sizeof(DevNode));
    const size_t o_regions = off;   off = align256(off + s->regions.size() * sizeof(DevRegion));
    const size_t o_mats = off;      off = align256(off + s->na_mats.size() * sizeof(DevMat));
    const size_t o_mus = off;       off = align256(off + s->na_mus.size() * sizeof(float));
    const size_t o_accel = off;     off = align256(off + s->accel.size() * sizeof(float));
    const size_t fixed = off;
    if (fixed > (size_t)SMCRT_SCRATCH_BYTES)
        return fail(SMCRT_ERANGE,
                    "the scene needs %zu bytes of scratch (program %zu, polygon accelerator %zu); the allocation-free "
                    "entry point has %d",
                    fixed, s->nodes.size() * sizeof(DevNode), s->accel.size() * sizeof(float), SMCRT_SCRATCH_BYTES);
    // per-XCD partial grids: as many as fit in the rest (fewer than 2: tally straight into d_jmean)
    const size_t nvox = (size_t)s->grid.nx * s->grid.ny * s->grid.nz;
    int copies = 0;
    if (d_jmean && !(o->flags & SMCRT_FLAG_SHARED_GRID) && n_xcc >= 2) {
        const size_t fit = ((size_t)SMCRT_SCRATCH_BYTES - fixed) / (nvox * sizeof(double));
        copies = (int)std::min(fit, (size_t)n_xcc);
        if (copies < 2) copies = 0;
    }
    unsigned long long* ctl = (unsigned long long*)(base + o_ctl);
    unsigned long long* setinfo = (unsigned long long*)(base + o_set);
    KCold* d_cold = (KCold*)(base + o_cold);
    double* xgrid = copies ? (double*)(base + fixed) : nullptr;

    // the scene's program, materials and accelerator (the scene's host vectors outlive the copies)
    HIPCHECK(hipMemcpyAsync(base + o_nodes, s->nodes.data(), s->nodes.size() * sizeof(DevNode), hipMemcpyHostToDevice, strm));
    HIPCHECK(hipMemcpyAsync(base + o_regions, s->regions.data(), s->regions.size() * sizeof(DevRegion),
                            hipMemcpyHostToDevice, strm));
    HIPCHECK(hipMemcpyAsync(base + o_mats, s->na_mats.data(), s->na_mats.size() * sizeof(DevMat), hipMemcpyHostToDevice, strm));
    HIPCHECK(hipMemcpyAsync(base + o_mus, s->na_mus.data(), s->na_mus.size() * sizeof(float), hipMemcpyHostToDevice, strm));
    if (!s->accel.empty())
        HIPCHECK(hipMemcpyAsync(base + o_accel, s->accel.data(), s->accel.size() * sizeof(float), hipMemcpyHostToDevice, strm));

    KParams P;
    NoallocInit init;
    memset(&init, 0, sizeof init);
    KCold& Cc = init.cold;
    const int saved_cu = s->num_cu;
    s->num_cu = num_cu;  // pick_shape sizes the persistent grid from it
    fill_common(s, P, Cc);
    P.nodes = (const DevNode*)(base + o_nodes);
    P.regions = (const DevRegion*)(base + o_regions);
    P.accel = s->accel.empty() ? nullptr : (const float*)(base + o_accel);
    P.mats = (const DevMat*)(base + o_mats);
    Cc.mat_mus = (const float*)(base + o_mus);
    P.jmean = d_jmean;
    P.xgrid = xgrid;
    P.xstride = copies ? nvox : 0;
    P.xcopies = copies;
    P.seed_lo = (uint32_t)o->seed;
    P.seed_hi = (uint32_t)(o->seed >> 32);
    P.total = o->n_packets;
    P.nsets = 1;
    init.setinfo[0] = 0ull;
    init.setinfo[1] = o->n_packets;
    init.setinfo[2] = o->first_packet;
    Cc.set_prefix = setinfo;
    Cc.set_first = setinfo + 2;
    P.work = ctl;
    Cc.stats = ctl + 8;
    Cc.moments = (double*)(ctl + 8 + NOUT + NEV);
    Cc.flags = o->flags;
    Cc.max_bounces = o->max_bounces > 0 ? o->max_bounces : 1000;
    Cc.n_det = 0;  // detector counts are not returned by the device-grid entry points
    P.cold = d_cold;
    // the scene's compiled kernels when they are already built on this device (building them
    // loads a module, which allocates); the library's otherwise.  The LDS tile when one has been
    // chosen (its pilot run allocates a grid); none otherwise.
    const bool compiled = jit_wanted(o) && s->jit.ok && s->device == dev;
    const bool use_tile = d_jmean && !(o->flags & SMCRT_FLAG_NO_LDS_TILE) && s->tile_ready;
    int tile_vox = 0;
    if (use_tile) {
        const int wi = register_blocks(s, false, compiled ? compiled_waves(o) : launch_waves(s, o, false), compiled) - 2;
        for (int a = 0; a < 3; ++a) { P.t0[a] = s->tile_t0[wi][a]; P.tn[a] = s->tile_tn[wi][a]; }
        tile_vox = P.tn[0] * P.tn[1] * P.tn[2];
    }
    LaunchShape sh = pick_shape(s, o, false, 1, tile_vox, P.nmat, compiled);
    s->num_cu = saved_cu;
    P.chunk = sh.chunk;
    P.dda_cap = sh.walk_cap;
    P.refill_min = sh.refill_min;

    k_noalloc_init<<<1, 64, 0, strm>>>(init, d_cold, setinfo, ctl, (int)n_ctl);
    HIPCHECK(hipGetLastError());
    if (copies) {
        k_fill_f64<<<(unsigned)fill_blocks((size_t)copies * nvox), 256, 0, strm>>>(xgrid, (size_t)copies * nvox);
        HIPCHECK(hipGetLastError());
    }
    HIPCHECK(launch_packet_kernel(s, false, sh, strm, P));
    HIPCHECK(hipGetLastError());
    if (copies) {
        k_reduce_copies<<<(unsigned)fill_blocks(nvox), 256, 0, strm>>>(d_jmean, xgrid, nvox, copies);
        HIPCHECK(hipGetLastError());
    }
    if (st) {
        double* d_sum = (double*)(ctl + 8 + NOUT + NEV + 24);
        if (d_jmean) {
            k_sum_f64<<<(unsigned)fill_blocks(nvox), 256, 0, strm>>>(d_jmean, nvox, d_sum);
            HIPCHECK(hipGetLastError());
        }
        unsigned long long h[8 + NOUT + NEV + 24 + 1];
        HIPCHECK(hipMemcpyAsync(h, ctl, sizeof h, hipMemcpyDeviceToHost, strm));
        HIPCHECK(hipStreamSynchronize(strm));
        const unsigned long long* c = h + 8;
        const unsigned long long* e = c + NOUT;
        memset(st, 0, sizeof *st);
        st->launched = c[O_LAUNCHED];
        st->absorbed = c[O_ABSORBED];
        st->escaped_r = c[O_R];
        st->escaped_t = c[O_T];
        st->killed_surface = c[O_KSURF];
        st->killed_bounce = c[O_KBOUNCE];
        st->scatters = c[O_SCATTERS];
        st->ev_sweeps = e[E_SWEEPS];
        st->ev_region_evals = e[E_REGION];
        st->ev_normals = e[E_NORMALS];
        st->ev_steps = e[E_STEPS];
        st->ev_crossings = e[E_CROSS];
        st->ev_interfaces = e[E_IFACE];
        st->ev_fresnel = e[E_FRESNEL];
        st->err_unreachable = e[E_UNREACH];
        st->err_nan = e[E_NAN];
        st->ev_atomics = e[E_ATOMICS];
        memcpy(st->moments, e + NEV, sizeof st->moments);
        st->kernel_ms = 0.0;  // no timing events (creating them allocates)
        if (d_jmean) memcpy(&st->tally_sum, e + NEV + 24, sizeof(double));
    }
    return SMCRT_OK;
}

// ------------------------------------------------------------------------------------------
// output helpers (src/writer.f90)
// ------------------------------------------------------------------------------------------
extern "C" int smcrt_normalise_fluence(const smcrt_grid* g, const double* raw, uint64_t n, float* out) {
    if (!g || !raw || !out || n == 0) return fail(SMCRT_EINVAL, "bad arguments");
    const double dx = 2.0 * g->xmax / g->nx, dy = 2.0 * g->ymax / g->ny, dz = 2.0 * g->zmax / g->nz;
    const double f = (2.0 * g->xmax * 2.0 * g->ymax) / ((double)n * dx * dy * dz);
    const size_t nv = (size_t)g->nx * g->ny * g->nz;
    for (size_t i = 0; i < nv; ++i) out[i] = (float)(raw[i] * f);
    return SMCRT_OK;
}

extern "C" int smcrt_write_nrrd(const char* path, const float* data, int32_t nx, int32_t ny, int32_t nz,
                                const char* meta) {
    if (!path || !data || nx < 1 || ny < 1 || nz < 1) return fail(SMCRT_EINVAL, "bad arguments");
    FILE* f = fopen(path, "wb");
    if (!f) return fail(SMCRT_EIO, "cannot open %s for writing", path);
    fprintf(f, "NRRD0004\ntype: float\ndimension: 3\nsizes: %d %d %d\nspace dimension: 3\nencoding: raw\nendian: little\n",
            nx, ny, nz);
    if (meta && *meta) {
        fputs(meta, f);
        size_t l = strlen(meta);
        if (meta[l - 1] != '\n') fputc('\n', f);
    }
    fputs("\n\n", f);  // `write(u,"(A)") new_line("C")`: a newline plus the record end
    size_t nv = (size_t)nx * ny * nz;
    size_t w = fwrite(data, sizeof(float), nv, f);
    int bad = (w != nv) | (fclose(f) != 0);
    if (bad) return fail(SMCRT_EIO, "short write to %s", path);
    return SMCRT_OK;
}

extern "C" const char* smcrt_scene_kernel(const smcrt_scene* s, uint32_t* scene_mask, uint32_t* variant_mask) {
    if (!s) return nullptr;
    const KernelVariant& v = kernel_variant(s);
    if (scene_mask) *scene_mask = s->kind_mask;
    if (variant_mask) *variant_mask = v.mask;
    smcrt_scene* m = const_cast<smcrt_scene*>(s);
    m->kernel_name = v.name;
    if (s->jit.ok) m->kernel_name += s->jit.cached ? "+compiled(cached)" : "+compiled";
    else if (s->jit.tried) m->kernel_name += "+compile-failed: " + s->jit.error;
    if (s->last_binned) m->kernel_name += "+binned";
    return m->kernel_name.c_str();
}
