/*
 * smcrt.h — C ABI of the MI355X photon-packet engine (libsmcrt.so).
 *
 * The library replaces the body of signedMCRT's packet loop: the OpenMP region of
 * `pathlength_scatter` (src/kernelsMod.f90:239-329) together with everything it calls per packet
 * (`tauint2`/`update_grids`, src/inttau2.f90:15-441; `scatter`, src/photon.f90:713-779; the
 * emitters, src/photon.f90:214-338,397-505; `reflect_refract`, src/surfaces.f90:14-120; the SDF
 * `evaluate`/`calcNormal` stack under src/sdfs/).  It is plain C: fixed-width integers, doubles
 * and pointers only, so Fortran binds it with `iso_c_binding` (signedmcrt_amd/fortran/smcrt.f90)
 * and Python with ctypes (signedmcrt_amd/_abi.py).
 *
 * Conventions
 *  - Every entry point returns an int status: SMCRT_OK (0) or a negative SMCRT_E* code.  The text
 *    of the last error on the calling thread is available from smcrt_last_error().  Nothing aborts
 *    the process (the reference's `error stop` paths become status codes or per-packet
 *    diagnostic counters, see smcrt_stats).
 *  - Host buffers are caller-owned; the library owns device memory and copies results back
 *    before returning.  smcrt_run_device() is the variant whose tally pointer is device memory.
 *  - Fluence grids are column-major like the reference's `jmean(nxg,nyg,nzg)`
 *    (src/iarray.f90:12): voxel (i,j,k), 0-based, lives at i + nx*(j + ny*k).
 *  - A scene handle is not re-entrant: one caller thread per handle.
 */
#ifndef SMCRT_H
#define SMCRT_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SMCRT_ABI_VERSION 3

/* ---- status codes ---------------------------------------------------------------------- */
enum {
    SMCRT_OK = 0,
    SMCRT_EINVAL = -1,   /* malformed descriptor / argument            */
    SMCRT_ENOMEM = -2,   /* host or device allocation failed           */
    SMCRT_EDEVICE = -3,  /* HIP runtime error or no usable GPU         */
    SMCRT_ERANGE = -4,   /* size beyond a documented limit             */
    SMCRT_EIO = -5       /* file could not be read or written          */
};

/* ---- SDF program ------------------------------------------------------------------------
 * A scene is the reference's `type(sdf)` array (src/sdfs/sdf_base.f90:29-39) flattened into a
 * postfix program.  Each array element is a *region*: a contiguous node range whose evaluation
 * leaves one signed distance.  Region order is array order; it decides the first-negative rule
 * of the interface logic (src/inttau2.f90:131-140) and the layer chosen at emission
 * (src/kernelsMod.f90:275-280).
 *
 * Node ops:
 *  PRIM       push prim(q), q = [x y z 1]·T with T the 4x3 `t` (src/vector_class.f90:292-304)
 *  CSG        d2 = pop, d1 = pop, push op(d1, d2, k); a `model` fold (src/sdfs/sdf_base.f90:145-160)
 *             is child1 child2 CSG child3 CSG ...
 *  MOD_BEGIN  push a modified evaluation position (twist, bend, revolution, elongate) or the
 *             same position (extrude, onion) — src/sdfs/sdfModifiers.f90:273-392
 *  MOD_END    pop that position and post-process the top distance (extrude, onion, elongate)
 * Value-stack depth <= SMCRT_MAX_VALUE_STACK, position-stack depth <= SMCRT_MAX_POS_STACK.
 */
#define SMCRT_MAX_VALUE_STACK 4
#define SMCRT_MAX_POS_STACK 3
#define SMCRT_MAX_REGIONS 64
#define SMCRT_MAX_NODES 4096
#define SMCRT_MAX_MATERIALS 64

enum { SMCRT_OP_PRIM = 1, SMCRT_OP_CSG = 2, SMCRT_OP_MOD_BEGIN = 3, SMCRT_OP_MOD_END = 4 };

/* primitives (src/sdfs/sdfs.f90:494-735); parameters p[] in order: */
enum {
    SMCRT_SPHERE = 1,   /* radius                                                         */
    SMCRT_BOX = 2,      /* half-lengths x,y,z (the constructor halves `lengths`, :510-525) */
    SMCRT_TORUS = 3,    /* oradius, iradius                                               */
    SMCRT_CYLINDER = 4, /* a.x a.y a.z b.x b.y b.z radius                                 */
    SMCRT_TRIPRISM = 5, /* h1, h2                                                         */
    SMCRT_SEGMENT = 6,  /* a.x a.y a.z b.x b.y b.z (tube thickness fixed at 0.1, :624)     */
    SMCRT_CAPSULE = 7,  /* a.x a.y a.z b.x b.y b.z r                                       */
    SMCRT_CONE = 8,     /* a.x a.y a.z b.x b.y b.z ra rb                                   */
    SMCRT_EGG = 9,      /* r1, r2, h                                                      */
    SMCRT_PLANE = 10,   /* a.x a.y a.z (unit normal)                                       */
    SMCRT_POLYGON = 11  /* first vertex, vertex count: a closed polygon in the xy plane from
                           smcrt_scene_desc.poly_vertices (no reference counterpart: the
                           reference's `segment` is a tube, not a region).  Signed 2-D distance
                           (z ignored; wrap it in SMCRT_EXTRUDE for a prism), evaluated through a
                           per-polygon uniform-grid accelerator built by smcrt_scene_create.
                           The polygon must be simple; either orientation.                  */
};
/* CSG operators (src/sdfs/sdfModifiers.f90:412-475) */
enum { SMCRT_UNION = 1, SMCRT_SMOOTH_UNION = 2, SMCRT_SUBTRACTION = 3, SMCRT_INTERSECTION = 4 };
/* modifiers (src/sdfs/sdfModifiers.f90:273-392); parameters p[]: */
enum {
    SMCRT_EXTRUDE = 1,    /* h            */
    SMCRT_REVOLUTION = 2, /* o            */
    SMCRT_ONION = 3,      /* thickness    */
    SMCRT_ELONGATE = 4,   /* size x,y,z   */
    SMCRT_TWIST = 5,      /* k            */
    SMCRT_BEND = 6        /* k            */
};

typedef struct smcrt_node {
    int32_t op;    /* SMCRT_OP_*                                                      */
    int32_t kind;  /* primitive / CSG / modifier kind                                 */
    double k;      /* CSG smoothing parameter (`model%k`)                             */
    double p[8];   /* shape or modifier parameters                                    */
    double t[12];  /* PRIM only: rows 1..4 of columns 1..3 of the 4x4 `transform`,
                      row-major: q_j = x*t[j] + y*t[3+j] + z*t[6+j] + t[9+j]          */
} smcrt_node;

typedef struct smcrt_region {
    int32_t first_node; /* index of the region's first node                          */
    int32_t n_nodes;    /* number of nodes                                           */
    int32_t material;   /* index into the material table                             */
    int32_t layer;      /* the reference's `layer` id (informational)                */
} smcrt_region;

/* `mono(mus, mua, hgg, n)` (src/opticalProps/opticalProperties.f90:107-125).  The derived
   kappa = mus + mua, albedo = mus/kappa (1 if mua < 1e-9) and g2 = hgg^2 are computed inside. */
typedef struct smcrt_material {
    double mus, mua, hgg, n;
} smcrt_material;

/* `cart_grid` (src/grid.f90:14-25, init_grid :53-93): counts and HALF sizes, centred at 0. */
typedef struct smcrt_grid {
    int32_t nx, ny, nz, pad;
    double xmax, ymax, zmax;
} smcrt_grid;

/* sources (src/photon.f90): pencil :464-505, uniform :397-461, circular :214-287, point :290-338 */
enum { SMCRT_SRC_PENCIL = 1, SMCRT_SRC_UNIFORM = 2, SMCRT_SRC_CIRCULAR = 3, SMCRT_SRC_POINT = 4 };

typedef struct smcrt_source {
    int32_t kind, pad;
    double pos[3];   /* `position` (pencil, point, circular — the disc is centred at -pos)  */
    double dir[3];   /* `direction` (pencil, uniform, circular)                           */
    double p1[3];    /* uniform: pos = p1 + u*p2 + v*p3                                   */
    double p2[3];
    double p3[3];
    double radius;   /* circular                                                          */
} smcrt_source;

/* ---- detectors (src/detectors/detectors.f90, src/detectors/detector_base.f90) ------------
 * Every packet, when it ends, is offered to every detector with the hit point the reference
 * builds (src/kernelsMod.f90:302-306): its final position and direction, value = sqrt(x^2+y^2)
 * of the final position, and its layer.  A detector counts it when `layer` matches and:
 *  circle   the ray from the final position along the direction meets the detector plane
 *           (normal `dir`, denominator > 1e-6) within `radius1` of `pos`, at t <= 5e-3
 *           (check_hit_circle, intersectCircle/intersectPlane in src/geometryMod.f90:217-271)
 *  annulus  radius1 <= |final position - pos| <= radius2 (check_hit_annulus)
 *  camera   the ray meets the plane through p1 with normal unit((p3-p1) x (p2-p1)) at t >= 0
 *           inside the p1/p2/p3 parallelogram (check_hit_camera)
 * 1-D detectors (circle, annulus) have nbins+1 bins of width maxval/(nbins-1) (circle) or
 * maxval/nbins (annulus) — 1 if nbins == 0 — and bin min(nint(value/width)+1, nbins+1);
 * the camera has (nbins+1)^2 bins of width maxval/(nbins+1), indexed by
 * (final z + p1.x, final y + p1.y) as record_hit_2D_sub does.  Counts are exact (64-bit). */
#define SMCRT_MAX_DETECTORS 16
enum { SMCRT_DET_CIRCLE = 1, SMCRT_DET_ANNULUS = 2, SMCRT_DET_CAMERA = 3 };

typedef struct smcrt_detector {
    int32_t kind;          /* SMCRT_DET_*                                                      */
    int32_t layer;         /* the reference's 1-based layer (region index + 1)                 */
    int32_t nbins;         /* as written in the input (the detector holds nbins+1 bins)        */
    int32_t track_history; /* accepted and ignored (history stacks are outside this path)      */
    double pos[3];         /* circle/annulus `position`; camera `p1`                          */
    double dir[3];         /* circle/annulus `direction`                                       */
    double p2[3], p3[3];   /* camera                                                           */
    double radius1, radius2, maxval;
} smcrt_detector;

typedef struct smcrt_scene_desc {
    uint32_t abi_version; /* SMCRT_ABI_VERSION                                           */
    int32_t n_nodes, n_regions, n_materials;
    const smcrt_node* nodes;
    const smcrt_region* regions;
    const smcrt_material* materials;
    smcrt_grid grid;
    smcrt_source source;
    int32_t n_detectors;  /* 0..SMCRT_MAX_DETECTORS                                       */
    int32_t n_poly_vertices; /* polygon vertices (x, y pairs) referenced by SMCRT_POLYGON nodes */
    const smcrt_detector* detectors;
    const double* poly_vertices; /* 2 * n_poly_vertices doubles (may be NULL if none)        */
} smcrt_scene_desc;

/* ---- run options, statistics, tallies -------------------------------------------------- */
enum {
    SMCRT_FLAG_MOMENTS = 1u << 0,   /* accumulate 1st/2nd moments of scatter sites 1..4
                                       (the `test_kernel` harness, src/kernelsMod.f90:335-448) */
    SMCRT_FLAG_END_EARLY = 1u << 1, /* terminate a packet at its 5th scatter (same harness)   */
    SMCRT_FLAG_NO_TALLY = 1u << 2,  /* skip the fluence tally                                  */
    SMCRT_FLAG_NO_LDS_TILE = 1u << 3, /* tally straight to HBM (no LDS-privatised tile)         */
    SMCRT_FLAG_WEIGHT = 1u << 4,    /* the `weight_scatter` estimator (src/kernelsMod.f90:12-163):
                                       at each interaction deposit w(1 - albedo) into its voxel
                                       (on top of the path-length tally, as the reference does),
                                       always scatter, Russian roulette below w = 0.01 (survival
                                       0.1, weight x10); Fresnel events scale w by R; detectors
                                       bin the final weight.  Roulette kills count as absorbed. */
    SMCRT_FLAG_SHARED_GRID = 1u << 5, /* tally every block into the one grid (default: one FP64
                                       partial grid per XCD, summed after the launch; same
                                       result up to FP64 summation order, ~2x faster where
                                       tally atomics dominate; costs n_xcd x grid memory)      */
    SMCRT_FLAG_DIRECT_TALLY = 1u << 6, /* chords outside the LDS tile go to FP64 atomics      */
    SMCRT_FLAG_BINNED_TALLY = 1u << 7  /* chords outside the LDS tile are staged per grid bin
                                       (16384 voxels) in LDS, written as plain stores into
                                       device pages and summed per bin after the launch (device
                                       memory: ~8 B per chord, within SMCRT_TALLY_POOL_MB, default
                                       half the free memory up to 64 GiB).  Neither flag: binned
                                       where the scene's pilot run puts less than 35 % of the
                                       path length into the tile and hands the global tally more
                                       than 1.5 entries per sweep.  Forward single-set launches
                                       of the run-time compiled kernels only (the library's
                                       kernels take the atomics); same result up to FP64
                                       summation order                                          */
};

/* Partial grids per XCD (flags bits 8-10 hold log2(copies per XCD) + 1; 0 -> the library's
   default: 1 per XCD for single-set launches, 8 for batches, whose chords have no LDS tile).  The
   tally's global FP64 atomics spread over the copies, so hot voxels contend less; each copy costs
   one grid of device memory and a clear and a sum per launch.  Up to 8 GiB of copies by default
   and 40 GiB when more than one per XCD is asked for, within half the free device memory.
   SMCRT_FLAG_GRID_COPIES(k): 2^k partial grids per XCD, k = 0..6.  Same result up to FP64
   summation order. */
#define SMCRT_FLAG_GRID_COPIES_SHIFT 8
#define SMCRT_FLAG_GRID_COPIES_MASK (7u << SMCRT_FLAG_GRID_COPIES_SHIFT)
#define SMCRT_FLAG_GRID_COPIES(k) ((uint32_t)((k) + 1) << SMCRT_FLAG_GRID_COPIES_SHIFT)


typedef struct smcrt_run_opts {
    uint64_t n_packets;    /* packets to launch (`nphotons`)                                 */
    uint64_t seed;         /* RNG key (`iseed`)                                              */
    uint64_t first_packet; /* global id of the first packet: packet j draws from stream
                              (seed, first_packet + j), so results do not depend on launch
                              shape and shards/CRN runs line up                              */
    uint32_t flags;        /* SMCRT_FLAG_*                                                   */
    int32_t max_bounces;   /* bounce cap, 0 -> 1000 (src/inttau2.f90:177)                    */
    int32_t block_size;    /* 0 -> library default (64..256, multiple of 64)                 */
    int32_t blocks_per_cu; /* 0 -> library default                                           */
    int32_t chunk;         /* packets a wavefront claims per refill, 0 -> default            */
    int32_t trace_packets; /* record the first `trace_packets` packets' events (0 = off)    */
    int32_t trace_events;  /* events per traced packet                                       */
    int32_t walk_cap;      /* voxel chords a lane walks per kernel-loop iteration, 0 -> default */
    int32_t refill_min;    /* regeneration policy: a wavefront refills its idle lanes once at
                              least this many are idle (1 = at once; 0 -> default)            */
    uint32_t grad_mask_lo; /* smcrt_grad: materials whose derivatives are computed, bit m =   */
    uint32_t grad_mask_hi; /* material m (0 and 0: every material); at most 16 set            */
    int32_t waves;         /* resident wavefronts per SIMD the packet kernel is built for: 2,
                              3 or 4 (0 -> 3).  3 exists for the kind-specialised kernels (fewer
                              registers, a few spilled; faster where latency dominates over
                              the tally); the general kernel always runs the 2-wave build.
                              4 exists only for the run-time compiled kernels (the library's
                              run 3 instead)                                                  */
} smcrt_run_opts;

/* How a packet ended.  A+R+T+K = launched. */
typedef struct smcrt_stats {
    uint64_t launched;
    uint64_t absorbed;       /* albedo test failed (src/kernelsMod.f90:287-297)              */
    uint64_t escaped_r;      /* left the world or the grid moving against the source dir     */
    uint64_t escaped_t;      /* ... moving with it                                           */
    uint64_t killed_surface; /* started a flight within eps of a surface (src/inttau2.f90:58) */
    uint64_t killed_bounce;  /* bounce cap (src/inttau2.f90:177-180)                          */
    uint64_t scatters;       /* `nscatt`                                                      */
    uint64_t ev_sweeps;      /* full SDF sweeps (every region evaluated once)                 */
    uint64_t ev_region_evals;/* region evaluations outside sweeps                             */
    uint64_t ev_normals;     /* normal (gradient) evaluations                                 */
    uint64_t ev_steps;       /* sphere-trace steps (each one a DDA walk)                      */
    uint64_t ev_crossings;   /* voxel chords tallied                                          */
    uint64_t ev_interfaces;  /* interface crossings resolved                                  */
    uint64_t ev_fresnel;     /* interfaces with n1 != n2                                      */
    uint64_t err_unreachable;/* the reference's "should not be reached" branch (:156)         */
    uint64_t err_nan;        /* non-finite state caught and the packet dropped                */
    uint64_t ev_atomics;     /* FP64 tally atomics issued (chords merged per voxel)           */
    double moments[2][4][3]; /* SMCRT_FLAG_MOMENTS: sum of pos and pos^2 at scatters 1..4    */
    double kernel_ms;        /* device time of the packet kernel (binned tally: through the
                                bin sums)                                                     */
    double tally_sum;        /* sum of the FP64 fluence accumulator (raw path length)        */
} smcrt_stats;

typedef struct smcrt_tallies {
    double* jmean;     /* nx*ny*nz FP64 raw path-length sums (may be NULL)                    */
    float* jmean_f32;  /* optional FP32 copy, the reference's `jmean` type (may be NULL)      */
    float* trace;      /* optional: trace_packets*trace_events*4 floats (x,y,z,event)         */
    uint64_t* detector_counts; /* optional: every detector's bins, concatenated in detector
                                  order (smcrt_detector_bins() each; may be NULL)              */
    smcrt_stats stats;
} smcrt_tallies;

typedef struct smcrt_scene smcrt_scene; /* opaque */

/* ---- entry points ---------------------------------------------------------------------- */

/* ABI version of the loaded library (SMCRT_ABI_VERSION). */
int smcrt_abi_version(void);

/* Copy the calling thread's last error message (NUL-terminated, truncated to len). */
int smcrt_last_error(char* buf, size_t len);

/* sizeof the named ABI struct ("node", "region", "material", "grid", "source", "scene_desc",
   "run_opts", "stats", "tallies", "grad_out", "batch_set"), -1 if unknown.  No reference
   counterpart: lets bindings (ctypes, iso_c_binding) check their struct layouts at load time. */
int64_t smcrt_sizeof(const char* struct_name);

/* Validate and flatten a scene.  Replaces `setup_simulation` + the SDF constructors
   (src/setup.f90:14-54, src/sdfs/sdfs.f90:158-492).  Host work only; device memory is set up
   on first use, so this succeeds without a GPU. */
int smcrt_scene_create(const smcrt_scene_desc* desc, smcrt_scene** out);
int smcrt_scene_destroy(smcrt_scene* scene);

/* Replace the material table of an existing scene (same count).  Used for finite differences
   and for gradient descent without re-flattening the geometry. */
int smcrt_scene_set_materials(smcrt_scene* scene, const smcrt_material* materials, int32_t n);

/* The packet loop.  Replaces the OpenMP region of `pathlength_scatter`
   (src/kernelsMod.f90:239-329): launches opts->n_packets packets, accumulates the path-length
   fluence estimator of `update_grids` (src/inttau2.f90:213-292) in FP64 and returns it in
   out->jmean (raw, un-normalised; smcrt_normalise_fluence applies src/writer.f90:25-50). */
int smcrt_run(smcrt_scene* scene, const smcrt_run_opts* opts, smcrt_tallies* out);

/* Same, with the FP64 accumulator in device memory (d_jmean, nx*ny*nz doubles, added to, not
   cleared) and `stream` a hipStream_t (NULL = default stream).  stats may be NULL.  Blocks until
   the launch has finished. */
int smcrt_run_device(smcrt_scene* scene, const smcrt_run_opts* opts, double* d_jmean,
                     smcrt_stats* stats, void* stream);

/* Bytes of the device array the allocation-free entry point works in (one per GPU, held from the
   library's load until the process ends). */
#define SMCRT_SCRATCH_BYTES (64 << 20)

/* smcrt_run_device without any allocation (no counterpart in the reference): for callers that
   cannot allocate device memory — code under HIP graph capture, tight serving loops, harnesses
   that forbid hipMalloc.  Same packets, counters and fluence as smcrt_run_device (up to FP64
   summation order); the caller passes the grid (d_jmean) and the stream (NULL = default stream).
   Every other buffer — counters, launch parameters, the scene's program, materials and polygon
   accelerator, and as many per-XCD partial grids as fit — lives in a file-scope device array of
   SMCRT_SCRATCH_BYTES per GPU; a scene whose program and accelerator do not fit is rejected
   (SMCRT_ERANGE), a grid whose partial copies do not fit is tallied directly.  It runs the
   scene's run-time compiled kernels only if an earlier call built them on this device, and the
   LDS tile only if an earlier call chose it.  Event tracing is rejected (SMCRT_EINVAL); detector
   bins are not recorded; stats->kernel_ms is 0.  With stats == NULL the call only enqueues work
   on `stream` and returns; with stats it waits for the launch.  Calls that overlap on one GPU share
   the array and race: serialise them. */
int smcrt_run_device_noalloc(smcrt_scene* scene, const smcrt_run_opts* opts, double* d_jmean,
                             smcrt_stats* stats, void* stream);

/* SDF evaluation on the device: the reference's `sdf%evaluate` (src/sdfs/sdf_base.f90:251-259)
   and `calcNormal` (:165-189) for region `region` at n host points (xyz interleaved).
   dist[n] and/or normals[3n] may be NULL.  FP32 on the device; normals are exact gradients
   (forward-mode derivatives) rather than finite differences. */
int smcrt_evaluate(smcrt_scene* scene, int32_t region, int64_t n, const float* points,
                   float* dist, float* normals);

/* ---- perturbation-MC adjoint (no reference counterpart) -------------------------------- */
typedef struct smcrt_grad_out {
    double* d_r_dmua;   /* n_materials: d(R_d)/d(mu_a) per material                          */
    double* d_r_dmus;   /* n_materials: d(R_d)/d(mu_s)                                      */
    double* d_w_dmua;   /* n_materials: d(sum_v w_v * raw_jmean_v / N)/d(mu_a)                */
    double* d_w_dmus;   /* n_materials: same for mu_s                                       */
    double r_d;         /* escaped_r / launched of this run                                 */
    double w_tally;     /* sum_v w_v * raw_jmean_v / N of this run                           */
    smcrt_stats stats;
} smcrt_grad_out;

/* Runs the packet loop with per-packet score accumulators (path length and scatter count per
   material) and returns score-function derivatives of R_d and of the weighted tally
   sum_v w_v*jmean_v (w = weights, nx*ny*nz floats on the host; NULL means w = 1, i.e. the
   fluence sum).  jmean may be NULL. */
int smcrt_grad(smcrt_scene* scene, const smcrt_run_opts* opts, const float* weights,
               double* jmean, smcrt_grad_out* out);

/* ---- batched launch (no reference counterpart) ----------------------------------------- */
typedef struct smcrt_batch_set {
    int32_t scene;       /* index into the scenes[] array of the batch                       */
    int32_t pad;
    uint64_t n_packets;  /* this set's budget                                                */
    uint64_t first_packet;
    const smcrt_material* materials; /* this set's material table (the scene's count)        */
} smcrt_batch_set;

typedef struct smcrt_batch smcrt_batch; /* opaque: device buffers + captured graph */

/* Build a batch of n_sets property sets over n_scenes geometries.  Every set gets its own
   FP64 grid (its scene's grid shape) and its own stats.  The launch is captured once into a
   HIP graph and replayed by smcrt_batch_run.  `shape` (may be NULL) supplies the launch shape
   of every geometry group: block_size, blocks_per_cu, chunk, walk_cap, refill_min (its other
   fields are ignored). */
int smcrt_batch_create(smcrt_scene* const* scenes, int32_t n_scenes, const smcrt_batch_set* sets,
                       int32_t n_sets, uint64_t seed, const smcrt_run_opts* shape,
                       smcrt_batch** out);
/* One persistent launch over every set (graph replay).  Clears the grids first. */
int smcrt_batch_run(smcrt_batch* batch);
/* Copy set `set`'s grid (raw FP64) and stats back; jmean may be NULL.  The outcome counters
   (launched .. scatters) are the set's own; the event and error counters (ev_*, err_*) are those
   of the launch of the set's geometry group, shared by every set on that geometry; kernel_ms is
   the device time of the last replay of the whole graph (every group's launch, the grid clears
   and the partial-grid reduction). */
int smcrt_batch_read(smcrt_batch* batch, int32_t set, double* jmean, smcrt_stats* stats);
int smcrt_batch_destroy(smcrt_batch* batch);

/* `render_sub` (src/sdfs/sdf_base.f90:302-346): voxelise the scene on the GPU.  With
   ns = nint(samples/2) and wid = grid half size / ns per axis, image(i,j,k) (1-based, column-
   major, samples[0] x samples[1] x samples[2] floats) = min over every region of its signed
   distance at ((i-ns_x) wid_x, (j-ns_y) wid_y, (k-ns_z) wid_z). */
int smcrt_render(smcrt_scene* scene, const int32_t samples[3], float* image);

/* Number of bins of detector `i` of a scene (0 if out of range). */
int64_t smcrt_detector_bins(const smcrt_scene* scene, int32_t i);

/* Diagnostics (no reference counterpart): the packet-kernel variant a scene runs.  The kernel
   is compiled for a few sets of primitive/operator/modifier kinds; a scene runs the smallest
   one covering the kinds its program uses.  Writes the scene's kind mask (bit k: primitive k;
   bit 15+k: operator k; bit 23+k: modifier k) and the variant's mask, and returns its name
   ("basic", "layers", "extruded" or "full"; NULL for a NULL scene), followed by "+compiled" or
   "+compiled(cached)" once the scene's kernels were compiled at run time, and "+binned" when the
   scene's last forward launch ran the binned tally (SMCRT_FLAG_BINNED_TALLY). */
const char* smcrt_scene_kernel(const smcrt_scene* scene, uint32_t* scene_mask, uint32_t* variant_mask);

/* ---- input: the reference's TOML schema (src/parse/parse.f90, parse_source.f90,
 * parse_detectors.f90) -------------------------------------------------------------------
 * smcrt_config_load reads a TOML input file into a flat, caller-owned struct, applying the
 * reference's defaults and its error checks (status SMCRT_EINVAL + message from
 * smcrt_last_error; SMCRT_EIO if the file cannot be read).  Strings are NUL-terminated.
 * Departure: a vector `direction` (e.g. [0.0, 0.0, -1.0]) is accepted together with the rest
 * of the source table; the reference returns early there and leaves the packet unset
 * (src/parse/parse_source.f90:99-108). */
typedef struct smcrt_config {
    /* [source] */
    char source_name[32];     /* "point"                                                 */
    int64_t nphotons;         /* 1000000                                                 */
    int32_t has_position, has_direction;
    double position[3], direction[3];
    double point1[3], point2[3], point3[3]; /* (-1,-1,1), (2,0,0), (0,2,0)                */
    double radius, beta, radius_hi;         /* 0.5, 5, 0.6                               */
    char annulus_type[32];                  /* "gaussian"                                */
    /* [grid] */
    int32_t nxg, nyg, nzg, pad0;            /* 200                                       */
    double xmax, ymax, zmax;                /* 1.0                                       */
    char units[16];                         /* "cm"                                      */
    /* [geometry] */
    char geom_name[64];                     /* "sphere"                                  */
    double tau, musb, muab, musc, muac, hgg;/* 10, 0, 0.01, 0, 0.01, 0.7                 */
    int32_t num_spheres, pad1;              /* 10                                        */
    /* [output] */
    char fluence[256], absorb[256], render[256];  /* fluence.nrrd, absorb.nrrd, geom_render.nrrd */
    int32_t render_geom, overwrite;
    int32_t render_size[3], pad2;           /* 200 200 200                               */
    /* [simulation] */
    int64_t iseed;                          /* 123456789                                 */
    int32_t tev, absorb_flag, load_checkpoint, checkpoint_every_n;
    char checkpoint_file[256];
    /* [[detectors]] */
    int32_t n_detectors, pad3;
    smcrt_detector detectors[SMCRT_MAX_DETECTORS];
    char history_file[256];                 /* "photPos.obj"                             */
} smcrt_config;

int smcrt_config_load(const char* path, smcrt_config* out);

/* ---- host-side output helpers (src/writer.f90) ----------------------------------------- */

/* `normalise_fluence` (src/writer.f90:25-50): out = raw * (2x*2y)/(N*dx*dy*dz). */
int smcrt_normalise_fluence(const smcrt_grid* grid, const double* raw, uint64_t n_packets,
                            float* out);

/* write_detected_photons (src/writer.f90:53-84): file `path` gets, for a circle detector, one
   line per bin j = 1..nbins+1 with j*bin_width and the count; annulus and camera detectors get
   an empty file, as in the reference. */
int smcrt_write_detector(const char* path, const smcrt_detector* det, const uint64_t* counts);

/* Write a 3-D FP32 array as NRRD the way `write_3d_r4_nrrd` does (src/writer.f90:220-251,
   294-334): header, optional metadata lines (`meta`, may be NULL), blank line, raw data. */
int smcrt_write_nrrd(const char* path, const float* data, int32_t nx, int32_t ny, int32_t nz,
                     const char* meta);

#ifdef __cplusplus
}
#endif
#endif /* SMCRT_H */
