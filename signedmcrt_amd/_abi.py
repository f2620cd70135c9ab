"""ctypes mirror of include/smcrt.h and the loader for libsmcrt.so.

The structures here must match the C header field for field; tests/test_abi.py checks every
size against the library's own `smcrt_sizeof_*` exports.  Loading fails loudly: there is no
fallback path when the HIP library is missing.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsmcrt.so")

ABI_VERSION = 3

# status codes
OK, EINVAL, ENOMEM, EDEVICE, ERANGE, EIO = 0, -1, -2, -3, -4, -5

# node ops / kinds (smcrt.h)
OP_PRIM, OP_CSG, OP_MOD_BEGIN, OP_MOD_END = 1, 2, 3, 4
SPHERE, BOX, TORUS, CYLINDER, TRIPRISM, SEGMENT, CAPSULE, CONE, EGG, PLANE, POLYGON = range(1, 12)
UNION, SMOOTH_UNION, SUBTRACTION, INTERSECTION = 1, 2, 3, 4
EXTRUDE, REVOLUTION, ONION, ELONGATE, TWIST, BEND = range(1, 7)
SRC_PENCIL, SRC_UNIFORM, SRC_CIRCULAR, SRC_POINT = 1, 2, 3, 4

FLAG_MOMENTS = 1 << 0
FLAG_END_EARLY = 1 << 1
FLAG_NO_TALLY = 1 << 2
FLAG_NO_LDS_TILE = 1 << 3
FLAG_WEIGHT = 1 << 4
FLAG_SHARED_GRID = 1 << 5
FLAG_DIRECT_TALLY = 1 << 6
FLAG_BINNED_TALLY = 1 << 7
FLAG_GRID_COPIES_SHIFT = 8  # flags bits 8-10: log2(partial grids per XCD) + 1, 0 = library default
FLAG_GRID_COPIES_MASK = 7 << FLAG_GRID_COPIES_SHIFT


def flag_grid_copies(per_xcd: int) -> int:
    """The flags field asking for `per_xcd` (a power of two, 1..64) partial grids per XCD."""
    k = int(per_xcd).bit_length() - 1
    if per_xcd < 1 or (1 << k) != per_xcd or k > 6:
        raise ValueError("partial grids per XCD: a power of two from 1 to 64")
    return (k + 1) << FLAG_GRID_COPIES_SHIFT

MAX_REGIONS = 64
MAX_NODES = 4096
MAX_MATERIALS = 64


class Node(C.Structure):
    _fields_ = [("op", C.c_int32), ("kind", C.c_int32), ("k", C.c_double),
                ("p", C.c_double * 8), ("t", C.c_double * 12)]


class Region(C.Structure):
    _fields_ = [("first_node", C.c_int32), ("n_nodes", C.c_int32),
                ("material", C.c_int32), ("layer", C.c_int32)]


class Material(C.Structure):
    _fields_ = [("mus", C.c_double), ("mua", C.c_double), ("hgg", C.c_double), ("n", C.c_double)]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("pad", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double)]


class Source(C.Structure):
    _fields_ = [("kind", C.c_int32), ("pad", C.c_int32),
                ("pos", C.c_double * 3), ("dir", C.c_double * 3),
                ("p1", C.c_double * 3), ("p2", C.c_double * 3), ("p3", C.c_double * 3),
                ("radius", C.c_double)]


DET_CIRCLE, DET_ANNULUS, DET_CAMERA = 1, 2, 3
MAX_DETECTORS = 16
SCRATCH_BYTES = 64 << 20  # the allocation-free entry point's device array per GPU


class Detector(C.Structure):
    _fields_ = [("kind", C.c_int32), ("layer", C.c_int32), ("nbins", C.c_int32),
                ("track_history", C.c_int32), ("pos", C.c_double * 3), ("dir", C.c_double * 3),
                ("p2", C.c_double * 3), ("p3", C.c_double * 3), ("radius1", C.c_double),
                ("radius2", C.c_double), ("maxval", C.c_double)]


class SceneDesc(C.Structure):
    _fields_ = [("abi_version", C.c_uint32), ("n_nodes", C.c_int32), ("n_regions", C.c_int32),
                ("n_materials", C.c_int32),
                ("nodes", C.POINTER(Node)), ("regions", C.POINTER(Region)),
                ("materials", C.POINTER(Material)),
                ("grid", Grid), ("source", Source),
                ("n_detectors", C.c_int32), ("n_poly_vertices", C.c_int32),
                ("detectors", C.POINTER(Detector)), ("poly_vertices", C.POINTER(C.c_double))]


class RunOpts(C.Structure):
    _fields_ = [("n_packets", C.c_uint64), ("seed", C.c_uint64), ("first_packet", C.c_uint64),
                ("flags", C.c_uint32), ("max_bounces", C.c_int32), ("block_size", C.c_int32),
                ("blocks_per_cu", C.c_int32), ("chunk", C.c_int32),
                ("trace_packets", C.c_int32), ("trace_events", C.c_int32),
                ("walk_cap", C.c_int32), ("refill_min", C.c_int32),
                ("grad_mask_lo", C.c_uint32), ("grad_mask_hi", C.c_uint32), ("waves", C.c_int32)]


STAT_COUNTERS = ["launched", "absorbed", "escaped_r", "escaped_t", "killed_surface",
                 "killed_bounce", "scatters", "ev_sweeps", "ev_region_evals", "ev_normals",
                 "ev_steps", "ev_crossings", "ev_interfaces", "ev_fresnel", "err_unreachable",
                 "err_nan", "ev_atomics"]


class Stats(C.Structure):
    _fields_ = [(n, C.c_uint64) for n in STAT_COUNTERS] + [
        ("moments", C.c_double * 24), ("kernel_ms", C.c_double), ("tally_sum", C.c_double)]

    def as_dict(self) -> dict:
        d = {n: int(getattr(self, n)) for n in STAT_COUNTERS}
        d["moments"] = [float(x) for x in self.moments]
        d["kernel_ms"] = float(self.kernel_ms)
        d["tally_sum"] = float(self.tally_sum)
        return d


class Tallies(C.Structure):
    _fields_ = [("jmean", C.POINTER(C.c_double)), ("jmean_f32", C.POINTER(C.c_float)),
                ("trace", C.POINTER(C.c_float)), ("detector_counts", C.POINTER(C.c_uint64)),
                ("stats", Stats)]


class GradOut(C.Structure):
    _fields_ = [("d_r_dmua", C.POINTER(C.c_double)), ("d_r_dmus", C.POINTER(C.c_double)),
                ("d_w_dmua", C.POINTER(C.c_double)), ("d_w_dmus", C.POINTER(C.c_double)),
                ("r_d", C.c_double), ("w_tally", C.c_double), ("stats", Stats)]


class BatchSet(C.Structure):
    _fields_ = [("scene", C.c_int32), ("pad", C.c_int32), ("n_packets", C.c_uint64),
                ("first_packet", C.c_uint64), ("materials", C.POINTER(Material))]


class Config(C.Structure):
    """`smcrt_config`: the TOML input read by smcrt_config_load."""
    _fields_ = [("source_name", C.c_char * 32), ("nphotons", C.c_int64),
                ("has_position", C.c_int32), ("has_direction", C.c_int32),
                ("position", C.c_double * 3), ("direction", C.c_double * 3),
                ("point1", C.c_double * 3), ("point2", C.c_double * 3), ("point3", C.c_double * 3),
                ("radius", C.c_double), ("beta", C.c_double), ("radius_hi", C.c_double),
                ("annulus_type", C.c_char * 32),
                ("nxg", C.c_int32), ("nyg", C.c_int32), ("nzg", C.c_int32), ("pad0", C.c_int32),
                ("xmax", C.c_double), ("ymax", C.c_double), ("zmax", C.c_double),
                ("units", C.c_char * 16),
                ("geom_name", C.c_char * 64),
                ("tau", C.c_double), ("musb", C.c_double), ("muab", C.c_double),
                ("musc", C.c_double), ("muac", C.c_double), ("hgg", C.c_double),
                ("num_spheres", C.c_int32), ("pad1", C.c_int32),
                ("fluence", C.c_char * 256), ("absorb", C.c_char * 256), ("render", C.c_char * 256),
                ("render_geom", C.c_int32), ("overwrite", C.c_int32),
                ("render_size", C.c_int32 * 3), ("pad2", C.c_int32),
                ("iseed", C.c_int64), ("tev", C.c_int32), ("absorb_flag", C.c_int32),
                ("load_checkpoint", C.c_int32), ("checkpoint_every_n", C.c_int32),
                ("checkpoint_file", C.c_char * 256),
                ("n_detectors", C.c_int32), ("pad3", C.c_int32),
                ("detectors", Detector * MAX_DETECTORS),
                ("history_file", C.c_char * 256)]


STRUCTS = {"node": Node, "region": Region, "material": Material, "grid": Grid,
           "source": Source, "scene_desc": SceneDesc, "run_opts": RunOpts, "stats": Stats,
           "tallies": Tallies, "grad_out": GradOut, "batch_set": BatchSet,
           "detector": Detector, "config": Config}

# every symbol include/smcrt.h declares (checked by tests/test_abi.py)
EXPORTS = ["smcrt_abi_version", "smcrt_last_error", "smcrt_sizeof", "smcrt_scene_create", "smcrt_scene_destroy",
           "smcrt_scene_set_materials", "smcrt_run", "smcrt_run_device", "smcrt_run_device_noalloc",
           "smcrt_evaluate",
           "smcrt_grad", "smcrt_batch_create", "smcrt_batch_run", "smcrt_batch_read",
           "smcrt_batch_destroy", "smcrt_normalise_fluence", "smcrt_write_nrrd",
           "smcrt_detector_bins", "smcrt_config_load", "smcrt_write_detector", "smcrt_render",
           "smcrt_scene_kernel"]

_lib = None


class SmcrtError(RuntimeError):
    """A non-zero status from libsmcrt, carrying smcrt_last_error()'s text."""

    def __init__(self, status: int, msg: str):
        super().__init__(f"smcrt status {status}: {msg}")
        self.status = status


def lib() -> C.CDLL:
    """The loaded libsmcrt.so (built by __graft_entry__.build()).  Raises if it is absent."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `python -c "
                          f"'import __graft_entry__ as g; g.build()'` (no CPU fallback exists)")
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    sig = {
        "smcrt_abi_version": (C.c_int, []),
        "smcrt_last_error": (C.c_int, [C.c_char_p, C.c_size_t]),
        "smcrt_scene_create": (C.c_int, [P(SceneDesc), P(C.c_void_p)]),
        "smcrt_scene_destroy": (C.c_int, [C.c_void_p]),
        "smcrt_scene_set_materials": (C.c_int, [C.c_void_p, P(Material), C.c_int32]),
        "smcrt_run": (C.c_int, [C.c_void_p, P(RunOpts), P(Tallies)]),
        "smcrt_run_device": (C.c_int, [C.c_void_p, P(RunOpts), C.c_void_p, P(Stats), C.c_void_p]),
        "smcrt_run_device_noalloc": (C.c_int, [C.c_void_p, P(RunOpts), C.c_void_p, P(Stats), C.c_void_p]),
        "smcrt_evaluate": (C.c_int, [C.c_void_p, C.c_int32, C.c_int64, P(C.c_float),
                                     P(C.c_float), P(C.c_float)]),
        "smcrt_grad": (C.c_int, [C.c_void_p, P(RunOpts), P(C.c_float), P(C.c_double), P(GradOut)]),
        "smcrt_batch_create": (C.c_int, [P(C.c_void_p), C.c_int32, P(BatchSet), C.c_int32,
                                         C.c_uint64, P(RunOpts), P(C.c_void_p)]),
        "smcrt_batch_run": (C.c_int, [C.c_void_p]),
        "smcrt_batch_read": (C.c_int, [C.c_void_p, C.c_int32, P(C.c_double), P(Stats)]),
        "smcrt_batch_destroy": (C.c_int, [C.c_void_p]),
        "smcrt_normalise_fluence": (C.c_int, [P(Grid), P(C.c_double), C.c_uint64, P(C.c_float)]),
        "smcrt_write_nrrd": (C.c_int, [C.c_char_p, P(C.c_float), C.c_int32, C.c_int32, C.c_int32,
                                       C.c_char_p]),
        "smcrt_sizeof": (C.c_int64, [C.c_char_p]),
        "smcrt_detector_bins": (C.c_int64, [C.c_void_p, C.c_int32]),
        "smcrt_render": (C.c_int, [C.c_void_p, P(C.c_int32), P(C.c_float)]),
        "smcrt_scene_kernel": (C.c_char_p, [C.c_void_p, P(C.c_uint32), P(C.c_uint32)]),
        "smcrt_config_load": (C.c_int, [C.c_char_p, P(Config)]),
        "smcrt_write_detector": (C.c_int, [C.c_char_p, P(Detector), P(C.c_uint64)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    if L.smcrt_abi_version() != ABI_VERSION:
        raise ImportError("libsmcrt.so ABI version mismatch")
    _lib = L
    return L


def last_error() -> str:
    buf = C.create_string_buffer(1024)
    lib().smcrt_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(status: int) -> None:
    if status != OK:
        raise SmcrtError(status, last_error())
