"""The C-ABI library without a GPU: it loads, exports every symbol include/smcrt.h declares, its
struct layouts match the ctypes/iso_c_binding mirrors, scene validation reports errors as
status codes (the reference's `error stop` paths become return values), and the host-side
output helpers reproduce src/writer.f90."""
from __future__ import annotations

import ctypes as C
import os
import re

import numpy as np
import pytest

from signedmcrt_amd import _abi as A
from signedmcrt_amd import scenes
from signedmcrt_amd.scene import Scene, cart_grid, source
from signedmcrt_amd.sdf import box, model, mono, sphere, union

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "smcrt.h")).read()
    return sorted(set(re.findall(r"^\s*(?:int|int64_t|const char\s*\*)\s*(smcrt_\w+)\s*\(", txt, re.M)))


def test_header_and_mirror_agree():
    assert header_functions() == sorted(A.EXPORTS)


def test_library_exports_every_symbol():
    L = A.lib()
    for name in header_functions():
        assert hasattr(L, name), name
    assert L.smcrt_abi_version() == A.ABI_VERSION


def test_struct_sizes_match_library():
    L = A.lib()
    for name, st in A.STRUCTS.items():
        assert L.smcrt_sizeof(name.encode()) == C.sizeof(st), name
    assert L.smcrt_sizeof(b"nonsense") == -1


def _create(scene):
    h = C.c_void_p()
    rc = A.lib().smcrt_scene_create(C.byref(scene.desc), C.byref(h))
    return rc, h


@pytest.mark.parametrize("name", ["slab", "refracting_sphere", "scat_test", "scat_test2", "aptran",
                                  "sphere_scene", "skin_vessels"])
def test_scene_create_without_gpu(name):
    sc, _ = scenes.build(name)
    rc, h = _create(sc)
    assert rc == A.OK, A.last_error()
    assert A.lib().smcrt_scene_destroy(h) == A.OK


def test_bad_abi_version_rejected():
    sc, _ = scenes.build("slab")
    sc.desc.abi_version = 99
    rc, _ = _create(sc)
    assert rc == A.EINVAL and "abi_version" in A.last_error()


def test_bad_material_rejected():
    sc = Scene([box((1, 1, 1), mono(1.0, 0.1, 1.5, 1.0), 1)], cart_grid(4, 4, 4), source("point"))
    rc, _ = _create(sc)
    assert rc == A.EINVAL and "hgg" in A.last_error()


def test_bad_program_rejected():
    sc = Scene([model([sphere(0.5, mono(1, 0, 0, 1), 1), sphere(0.4, mono(1, 0, 0, 1), 1)], union)],
               cart_grid(4, 4, 4), source("point"))
    sc._nodes[2].op = 7  # corrupt the CSG node
    rc, _ = _create(sc)
    assert rc == A.EINVAL and "op/kind" in A.last_error()
    sc._nodes[2].op = A.OP_PRIM  # now the region leaves three distances
    rc, _ = _create(sc)
    assert rc == A.EINVAL and "leaves 3 distances" in A.last_error()


def test_non_unit_direction_rejected():
    sc = Scene([box((1, 1, 1), mono(1, 0, 0, 1), 1)], cart_grid(4, 4, 4),
               source("pencil", (0, 0, 0), (0, 0, -2)))
    rc, _ = _create(sc)
    assert rc == A.EINVAL and "unit vector" in A.last_error()


def test_normalise_fluence_formula():
    """src/writer.f90:25-50: jmean * (2x*2y)/(N*dx*dy*dz)."""
    g = A.Grid()
    g.nx, g.ny, g.nz, g.xmax, g.ymax, g.zmax = 4, 5, 6, 1.0, 2.0, 0.5
    raw = np.arange(120, dtype=np.float64)
    out = np.empty(120, dtype=np.float32)
    A.check(A.lib().smcrt_normalise_fluence(C.byref(g), raw.ctypes.data_as(C.POINTER(C.c_double)), 1000,
                                            out.ctypes.data_as(C.POINTER(C.c_float))))
    f = (2 * 1.0 * 2 * 2.0) / (1000 * (2.0 / 4) * (4.0 / 5) * (1.0 / 6))
    np.testing.assert_allclose(out, (raw * f).astype(np.float32), rtol=1e-7)


def test_nrrd_layout(tmp_path):
    """src/writer.f90:220-251,294-334: header, metadata lines, blank line(s), raw LE float32."""
    data = np.arange(24, dtype=np.float32).reshape((2, 3, 4), order="F")
    p = tmp_path / "f.nrrd"
    A.check(A.lib().smcrt_write_nrrd(str(p).encode(), np.ascontiguousarray(data.reshape(-1, order="F")).ctypes.data_as(C.POINTER(C.c_float)),
                                     2, 3, 4, b'nphotons = 10\nsource = "pencil"'))
    raw = p.read_bytes()
    head = raw[: -24 * 4].decode()
    assert head.startswith("NRRD0004\ntype: float\ndimension: 3\nsizes: 2 3 4\nspace dimension: 3\n"
                           "encoding: raw\nendian: little\nnphotons = 10\n")
    assert head.endswith("\n\n\n")
    np.testing.assert_array_equal(np.frombuffer(raw[-96:], dtype="<f4"), data.reshape(-1, order="F"))


@pytest.mark.parametrize("name, kw, want", [
    ("refracting_sphere", {}, "basic"), ("slab", {}, "basic"), ("scat_test", {}, "basic"),
    ("skin_vessels", {"nxyz": 16}, "layers"),
    ("polygon_regions", {"n_side": 2, "n_edges": 12, "nxy": 16, "nz": 4}, "extruded"),
])
def test_scene_kernel_variant(name, kw, want, monkeypatch):
    """Each scene runs the smallest packet-kernel variant covering its kinds; SMCRT_KERNEL=full
    forces the general one (no GPU needed: the choice is made on the host)."""
    sc, _ = scenes.build(name, **kw)
    lib = A.lib()
    h = C.c_void_p()
    assert lib.smcrt_scene_create(C.byref(sc.desc), C.byref(h)) == A.OK
    sm, vm = C.c_uint32(), C.c_uint32()
    try:
        assert lib.smcrt_scene_kernel(h, C.byref(sm), C.byref(vm)).decode() == want
        assert sm.value & ~vm.value == 0 and sm.value != 0
        monkeypatch.setenv("SMCRT_KERNEL", "full")
        assert lib.smcrt_scene_kernel(h, C.byref(sm), C.byref(vm)).decode() == "full"
    finally:
        lib.smcrt_scene_destroy(h)
    assert lib.smcrt_scene_kernel(None, None, None) is None


def test_general_kinds_run_the_full_variant():
    from signedmcrt_amd.sdf import torus
    sc = Scene([torus(0.5, 0.2, mono(1, 0.1, 0, 1), 1), box((2, 2, 2), mono(0, 0, 0, 1), 2)],
               cart_grid(4, 4, 4, 1, 1, 1), source("point"))
    lib = A.lib()
    h = C.c_void_p()
    assert lib.smcrt_scene_create(C.byref(sc.desc), C.byref(h)) == A.OK
    sm = C.c_uint32()
    try:
        assert lib.smcrt_scene_kernel(h, C.byref(sm), None).decode() == "full"
        assert sm.value == (1 << 3) | (1 << 2)  # torus, box
    finally:
        lib.smcrt_scene_destroy(h)
