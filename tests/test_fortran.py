"""The Fortran host over the C ABI (signedmcrt_amd/fortran/): the iso_c_binding layouts against
the library, the driver's error paths (CPU), and on the GPU the driver's output files against the
Python host's for the same input (same packets: identical detector counts, fluence equal to the
FP32 tally-tile accumulation order)."""
from __future__ import annotations

import os
import subprocess

import numpy as np
import pytest

from signedmcrt_amd import config

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
FDIR = os.path.join(ROOT, "signedmcrt_amd", "fortran")
RUN = os.path.join(FDIR, "smcrt_run")
CHECK = os.path.join(FDIR, "smcrt_abi_check")
SLAB_DECTS = os.path.join(ROOT, "tests", "golden", "toml", "slab_dects.toml")

needs_fortran = pytest.mark.skipif(not (os.path.exists(RUN) and os.path.exists(CHECK)),
                                   reason="Fortran host not built (no flang)")


@needs_fortran
def test_iso_c_binding_layouts_match_library():
    p = subprocess.run([CHECK], capture_output=True, text=True, timeout=60)
    assert p.returncode == 0, p.stdout + p.stderr
    names = [line.split()[0] for line in p.stdout.splitlines()]
    assert {"node", "scene_desc", "run_opts", "stats", "tallies", "detector", "config"} <= set(names)


@needs_fortran
@pytest.mark.parametrize("text, msg", [
    ('[source]\nname = "point"\nposition = [0, 0, 0]\n[geometry]\n[output]\n[simulation]\n',
     "Need grid table in input param file"),
    ('[source]\nname = "point"\nposition = [0, 0, 0]\n[grid]\n[geometry]\ngeom_name = "logo"\n'
     '[output]\n[simulation]\n', "no such routine: logo"),
])
def test_driver_reports_input_errors(tmp_path, text, msg):
    f = tmp_path / "in.toml"
    f.write_text(text)
    p = subprocess.run([RUN, str(f)], capture_output=True, text=True, timeout=60, cwd=tmp_path)
    assert p.returncode == 1
    assert msg in p.stdout + p.stderr


def _read_nrrd(path):
    """(header lines, data): the raw little-endian float32 block is the file's last
    nx*ny*nz*4 bytes, after the header and its blank line(s)."""
    raw = open(path, "rb").read()
    head = raw.split(b"\n\n", 1)[0].decode().splitlines()
    sizes = [int(v) for v in next(h for h in head if h.startswith("sizes:")).split()[1:]]
    nbytes = 4 * int(np.prod(sizes))
    assert raw[:-nbytes].endswith(b"\n\n\n")
    return head, np.frombuffer(raw[-nbytes:], dtype="<f4")


@needs_fortran
@pytest.mark.gpu
def test_driver_matches_python_host(gpu, tmp_path):
    fdir, pdir = tmp_path / "fortran", tmp_path / "python"
    fdir.mkdir()
    pdir.mkdir()
    p = subprocess.run([RUN, SLAB_DECTS], capture_output=True, text=True, timeout=300, cwd=fdir)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "Photons/s" in p.stdout and "Average # of scatters per photon" in p.stdout
    out = config.run_toml(SLAB_DECTS, outdir=str(pdir))
    hf, df = _read_nrrd(fdir / "data" / "jmean" / "slab.nrrd")
    hp, dp = _read_nrrd(out["fluence_path"])
    assert hf[:7] == hp[:7] == ["NRRD0004", "type: float", "dimension: 3", "sizes: 50 50 50",
                                "space dimension: 3", "encoding: raw", "endian: little"]
    assert df.size == dp.size == 50 ** 3
    np.testing.assert_allclose(df, dp, rtol=2e-6, atol=1e-6 * float(dp.max()))
    for i in (1, 2, 3):
        a = (fdir / "data" / "detectors" / f"detector_{i}.dat").read_text()
        b = open(out["detector_paths"][i - 1]).read()
        assert a == b, i
    # render_geom: both hosts voxelise the scene with smcrt_render (same floats)
    hr, dr = _read_nrrd(fdir / "data" / "geom.nrrd")
    hq, dq = _read_nrrd(out["render_path"])
    assert hr[3] == hq[3] == "sizes: 20 20 10"
    np.testing.assert_array_equal(dr, dq)
    # the circle detector's file lists bins 1..nbins+1 with j*bin_width and the count
    rows = [line.split() for line in (fdir / "data" / "detectors" / "detector_1.dat").read_text().splitlines()]
    assert len(rows) == 21
    assert float(rows[0][0]) == pytest.approx(1.0 / 19.0)
    assert sum(float(r[1]) for r in rows) == int(out["result"].detectors[0].sum())
