"""The TOML input reader (`smcrt_config_load`, the reference's src/parse/*.f90 schema): defaults,
the forms of `direction`, required tables and the reference's error messages, detector order,
and the reference's own input files (when the reference checkout is present)."""
from __future__ import annotations

import glob
import os

import pytest

from signedmcrt_amd import _abi as A
from signedmcrt_amd import config

HERE = os.path.dirname(os.path.abspath(__file__))
TOML = os.path.join(HERE, "golden", "toml")


def _write(tmp_path, text, name="in.toml"):
    p = tmp_path / name
    p.write_text(text)
    return str(p)


BASE = """
[grid]
[geometry]
[output]
[simulation]
"""


def test_defaults_follow_the_reference():
    c = config.load(os.path.join(TOML, "minimal.toml"))
    # src/parse/parse.f90 and parse_source.f90 defaults
    assert c["source"]["nphotons"] == 1000000
    assert c["source"]["radius"] == 0.5 and c["source"]["beta"] == 5.0 and c["source"]["radius_hi"] == 0.6
    assert c["grid"] == {"nxg": 200, "nyg": 200, "nzg": 200, "xmax": 1.0, "ymax": 1.0, "zmax": 1.0,
                         "units": "cm"}
    g = c["geometry"]
    assert g["geom_name"] == "sphere" and g["tau"] == 10.0 and g["num_spheres"] == 10
    assert (g["musb"], g["muab"], g["musc"], g["muac"], g["hgg"]) == (0.0, 0.01, 0.0, 0.01, 0.7)
    o = c["output"]
    assert (o["fluence"], o["absorb"], o["render"]) == ("fluence.nrrd", "absorb.nrrd", "geom_render.nrrd")
    assert o["render_size"] == [200, 200, 200] and not o["render_geom"] and not o["overwrite"]
    s = c["simulation"]
    assert s["iseed"] == 123456789 and s["checkpoint_file"] == "check.ckpt"
    assert s["checkpoint_every_n"] == 1000000
    assert c["detectors"] == [] and c["history_file"] == "photPos.obj"
    assert c["source"]["direction"] is None  # a point source needs none


def test_toml_syntax_forms():
    c = config.load(os.path.join(TOML, "syntax.toml"))
    s = c["source"]
    assert s["name"] == "circular" and s["nphotons"] == 1000
    assert s["position"] == [0.0, 0.0, -0.999] and s["direction"] == [0.0, 0.0, -1.0]
    assert s["radius"] == pytest.approx(0.3)
    assert (c["grid"]["nxg"], c["grid"]["nyg"], c["grid"]["nzg"]) == (64, 32, 16)
    assert (c["grid"]["xmax"], c["grid"]["ymax"], c["grid"]["zmax"]) == (2.0, 1.5, 0.75)
    assert c["geometry"]["geom_name"] == "refracting_sphere" and c["geometry"]["tau"] == 2.5
    assert c["output"]["fluence"] == "a b.nrrd" and c["output"]["render_size"] == [10, 20, 30]
    assert c["output"]["render_geom"] is True
    assert c["simulation"]["iseed"] == -7


@pytest.mark.parametrize("d, want", [('"x"', [1, 0, 0]), ('"-x"', [-1, 0, 0]), ('"y"', [0, 1, 0]),
                                     ('"-y"', [0, -1, 0]), ('"z"', [0, 0, 1]), ('"-z"', [0, 0, -1]),
                                     ("[0.6, 0.0, -0.8]", [0.6, 0.0, -0.8])])
def test_direction_cardinal_or_vector(tmp_path, d, want):
    p = _write(tmp_path, f'[source]\nname = "pencil"\nposition = [0, 0, 0]\ndirection = {d}\nradius = 0.25\n' + BASE)
    c = config.load(p)
    assert c["source"]["direction"] == [float(x) for x in want]
    # with a vector direction the rest of the source table is still read (the reference returns
    # early there, src/parse/parse_source.f90:99-108)
    assert c["source"]["radius"] == 0.25


@pytest.mark.parametrize("src, msg", [
    ('name = "pencil"\nposition = [0, 0, 0]\ndirection = "w"', "Direction needs a cardinal direction"),
    ('name = "pencil"\nposition = [0, 0, 0]', "Need to specify direction for source type!"),
    ('name = "pencil"\ndirection = "z"', "Expected vector of size 3 for position"),
    ('name = "pencil"\nposition = [0, 0]\ndirection = "z"', "Expected vector of size 3 for position"),
    ('name = "uniform"\ndirection = "z"\npoint1 = [0, 0, 0]', "Uniform source requires point2 variable"),
    ('name = "uniform"\ndirection = "z"\npoint1 = [0, 0]', "Need a matrix row for points"),
])
def test_source_errors(tmp_path, src, msg):
    p = _write(tmp_path, "[source]\n" + src + "\n" + BASE)
    with pytest.raises(A.SmcrtError, match=msg):
        config.load(p)


@pytest.mark.parametrize("drop, msg", [("source", "Simulation needs Source table"),
                                       ("grid", "Need grid table in input param file"),
                                       ("geometry", "Need geometry table in input param file"),
                                       ("output", "Need output table in input param file"),
                                       ("simulation", "Need simulation table in input param file")])
def test_required_tables(tmp_path, drop, msg):
    tables = {"source": '[source]\nname = "point"\nposition = [0, 0, 0]', "grid": "[grid]",
              "geometry": "[geometry]", "output": "[output]", "simulation": "[simulation]"}
    text = "\n".join(v for k, v in tables.items() if k != drop)
    with pytest.raises(A.SmcrtError, match=msg):
        config.load(_write(tmp_path, text))


@pytest.mark.parametrize("bad", ["nxg = 1.5", "nxg = [1, 2]", 'xmax = "big"', "tev = 1", "[grid]",
                                 "nxg = 1\nnxg = 2", "key = {a = 1}", 'name = "unterminated'])
def test_malformed_input_is_an_error(tmp_path, bad):
    text = '[source]\nname = "point"\nposition = [0, 0, 0]\n[grid]\n' + bad + "\n[geometry]\n[output]\n[simulation]\ntev = false\n"
    if bad == "tev = 1":
        text = text.replace("tev = false", "tev = 1")
    with pytest.raises(A.SmcrtError):
        config.load(_write(tmp_path, text))


def test_missing_file_is_an_io_error(tmp_path):
    with pytest.raises(A.SmcrtError) as e:
        config.load(str(tmp_path / "nope.toml"))
    assert e.value.status == A.EIO


def test_detectors_ordered_by_type_and_validated(tmp_path):
    c = config.load(os.path.join(TOML, "slab_dects.toml"))
    kinds = [d.kind for d in c["detectors"]]
    # the reference's array holds circles, then annuli, then cameras (src/parse/parse_detectors.f90:96-109)
    assert kinds == [A.DET_CIRCLE, A.DET_ANNULUS, A.DET_CAMERA]
    circ, ann, cam = c["detectors"]
    assert list(circ.dir) == [0.0, 0.0, 2.0] and circ.radius1 == 2.0 and circ.nbins == 20
    assert ann.radius2 == 0.3 and ann.layer == 1
    assert list(cam.pos) == [-1.0, -1.0, -1.0] and cam.nbins == 7 and cam.maxval == 2.0
    bad = ('[[detectors]]\ntype = "annulus"\nposition = [0, 0, 0]\nradius1 = 1.0\nradius2 = 0.5\n')
    with pytest.raises(A.SmcrtError, match="Radii are invalid"):
        config.load(_write(tmp_path, '[source]\nname = "point"\nposition = [0, 0, 0]\n' + bad + BASE))
    bad = '[[detectors]]\ntype = "square"\n'
    with pytest.raises(A.SmcrtError, match="Invalid detector type"):
        config.load(_write(tmp_path, '[source]\nname = "point"\nposition = [0, 0, 0]\n' + bad + BASE))


def test_scene_from_toml_uses_file_grid_source_and_detectors():
    scene, cfg = config.scene_from_toml(os.path.join(TOML, "slab_dects.toml"))
    assert scene.grid.shape == (50, 50, 50) and scene.grid.zmax == 0.25
    assert scene.source.name == "pencil" and list(scene.source.direction) == [0.0, 0.0, -1.0]
    assert len(scene.detectors) == 3 and scene.desc.n_detectors == 3
    h = __import__("ctypes").c_void_p()
    assert A.lib().smcrt_scene_create(__import__("ctypes").byref(scene.desc), __import__("ctypes").byref(h)) == A.OK
    nb = [A.lib().smcrt_detector_bins(h, i) for i in range(4)]
    assert nb == [21, 13, 64, 0]
    A.lib().smcrt_scene_destroy(h)


REF_RES = "/root/reference/res"
# input files the reference's own parser rejects too (source without a position, or a uniform
# source with a string direction and no point1)
REF_REJECTED = {"exp.toml": "Expected vector of size 3 for position",
                "slab_test.toml": "Expected vector of size 3 for position",
                "skin.toml": "Uniform source requires point1 variable"}


@pytest.mark.skipif(not os.path.isdir(REF_RES), reason="reference checkout not present")
def test_reference_input_files_parse():
    files = sorted(glob.glob(os.path.join(REF_RES, "*.toml")))
    assert files
    for f in files:
        name = os.path.basename(f)
        if name in REF_REJECTED:
            with pytest.raises(A.SmcrtError, match=REF_REJECTED[name]):
                config.load(f)
            continue
        c = config.load(f)
        assert c["geometry"]["geom_name"], name
    c = config.load(os.path.join(REF_RES, "test_dects.toml"))
    assert [d.kind for d in c["detectors"]] == [A.DET_CIRCLE, A.DET_ANNULUS, A.DET_CAMERA]
    c = config.load(os.path.join(REF_RES, "aptran.toml"))
    assert c["source"]["name"] == "uniform" and c["source"]["direction"] == [0.0, 0.0, -1.0]


def test_geometry_table_reaches_only_the_builders_that_read_it(tmp_path):
    """[geometry] defaults (hgg = 0.7, tau = 10) must not leak into scenes whose reference
    setup routine does not read them; scat_test2 takes tau and hgg."""
    base = '[source]\nname = "pencil"\nposition = [0, 0, 0.2]\ndirection = "-z"\n[grid]\n[output]\n[simulation]\n'
    scene, _ = config.scene_from_toml(_write(tmp_path, base + '[geometry]\ngeom_name = "slab"\n', "a.toml"))
    assert scene.materials[0].hgg == 0.9 and scene.materials[0].mus == 100.0
    scene, _ = config.scene_from_toml(_write(tmp_path, base + '[geometry]\ngeom_name = "scat_test2"\ntau = 3.0\nhgg = 0.5\n', "b.toml"))
    assert scene.materials[0].mus == 3.0 and scene.materials[0].hgg == 0.5
