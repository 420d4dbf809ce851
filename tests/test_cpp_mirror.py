"""The C++ mirror (pathguidingdissertation_amd/include/pgd/pgd.hpp) as a host program
uses it: pgd_render, built by build.py against libpgtrace.so."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT, make_scene_dir

EXE = ROOT / "pathguidingdissertation_amd" / "pgd_render"


def _exe():
    if not EXE.exists():
        raise RuntimeError(f"{EXE} missing: run __graft_entry__.build()")
    return str(EXE)


def test_mirror_selftest_and_no_device_error():
    r = subprocess.run([_exe(), "--selftest"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "selftest ok" in r.stdout


@pytest.mark.gpu
def test_mirror_render_equals_python_path(pg, device, cornell, tmp_path):
    out = tmp_path / "film.bin"
    r = subprocess.run([_exe(), "1", "3", "5", str(out)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    raw = np.fromfile(out, np.int32, 3)
    w, h, spp = (int(v) for v in raw)
    film = np.fromfile(out, np.float32, offset=12).reshape(h, w, 3)
    sc = pg.Scene(device, cornell)
    rt = pg.RayTracer(sc, seed=5)
    rt.clear()
    rt.render(3)
    ref, n = rt.film()
    assert spp == n == 3
    assert np.array_equal(film, ref)


@pytest.mark.gpu
def test_mirror_renders_scene_directory(pg, device, tmp_path):
    """main.cpp's flow through the C++ mirror: loadScene on a scene directory, render,
    savePNG / saveHDR; the files equal the Python path's tonemap and image."""
    root = make_scene_dir(pg, tmp_path / "scene", width=64, height=48)
    png, hdr = tmp_path / "out.png", tmp_path / "out.hdr"
    r = subprocess.run([_exe(), "--scene", str(root), "3", "7", str(png), str(hdr)], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "velvet" in r.stdout  # the skipped instance is reported
    data = pg.load_scene(root)
    rt = pg.RayTracer(pg.Scene(device, data), seed=7)
    rt.clear()
    rt.render(3)
    assert np.array_equal(pg.read_image(png), rt.tonemap().astype(np.float32))
    img = rt.image()
    assert np.all(np.abs(pg.read_image(hdr) - img) <= img.max(-1, keepdims=True) / 128 + 1e-30)
