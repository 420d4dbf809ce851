"""The C++ mirror (pathguidingdissertation_amd/include/pgd/pgd.hpp) as a host program
uses it: pgd_render, built by build.py against libpgtrace.so."""
import subprocess

import numpy as np
import pytest

from conftest import ROOT

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
