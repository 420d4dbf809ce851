"""The device BSDFs (device_shading.h: bsdf_sample / bsdf_eval / bsdf_invert) against
the reference's own BSDF classes (Materials.h:121-897), through pg_bsdf_query with the
random numbers injected as the reference's test sampler does (main.cpp:13-72).

Vectors: tests/golden/ref_bsdf.npz (Diffuse, Mirror, Conductor smooth and rough,
Glass, Dielectric smooth and rough) and ref_bsdf_ext.npz (Oren-Nayar, Plastic and
Layered over four bases), 6,656 queries. The sampled direction, its value and pdf,
evaluate() and PDF() at a second direction and invert() of the sampled direction
agree to 2e-5 relative (1e-6 absolute): a few ulp of the device's acosf / cosf /
sinf / atan2f chains against the reference platform's. The flags are exact."""
import numpy as np
import pytest

from conftest import GOLDEN

pytestmark = pytest.mark.gpu


def _close(got, exp):
    err = np.abs(got - exp)
    return (err <= 2e-5 * np.abs(exp)) | (err <= 1e-6) | (np.isnan(got) & np.isnan(exp))


@pytest.mark.parametrize("name", ["ref_bsdf.npz", "ref_bsdf_ext.npz"])
def test_device_bsdfs_match_reference(pg, device, name):
    g = np.load(GOLDEN / name)
    mat, sd, u, we, ref = g["mat"], g["sd"], g["u"], g["wi_eval"], g["out"]
    # one call per distinct material (the golden groups cases by material)
    keys = np.unique(mat, axis=0, return_inverse=True)[1].reshape(-1)
    got = np.empty((mat.shape[0], 15), np.float32)
    for k in np.unique(keys):
        sel = np.nonzero(keys == k)[0]
        got[sel] = device.bsdf_query(mat[sel[0]], sd[sel], u[sel], we[sel])
    base = (mat[:, 0].astype(int)) % 8
    # sample: wi, f, pdf; evaluate / PDF at wi_eval
    ok = np.all(_close(got[:, :11], ref[:, :11]), axis=1)
    bad = np.nonzero(~ok)[0]
    assert ok.mean() > 0.998, [(int(i), int(mat[i, 0]), got[i, :11], ref[i, :11]) for i in bad[:3]]
    for t in np.unique(base):
        assert ok[base == t].mean() > 0.99, (t, ok[base == t].mean())
    # invert of the sampled direction where the device BSDF has one (Diffuse, Oren-Nayar,
    # rough Conductor) and the sample lies above the surface
    invertible = ((base == 0) | (base == 5) | ((base == 2) & (mat[:, 15] > 0))) & (ref[:, 13] == 0) & ok
    has = invertible & ((got[:, 11] != 0) | (got[:, 12] != 0))
    assert has.sum() > 500
    assert np.all(np.abs(got[has, 11:13] - ref[has, 11:13]) <= 2e-5 + 2e-5 * np.abs(ref[has, 11:13])), \
        np.abs(got[has, 11:13] - ref[has, 11:13]).max()
    # flags
    assert np.array_equal(got[:, 13], ref[:, 13])
    if ref.shape[1] > 14:
        assert np.array_equal(got[:, 14], ref[:, 14])
