"""Sign test for same-stream comparisons of the GPU film with the oracle's.

Both sides draw the same random numbers (counter RNG keyed by seed, sample, pixel and
dimension), so most pixels agree to float rounding. Where a few-ulp difference in a
transcendental flips a discrete decision (a Fresnel choice, a roulette test, a box hit)
the two paths continue as two samples of the same integrand, and those pixels differ
by ordinary Monte-Carlo noise. When both sides are the same estimator, (gpu, oracle)
is exchangeable in every pixel, so the difference d = gpu - oracle is symmetric about
zero, and for independent symmetric d_i the self-normalised sum
z = sum(d) / sqrt(sum(d^2)) obeys P(|z| >= t) <= 2 exp(-t^2 / 2) whatever the
distribution of the d_i (Efron, "Student's t-test under symmetry conditions", JASA
1969). A bias of one side (a wrong term in the glass or conductor BSDF, a light pdf off
by a factor) makes the differing pixels lean to one sign and drives |z| up with the
square root of their count. The test is applied per class of the pixel's primary hit
(environment, emitter, each BSDF type) and over the whole image.
"""
import numpy as np

CLASS_NAMES = {-1: "environment", -2: "emitter"}


def self_normalised_z(d: np.ndarray) -> float:
    d = np.asarray(d, np.float64).reshape(-1)
    s2 = float((d * d).sum())
    return 0.0 if s2 == 0.0 else float(d.sum() / np.sqrt(s2))


def primary_classes(pg, sc, data) -> np.ndarray:
    """(H, W) int labels: the BSDF type of the triangle the pixel-centre camera ray hits,
    -1 for the environment, -2 for an emitter."""
    W, H = data.width, data.height
    ys, xs = np.mgrid[0:H, 0:W]
    pxy = np.stack([xs.reshape(-1) + 0.5, ys.reshape(-1) + 0.5], 1).astype(np.float32)
    return classes_of_hits(data, sc.traverse(sc.generate_rays(pxy))[0]).reshape(H, W)


def classes_of_hits(data, tri_ids: np.ndarray) -> np.ndarray:
    tri_ids = np.asarray(tri_ids, np.int64)
    miss = tri_ids >= data.n_tris
    safe = np.where(miss, 0, tri_ids)
    mats = data.materials[data.material_index[safe]]
    lab = (mats["type"].astype(np.int64) & 7)
    lab = np.where(mats["emission"].max(-1) > 0, -2, lab)
    return np.where(miss, -1, lab)


def z_by_class(a: np.ndarray, b: np.ndarray, labels: np.ndarray, tol: float = 1e-3) -> dict:
    """Self-normalised z of the per-pixel channel sums of a - b over the pixels that
    differ by more than float rounding (some channel off by more than tol relative),
    over the image and per primary-hit class (classes without such a pixel give 0).
    Pixels within rounding are left out: a transcendental that rounds one way more often
    than the other (environment lookups differ by up to a few ulp between device libm
    and glibc) would otherwise make a harmless ulp-level lean significant."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    differs = np.any(np.abs(a - b) > tol * (np.abs(b) + tol), axis=-1)
    d = np.where(differs, (a - b).sum(-1), 0.0)
    out = {"all": self_normalised_z(d)}
    for c in np.unique(labels):
        out[CLASS_NAMES.get(int(c), f"bsdf{int(c)}")] = self_normalised_z(d[labels == c])
    return out


def assert_exchangeable(a, b, labels, t: float = 4.5, tol: float = 1e-3) -> dict:
    """Every |z| below t (false alarm at most 2 exp(-t^2/2) = 8e-5 per class at 4.5)."""
    z = z_by_class(a, b, labels, tol)
    bad = {k: round(v, 2) for k, v in z.items() if abs(v) >= t}
    assert not bad, (bad, {k: round(v, 2) for k, v in z.items()})
    return z
