"""The time-to-relMSE bookkeeping of bench.py on synthetic images: the reference's own
error estimated from its two halves, the trimmed relMSE, the final-crossing time and
the per-ray algorithmic bytes of the roofline."""
import numpy as np

import bench

EPS = 1e-2


def _relmse(img, ref):
    # mean over pixels and RGB of (I - R)^2 / (R^2 + 1e-2), in FP64
    img = img.astype(np.float64)
    ref = ref.astype(np.float64)
    return float(np.mean((img - ref) ** 2 / (ref ** 2 + EPS)))


def test_floor_estimates_the_reference_error_and_separates_it():
    rng = np.random.default_rng(3)
    truth = rng.uniform(0.5, 2.0, (96, 128, 3))
    sigma = 0.05 * truth  # per-pixel noise of one half
    halves = [truth + sigma * rng.standard_normal(truth.shape) for _ in range(2)]
    ref = (halves[0] + halves[1]) / 2
    floor = bench.relmse_floor(ref, halves)
    true_ref_error = _relmse(ref, truth)
    assert abs(floor - true_ref_error) < 0.05 * true_ref_error, (floor, true_ref_error)
    # an independent image: relMSE against the reference less the floor is its error
    # against the truth
    img = truth + 0.2 * truth * rng.standard_normal(truth.shape)
    assert abs((_relmse(img, ref) - floor) - _relmse(img, truth)) < 0.03 * _relmse(img, truth)


def test_floor_of_identical_halves_is_zero():
    a = np.full((4, 4, 3), 0.3)
    assert bench.relmse_floor(a, (a, a.copy())) == 0.0


def test_trimmed_relmse_drops_only_the_worst_pixels():
    rng = np.random.default_rng(5)
    ref = rng.uniform(0.2, 1.0, (100, 100, 3))
    img = ref * (1 + 0.1 * rng.standard_normal(ref.shape))
    assert abs(bench.trimmed_relmse(img, ref, drop=0.0) - _relmse(img, ref)) < 1e-12
    fire = img.copy()
    fire[7, 11] = 1e4  # one firefly among 10,000 pixels
    assert _relmse(fire, ref) > 100 * _relmse(img, ref)
    # dropping 0.01% removes exactly that pixel: the rest is the unaffected error
    keep = np.ones((100, 100), bool)
    keep[7, 11] = False
    per_pixel = ((img - ref) ** 2 / (ref ** 2 + EPS)).mean(-1)
    assert abs(bench.trimmed_relmse(fire, ref, drop=1e-4) - per_pixel[keep].mean()) < 1e-12


def test_time_staying_below_counts_from_the_final_crossing():
    # columns: spp, seconds, relmse
    curve = [(1, 0.1, 0.5), (2, 0.2, 0.009), (4, 0.4, 0.02), (8, 0.8, 0.008), (16, 1.6, 0.006)]
    assert bench.time_staying_below(curve, 0.01) == (0.8, 8)
    assert bench.time_staying_below(curve + [(32, 3.2, 0.011)], 0.01) == (None, None)
    assert bench.time_staying_below(curve, 1.0) == (0.1, 1)


def test_algorithmic_bytes_per_ray():
    # 32-B nodes, 36 B of vertices + 4 B of triangle index per triangle, 32-B ray in,
    # 20-B hit or 4-B visibility out
    assert bench.algorithmic_bytes(0, 0, 20) == 52
    assert bench.algorithmic_bytes(10, 0, 4) == 32 * 10 + 36
    assert bench.algorithmic_bytes(0, 3, 20) == 3 * 40 + 52
    assert abs(bench.algorithmic_bytes(86.1, 25.51, 20) - 3827.6) < 0.1


def test_bench_refuses_more_than_one_gpu():
    import subprocess
    import sys
    from pathlib import Path
    root = Path(bench.__file__).resolve().parent
    r = subprocess.run([sys.executable, str(root / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=root)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
