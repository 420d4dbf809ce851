"""Benchmark: unguided wavefront path tracing of the ~1.09M-triangle procedural interior
(1280x720, max depth 12, env map through an alias table) on one MI355X.

A step is one sample pass over the full frame (921,600 paths, one per pixel). The
default run times 64 steps (the workload's 64 SPP) after 4 warmup steps and prints one
JSON line: Mrays/s (closest-hit + shadow rays traced), the roofline of the dominant
kernel (extend: closest-hit traversal) from HIP-event timings taken inside the timed
region, and the CPU oracle timed on the host cores on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_HBM_GBS = 8000.0  # MI355X HBM3E


def algorithmic_bytes(nodes_per_ray: float, tris_per_ray: float, out_bytes: int) -> float:
    """Per ray: 32-B nodes visited, 36 B of vertices + 4 B triangle index per triangle
    tested, 32 B ray in, the hit (20 B) or visibility (4 B) out."""
    return 32.0 * nodes_per_ray + 40.0 * tris_per_ray + 32.0 + out_bytes


def profiled_traffic(config: int) -> dict:
    """HBM bytes per dispatch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE runs of this
    benchmark (scaled by a plain copy kernel), as committed under profiles/."""
    path = ROOT / "profiles" / f"pmc_traffic_config{config}.txt"
    out = {}
    if path.exists():
        for line in path.read_text().splitlines():
            if line and not line.startswith("#"):
                k, v = line.split()
                out[k] = float(v)
    return out


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=64)
    ap.add_argument("--warmup", type=int, default=4)
    ap.add_argument("--config", type=int, default=int(os.environ.get("PG_BENCH_CONFIG", "2")))
    ap.add_argument("--cpu-spp", type=int, default=int(os.environ.get("PG_BENCH_CPU_SPP", "40")))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-check", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM reference point / counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the path shards by pixels/samples/views with a single "
              "film reduction (see DESIGN.md); multi-GPU runs are not part of this build.", file=sys.stderr)
        return 2

    import pathguidingdissertation_amd as pg

    t_setup = time.time()
    data = pg.generate_scene(a.config, 0)
    dev = pg.Device(0)
    sc = pg.Scene(dev, data)
    bvh = sc.bvh_info()
    max_depth = 12
    rt = pg.RayTracer(sc, seed=1, max_depth=max_depth)
    W, H = data.width, data.height
    print(f"bench: config {a.config}, {data.n_tris} tris, {W}x{H}, BVH {bvh['nodes']} nodes "
          f"depth {bvh['max_depth']} ({bvh['build_ms']:.0f} ms), setup {time.time() - t_setup:.1f}s",
          file=sys.stderr, flush=True)

    # calibration: count node visits / triangle tests per ray on one pass
    rt.set_counting(True)
    rt.clear()
    rt.reset_stats()
    rt.render(1)
    cal = rt.stats()
    rt.set_counting(False)
    nodes_c = cal["nodes_visited"] / max(cal["closest_rays"], 1)
    tris_c = cal["tris_tested"] / max(cal["closest_rays"], 1)
    nodes_s = cal["shadow_nodes_visited"] / max(cal["shadow_rays"], 1)
    tris_s = cal["shadow_tris_tested"] / max(cal["shadow_rays"], 1)
    bytes_c = algorithmic_bytes(nodes_c, tris_c, 20)
    bytes_s = algorithmic_bytes(nodes_s, tris_s, 4)

    rt.clear()
    if a.warmup:
        rt.render(a.warmup)
    rt.clear()
    rt.reset_stats()
    t0 = time.perf_counter()
    rt.render(a.steps)  # the call returns after the stream has drained
    t1 = time.perf_counter()
    st = rt.stats()
    wall = t1 - t0
    rays = st["closest_rays"] + st["shadow_rays"]
    mrays = rays / wall / 1e6
    ext_ms = st["kernel_ms_extend"]
    launches = max(st["extend_launches"], 1)
    ext_bytes_total = bytes_c * st["closest_rays"]
    achieved = ext_bytes_total / (ext_ms * 1e-3) / 1e9 if ext_ms > 0 else 0.0
    sh_ms = st["kernel_ms_shadow"]
    sh_achieved = bytes_s * st["shadow_rays"] / (sh_ms * 1e-3) / 1e9 if sh_ms > 0 else 0.0

    img = rt.image()
    finite = bool(np.isfinite(img).all())
    traffic = profiled_traffic(a.config).get("k_extend")

    copy = None
    if a.copy_check:
        import ctypes as C
        ms = C.c_double()
        nbytes = 1 << 30
        pg._lib.check(dev.lib.pg_copy_benchmark(dev.ctx, nbytes, 10, C.byref(ms)), dev.ctx)
        copy = {"bytes": nbytes, "ms": round(ms.value, 4), "gbs": round(2 * nbytes / (ms.value * 1e-3) / 1e9, 1)}

    cpu = None
    if not a.no_cpu:
        from oracle.pg_oracle import OracleScene
        threads = min(os.cpu_count() or 1, 16)
        osc = OracleScene(data)
        film, cst = osc.render(a.cpu_spp, seed=1, max_depth=max_depth, accum_f64=False, threads=threads, env_mode=1)
        cpu_rays = cst["closest_rays"] + cst["shadow_rays"]
        cpu = {"value": round(cpu_rays / cst["seconds"] / 1e6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"{a.cpu_spp} SPP of the same {W}x{H} frame ({cst['paths']} paths, {cpu_rays} rays, "
                         f"{cst['seconds']:.1f} s); oracle restatement with std::thread tiles (Renderer.h:648-747)",
               "samples_per_s": round(cst["paths"] / cst["seconds"], 1)}

    line = {
        "metric": "Mrays/s; time-to-relMSE 0.01: learned vs QuadTree vs unguided, 1 MI355X",
        "value": round(mrays, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": round(wall * 1e3 / a.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (procedural scene, seed 0)",
        "config": {"workload": f"unguided wavefront path tracer, config {a.config}: {data.n_tris} tris, "
                               f"{W}x{H}, {a.steps} SPP, max depth {max_depth}, env map alias table",
                   "tris": data.n_tris, "width": W, "height": H, "spp": a.steps, "max_depth": max_depth},
        "samples_per_s": round(st["paths"] / wall, 1),
        "closest_rays": st["closest_rays"],
        "shadow_rays": st["shadow_rays"],
        "kernel_ms": {"extend": round(ext_ms, 3), "shade": round(st["kernel_ms_shade"], 3),
                      "shadow": round(sh_ms, 3), "gpu_total": round(st["kernel_ms_total"], 3)},
        "roofline": {"bound": "hbm", "kernel": "k_extend", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": "profiles/pmc_traffic_config%d.txt (bytes per launch)" % a.config,
                     "bytes_per_ray": round(bytes_c, 1), "nodes_per_ray": round(nodes_c, 2),
                     "tris_per_ray": round(tris_c, 2), "launches": launches,
                     "avg_launch_ms": round(ext_ms / launches, 4)},
        "shadow_roofline": {"achieved": round(sh_achieved, 1), "bytes_per_ray": round(bytes_s, 1),
                            "nodes_per_ray": round(nodes_s, 2), "tris_per_ray": round(tris_s, 2)},
        "cpu_baseline": cpu,
        "copy_reference": copy,
        "image_finite": finite,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
