"""Benchmark: unguided wavefront path tracing of the ~1.09M-triangle procedural interior
(1280x720, 64 SPP, max depth 12, env map through an alias table) on one MI355X.

A step is one full frame of the workload: 64 samples per pixel over 1280x720
(58,982,400 camera paths), traced as one wavefront pass of 64 samples per pixel (the
library's automatic samples_per_pass). The run times --steps frames after --warmup
frames and prints one JSON line: Mrays/s (closest-hit + shadow rays traced), the
roofline of the dominant kernel (extend: closest-hit traversal) from HIP-event timings
taken inside the timed region, and the CPU oracle timed on the host cores on a bounded
sample (all threads of the host's share and one thread). Further parts: the batched
views of the config-5 family at its specification (48 views, 32 SPP; --batch-views 0
skips it) and time-to-relMSE of the unguided, quadtree-guided and learned-guided
renderers against a 4096-SPP reference, on the open interior and on the window-lit room
(--ttr-config picks the scenes; --no-ttr skips it).
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


# Shade step (one closest hit): hit 20 B, shading attributes 64 B, material 32 B, path
# state read + write 48 B, new ray 32 B; a shadow ray adds the shadow ray 48 B, the
# alias-table entry 16 B and the bilinear environment texels 48 B.
SHADE_BYTES_PER_HIT = 20 + 64 + 32 + 48 + 32
SHADE_BYTES_PER_SHADOW = 48 + 16 + 48


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


def reference_image(pg, sc, data, spp: int, max_depth: int, env_visibility: int = 0) -> np.ndarray:
    """Unguided GPU render of `spp` samples per pixel accumulated in FP64 on the host
    (64-SPP film reads summed in double; chunk k has seed 1000 + k). The GPU path tracer is the same
    estimator as the CPU oracle (pixel agreement is tested), so this stands in for the
    4096-SPP FP64 CPU image, which would take ~20 minutes on 16 host cores."""
    acc = [np.zeros((data.height, data.width, 3), np.float64) for _ in range(2)]
    cnt = [0, 0]
    done = 0
    chunk = 0
    while done < spp:
        # pg_film_clear restarts the sample index, so every 64-SPP chunk takes its own
        # seed: the chunks are independent samples of the estimator
        rt = pg.RayTracer(sc, seed=1000 + chunk, max_depth=max_depth, env_visibility=env_visibility)
        rt.clear()
        rt.render(min(64, spp - done))
        film, n = rt.film()
        acc[chunk % 2] += film  # alternate chunks into two independent halves
        cnt[chunk % 2] += n
        done += n
        chunk += 1
    ref = (acc[0] + acc[1]) / done
    halves = (acc[0] / max(cnt[0], 1), acc[1] / max(cnt[1], 1))
    return ref, halves


def relmse_floor(ref, halves, eps=1e-2) -> float:
    """The reference's own share of relMSE(I, ref): with two independent halves R1, R2 of
    the reference, (R1 - R2)^2 / 4 estimates the variance of their mean per pixel, so
    mean((R1 - R2)^2 / 4 / (ref^2 + eps)) estimates E[relMSE(ref, truth)], and
    relMSE(I, ref) - floor estimates relMSE(I, truth) (the cross term vanishes for
    independent images)."""
    d = (halves[0] - halves[1]) ** 2 / 4.0
    return float(np.mean(d / (ref.astype(np.float64) ** 2 + eps)))


def trimmed_relmse(img, ref, drop=1e-4, eps=1e-2) -> float:
    """relMSE over the pixels left after dropping the fraction `drop` with the largest
    per-pixel error (0.01%: 92 of 921,600 pixels at 1280x720), in FP64: the part of
    the error that is not carried by a handful of fireflies."""
    e = ((img.astype(np.float64) - ref) ** 2 / (ref.astype(np.float64) ** 2 + eps)).reshape(-1, 3).mean(1)
    k = int(e.size * drop)
    if k:
        e = np.partition(e, e.size - k)[: e.size - k]
    return float(e.mean())


def progressive(pg, label, ref, setup, advance, cap_s, target, log, floor=0.0):
    """Run one method until relMSE <= target or cap_s seconds of its own wall time.
    The clock covers setup and every render/training call; relMSE checkpoints (film
    read-back + FP64 metric on the host) are outside it. Increments double up to 64
    SPP, then hold about one second of work each."""
    t_total = 0.0
    t0 = time.perf_counter()
    setup()
    t_total += time.perf_counter() - t0
    spp, k, curve, hit = 0, 1, [], None
    while t_total < cap_s:
        t0 = time.perf_counter()
        img = advance(k)
        dt = time.perf_counter() - t0
        t_total += dt
        spp += k
        e = pg.relmse(img, ref)
        curve.append((spp, round(t_total, 3), e, trimmed_relmse(img, ref), e - floor))
        if e <= target:
            hit = (spp, t_total)
            break
        per_spp = dt / k
        k = k * 2 if spp < 64 else max(1, min(256, int(1.0 / max(per_spp, 1e-4))))
    log(f"ttr {label}: {len(curve)} checkpoints, {spp} SPP in {t_total:.1f}s, relMSE {curve[-1][2]:.4f}"
        f" (trimmed {curve[-1][3]:.4f}, less the reference's own {curve[-1][4]:.4f})"
        + (f", reached {target} at {hit[1]:.2f}s" if hit else ""))
    return curve, hit


def time_staying_below(curve, e, col=2):
    """(seconds, spp) of the first checkpoint from which every later checkpoint has
    error <= e; (None, None) when the last one is above e. A curve that dips below e
    and rises again (a firefly arriving later) only counts from its final crossing."""
    hit = (None, None)
    for c in curve:
        if c[col] <= e:
            if hit[0] is None:
                hit = (c[1], c[0])
        else:
            hit = (None, None)
    return hit


def time_to_relmse(pg, dev, sc, data, max_depth, ref_spp, cap_s, log, target=0.01, env_visibility=1, repeats=1,
                   with_pss=False, train_every=1, train_steps=1, record_samples=1, guide_depth=0,
                   train_batch=1 << 18):
    """Time-to-relMSE study: unguided vs quadtree guide vs learned guide. Besides the
    time to relMSE <= target (null when a method does not reach it within cap_s), the
    learned guide is compared with each baseline at that baseline's own final relMSE:
    speedup = time the baseline needed to stay below it / time the learned guide needed
    to stay below it, on the full relMSE and on relMSE without the worst 0.01% of pixels.
    With repeats > 1 every method runs again with new seeds against the same reference;
    the comparisons of each run and their medians are reported (curves of run 0)."""
    t0 = time.perf_counter()
    ev = dict(env_visibility=env_visibility)
    gd = dict(guide_depth=guide_depth)  # the guided methods guide the first bounces only
    ref, halves = reference_image(pg, sc, data, ref_spp, max_depth, env_visibility)
    floor = relmse_floor(ref, halves)
    log(f"ttr reference: {ref_spp} SPP in {time.perf_counter() - t0:.1f}s, mean {ref.mean():.5f}, "
        f"own relMSE {floor:.4f}")

    def adv(rt, fn):
        def f(k):
            fn(k)
            return rt.image()
        return f

    def run_once(r):
        un = pg.RayTracer(sc, seed=2 + 10 * r, max_depth=max_depth, **ev)
        qt = pg.RayTracer(sc, seed=3 + 10 * r, max_depth=max_depth, integrator=2, guide_prob=0.5, **ev, **gd)
        ps = pg.RayTracer(sc, seed=4 + 10 * r, max_depth=max_depth, integrator=4, guide_prob=0.5, **ev, **gd)
        pf = pg.RayTracer(sc, seed=8 + 10 * r, max_depth=max_depth, integrator=4, guide_prob=0.5, knn_k=0, **ev,
                          **gd)
        ml = pg.RayTracer(sc, seed=5 + 10 * r, max_depth=max_depth, record_nee=1, **ev, **gd)

        def cache_setup(rt, seed):
            # pass 1: 1 SPP record pass (Li with NEE), then radix sort + point BVH + quadtrees
            dev.clear_records()
            rec = pg.RayTracer(sc, seed=seed, max_depth=max_depth, integrator=1, record=1, record_nee=1, **ev)
            rec.clear()
            rec.render(1)
            dev.cache_build(cell_size=4096)
            rt.clear()

        def qt_setup():
            cache_setup(qt, 1 + 10 * r)

        def ps_setup():
            cache_setup(ps, 6 + 10 * r)

        def pf_setup():
            cache_setup(pf, 9 + 10 * r)

        def ml_setup():
            dev.mlp_init(seed=7 + r, batch=train_batch)
            ml.clear()
            ml.reset_stats()

        methods = {
            "unguided": (lambda: un.clear(), adv(un, lambda k: un.render(k))),
            "quadtree": (qt_setup, adv(qt, lambda k: qt.render(k))),
        }
        if with_pss:
            methods["pss"] = (ps_setup, adv(ps, lambda k: ps.render(k)))
            methods["pss_approx"] = (pf_setup, adv(pf, lambda k: pf.render(k)))
        methods["learned"] = (ml_setup, adv(ml, lambda k: ml.render_online(k, steps=train_steps, window=1,
                                                                          train_every=train_every,
                                                                          record_samples=record_samples)))
        res = {}
        for name, (setup, advance) in methods.items():
            curve, hit = progressive(pg, name if r == 0 else f"{name} (run {r})", ref, setup, advance, cap_s,
                                     target, log, floor)
            t_c = time_staying_below(curve, target, 4)
            res[name] = {"seconds": None if hit is None else round(hit[1], 3),
                         "spp": curve[-1][0], "final_relmse": round(curve[-1][2], 5),
                         "final_trimmed_relmse": round(curve[-1][3], 5),
                         "final_relmse_less_reference": round(curve[-1][4], 5),
                         "seconds_to_target_less_reference": t_c[0], "curve": curve}
            if name == "learned":
                st = ml.stats()
                res[name]["kernel_ms_per_spp"] = {
                    k: round(st[f"kernel_ms_{k}"] / max(curve[-1][0], 1), 3)
                    for k in ("extend", "shade", "shadow", "guide", "total")}

        def speedup(fast, slow, col=2):
            # time `slow` needs to stay below its own final relMSE over the time `fast`
            # needs for it (null: `fast` ends above it)
            tgt = res[slow]["curve"][-1][col]
            t_slow = time_staying_below(res[slow]["curve"], tgt, col)[0]
            t_fast = time_staying_below(res[fast]["curve"], tgt, col)[0]
            return {"target_relmse": round(tgt, 5), f"{slow}_s": t_slow, f"{fast}_s": t_fast,
                    "speedup": None if not t_fast else round(t_slow / t_fast, 3)}

        def ratio(a, b, key):
            return round(res[a][key] / res[b][key], 3)

        def at_target(fast, slow):
            # both reach relMSE <= target once the reference's own share is removed
            a, b = res[fast]["seconds_to_target_less_reference"], res[slow]["seconds_to_target_less_reference"]
            return {f"{slow}_s": b, f"{fast}_s": a, "speedup": round(b / a, 3) if a and b else None}

        cmp = {"learned_vs_unguided": speedup("learned", "unguided"),
               "learned_vs_quadtree": speedup("learned", "quadtree"),
               "less_reference": {"learned_vs_unguided": speedup("learned", "unguided", 4),
                                  "learned_vs_quadtree": speedup("learned", "quadtree", 4),
                                  "at_target_learned_vs_unguided": at_target("learned", "unguided"),
                                  "at_target_learned_vs_quadtree": at_target("learned", "quadtree")},
               "trimmed": {"learned_vs_unguided": speedup("learned", "unguided", 3),
                           "learned_vs_quadtree": speedup("learned", "quadtree", 3)},
               "relmse_ratio_at_cap": {"unguided_over_learned": ratio("unguided", "learned", "final_relmse"),
                                       "quadtree_over_learned": ratio("quadtree", "learned", "final_relmse")},
               "trimmed_relmse_ratio_at_cap": {
                   "unguided_over_learned": ratio("unguided", "learned", "final_trimmed_relmse"),
                   "quadtree_over_learned": ratio("quadtree", "learned", "final_trimmed_relmse")}}
        # Monte Carlo relMSE falls as 1/time: the last point of each run (reference noise
        # removed) projects its time to the target; the ratios of those times against
        # the target factors (4x unguided, 2x quadtree) give the shortfall
        def projected(name):
            c = res[name]["curve"][-1]
            return round(c[4] * c[1] / target, 1) if c[4] > 0 else None
        proj = {m: projected(m) for m in res}
        cmp["projected_seconds_to_target"] = proj

        def proj_speedup(fast, slow, factor):
            if not proj[fast] or not proj[slow]:
                return {"speedup": None, "target_factor": factor, "fraction_of_target": None}
            sp = proj[slow] / proj[fast]
            return {"speedup": round(sp, 3), "target_factor": factor, "fraction_of_target": round(sp / factor, 3)}
        cmp["projected_speedup"] = {"learned_vs_unguided": proj_speedup("learned", "unguided", 4.0),
                                    "learned_vs_quadtree": proj_speedup("learned", "quadtree", 2.0)}

        # time to a given error = (cost per sample) x (error x samples): a method's
        # speedup over another is its per-sample error gain divided by its cost ratio
        def per_sample(name):
            spp, sec, e, et, _ = res[name]["curve"][-1]
            return {"ms_per_spp": round(sec * 1e3 / spp, 3), "relmse_x_spp": round(e * spp, 2),
                    "trimmed_relmse_x_spp": round(et * spp, 3)}
        cost = {m: per_sample(m) for m in res}

        def decompose(fast, slow):
            a, b = cost[fast], cost[slow]
            c = a["ms_per_spp"] / b["ms_per_spp"]
            g = b["relmse_x_spp"] / a["relmse_x_spp"]
            gt = b["trimmed_relmse_x_spp"] / a["trimmed_relmse_x_spp"]
            return {"cost_ratio": round(c, 3), "per_sample_gain": round(g, 3), "per_sample_gain_trimmed": round(gt, 3),
                    "equal_time_gain": round(g / c, 3), "equal_time_gain_trimmed": round(gt / c, 3)}
        cmp["per_sample"] = {"methods": cost, "learned_vs_unguided": decompose("learned", "unguided"),
                             "learned_vs_quadtree": decompose("learned", "quadtree")}
        if with_pss:
            for v in ("pss", "pss_approx"):
                cmp[f"{v}_vs_unguided"] = speedup(v, "unguided")
                cmp["trimmed"][f"{v}_vs_unguided"] = speedup(v, "unguided", 3)
                cmp["relmse_ratio_at_cap"][f"unguided_over_{v}"] = ratio("unguided", v, "final_relmse")
                cmp["trimmed_relmse_ratio_at_cap"][f"unguided_over_{v}"] = ratio("unguided", v, "final_trimmed_relmse")
        return res, cmp

    res, cmp = run_once(0)
    info = dev.mlp_info()
    runs = [cmp]
    for r in range(1, repeats):
        runs.append(run_once(r)[1])
    for rr in res.values():
        c = rr["curve"]
        step = max(1, len(c) // 16)
        rr["curve"] = [[s, t, round(e, 5), round(et, 5), round(ec, 5)]
                       for s, t, e, et, ec in (c[::step] + ([c[-1]] if (len(c) - 1) % step else []))]
    # network FLOPs over the used layer sizes (80 -> 64 -> 64 -> 64 -> 41): 2 * sum(in * out)
    # per query forward; a training sample adds the input and weight gradients (3x)
    fwd = 2 * (80 * 64 + 64 * 64 + 64 * 64 + 64 * 41)
    batch = train_batch
    train_ms = float(info["train_ms"])  # GPU time per Adam step
    mlp = {"forward_flop_per_query": fwd, "train_flop_per_sample": 3 * fwd, "batch": batch,
           "train_ms_per_step": round(train_ms, 3),
           "train_tflops": round(3 * fwd * batch / (train_ms * 1e-3) / 1e12, 2) if train_ms > 0 else None,
           "peak_tflops_bf16_dense": 2500.0,
           "mfma_frac": round(3 * fwd * batch / (train_ms * 1e-3) / 2.5e15, 5) if train_ms > 0 else None,
           "inference_ms_per_spp": res["learned"]["kernel_ms_per_spp"]["guide"],
           "bound": "hash-grid gradient atomics (training) and gathers (inference), not MFMA"}
    out = {"target": target, "cap_s": cap_s,
           "reference": f"GPU unguided {ref_spp} SPP, FP64 host accumulation, 64-SPP chunks with seeds 1000 + k",
           "env_visibility": ("environment shadow rays along the light direction for every method and the "
                              "reference image (the reference's target sceneCentre + wi * sceneRadius, "
                              "Renderer.h:331-332, loses the direct sun behind the window, and guided and "
                              "unguided MIS estimators then converge to different images)"
                              if env_visibility else "Renderer.h:331-332 scene-bound target"),
           "curve_columns": ["spp", "seconds", "relmse", "relmse_without_top_0.01pct_pixels",
                             "relmse_less_reference_own"],
           "reference_own_relmse": round(floor, 5),
           "reference_own_relmse_note": ("mean((R1 - R2)^2 / 4 / (R^2 + 1e-2)) over two independent halves of the "
                                         "reference: its own noise, subtracted in the less_reference column"),
           "methods": {k: v for k, v in {"unguided": "path tracer with NEE + MIS (Renderer.h:531-618)",
                       "quadtree": "1-SPP record pass, point-BVH + per-cell quadtrees, 50/50 one-sample MIS, kNN k=16",
                       "pss": "1-SPP record pass, point-BVH; per vertex the 16 nearest records inverted into the "
                              "BSDF's primary sample space (Renderer.h:398-414), 4x4 flux histogram, 50/50 MIS",
                       "pss_approx": "as pss with approximate neighbours (16 nearest of the 64 records under a "
                                     "greedy point-BVH descent)",
                       "learned": f"online: a {record_samples}-SPP guided record pass and {train_steps} Adam step(s) "
                                  f"of {train_batch} samples every {train_every} SPP, guided passes between"}.items()
                       if k in res},
           "guide_depth": (f"the guided methods guide the first {guide_depth} bounces and sample the BSDF alone "
                           "deeper" if guide_depth else "the guided methods guide every bounce"),
           **res,
           "learned_diverged": bool(info["diverged"]),
           "learned_network": mlp,
           "comparison": cmp}
    if repeats > 1:
        def med(path):
            vals = []
            for c in runs:
                v = c
                for k in path:
                    v = v[k]
                vals.append(v)
            ok = [v for v in vals if v is not None]
            return {"runs": vals, "median": float(np.median(ok)) if len(ok) == len(vals) else None}
        out["repeats"] = {
            "n": repeats,
            "learned_vs_unguided": med(["learned_vs_unguided", "speedup"]),
            "learned_vs_quadtree": med(["learned_vs_quadtree", "speedup"]),
            "trimmed_learned_vs_unguided": med(["trimmed", "learned_vs_unguided", "speedup"]),
            "trimmed_learned_vs_quadtree": med(["trimmed", "learned_vs_quadtree", "speedup"]),
            "relmse_ratio_unguided_over_learned": med(["relmse_ratio_at_cap", "unguided_over_learned"]),
            "relmse_ratio_quadtree_over_learned": med(["relmse_ratio_at_cap", "quadtree_over_learned"]),
            "trimmed_ratio_unguided_over_learned": med(["trimmed_relmse_ratio_at_cap", "unguided_over_learned"]),
            "trimmed_ratio_quadtree_over_learned": med(["trimmed_relmse_ratio_at_cap", "quadtree_over_learned"])}
    return out


def batch_study(pg, dev, n_views: int, spp: int, max_depth: int, log, shape: dict | None = None,
                guided: bool = True, guide_depth: int = 0) -> dict:
    """Batched views of the config-5 scene family (seed i for view i, 100k-1M triangles),
    resolutions drawn with a seeded RNG from {128x128, 256x512, 640x360, 1280x720}: one
    hipGraph launch per SPP after the autotuner has picked the launch configuration
    (`shape`: a fixed configuration instead, for profiling runs; `guided`: also the
    learned-guided batch)."""
    rng = np.random.default_rng(12345)
    sizes = [(128, 128), (256, 512), (640, 360), (1280, 720)]
    b = pg.Batch(dev)
    t0 = time.perf_counter()
    tris = pixels = 0
    for i in range(n_views):
        d = pg.generate_scene(5, i)
        tris += d.n_tris
        sid = b.add_scene(d)
        w, h = sizes[int(rng.integers(0, 4))]
        cam = dict(d.camera)
        cam.update(width=w, height=h)
        b.add_view(sid, cam)
        pixels += w * h
    setup = time.perf_counter() - t0
    t0 = time.perf_counter()
    if shape:
        best, tuned_ms = dict(shape), None
        b.configure(**best)
    else:
        best, tuned_ms = b.autotune(max_depth=max_depth)
    tune = time.perf_counter() - t0
    b.clear()
    s0 = b.stats()
    t0 = time.perf_counter()
    b.render(spp, max_depth=max_depth)
    wall = time.perf_counter() - t0
    s1 = b.stats()
    rays = (s1["closest_rays"] - s0["closest_rays"]) + (s1["shadow_rays"] - s0["shadow_rays"])
    log(f"batch: {n_views} views, {tris} tris, {pixels} pixels, {wall * 1e3 / spp:.2f} ms/SPP, config {best}")
    # the same launch configuration with each pass replayed as one captured hipGraph
    graph_cfg = dict(best, use_graph=1)
    b.configure(**graph_cfg)
    b.clear()
    b.render(spp, max_depth=max_depth)  # captures the pass graph (and warms it)
    b.clear()
    t0 = time.perf_counter()
    b.render(spp, max_depth=max_depth)
    gr_wall = time.perf_counter() - t0
    gs = b.stats()
    graph = {"ms_per_spp": round(gr_wall * 1e3 / spp, 3), "graph_nodes": gs["graph_nodes"],
             "config": graph_cfg,
             "vs_tuned": round(wall / gr_wall, 3)}
    log(f"batch graph replay: {graph['ms_per_spp']} ms/SPP, {gs['graph_nodes']} graph nodes")
    b.configure(**best)
    guided_res = None
    if guided:
        guided_res = _batch_guided(pg, dev, b, n_views, spp, max_depth, pixels, log, guide_depth)
    b.close()
    return {"views": n_views, "scenes_tris": tris, "pixels": pixels, "spp": spp, "max_depth": max_depth,
            "resolutions": "drawn from {128x128, 256x512, 640x360, 1280x720} with a seeded RNG",
            "learned_guided": guided_res, "graph_replay": graph,
            "ms_per_spp": round(wall * 1e3 / spp, 3), "samples_per_s": round(pixels * spp / wall, 1),
            "mrays_per_s": round(rays / wall / 1e6, 2), "gpu_ms_per_spp": round((s1["gpu_ms"] - s0["gpu_ms"]) / spp, 3),
            "graph_nodes": s1["graph_nodes"], "chunks": s1["chunks"],
            "tuned_config": best, "tuned_ms_per_spp": round(tuned_ms, 3) if tuned_ms is not None else None,
            "config_source": "--batch-shape" if shape else "pg_batch_autotune",
            "setup_s": round(setup, 1), "tune_s": round(tune, 1)}


def _batch_guided(pg, dev, b, n_views: int, spp: int, max_depth: int, pixels: int, log, guide_depth: int = 0) -> dict:
    """The batch's views guided by per-scene learned networks (each trained online on its
    scene's own camera first: 16 guided record passes with one Adam step each), guiding
    every bounce and, with guide_depth > 0, the first guide_depth bounces."""
    t0 = time.perf_counter()
    for i in range(n_views):
        d = pg.generate_scene(5, i)
        sc_i = pg.Scene(dev, d)
        dev.mlp_init(seed=7 + i)
        pg.RayTracer(sc_i, seed=1 + i, max_depth=max_depth, record_nee=1).render_online(16, steps=1)
        p_i, g_i = dev.mlp_params()
        b.set_guide(i, p_i, g_i, seed=7 + i)
    train = time.perf_counter() - t0
    out = {"train_s": round(train, 1),
           "guides": "one network per scene, trained online on the scene's own camera (16 SPP)"}
    for depth in (0, guide_depth) if guide_depth else (0,):
        # an untimed render first (it captures the pass graph when the tuner chose one)
        b.clear()
        b.render(spp, max_depth=max_depth, integrator=3, guide_depth=depth)
        b.clear()
        g0 = b.stats()
        t0 = time.perf_counter()
        b.render(spp, max_depth=max_depth, integrator=3, guide_depth=depth)
        gwall = time.perf_counter() - t0
        g1 = b.stats()
        grays = (g1["closest_rays"] - g0["closest_rays"]) + (g1["shadow_rays"] - g0["shadow_rays"])
        r = {"ms_per_spp": round(gwall * 1e3 / spp, 3), "samples_per_s": round(pixels * spp / gwall, 1),
             "mrays_per_s": round(grays / gwall / 1e6, 2)}
        log(f"batch guided (guide_depth {depth}): {r['ms_per_spp']:.2f} ms/SPP after {train:.1f}s of per-scene training")
        if depth == 0:
            out.update(r)
        else:
            out[f"guide_depth_{depth}"] = r
    return out


def host_cpu() -> dict:
    """The host's CPU: model string, logical cores (std::thread::hardware_concurrency's
    count, os.cpu_count) and the cores this process may run on (nproc)."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        nproc = len(os.sched_getaffinity(0))
    except AttributeError:
        nproc = os.cpu_count() or 1
    return {"cpu_model": model, "hardware_concurrency": os.cpu_count() or 1, "nproc": nproc}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=8, help="timed frames (64 SPP each)")
    ap.add_argument("--warmup", type=int, default=1, help="untimed frames before the timed ones")
    ap.add_argument("--config", type=int, default=int(os.environ.get("PG_BENCH_CONFIG", "2")))
    ap.add_argument("--cpu-spp", type=int, default=int(os.environ.get("PG_BENCH_CPU_SPP", "32")),
                    help="SPP of the CPU oracle's multi-threaded sample")
    ap.add_argument("--cpu-spp-1t", type=int, default=int(os.environ.get("PG_BENCH_CPU_SPP_1T", "2")),
                    help="SPP of the CPU oracle's single-thread sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-ttr", action="store_true", help="skip the time-to-relMSE study")
    ap.add_argument("--batch-views", type=int, default=int(os.environ.get("PG_BENCH_BATCH_VIEWS", "48")),
                    help="views of the batched-render measurement (0 skips it)")
    ap.add_argument("--batch-spp", type=int, default=int(os.environ.get("PG_BENCH_BATCH_SPP", "32")))
    ap.add_argument("--batch-shape", default="",
                    help="fixed batch configuration instead of the batch tuner's pick, as comma-separated "
                         "pg_batch_config fields, e.g. queue_paths=1048576,samples_per_pass=64 (profiling runs)")
    ap.add_argument("--batch-no-guided", action="store_true", help="skip the learned-guided batch")
    ap.add_argument("--ttr-cap", type=float, default=float(os.environ.get("PG_BENCH_TTR_CAP", "30")),
                    help="wall-clock seconds per method in the time-to-relMSE study")
    ap.add_argument("--ttr-ref-spp", type=int, default=int(os.environ.get("PG_BENCH_TTR_REF", "4096")))
    ap.add_argument("--ttr-config", default=os.environ.get("PG_BENCH_TTR_CONFIG", "2,3"),
                    help="scenes of the time-to-relMSE study, comma-separated (2: the open interior; 3: the "
                         "closed room lit through one window, where the guiding configurations are defined)")
    ap.add_argument("--ttr-repeats", type=int, default=1,
                    help="independent runs of each method in the study (medians reported when > 1)")
    ap.add_argument("--ttr-env-visibility", type=int, choices=(0, 1), default=1,
                    help="environment shadow rays in the study: 1 along the light direction (default), "
                         "0 the reference's scene-bound target (Renderer.h:331-332)")
    ap.add_argument("--ttr-train-every", type=int, default=int(os.environ.get("PG_BENCH_TTR_TRAIN_EVERY", "32")),
                    help="learned guide: a record pass + training steps every N SPP, guided passes between")
    ap.add_argument("--ttr-train-steps", type=int, default=int(os.environ.get("PG_BENCH_TTR_TRAIN_STEPS", "8")),
                    help="learned guide: Adam steps after each record pass")
    ap.add_argument("--ttr-train-batch", type=int, default=int(os.environ.get("PG_BENCH_TTR_TRAIN_BATCH", "131072")),
                    help="learned guide: training samples per Adam step")
    ap.add_argument("--ttr-record-samples", type=int, default=int(os.environ.get("PG_BENCH_TTR_RECORD_SAMPLES", "2")),
                    help="learned guide: samples per pixel of each record pass (traced as one wavefront pass)")
    ap.add_argument("--ttr-guide-depth", type=int, default=int(os.environ.get("PG_BENCH_TTR_GUIDE_DEPTH", "2")),
                    help="guided methods guide the bounces below this depth and sample the BSDF alone deeper "
                         "(0: every bounce); the learned-guided batch is timed at 0 and at this depth")
    ap.add_argument("--ttr-pss", action="store_true",
                    help="add the primary-sample-space guide (exact and approximate neighbours) to the study")
    ap.add_argument("--shape", default="",
                    help="fixed launch shape instead of the tuner's pick, as comma-separated render "
                         "parameters, e.g. stack_lds=16,dyn_fetch=0,shade_blocks_per_cu=3 (profiling runs)")
    ap.add_argument("--copy-check", action="store_true",
                    help="also time a plain 1 GiB device copy (HBM reference point / counter scaling)")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py measures one GPU only: the path would shard by pixels, samples or views with a "
              "single film reduction; multi-GPU runs are not built.", file=sys.stderr)
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

    frame_spp = 64  # the workload's SPP: one step is one whole frame
    # launch shape (LDS stack, traversal / shade blocks per CU) picked by the library's
    # tuner on passes of the frame's shape: all 64 samples in one pass when the paths
    # fit one pass (2^26), as the automatic split renders a frame
    if W * H * frame_spp <= (1 << 26):
        rt.params.samples_per_pass = frame_spp
    t_tune = time.perf_counter()
    if a.shape:
        for kv in a.shape.split(","):
            k, v = kv.split("=")
            setattr(rt.params, k.strip(), int(v))
        tuned_ms = float("nan")
    else:
        tuned_ms = rt.tune()
    tune_s = time.perf_counter() - t_tune
    shape = {"stack_lds": rt.params.stack_lds or 16,
             "trace_blocks_per_cu": rt.params.trace_blocks_per_cu or "lds-max",
             "shade_blocks_per_cu": rt.params.shade_blocks_per_cu or 3,
             "extend_kernel": "k_extend_dyn" if rt.params.dyn_fetch else "k_extend",
             "samples_per_pass": rt.params.samples_per_pass or "auto", "tune_s": round(tune_s, 2),
             "tuned_ms_per_spp": round(tuned_ms, 3) if tuned_ms == tuned_ms else None,
             "source": "--shape" if a.shape else "pg_render_tune"}
    print(f"bench: launch shape {shape}", file=sys.stderr, flush=True)

    rt.clear()
    for _ in range(a.warmup):
        rt.render(frame_spp)
    rt.clear()
    rt.reset_stats()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        rt.render(frame_spp)  # each call returns after the stream has drained
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
    ext_name = "k_extend_dyn" if rt.params.dyn_fetch else "k_extend"
    prof = profiled_traffic(a.config)
    traffic = prof.get(ext_name)
    avg_launch_s = ext_ms * 1e-3 / launches
    shade_ms = st["kernel_ms_shade"]
    shade_bytes = SHADE_BYTES_PER_HIT * st["closest_rays"] + SHADE_BYTES_PER_SHADOW * st["shadow_rays"]
    shade_traffic = prof.get("k_shade")

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
        host = host_cpu()
        # one GPU's share of the box's host cores is 16 threads, whatever the host shows
        threads = min(host["hardware_concurrency"], host["nproc"], 16)
        osc = OracleScene(data)
        film, cst = osc.render(a.cpu_spp, seed=1, max_depth=max_depth, accum_f64=False, threads=threads, env_mode=1)
        cpu_rays = cst["closest_rays"] + cst["shadow_rays"]
        _, c1 = osc.render(a.cpu_spp_1t, seed=1, max_depth=max_depth, accum_f64=False, threads=1, env_mode=1)
        rays_1t = c1["closest_rays"] + c1["shadow_rays"]
        cpu = {"value": round(cpu_rays / cst["seconds"] / 1e6, 3), "unit": "Mrays/s", "cores": threads,
               "kind": "port",
               "sample": f"{a.cpu_spp} SPP of the same {W}x{H} frame ({cst['paths']} paths, {cpu_rays} rays, "
                         f"{cst['seconds']:.1f} s) on {threads} threads; oracle restatement with std::thread "
                         f"tiles (Renderer.h:648-747)",
               "samples_per_s": round(cst["paths"] / cst["seconds"], 1),
               "single_thread": {"value": round(rays_1t / c1["seconds"] / 1e6, 3), "unit": "Mrays/s",
                                 "sample": f"{a.cpu_spp_1t} SPP of the same frame ({c1['paths']} paths, "
                                           f"{rays_1t} rays, {c1['seconds']:.1f} s)"},
               **host,
               "threads_note": "16 threads: one GPU's share of the host's cores"}

    batch = None
    if a.batch_views > 0:
        bshape = None
        if a.batch_shape:
            bshape = {k.strip(): int(v) for k, v in (kv.split("=") for kv in a.batch_shape.split(","))}
        batch = batch_study(pg, dev, a.batch_views, a.batch_spp, max_depth,
                            lambda m: print(m, file=sys.stderr, flush=True), shape=bshape,
                            guided=not a.batch_no_guided, guide_depth=a.ttr_guide_depth)

    ttr = None
    if not a.no_ttr:
        ttr = {}
        for tc in (int(v) for v in str(a.ttr_config).split(",") if v.strip()):
            tdata = data if tc == a.config else pg.generate_scene(tc, 0)
            # a fresh upload: the batch study above leaves the device holding its last scene
            tsc = pg.Scene(dev, tdata)
            print(f"ttr: config {tc}", file=sys.stderr, flush=True)
            r = time_to_relmse(pg, dev, tsc, tdata, max_depth, a.ttr_ref_spp, a.ttr_cap,
                               lambda m: print(m, file=sys.stderr, flush=True), env_visibility=a.ttr_env_visibility,
                               repeats=a.ttr_repeats, with_pss=a.ttr_pss, train_every=a.ttr_train_every,
                               train_steps=a.ttr_train_steps, record_samples=a.ttr_record_samples,
                               guide_depth=a.ttr_guide_depth, train_batch=a.ttr_train_batch)
            ttr[f"config{tc}"] = {"workload": f"config {tc}: {tdata.n_tris} tris, {tdata.width}x{tdata.height}, "
                                              f"max depth {max_depth}", **r}
            ps = r["comparison"]["per_sample"]
            print(f"ttr config {tc}: learned vs unguided {ps['learned_vs_unguided']}, "
                  f"vs quadtree {ps['learned_vs_quadtree']}", file=sys.stderr, flush=True)
        # one line per scene: the learned guide's time-to-relMSE factors against the targets
        # (4x unguided, 2x quadtree), measured at each baseline's final relMSE (full and
        # without the worst 0.01% of pixels)
        ttr["summary"] = {k: {"learned_vs_unguided": v["comparison"]["learned_vs_unguided"]["speedup"],
                              "learned_vs_quadtree": v["comparison"]["learned_vs_quadtree"]["speedup"],
                              "trimmed_learned_vs_unguided": v["comparison"]["trimmed"]["learned_vs_unguided"]["speedup"],
                              "trimmed_learned_vs_quadtree": v["comparison"]["trimmed"]["learned_vs_quadtree"]["speedup"],
                              "targets": {"learned_vs_unguided": 4.0, "learned_vs_quadtree": 2.0},
                              "reached_relmse_0.01_s": {m: v[m]["seconds"] for m in ("unguided", "quadtree", "learned")}}
                          for k, v in ttr.items()}

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
                               f"{W}x{H}, {frame_spp} SPP per step (one frame), max depth {max_depth}, "
                               f"env map alias table",
                   "tris": data.n_tris, "width": W, "height": H, "spp_per_step": frame_spp,
                   "paths_per_step": W * H * frame_spp, "max_depth": max_depth},
        "samples_per_s": round(st["paths"] / wall, 1),
        "launch_shape": shape,
        "closest_rays": st["closest_rays"],
        "shadow_rays": st["shadow_rays"],
        "kernel_ms": {"extend": round(ext_ms, 3), "shade": round(st["kernel_ms_shade"], 3),
                      "shadow": round(sh_ms, 3), "gpu_total": round(st["kernel_ms_total"], 3)},
        "roofline": {"bound": "hbm", "kernel": ext_name, "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS,
                     "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "traffic_source": "profiles/pmc_traffic_config%d.txt (bytes per launch)" % a.config,
                     "bytes_per_ray": round(bytes_c, 1), "nodes_per_ray": round(nodes_c, 2),
                     "tris_per_ray": round(tris_c, 2), "launches": launches,
                     "avg_launch_ms": round(ext_ms / launches, 4),
                     "traffic_frac": None if traffic is None else round(traffic / avg_launch_s / 1e9 / PEAK_HBM_GBS, 4),
                     "note": "algorithmic node and triangle bytes come from L2 (hit rate 0.97) and the Infinity "
                             "Cache, so frac can exceed 1; traffic_frac is the measured HBM share; the measured "
                             "limiter is the L1 data path (profiles/k_extend_config2_1280x720.md)"},
        "shade_roofline": {"bound": "hbm", "kernel": "k_shade",
                           "achieved": round(shade_bytes / (shade_ms * 1e-3) / 1e9, 1) if shade_ms > 0 else None,
                           "peak": PEAK_HBM_GBS, "unit": "GB/s",
                           "measured_hbm_gbs": None if shade_traffic is None or shade_ms <= 0 else
                           round(shade_traffic * launches / (shade_ms * 1e-3) / 1e9, 1),
                           "traffic": None if shade_traffic is None else round(shade_traffic)},
        "shadow_roofline": {"achieved": round(sh_achieved, 1), "bytes_per_ray": round(bytes_s, 1),
                            "nodes_per_ray": round(nodes_s, 2), "tris_per_ray": round(tris_s, 2)},
        "cpu_baseline": cpu,
        "time_to_relmse": ttr,
        "batch": batch,
        "copy_reference": copy,
        "image_finite": finite,
    }
    print(json.dumps(line), flush=True)
    dev.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
