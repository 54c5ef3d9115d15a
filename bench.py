#!/usr/bin/env python3
"""Benchmark of the vkvg MI355X backend.

Headline: the dense 4096x4096 scene (20,000 paths of 1-40 cubics, 60% non-zero fills / 40%
strokes, 30% gradient sources). One step = one frame through the whole kernel pipeline
(flatten -> stroke -> edges -> binning -> coarse -> fine composite) with the frame's inputs
already resident in HBM. value = paths per second (Mpaths/s); Gpix/s composited is reported
beside it. Also measured: the batched form (256 uneven surfaces in one frame vs 256 plain
per-surface frames) and, when built, the differentiable renderer's fwd+bwd vs forward time.

Prints ONE JSON line.
"""
from __future__ import annotations

import argparse
import gc
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s
METRIC = "Mpaths/s + Gpix/s composited (4K scene); batched & fwd+bwd ms vs plain"


def _sync():
    try:
        import torch
        if torch.cuda.is_available():
            torch.cuda.synchronize()
    except Exception:
        pass


def _read_pmc(name: str, key: str):
    """A per-launch figure of the dominant kernel from the committed rocprofv3 --pmc summary."""
    p = os.path.join(ROOT, "profiles", name)
    if not os.path.exists(p):
        return None
    try:
        with open(p) as f:
            return json.load(f).get(key)
    except Exception:
        return None


def _read_traffic(name: str):
    """HBM bytes per launch of the dominant kernel."""
    return _read_pmc(name, "hbm_bytes_per_launch")


def _timed(dev, vkvg_amd, steps):
    """Executes the resident frame `steps` times; returns wall ms per frame and per-run stats.
    Python's garbage collector is run before and held off during the timed loop (the scene's
    command stream is a large Python container, and a collection inside the loop would be timed)."""
    gc.collect()
    gc.disable()
    try:
        _sync()
        runs = []
        t0 = time.perf_counter()
        for _ in range(steps):
            vkvg_amd.execute(dev)
            runs.append(dev.stats())
        _sync()
        return (time.perf_counter() - t0) * 1e3 / steps, runs
    finally:
        gc.enable()


def bench_c3(args, dev, vkvg_amd, scenes):
    sc = scenes.config3()
    surf = vkvg_amd.Surface(dev, sc.width, sc.height)
    surf.upload(sc.words())  # frame resident in HBM
    tune = dev.autotune(reps=2) if not args.no_tune else None
    # eager executions (per-stage events, fine-kernel time for the roofline)
    for _ in range(args.warmup):
        vkvg_amd.execute(dev)
    ms_eager, runs = _timed(dev, vkvg_amd, args.steps)
    fine_ms = float(np.mean([r["ms_fine"] for r in runs]))
    stage = {k: float(np.mean([r["ms_stage"][k] for r in runs])) for k in runs[0]["ms_stage"]}
    st = runs[-1]
    # headline: the same frame replayed as one captured graph per step
    ms_step, dev_ms = ms_eager, float(np.mean([r["ms_device"] for r in runs]))
    step_ms = [round(r["ms_total"], 3) for r in runs]
    if not args.no_graph:
        dev.set_graph(True)
        vkvg_amd.execute(dev)  # sizing run + capture
        for _ in range(args.warmup):
            vkvg_amd.execute(dev)
        ms_step, g_runs = _timed(dev, vkvg_amd, args.steps)
        assert all(r["graph"] == 1 for r in g_runs)
        step_ms = [round(r["ms_total"], 3) for r in g_runs]
        dev_ms = float(np.mean([r["ms_device"] for r in g_runs]))
        dev.set_graph(False)
    surf.close()
    return sc, ms_step, ms_eager, fine_ms, dev_ms, stage, st, tune, step_ms


def bench_resident(args, dev, vkvg_amd, sc, steps):
    """One scene's frame with its inputs resident in HBM, tile 16, replayed as a captured graph:
    (ms per frame, per-stage ms of the eager runs)."""
    surf = vkvg_amd.Surface(dev, sc.width, sc.height)
    try:
        surf.upload(sc.words())
        for _ in range(max(1, args.warmup)):
            vkvg_amd.execute(dev)
        _, runs = _timed(dev, vkvg_amd, max(3, steps // 2))
        stage = {k: float(np.mean([r["ms_stage"][k] for r in runs])) for k in runs[0]["ms_stage"]}
        ms = float(np.mean([r["ms_total"] for r in runs]))
        if not args.no_graph:
            dev.set_graph(True)
            vkvg_amd.execute(dev)
            for _ in range(max(1, args.warmup)):
                vkvg_amd.execute(dev)
            ms, _ = _timed(dev, vkvg_amd, steps)
            dev.set_graph(False)
    finally:
        surf.close()
    return ms, stage


def bench_api(dev, vkvg_amd, sc, reps: int = 5):
    """One frame through the C API as a client renders it: vkvg_create, every path and paint call of
    the scene, vkvg_destroy (host recording, one H2D copy of the frame, the kernel pipeline, the
    wait). The image stays on the device. Median of `reps` after one warm-up."""
    surf = vkvg_amd.Surface(dev, sc.width, sc.height)
    words = sc.words()
    try:
        surf.replay(words)
        ts = []
        gc.collect()
        gc.disable()
        try:
            for _ in range(reps):
                _sync()
                t = time.perf_counter()
                surf.replay(words)
                ts.append((time.perf_counter() - t) * 1e3)
        finally:
            gc.enable()
    finally:
        surf.close()
    return float(np.median(ts))


def bench_configs(args, dev, vkvg_amd, scenes, c3_scene, c3_ms):
    """Frame times of the other configurations and of the reference's own random-shape benchmark
    (tests/perfs/randoms.c: 500 shapes on 512x512), resident and through the API."""
    out = {}
    for key, sc in (("c1_512_nz_cubics", scenes.config1()), ("c2_1080p_stroke_eo", scenes.config2()),
                    ("randoms_random_fill_stroke_512", scenes.randoms(scenes.SHAPE_RANDOM, "fill_stroke")),
                    ("randoms_circles_fill_stroke_512", scenes.randoms(scenes.SHAPE_CIRCLE, "fill_stroke")),
                    ("randoms_rounded_rects_stroke_4096", scenes.randoms(scenes.SHAPE_ROUNDED_RECTANGLE, "stroke",
                                                                        n=20000, size=4096))):
        ms, stage = bench_resident(args, dev, vkvg_amd, sc, max(5, args.steps // 2))
        api = bench_api(dev, vkvg_amd, sc)
        out[key] = {"surface": f"{sc.width}x{sc.height}", "paths": sc.n_paths, "ms_per_frame": ms,
                    "mpaths_per_s": sc.n_paths / ms / 1e3, "gpix_per_s": sc.width * sc.height / ms / 1e6,
                    "stages_ms": stage, "api_ms_per_frame": api, "api_mpaths_per_s": sc.n_paths / api / 1e3}
    api = bench_api(dev, vkvg_amd, c3_scene, reps=3)
    out["c3_4k_dense"] = {"surface": f"{c3_scene.width}x{c3_scene.height}", "paths": c3_scene.n_paths,
                          "ms_per_frame": c3_ms, "mpaths_per_s": c3_scene.n_paths / c3_ms / 1e3,
                          "api_ms_per_frame": api, "api_mpaths_per_s": c3_scene.n_paths / api / 1e3}
    dev.set_tile_size(args.tile)
    return out


def bench_c4(args, dev, vkvg_amd, scenes):
    subs = scenes.config4()
    surfs = [vkvg_amd.Surface(dev, s.width, s.height) for s in subs]
    words = [s.words() for s in subs]
    # plain: one frame per surface (each execute timed on resident inputs), default configuration
    dev.set_tile_size(16)
    dev.set_fine_words(4096)
    plain = 0.0
    gc.collect()
    gc.disable()
    try:
        for s, w in zip(surfs, words):
            s.upload(w)
            vkvg_amd.execute(dev)  # warm
            _sync()
            t = time.perf_counter()
            vkvg_amd.execute(dev)
            plain += time.perf_counter() - t
    finally:
        gc.enable()
    # batched: one frame over all 256 surfaces, autotuned, replayed as one graph launch
    vkvg_amd.upload_batch(dev, surfs, words)
    tune = dev.autotune(reps=1) if not args.no_tune else None
    reps = max(3, args.steps // 4)
    for _ in range(max(1, args.warmup)):
        vkvg_amd.execute(dev)
    eager, runs = _timed(dev, vkvg_amd, reps)
    stage = {k: float(np.mean([r["ms_stage"][k] for r in runs])) for k in runs[0]["ms_stage"]}
    st = runs[-1]
    batched = eager
    if not args.no_graph:
        dev.set_graph(True)
        vkvg_amd.execute(dev)
        batched, _ = _timed(dev, vkvg_amd, reps)
        dev.set_graph(False)
    for s in surfs:
        s.close()
    dev.set_tile_size(16)
    dev.set_fine_words(4096)
    return {"plain_ms": plain * 1e3, "batched_ms": batched, "batched_eager_ms": eager,
            "ratio": plain * 1e3 / batched, "surfaces": len(subs), "paths": int(sum(s.n_paths for s in subs)),
            "pixels": int(sum(s.width * s.height for s in subs)),
            "tuned": tune["best"] if tune else None, "stages_ms": stage,
            "counts": {k: st[k] for k in ("points", "stroke_triangles", "edges", "path_tiles", "segments", "commands")}}


def bench_fwd_bwd(args):
    try:
        from vkvg_amd import diff
    except Exception:
        return None
    if not hasattr(diff, "bench"):
        return None
    return diff.bench(steps=max(5, args.steps), warmup=max(1, args.warmup))


def _host_cpu():
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return cores, model


def cpu_baseline(n_paths: int = 15000):
    """The CPU restatement of vkvg (the oracle) timed on this host, lavapipe being unavailable on these
    machines: one thread on the first n_paths paths of the 4096x4096 dense scene, then all host
    cores (up to 16) rasterizing bands of rows of the whole 20000-path scene, each band repeating the
    geometry; also the 512x512 (200 paths) and 1920x1080 (7000 paths) scenes, single thread and all
    cores."""
    from oracle import ovg
    from vkvg_amd import scenes
    cores, model = _host_cpu()
    threads = max(1, min(16, cores))
    sub = scenes.config3(n_paths=n_paths)
    t = time.perf_counter()
    ovg.render(sub.width, sub.height, sub.words(), with_sig=False)
    dt1 = time.perf_counter() - t
    full = scenes.config3()
    t = time.perf_counter()
    ovg.render_bands(full.width, full.height, full.words(), threads, with_sig=False)
    dtn = time.perf_counter() - t
    others = {}
    for key, sc in (("c1_512_nz_cubics", scenes.config1()), ("c2_1080p_stroke_eo", scenes.config2())):
        w = sc.words()
        t = time.perf_counter()
        ovg.render(sc.width, sc.height, w, with_sig=False)
        a = time.perf_counter() - t
        t = time.perf_counter()
        ovg.render_bands(sc.width, sc.height, w, threads, with_sig=False)
        b = time.perf_counter() - t
        others[key] = {"ms_1_thread": a * 1e3, f"ms_{threads}_threads": b * 1e3,
                       "mpaths_per_s_1_thread": sc.n_paths / a / 1e6,
                       f"mpaths_per_s_{threads}_threads": sc.n_paths / b / 1e6}
    return {"value": full.n_paths / dtn / 1e6, "unit": "Mpaths/s", "cores": threads, "kind": "port",
            "sample": f"the whole 20000-path 4096x4096 dense scene as {threads} bands of rows on {threads} "
                      f"threads ({dtn:.1f} s; geometry repeated per band), CPU restatement of vkvg",
            "single_thread": {"value": n_paths / dt1 / 1e6, "unit": "Mpaths/s", "cores": 1,
                              "sample": f"first {n_paths} of the 20000 paths, 1 thread, {dt1:.1f} s"},
            "other_configs": others, "host_cores": cores, "cpu_model": model}


def copy_probe(mib: int) -> int:
    """A plain elementwise copy of known size (reads and writes mib MiB once per launch) used to
    calibrate the FETCH_SIZE / WRITE_SIZE counters of rocprofv3 --pmc."""
    import torch
    n = mib * (1 << 20) // 4
    src = torch.rand(n, device="cuda")
    dst = torch.empty_like(src)
    for _ in range(5):
        torch.mul(src, 1.0, out=dst)
    torch.cuda.synchronize()
    print(json.dumps({"copy_probe_bytes_read": n * 4, "copy_probe_bytes_written": n * 4, "launches": 5}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--tile", type=int, default=16)
    ap.add_argument("--no-batch", action="store_true")
    ap.add_argument("--no-configs", action="store_true", help="skip the other configurations and API timings")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-paths", type=int, default=15000)
    ap.add_argument("--no-tune", action="store_true", help="skip the autotuner (tile 16, 4096 words)")
    ap.add_argument("--no-graph", action="store_true", help="time eager executions instead of graph replays")
    ap.add_argument("--copy-probe", type=int, default=0, metavar="MIB",
                    help="only run a plain copy kernel of MIB MiB (counter calibration)")
    args = ap.parse_args()
    if args.copy_probe:
        return copy_probe(args.copy_probe)
    if args.gpus != 1:
        print(json.dumps({"error": "this backend renders on one GPU (how frames would shard over several is "
                                   "a design, not built): run with --gpus 1"}))
        return 2
    try:  # torch (used by the differentiable part) must bring up its device before the library does
        import torch
        torch.cuda.init()
    except Exception:
        pass
    import vkvg_amd
    from vkvg_amd import scenes

    dev = vkvg_amd.Device()
    dev.set_tile_size(args.tile)
    sc, ms_step, ms_eager, fine_ms, dev_ms, stage, st, tune, step_ms = bench_c3(args, dev, vkvg_amd, scenes)
    paths_per_s = sc.n_paths / (ms_step / 1e3)
    gpix = sc.width * sc.height / (ms_step / 1e3) / 1e9
    # dominant kernel: the fine coverage + composite kernel; its algorithmic bytes per launch
    # are the frame's image write plus per-path style and gradient records (geometry reads are
    # intermediates and do not count)
    fine_bytes = 4 * sc.width * sc.height + 32 * sc.n_paths + 24 * sc.gradients + 20 * sc.gradient_stops
    achieved = fine_bytes / (fine_ms / 1e3) / 1e9
    traffic = _read_traffic("fine_c3_pmc.json")
    # The fine kernel is limited by vector instruction issue, not by HBM: a wave64 VALU instruction
    # occupies a 16-lane SIMD for 4 cycles, so the chip issues at most 1024 SIMDs x 2.4 GHz / 4 wave
    # instructions per second. Its VALU instructions per launch come from the same counter summary.
    valu = _read_pmc("fine_c3_pmc.json", "valu_insts_per_launch")
    issue_peak = 1024 * 2.4e9 / 4
    issue = None if valu is None else {"achieved": valu / (fine_ms / 1e3), "peak": issue_peak,
                                       "unit": "wave64 VALU instructions/s",
                                       "frac": valu / (fine_ms / 1e3) / issue_peak,
                                       "valu_per_launch": valu,
                                       "model": "4 cycles per counted wave64 VALU instruction on a 16-lane "
                                                "SIMD; packed 16-bit instructions issue faster, so frac "
                                                "near or above 1 means the kernel is at the issue limit"}
    # LDS from the same counter summary: SQ_LDS_IDX_ACTIVE over (GRBM_GUI_ACTIVE per XCD x 256 CUs)
    # is the fraction of CU-cycles the LDS served indexed accesses (peak 1); LDS atomics per second
    # against one wave64 LDS instruction per CU per cycle
    lds_act = _read_pmc("fine_c3_pmc.json", "lds_idx_active_cycles")
    lds_bank = _read_pmc("fine_c3_pmc.json", "lds_bank_conflict_cycles")
    gui = _read_pmc("fine_c3_pmc.json", "grbm_gui_active_cycles")
    lds_atom = _read_pmc("fine_c3_pmc.json", "lds_atomic_insts_per_launch")
    lds = None
    if None not in (lds_act, lds_bank, gui, lds_atom):
        lds = {"util": lds_act / (gui / 8 * 256), "bank_conflict_share": lds_bank / lds_act,
               "atomic_insts_per_s": lds_atom / (fine_ms / 1e3), "atomic_peak": 256 * 2.4e9,
               "atomic_frac": lds_atom / (fine_ms / 1e3) / (256 * 2.4e9)}
    out = {
        "metric": METRIC,
        "value": paths_per_s / 1e6,
        "unit": "Mpaths/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u8",
        "data": "synthetic",
        "config": {"workload": "dense 4096x4096 RGBA8 scene: 20000 paths of 1-40 cubics, 16-blob Gaussian "
                               "mixture, 60% non-zero fill / 40% stroke 0.5-8 px, 30% linear/radial gradients",
                   "surface": "4096x4096", "paths": sc.n_paths, "control_points": sc.ctrl_points,
                   "gradients": sc.gradients, "tile": st["tile_size"], "fine_words": st["fine_words"],
                   "inputs": "resident in HBM"},
        "gpix_per_s": gpix,
        "frame_algorithmic_bytes": sc.algorithmic_bytes(),
        "frame_hbm_gbs": sc.algorithmic_bytes() / (ms_step / 1e3) / 1e9,
        "device_ms": dev_ms,
        "eager_ms_per_step": ms_eager,
        "step_host_ms": step_ms,
        "graph_replay": not args.no_graph,
        "tuned": tune["best"] if tune else {"tile": st["tile_size"], "fine_words": st["fine_words"]},
        "stages_ms": stage,
        "counts": {k: st[k] for k in ("points", "stroke_triangles", "edges", "path_tiles", "segments", "commands")},
        "roofline": {"kernel": f"k_fine<{st['tile_size']},{st['fine_words']},false>", "bound": "hbm",
                     "achieved": achieved, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "algorithmic_bytes_per_launch": fine_bytes, "avg_launch_ms": fine_ms,
                     "vector_issue": issue, "lds": lds},
    }
    if not args.no_configs:
        out["configs"] = bench_configs(args, dev, vkvg_amd, scenes, sc, ms_step)
    if not args.no_batch:
        out["batched"] = bench_c4(args, dev, vkvg_amd, scenes)
    fb = bench_fwd_bwd(args)
    if fb is not None:
        out["fwd_bwd"] = fb
    dev.close()
    if not args.no_cpu:
        cb = cpu_baseline(args.cpu_paths)
        out["cpu_baseline"] = cb
        out["speedup_vs_cpu"] = out["value"] / cb["value"]
        out["speedup_vs_cpu_1_thread"] = out["value"] / cb["single_thread"]["value"]
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
