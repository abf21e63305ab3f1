#!/usr/bin/env python3
"""Throughput of the renderer on the 1920x1080 ~100k-triangle workload.

A step is one pass of the path tracer over the full frame (one sample per
pixel, GPURendererKernel.run() for every pixel) with inputs resident on the
GPU.  `value` is Mrays/s over the timed steps, where a ray is one
intersection query (closest-hit or any-hit), counted on the device.  The
timed steps launch the frame loop's kernels eagerly (the renderer's default)
with no per-launch instrumentation; the same passes replayed from captured
HIP graphs and a separate run with HIP events around every launch (the
per-stage times and the roofline's launch durations) follow.

Besides the contract fields the JSON line carries:
  roofline     - the dominant kernel's (the stage with the most kernel time)
                 algorithmic rate: its per-launch work counts x per-unit byte
                 and FLOP figures over its average launch duration, both
                 measured live (device counters; HIP events around every
                 launch on the renderer's stream), against the HBM peak when
                 its FLOP per byte is below the FP32 / HBM balance point,
                 else against the FP32 peak; `measured` sets the same rates
                 against this GPU's measured roofs (df_measure_roof: L2,
                 MALL and HBM read bandwidth, unpacked FP32 FMA rate) and,
                 from the committed rocprof issue counters, the kernel's
                 VALU issue rate and active-lane fraction;
  stages       - kernel time, launches and ray queries per step of each stage;
  batch        - the batch mix (256 scenes of 64^2..512^2 pixels and 8..128 samples per
                 pixel): Mrays/s rendering the scenes one after another (plain)
                 and in one batched launch, and their ratio;
  adjoint      - the differentiation workload (512^2, 32 spp, 347,680 triangles):
                 device time of the forward render and of the path-replay
                 adjoint, and their ratio;
  cpu_baseline - the CPU oracle (the reference's per-pixel kernel restated in
                 C++ with the reference's thread-pool semantics) on a band of
                 rows of the same frame, on this host's CPU share (at most 16
                 threads; the host's logical CPU count is reported beside).

`traced_mrays_s` in `config` counts only the queries that are traced: the
Russian-roulette exit re-tests its just-hit ray with the any-hit query, which
always hits, so the device counts it as a ray without tracing it.

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""

import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "dayflower-path-tracer_amd"))
sys.path.insert(0, REPO)

METRIC = "Mrays/s on 100k-tri 1080p scene, batched vs plain; adjoint/forward time ratio"

# Per-unit algorithmic work (hand-counted from the reference kernel lines):
#   box test   (GPURendererKernel.java:1108-1134): 6 sub, 6 mul, 12 min/max;   24 B bounds + 8 B links
#   triangle   (:1541-1594): ~45 FLOP incl. one division;                       36 B (9 floats)
#   primitive  ray transform (:1085-1095): ~36 FLOP incl. 3 divisions;          48 B (12 floats)
#   sphere     (:1433-1467): ~25 FLOP incl. sqrt and 2 divisions;               16 B
#   terrain march step (:1484-1501): 2 sin + 12 FLOP, a sin counted as one op;  0 B
#   sky        (:2127-2236): ~15 double transcendentals + ~45 FLOP;             0 B
FLOP_PER = {"box_tests": 24, "tri_tests": 45, "prim_tests": 36, "sphere_tests": 25, "terrain_steps": 14,
            "sky_evals": 60}
BYTES_PER = {"box_tests": 32, "tri_tests": 36, "prim_tests": 48, "sphere_tests": 16}
PEAK_FP32_TFLOPS = 157.3
# the timed passes run the trace kernels without the unit counters (kUnits = false), node walk (kPairs = false);
# a stage's launches may run several instantiations of one kernel: the bounce traces run the seven-wave kernel
# before the schedule's tail bounce and the four-wave kernel from it on
KERNEL_SYMBOLS = {"megakernel": ["dfk::render_kernel<0, 2>"], "raygen": ["dfk::wf_raygen<0>"],
                  "trace": ["dfk::wf_trace<0, false, false, false, 7>", "dfk::wf_trace<0, false, false, false, 4>"],
                  "shade": ["dfk::wf_shade<0, false>"], "escape": ["dfk::wf_trace<0, true, false, false, 4>"],
                  "film": ["dfk::wf_film<0>"]}
KERNEL_NAMES = {"megakernel": "render_kernel (one work-item per pixel)", "raygen": "wf_raygen",
                "trace": "wf_trace<0,false,false,false,7|4> (a bounce's closest-hit + sun any-hit queries; seven "
                         "waves per SIMD before the tail bounce, four from it on)", "shade": "wf_shade",
                "escape": "wf_trace<0,true,false,false,4> (escape tests + the last sun queries)", "film": "wf_film"}
PEAK_HBM_GBS = 8000.0


def measured_roofs():
    """This GPU's sustained rates (df_measure_roof): GB/s read from L2, MALL and HBM, TFLOP/s of unpacked FP32
    FMA, and the VALU wave-instruction issue rate that implies (one v_fma_f32 = 64 lanes x 2 FLOP)."""
    import dayflower_amd as dfa

    out = {}
    for kind, key, scale in (("l2_read", "l2_gbs", 1e-9), ("mall_read", "mall_gbs", 1e-9),
                             ("hbm_read", "hbm_gbs", 1e-9), ("valu_fp32", "valu_fp32_tflops", 1e-12)):
        rate, _ = dfa.measure_roof(kind)
        out[key] = round(rate * scale, 2)
    out["valu_winst_per_s"] = out["valu_fp32_tflops"] * 1e12 / 128.0
    return out


def _per_dispatch(entries, keys):
    """Means per dispatch over several kernels' entries ({counter: mean per dispatch, "dispatches": n}),
    weighted by their dispatches, or None when one of them is missing."""
    if not entries or any(e is None for e in entries):
        return None
    n = sum(e["dispatches"] for e in entries)
    return {k: sum(e[k] * e["dispatches"] for e in entries) / n for k in keys if all(k in e for e in entries)}


def issue_from_profiles(workload, kernel_symbols):
    """VALU issue counters per dispatch of the kernels (one stage's launches) from
    profiles/issue_workload<N>.json, or None."""
    path = os.path.join(REPO, "profiles", "issue_workload%d.json" % workload)
    if not os.path.exists(path):
        return None
    kernels = json.load(open(path)).get("kernels", {})
    r = _per_dispatch([kernels.get(k) for k in kernel_symbols],
                      ("SQ_INSTS_VALU", "SQ_ACTIVE_INST_VALU", "SQ_THREAD_CYCLES_VALU"))
    if r and r.get("SQ_ACTIVE_INST_VALU") and "SQ_THREAD_CYCLES_VALU" in r:
        r["active_lane_fraction"] = round(r["SQ_THREAD_CYCLES_VALU"] / (64.0 * r["SQ_ACTIVE_INST_VALU"]), 4)
    return r


def issue_summary(out_path, csv_paths, workload):
    """VALU issue counters per dispatch of each renderer kernel from rocprofv3 --pmc counter_collection.csv
    files (SQ_INSTS_VALU, SQ_ACTIVE_INST_VALU, SQ_THREAD_CYCLES_VALU, SQ_WAVES and the like): their means per
    dispatch and the active-lane fraction SQ_THREAD_CYCLES_VALU / (64 SQ_ACTIVE_INST_VALU)."""
    import csv

    per = {}
    for path in csv_paths:
        for r in csv.DictReader(open(path)):
            name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
            if not name.startswith(("dfk::", "dfk_roof::")):
                continue
            per.setdefault(name, {}).setdefault(r["Counter_Name"], []).append(float(r["Counter_Value"]))
    kernels = {}
    for name, counters in per.items():
        k = {c: round(sum(v) / len(v), 1) for c, v in counters.items()}
        k["dispatches"] = max(len(v) for v in counters.values())
        if k.get("SQ_ACTIVE_INST_VALU"):
            if "SQ_THREAD_CYCLES_VALU" in k:
                k["active_lane_fraction"] = round(k["SQ_THREAD_CYCLES_VALU"] / (64.0 * k["SQ_ACTIVE_INST_VALU"]), 4)
        kernels[name] = k
    summary = {"workload": workload, "files": [os.path.basename(os.path.dirname(p)) for p in csv_paths],
               "kernels": kernels}
    with open(out_path, "w") as fh:
        json.dump(summary, fh, indent=1)
        fh.write("\n")
    print(json.dumps(summary))


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(scene, params, seed, target_seconds, threads):
    """The oracle on a band of rows of one pass of the same frame."""
    import numpy as np

    import dayflower_amd as dfa
    from oracle import oracle_ffi

    w, h = scene.width, scene.height
    seeds = dfa.pixel_seeds(seed, w * h)
    start_row = h // 2 - 8
    st = oracle_ffi.OracleState(w, h, seeds)
    t0 = time.perf_counter()
    oracle_ffi.render(scene, params, st, 1, threads=threads, pixel_begin=start_row * w, pixel_end=(start_row + 4) * w)
    probe = max(time.perf_counter() - t0, 1e-3)
    rows = int(max(4, min(h, 4 * target_seconds / probe)))
    start_row = max(0, h // 2 - rows // 2)
    rows = min(rows, h - start_row)
    st = oracle_ffi.OracleState(w, h, seeds)
    t0 = time.perf_counter()
    oracle_ffi.render(scene, params, st, 1, threads=threads, pixel_begin=start_row * w,
                      pixel_end=(start_row + rows) * w)
    dt = time.perf_counter() - t0
    rays = st.counters["rays"]
    return {
        "value": round(rays / dt / 1e6, 4),
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": "rows %d-%d (%d pixels) of one 1920x1080 pass, same seeds and parameters: %d rays in %.2f s"
                  % (start_row, start_row + rows - 1, rows * w, rays, dt),
    }, np.float64(dt)


def _pmc_per_kernel(path):
    """{kernel name (no arguments): [counter value per dispatch]} from a rocprofv3 counter_collection.csv."""
    import csv

    out = {}
    for r in csv.DictReader(open(path)):
        name = r["Kernel_Name"].split("(")[0].replace("void ", "").strip()
        out.setdefault(name, []).append(float(r["Counter_Value"]))
    return out


def traffic_summary(out_path, fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, workload):
    """HBM bytes per dispatch of each renderer kernel: FETCH_SIZE and WRITE_SIZE (KiB per dispatch), each
    scaled by bytes / counter of the 1 GiB calibration copy (which reads and writes 2^30 bytes)."""
    gib_kib = float(1 << 20)
    cal_f = max(v for vals in _pmc_per_kernel(cal_fetch_csv).values() for v in vals)
    cal_w = max(v for vals in _pmc_per_kernel(cal_write_csv).values() for v in vals)
    f_scale, w_scale = gib_kib / cal_f, gib_kib / cal_w
    fetch, write = _pmc_per_kernel(fetch_csv), _pmc_per_kernel(write_csv)
    kernels = {}
    for name in fetch:
        if not name.startswith("dfk::") or name not in write:
            continue
        f = sum(fetch[name]) / len(fetch[name])
        w = sum(write[name]) / len(write[name])
        kernels[name] = {"dispatches": len(fetch[name]), "fetch_kib": round(f, 1), "write_kib": round(w, 1),
                         "hbm_bytes_per_dispatch": round((f * f_scale + w * w_scale) * 1024.0)}
    summary = {"workload": workload, "fetch_scale": round(f_scale, 4), "write_scale": round(w_scale, 4),
               "calibration": "torch.neg over 2^28 float32 (2^30 bytes read, 2^30 written): FETCH_SIZE %.1f KiB, "
                              "WRITE_SIZE %.1f KiB per dispatch" % (cal_f, cal_w),
               "kernels": kernels}
    with open(out_path, "w") as fh:
        json.dump(summary, fh, indent=1)
        fh.write("\n")
    print(json.dumps(summary))


def traffic_from_profiles(workload, kernel_symbols):
    """HBM bytes per dispatch of the kernels (one stage's launches) from profiles/traffic_workload<N>.json, or
    None."""
    path = os.path.join(REPO, "profiles", "traffic_workload%d.json" % workload)
    if not os.path.exists(path):
        return None
    kernels = json.load(open(path)).get("kernels", {})
    r = _per_dispatch([kernels.get(k) for k in kernel_symbols], ("hbm_bytes_per_dispatch",))
    return round(r["hbm_bytes_per_dispatch"]) if r else None


def adjoint_measure():
    """The differentiation workload (512x512, 32 samples per pixel, depth 6, 347,680 triangles): the forward
    render and the path-replay adjoint of L = sum (film - target)^2 with respect to 64 RGB albedo groups, 4
    emissions, the checkerboard ground's two RGB colours, turbidity, sun elevation and sun azimuth (target: a
    render at the hidden parameters of seed 500; current parameters from seed 501), device time of each, after
    the schedule autotuner has settled."""
    import numpy as np

    import dayflower_amd as dfa

    sc = dfa.build_config(5)
    sa, se, na, ne = dfa.diff_maps(sc)
    sk, nc, _ = dfa.diff_checker_map(sc)

    def draw(seed):
        r = np.random.default_rng(seed)
        return (r.uniform(0.1, 0.9, (na, 3)).astype(np.float32), r.uniform(1.0, 20.0, ne).astype(np.float32),
                r.uniform(0.1, 0.9, (nc, 2, 3)).astype(np.float32), float(r.uniform(2.0, 6.0)),
                float(r.uniform(10.0, 70.0)), float(r.uniform(0.0, 90.0)))

    def set_diff(k, th):
        a, e, c, t, el, az = th
        k.setDifferentiable(sa, se, a, e, t, el, az, surface_checker=sk, checker=c)

    k = dfa.HIPRendererKernel(sc)
    try:
        k.setParams(sc.copy_params())  # depth 6, Russian roulette off
        k.update(sc.width, sc.height)
        set_diff(k, draw(500))
        k.seed(5)
        k.lossGradient(np.zeros(sc.width * sc.height * 3, np.float32), sc.spp, gradient=False)
        target = k.getFilm()[0].copy()
        set_diff(k, draw(501))
        runs = []
        for _ in range(3):
            k.seed(7)
            loss, g, fms, ams = k.lossGradient(target, sc.spp, gradient=True)
            runs.append((fms, ams))
        engine = k.engine()
    finally:
        k.dispose()
    fms, ams = runs[-1]
    return {"workload": "512x512, %d spp, depth 6, %d triangles, %d parameters" % (sc.spp, dfa.scene.stats(sc)["triangles"],
                                                                                   g.size),
            "forward_ms": round(fms, 3), "adjoint_ms": round(ams, 3), "adjoint_over_forward": round(ams / fms, 3),
            "schedule": "wavefront" if engine == dfa.PT_WAVEFRONT else "megakernel", "loss": loss}


def batch_measure(n_scenes):
    """The uneven batch mix (scene seeds 100.., each at its own resolution and samples per pixel): every scene
    rendered alone one after another (plain, the best single-scene schedule), then all in one batched launch.
    Before the plain leg's clock starts, every scene has settled its schedule (the autotuner's timing passes,
    then the chosen schedule fixed) and allocated its path state, and its seeds and film are reset, so the
    timed region holds only the renders."""
    import dayflower_amd as dfa

    scenes = [dfa.build_batch_scene(100 + i) for i in range(n_scenes)]
    b = dfa.HIPBatchRenderer()
    try:
        for i, sc in enumerate(scenes):
            b.add(sc, passes=sc.spp, seed=1 + i)
        jobs = [b.job(i) for i in range(n_scenes)]
        engines = {dfa.PT_MEGAKERNEL: 0, dfa.PT_WAVEFRONT: 0}
        for i, j in enumerate(jobs):
            while j.engine() == dfa.PT_AUTO:
                j.execute(1)
            e = j.engine()
            engines[e] += 1
            p = j.getParams()
            p.pt_engine = e
            j.setParams(p)
            j.execute(1)  # the fixed schedule's first pass (allocations), untimed
            j.seed(1 + i)
            j.clear()
            j.synchronize()
        for j in jobs:
            j.resetCounters()
        t0 = time.perf_counter()
        for j, sc in zip(jobs, scenes):
            j.execute(sc.spp)
            j.synchronize()
        t1 = time.perf_counter()
        rays_plain = sum(j.counters()["rays"] for j in jobs)
        for j in jobs:
            j.resetCounters()
        t2 = time.perf_counter()
        b.execute()
        t3 = time.perf_counter()
        rays_batched = sum(j.counters()["rays"] for j in jobs)
    finally:
        b.dispose()
    plain = rays_plain / (t1 - t0) / 1e6
    batched = rays_batched / (t3 - t2) / 1e6
    return {"scenes": n_scenes, "samples": int(sum(sc.width * sc.height * sc.spp for sc in scenes)),
            "plain_mrays_s": round(plain, 2), "plain_s": round(t1 - t0, 3), "batched_mrays_s": round(batched, 2),
            "batched_s": round(t3 - t2, 3), "batched_over_plain": round(batched / plain, 3),
            "rays_plain": rays_plain, "rays_batched": rays_batched,
            "plain_schedules": {"megakernel": engines[dfa.PT_MEGAKERNEL], "wavefront": engines[dfa.PT_WAVEFRONT]}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", type=int, default=3, help="benchmark workload id (3 = 1920x1080 mesh + terrain)")
    ap.add_argument("--math", choices=["fast", "exact"], default="fast")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target length of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint / forward time measurement")
    ap.add_argument("--batch-scenes", type=int, default=256,
                    help="scenes of the batch mix (seeds 100, 101, ...) rendered plain and batched; 0 skips it")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="run only a plain 1 GiB float copy kernel (negation), the reference for scaling the "
                         "FETCH_SIZE / WRITE_SIZE counters to bytes")
    ap.add_argument("--calibrate-roofs", action="store_true",
                    help="run only the roof microbenchmarks (df_measure_roof) and print their rates")
    ap.add_argument("--issue-summary", nargs="+", metavar="PATH",
                    help="OUT CSV...: write the VALU issue summary OUT from rocprofv3 --pmc counter_collection.csv "
                         "files of this bench (SQ_* counters)")
    ap.add_argument("--traffic-summary", nargs=5, metavar=("OUT", "FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV",
                                                           "CAL_WRITE_CSV"),
                    help="write the HBM-traffic summary OUT from rocprofv3 --pmc counter_collection.csv files of "
                         "this bench (FETCH_SIZE, WRITE_SIZE) and of --calibrate-copy (same counters)")
    args = ap.parse_args()
    if args.traffic_summary:
        traffic_summary(*args.traffic_summary, workload=args.workload)
        return
    if args.issue_summary:
        issue_summary(args.issue_summary[0], args.issue_summary[1:], workload=args.workload)
        return

    if args.gpus != 1:
        print("bench.py measures one GPU: this build runs on a single MI355X and has no multi-GPU path "
              "(--gpus %d requested)" % args.gpus, file=sys.stderr)
        sys.exit(2)

    import torch

    import dayflower_amd as dfa

    if not torch.cuda.is_available():
        log("no HIP device visible")
        sys.exit(1)
    if args.calibrate_copy:
        n = 1 << 28  # 1 GiB of float32 read and 1 GiB written per launch
        x = torch.ones(n, device="cuda", dtype=torch.float32)
        y = torch.empty_like(x)
        for _ in range(3):
            torch.neg(x, out=y)
        torch.cuda.synchronize()
        print(json.dumps({"calibration": "torch.neg float32", "bytes_read_per_launch": n * 4,
                          "bytes_written_per_launch": n * 4, "launches": 3}), flush=True)
        return
    if args.calibrate_roofs:
        print(json.dumps(measured_roofs()), flush=True)
        return

    scene = dfa.build_config(args.workload)
    params = scene.copy_params()
    params.math_mode = dfa.MATH_FAST if args.math == "fast" else dfa.MATH_EXACT
    stats = dfa.scene.stats(scene)
    seed = 42
    k = dfa.HIPRendererKernel(scene, device=0)
    k.setParams(params)
    k.update(scene.width, scene.height)
    k.updateCamera()
    k.seed(seed)
    tune_passes = 0
    while k.engine() == dfa.PT_AUTO:  # the autotuner times both schedules on the first passes
        k.execute(1)
        tune_passes += 1
    # the timed passes leave the unit counters off (rays and samples are still counted); the passes whose
    # kernel times the roofline uses are counted again afterwards from the same seeds
    k.setUnitCounters(False)
    for _ in range(args.warmup):
        k.execute(1)
    k.synchronize()
    k.resetCounters()

    # the timed steps: eager launches (the default), no per-launch instrumentation
    torch.cuda.synchronize()
    k.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        k.execute(1)
    k.synchronize()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    c = k.counters()
    stage_counts = k.stageStats()  # per-stage device counters of the timed steps (no times: not instrumented)
    k_engine = k.engine()

    # the same passes replayed from the frame loop's captured graphs, for the graphs' own effect
    eager_steps = max(1, min(args.steps, 10))
    k.setGraphs(True)
    k.execute(1)  # captures
    k.synchronize()
    te0 = time.perf_counter()
    for _ in range(eager_steps):
        k.execute(1)
    k.synchronize()
    te1 = time.perf_counter()
    k.setGraphs(False)

    # per-stage kernel times: a separate run with HIP events around every launch, then the same passes again
    # from the same seeds with the unit counters on (the work of a pass is a function of the scene, the
    # parameters and the seeds), for the counts the roofline divides by those times
    stage_steps = eager_steps
    seeds_stage = k.getSeeds()
    k.setStageTiming(True)
    k.resetCounters()
    for _ in range(stage_steps):
        k.execute(1)
    k.synchronize()
    stages = k.stageStats()
    k.setStageTiming(False)
    k.setSeeds(seeds_stage)
    k.setUnitCounters(True)
    k.resetCounters()
    for _ in range(stage_steps):
        k.execute(1)
    k.synchronize()
    counted = k.stageStats()
    for n in stages:
        assert counted[n]["counters"]["rays"] == stages[n]["counters"]["rays"], n  # the same passes
        stages[n]["counters"] = counted[n]["counters"]
    k.dispose()

    wall = t1 - t0
    steps = args.steps
    value = c["rays"] / wall / 1e6
    # the dominant kernel: the stage with the most kernel time
    dom = max((n for n in stages if stages[n]["launches"] > 0), key=lambda n: stages[n]["ms"])
    st = stages[dom]
    sc_ = st["counters"]
    avg_launch_s = st["ms"] / st["launches"] / 1e3
    flop = sum(FLOP_PER[n] * sc_[n] for n in FLOP_PER) / st["launches"]
    alg_bytes = sum(BYTES_PER[n] * sc_[n] for n in BYTES_PER) / st["launches"]
    achieved_tflops = flop / avg_launch_s / 1e12
    achieved_gbs = alg_bytes / avg_launch_s / 1e9
    # the roofline that bounds the kernel: HBM below the FP32 / HBM balance point
    intensity = flop / alg_bytes if alg_bytes else float("inf")
    balance = PEAK_FP32_TFLOPS * 1e12 / (PEAK_HBM_GBS * 1e9)
    if intensity < balance:
        bound, achieved, peak, unit = "hbm", achieved_gbs, PEAK_HBM_GBS, "GB/s"
        note = ("%.3g FLOP per algorithmic byte, below the FP32 / HBM balance of %.3g: the HBM roofline bounds it "
                "(FP32 side: %.4g TFLOP/s = %.4f of %.4g; no matrix work)"
                % (intensity, balance, achieved_tflops, achieved_tflops / PEAK_FP32_TFLOPS, PEAK_FP32_TFLOPS))
    else:
        bound, achieved, peak, unit = "mfma", achieved_tflops, PEAK_FP32_TFLOPS, "TFLOP/s"
        note = ("%.3g FLOP per algorithmic byte, above the FP32 / HBM balance of %.3g: the FP32 peak bounds it "
                "(no matrix work)" % (intensity, balance))
    roofs = measured_roofs()
    issue = issue_from_profiles(args.workload, KERNEL_SYMBOLS.get(dom, []))
    measured = {
        "roofs": roofs,
        "alg_bytes_over_l2": round(achieved_gbs / roofs["l2_gbs"], 4),
        "alg_bytes_over_mall": round(achieved_gbs / roofs["mall_gbs"], 4),
        "alg_flops_over_valu": round(achieved_tflops / roofs["valu_fp32_tflops"], 4),
    }
    if issue and issue.get("SQ_INSTS_VALU"):
        measured["valu_insts_per_launch"] = issue["SQ_INSTS_VALU"]
        measured["valu_issue_over_peak"] = round(issue["SQ_INSTS_VALU"] / avg_launch_s / roofs["valu_winst_per_s"], 4)
        if "active_lane_fraction" in issue:
            measured["active_lane_fraction"] = issue["active_lane_fraction"]
        measured["issue_counters"] = "profiles/issue_workload%d.json" % args.workload
    roofline = {
        "bound": bound,
        "achieved": round(achieved, 4),
        "peak": peak,
        "unit": unit,
        "frac": round(achieved / peak, 6),
        "traffic": traffic_from_profiles(args.workload, KERNEL_SYMBOLS.get(dom, [])),
        "kernel": KERNEL_NAMES[dom],
        "note": note,
        "avg_launch_ms": round(avg_launch_s * 1e3, 4),
        "launches_per_step": st["launches"] / stage_steps,
        "flop_per_launch": flop,
        "bytes_per_launch": alg_bytes,
        "measured": measured,
    }
    stage_summary = {n: {"ms_per_step": round(stages[n]["ms"] / stage_steps, 3),
                         "launches_per_step": stages[n]["launches"] / stage_steps,
                         "rays_per_step": stages[n]["counters"]["rays"] / stage_steps}
                     for n in stages if stages[n]["launches"] > 0}

    batch = batch_measure(args.batch_scenes) if args.batch_scenes > 0 else None
    adjoint = None if args.no_adjoint else adjoint_measure()

    cpu = None
    if not args.no_cpu:
        # the host's CPU share for one GPU is 16 threads, whatever the host's own count
        visible = len(os.sched_getaffinity(0))
        threads = min(16, visible)
        cpu, _ = cpu_baseline(scene, params, seed, args.cpu_seconds, threads)
        cpu["gpu_over_cpu"] = round(value / cpu["value"], 2) if cpu["value"] > 0 else None
        cpu["host_logical_cpus"] = os.cpu_count()
        cpu["affinity_cpus"] = visible

    out = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: procedural workload %d (%s), seeds documented in the scene compiler" % (args.workload,
                                                                                                  scene.name),
        "config": {
            "workload": "%s: %dx%d, 1 sample per pixel per step, depth %d, %d triangles, %d spheres, %d terrain, "
                        "%s math" % (scene.name, scene.width, scene.height, params.pt_depth_maximum,
                                     stats["triangles"], stats["spheres"], stats["terrains"], args.math),
            "samples_per_step": scene.width * scene.height,
            "schedule": "wavefront" if k_engine == dfa.PT_WAVEFRONT else "megakernel",
            "autotune_passes": tune_passes,
            "rays_per_step": c["rays"] / steps,
            "traced_mrays_s": round((c["rays"] - stage_counts["shade"]["counters"]["rays"]) / wall / 1e6, 3),
            "frame_loop": "eager launches",
            "ms_per_step_graph_replay": round((te1 - te0) / eager_steps * 1e3, 3),
            "ms_per_step_stage_timed": round(sum(stages[n]["ms"] for n in stages) / stage_steps, 3),
            "samples_per_s": round(c["samples"] / wall, 1),
        },
        "roofline": roofline,
        "stages": stage_summary,
        "batch": batch,
        "adjoint": adjoint,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
