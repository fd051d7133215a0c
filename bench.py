#!/usr/bin/env python3
"""Benchmark of the per-plane path on one MI355X.

A step is one batched launch sequence of a workload over F distinct frames
already resident in HBM (F = 64 by default, so each step moves more than the
256 MB MALL holds). Workloads (sizes and inputs in zimg_amd/workloads.py):

  fhd_709_to_2020  FHD 1920x1080 YUV 4:2:0 10-bit limited, BT.709 -> BT.2020 NCL,
                   bicubic chroma, FP32 intermediates          (primary: `value`)
  uhd_lanczos      WORD 1920x1080 -> 3840x2160 Lanczos3
  fp32_down        FP32 3840x2160 -> 1920x1080 Lanczos3   (exact, and FP16 hi+lo MFMA)
  fp32_up          FP32 1280x720 -> 3840x2160 bicubic     (exact, and FP16 hi+lo MFMA)
  batch64          64 uneven 4:2:0 jobs (seeded list), one persistent launch
  uhd_pq_to_hlg    UHD RGB FP32 PQ -> HLG -> 12-bit error diffusion

Each workload reports frames/s over K timed steps (HIP events on the launch
stream) and a roofline object: `achieved` is the workload's algorithmic bytes
per frame (input planes read once, output planes written once) times the
frames one launch of the dominant kernel processes, divided by that kernel's
average duration (HIP events around each stage, on the same stream);
`whole_graph` divides the same bytes by the whole step. `traffic` is the
dominant kernel's HBM bytes per launch from the committed rocprofv3 counter
summary (profiles/<workload>_counters.txt); `binding` names what bounds the
kernel when it is not HBM (VALU issue, from the same counters).

cpu_baseline: the oracle (the project's C restatement of the reference's C
path) on this host: single-thread median of >= 10 runs, and frame-parallel
over the usable cores (one frame per thread, the reference harness pattern of
graphapp.cpp:355-426), with the CPU model. For batch64: jobs/s of
single-thread passes over the 64-job list and of the jobs spread over the
cores.

Two-stream workloads alternate batches over one pair of streams created once
per process. Before the timed region each graph's tiled resize steps are
autotuned on the workload's own buffers (zimg_filter_graph_autotune_device:
candidate tile shapes timed per step, the fastest kept; outputs unchanged).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME [--variant FORM]] [--no-cpu]
"""
import argparse
import json
import os
import statistics
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
METRIC = "frames/s + HBM GB/s, UHD Lanczos & 709→2020; MFMA/learned vs plain at ≤0.5 LSB"
WORKLOADS = ["fhd_709_to_2020", "uhd_lanczos", "fp32_down", "fp32_up", "batch64", "uhd_pq_to_hlg"]
# Streams the timed batches alternate on: the error-diffusion stage of
# uhd_pq_to_hlg is latency-bound (one dependent wavefront per plane), so the
# next batch's colourspace stage runs beside it; in fhd_709_to_2020 the next
# batch's first launches fill the tails of the current batch's five launches.
DEFAULT_STREAMS = {"uhd_pq_to_hlg": 2, "fhd_709_to_2020": 2}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--frames", type=int, default=None, help="distinct frames per step (default 64)")
    ap.add_argument("--workload", default=None, choices=WORKLOADS,
                    help="measure one workload only (its line becomes the primary); default: all, fhd_709_to_2020 primary")
    ap.add_argument("--variant", default="all", choices=["all", "exact", "fp16", "bf16", "approx", "learned"],
                    help="with --workload: measure only this form (exact path, FP16/BF16 matrix-core resize, "
                         "approximate gamma, learned transfers); default: the exact path and every form the workload has")
    ap.add_argument("--streams", type=int, default=0,
                    help="streams the timed batches alternate on (0: the workload's default, 2 where a latency-bound stage "
                         "can run beside the next batch's compute-bound stage)")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baselines")
    ap.add_argument("--cpu-runs", type=int, default=10, help="runs per CPU baseline mode (median reported)")
    return ap.parse_args()


def device_batch(torch, graph, planes, out_shapes, frames, stream=None, inputs=None):
    """Upload `frames` distinct frames per plane (or reuse `inputs`, the
    tensors of another launcher); return a launcher closure on `stream`
    (default: torch's current stream) with its own outputs and tmp, and its tensors."""
    import zimg_amd as z
    keep = []
    src_bufs, src_fs = [None] * 4, [0] * 4
    ins = []
    for p, stack in enumerate(planes):
        if stack is None:
            continue
        t = inputs[len(ins)] if inputs is not None else torch.from_numpy(np.ascontiguousarray(stack)).to("cuda")
        ins.append(t)
        keep.append(t)
        F, h, w = stack.shape
        src_bufs[p] = (t.data_ptr(), w * stack.itemsize)
        src_fs[p] = h * w * stack.itemsize
    dst_bufs, dst_fs = [None] * 4, [0] * 4
    for p, spec in enumerate(out_shapes):
        if spec is None:
            continue
        h, w, dt = spec
        t = torch.empty((frames, h, w * np.dtype(dt).itemsize), dtype=torch.uint8, device="cuda")
        keep.append(t)
        dst_bufs[p] = (t.data_ptr(), w * np.dtype(dt).itemsize)
        dst_fs[p] = h * w * np.dtype(dt).itemsize
    tmp = torch.empty(max(graph.device_tmp_size(frames), 1), dtype=torch.uint8, device="cuda")
    keep.append(tmp)
    src = z.image_buffer(src_bufs)
    dst = z.image_buffer(dst_bufs)
    hs = (stream or torch.cuda.current_stream()).cuda_stream

    def launch():
        graph.process_device_batch(frames, src, src_fs, dst, dst_fs, tmp.data_ptr(), hs)

    def profiled():
        return graph.process_device_profiled(frames, src, src_fs, dst, dst_fs, tmp.data_ptr(), hs)

    def autotune():
        return graph.autotune_device(frames, src, src_fs, dst, dst_fs, tmp.data_ptr(), hs)

    launch.profiled = profiled
    launch.autotune = autotune
    launch.inputs = ins
    return launch, keep


_PIPE_STREAMS = []


def pipe_streams(torch, n):
    """The streams batches alternate on, created once per process and reused
    by every workload: each new stream is bound to one of the device's few
    hardware queues when it is created, and two streams that land on the
    same queue run one after the other, so later workloads must not get a
    fresh pair."""
    while len(_PIPE_STREAMS) < n:
        _PIPE_STREAMS.append(torch.cuda.Stream())
    return _PIPE_STREAMS[:n]


def time_steps_pipelined(torch, launches, streams, steps, warmup):
    """Seconds for `steps` launches issued alternately on len(streams)
    streams (each launcher its own outputs and tmp): consecutive batches are
    independent, so one batch's latency-bound stage runs beside the next
    batch's compute-bound stage. Events on the first stream bracket all of it."""
    n = len(streams)
    for k in range(warmup):
        with torch.cuda.stream(streams[k % n]):
            launches[k % n]()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record(streams[0])
    for st in streams[1:]:
        st.wait_event(start)
    for k in range(steps):
        launches[k % n]()
    for st in streams[1:]:
        ev = torch.cuda.Event()
        ev.record(st)
        streams[0].wait_event(ev)
    end.record(streams[0])
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1000.0


def time_steps(torch, launch, steps, warmup):
    """Seconds for `steps` launches, HIP events on the current (launch) stream."""
    for _ in range(warmup):
        launch()
    torch.cuda.synchronize()
    start = torch.cuda.Event(enable_timing=True)
    end = torch.cuda.Event(enable_timing=True)
    start.record()
    for _ in range(steps):
        launch()
    end.record()
    torch.cuda.synchronize()
    return start.elapsed_time(end) / 1000.0


def build_workload(name, frames, approx=False):
    """(graph, input stacks, output shapes, algorithmic bytes per frame, dtype).
    approx: allow_approximate_gamma (the reference's lookup-table gamma)."""
    import zimg_amd as z
    from zimg_amd import workloads as W
    if name in ("fp32_down", "fp32_up"):
        sw, sh, dw, dh, filt = (3840, 2160, 1920, 1080, z.RESIZE_LANCZOS) if name == "fp32_down" else (1280, 720, 3840, 2160, z.RESIZE_BICUBIC)
        s, d, p = W.fp32_formats(sw, sh, dw, dh, filt, 3.0 if filt == z.RESIZE_LANCZOS else float("nan"))
        g = z.FilterGraph(s, d, p)
        base = W.fp32_plane(sw, sh, seed=8)
        planes = [np.stack([np.roll(base, 17 * k, axis=1) for k in range(frames)]), None, None, None]
        return g, planes, [(dh, dw, np.float32), None, None, None], (sw * sh + dw * dh) * 4, "f32"
    if name == "uhd_pq_to_hlg":
        s, d, p = W.uhd_pq_to_hlg_formats()
        g = z.FilterGraph(s, d, p)
        R, G, B = W.pq_rgb(3840, 2160, seed=9)
        planes = [np.stack([np.roll(c, 29 * k, axis=1) for k in range(frames)]) for c in (R, G, B)] + [None]
        return g, planes, [(2160, 3840, np.uint16)] * 3 + [None], 3840 * 2160 * 3 * (4 + 2), "f32 in, u16 (12-bit) out"
    if name == "fhd_709_to_2020":
        s, d, p = W.fhd_709_to_2020_formats()
        if approx:
            p.allow_approximate_gamma = b"\x01"
        g = z.FilterGraph(s, d, p)
        Ys, Us, Vs = [], [], []
        for k in range(frames):
            Y, U, V = W.yuv420_10bit(1920, 1080, seed=1000 + k)
            Ys.append(Y), Us.append(U), Vs.append(V)
        planes = [np.stack(Ys), np.stack(Us), np.stack(Vs), None]
        outs = [(1080, 1920, np.uint16), (540, 960, np.uint16), (540, 960, np.uint16), None]
        return g, planes, outs, 2 * (1920 * 1080 * 2 + 2 * 960 * 540 * 2), "u16/f32 (10-bit in/out, FP32 intermediates)"
    s, d, p = W.uhd_lanczos_formats()
    g = z.FilterGraph(s, d, p)
    base = W.word_plane(1920, 1080, seed=7)
    planes = [np.stack([np.roll(base, 13 * k, axis=1) for k in range(frames)]), None, None, None]
    return g, planes, [(2160, 3840, np.uint16), None, None, None], 1920 * 1080 * 2 + 3840 * 2160 * 2, "u16 (Q1.14 integer taps)"


def counters_of(key):
    """Key: value lines of the committed rocprofv3 counter summary profiles/<key>_counters.txt."""
    path = os.path.join(ROOT, "profiles", f"{key}_counters.txt")
    out = {}
    if not os.path.exists(path):
        return out
    for line in open(path):
        if ":" in line and not line.startswith((" ", "#")):
            k, v = line.split(":", 1)
            try:
                out[k.strip()] = float(v)
            except ValueError:
                out[k.strip()] = v.strip()
    return out


def train_chain_surrogates():
    """The PQ EOTF and HLG OETF surrogates of uhd_pq_to_hlg, validated per bin
    against the tolerances the 12-bit 0.5 LSB budget gives them (trained
    before, and outside, the timed region)."""
    import zimg_amd as z
    t0 = time.perf_counter()
    tol_pq, tol_hlg = z.pq_to_hlg_surrogate_tolerances(bits=12, lsb=0.5)
    pq = z.TransferSurrogate(z.TRANSFER_ST2084, True, tol_pq, True, steps=20000, seed=7)
    hlg = z.TransferSurrogate(z.TRANSFER_ARIB_B67, False, tol_hlg, False, steps=20000, seed=7)
    info = {"train_s": round(time.perf_counter() - t0, 2), "tol_pq_rel": tol_pq, "tol_hlg_abs": tol_hlg,
            "pq_coverage": round(pq.info()["coverage"], 4), "hlg_coverage": round(hlg.info()["coverage"], 4)}
    return [pq, hlg], info


def measure(torch, name, frames, steps, warmup, precision=0, approx=False, surrogates=None, streams=1):
    g, planes, outs, algo_bytes, dtype = build_workload(name, frames, approx)
    if approx:
        dtype += " (approximate gamma: lookup tables)"
    if surrogates:
        g.set_transfer_surrogates(surrogates[0])
        dtype += " (PQ EOTF and HLG OETF on the learned surrogate, exact fallback)"
    if precision:
        g.set_resize_precision(precision)
        dtype += f" (resize on MFMA, {'fp16' if precision == 1 else 'bf16'} hi+lo, FP32 accumulate)"
    launch, keep = device_batch(torch, g, planes, outs, frames)
    # Tile shapes of the tiled resize steps timed on these buffers before the
    # timed region (zimg_filter_graph_autotune_device); outputs are the same.
    tuned = launch.autotune()
    elapsed = time_steps(torch, launch, steps, warmup)
    single = elapsed
    if streams > 1:
        # The same input frames, each stream its own outputs and tmp.
        sts = pipe_streams(torch, streams)
        launches = [device_batch(torch, g, planes, outs, frames, stream=st, inputs=launch.inputs) for st in sts]
        elapsed = time_steps_pipelined(torch, [l for l, _ in launches], sts, steps, warmup)
        del launches
    per_step = elapsed / steps
    stages = g.device_stages()
    acc = None
    for _ in range(steps):
        st = launch.profiled()
        acc = [a + b[0] for a, b in zip(acc, st)] if acc else [b[0] for b in st]
    stage_ms = [ms / steps for ms in acc]
    k = int(np.argmax(stage_ms))
    key = name if not precision else f"{name}_{'fp16' if precision == 1 else 'bf16'}"
    if approx:
        key += "_approx"
    if surrogates:
        key += "_learned"
    ctr = counters_of(key)
    dom_ms = stage_ms[k]
    achieved = algo_bytes * frames / (dom_ms * 1e-3) / 1e9
    whole = algo_bytes * frames / per_step / 1e9
    roof = {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(achieved / PEAK_HBM_GBS, 4),
            "traffic": ctr.get("traffic_bytes_per_launch"), "kernel": stages[k] if k < len(stages) else str(k),
            "kernel_ms_per_launch": round(dom_ms, 4), "algorithmic_bytes_per_launch": algo_bytes * frames,
            "whole_graph": {"achieved": round(whole, 1), "frac": round(whole / PEAK_HBM_GBS, 4)},
            "stages": [{"ms": round(ms, 4), "stage": (stages[i] if i < len(stages) else str(i))[:72]} for i, ms in enumerate(stage_ms)]}
    if "binding" in ctr:
        roof["binding"] = {kk: ctr[kk] for kk in ("binding", "valu_insts_per_px", "valu_issue_frac", "mfma_busy_frac") if kk in ctr}
    del keep
    torch.cuda.empty_cache()
    r = {"fps": round(frames * steps / elapsed, 1), "ms_per_step": round(per_step * 1e3, 4), "frames": frames, "launches": len(stages),
         "dtype": dtype, "roofline": roof, "autotuned_steps": tuned}
    if streams > 1:
        r["streams"] = streams
        r["one_stream_fps"] = round(frames * steps / single, 1)
    if surrogates:
        r["surrogates"] = surrogates[1]
    return r


def measure_batch(torch, steps, warmup):
    """64 uneven jobs: one persistent launch (autotuned, replayed from a HIP graph)
    against each job's own launches."""
    import zimg_amd as z
    from zimg_amd import workloads as W
    jobs = W.batch_jobs(64)
    stream = torch.cuda.current_stream().cuda_stream
    cache, keep, graphs, srcs, dsts = {}, [], [], [], []
    algo = 0
    for j in jobs:
        key = (tuple(j["src"]), j["bits"])
        if key not in cache:
            cache[key] = W.batch_job_planes(j, seed=900 + len(cache))
        pl = cache[key]
        graphs.append(z.FilterGraph(*W.batch_job_graph(j)))
        sb, db = [], []
        (dw, dh) = j["dst"]
        for p, a in enumerate(pl):
            t = torch.from_numpy(np.ascontiguousarray(a).view(np.uint8).reshape(-1)).to("cuda")
            keep.append(t)
            sb.append((t.data_ptr(), a.shape[1] * a.itemsize))
            w, h = (dw, dh) if p == 0 else (dw // 2, dh // 2)
            o = torch.empty((h * w,), dtype=torch.uint8, device="cuda")
            keep.append(o)
            db.append((o.data_ptr(), w))
        srcs.append(z.image_buffer(sb + [None]))
        dsts.append(z.image_buffer(db + [None]))
        sw, sh = j["src"]
        algo += int(1.5 * (sw * sh * (1 if j["bits"] == 8 else 2) + dw * dh))

    def per_job():
        for g, s_, d_ in zip(graphs, srcs, dsts):
            g.process_device(s_, d_, None, stream)

    t_job = time_steps(torch, per_job, steps, warmup)
    jb = z.DeviceJobs(graphs, srcs, dsts, None, z.DEVICE_JOBS_AUTOTUNE, stream)
    t_pers = time_steps(torch, lambda: jb.run(stream), steps, warmup)
    ctr = counters_of("batch64")
    achieved = algo * steps / t_pers / 1e9
    per_job_launches = sum(len(g.device_stages()) for g in graphs)
    return {"fps": round(64 * steps / t_pers, 1), "unit": "jobs/s", "ms_per_step": round(t_pers / steps * 1e3, 4), "frames": 64,
            "launches": 1, "dtype": "u8/u16 in, WORD Q1.14 taps, u8 out (Bayer dither)",
            "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                         "frac": round(achieved / PEAK_HBM_GBS, 4), "traffic": ctr.get("traffic_bytes_per_launch"),
                         "kernel": jb.describe().strip(), "kernel_ms_per_launch": round(t_pers / steps * 1e3, 4),
                         "algorithmic_bytes_per_launch": algo},
            "per_job_launches": {"fps": round(64 * steps / t_job, 1), "achieved_gbs": round(algo * steps / t_job / 1e9, 1),
                                 "launches_per_step": per_job_launches},
            "speedup_vs_per_job": round(t_job / t_pers, 3)}


# ---- CPU baseline --------------------------------------------------------------

def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def usable_cores():
    try:
        n = len(os.sched_getaffinity(0))
    except AttributeError:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))


def cpu_frame_fn(name):
    """One whole frame of the workload through the oracle (tests/pipelines.py)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines as P
    from zimg_amd import workloads as W
    if name == "fhd_709_to_2020":
        Y, U, V = W.yuv420_10bit(1920, 1080, seed=99)
        return lambda: P.yuv420_10_709_to_2020(Y, U, V)
    if name == "uhd_lanczos":
        src = W.word_plane(1920, 1080, seed=99)
        return lambda: P.word_resize(src, 3840, 2160)
    if name in ("fp32_down", "fp32_up"):
        import zimg_amd as z
        sw, sh, dw, dh, ft, a, b = (3840, 2160, 1920, 1080, z.RESIZE_LANCZOS, 3.0, 0.0) if name == "fp32_down" else \
            (1280, 720, 3840, 2160, z.RESIZE_BICUBIC, 0.0, 0.5)
        src = W.fp32_plane(sw, sh, seed=99)
        return lambda: P.float_resize(src, dw, dh, ft, a, b)
    if name == "uhd_pq_to_hlg":
        R, G, B = W.pq_rgb(3840, 2160, seed=99)
        return lambda: P.rgb_pq_to_hlg_12bit_ed(R, G, B)
    return None


def cpu_baseline(name, runs, full):
    """Oracle on this host: median seconds per frame of `runs` single-thread runs;
    with `full`, also frame-parallel runs over the usable cores (median of `runs`)."""
    fn = cpu_frame_fn(name)
    if fn is None:
        return None
    fn()  # warm-up
    n1 = runs if full else max(3, runs // 3)
    single = []
    for _ in range(n1):
        t0 = time.perf_counter()
        fn()
        single.append(time.perf_counter() - t0)
    med1 = statistics.median(single)
    out = {"value": round(1.0 / med1, 2), "unit": "frames/s", "cores": 1, "kind": "port",
           "sample": f"{name}: median of {n1} single-thread whole frames through the oracle (tests/pipelines.py)",
           "cpu": cpu_model()}
    if full:
        from concurrent.futures import ThreadPoolExecutor
        n = usable_cores()
        par = []
        with ThreadPoolExecutor(max_workers=n) as ex:
            list(ex.map(lambda _: fn(), range(n)))
            for _ in range(runs):
                t0 = time.perf_counter()
                list(ex.map(lambda _: fn(), range(n)))
                par.append(time.perf_counter() - t0)
        medn = statistics.median(par)
        out["frame_parallel"] = {"value": round(n / medn, 2), "cores": n,
                                 "sample": f"median of {runs} runs of {n} frames, one frame per thread"}
    return out


def cpu_batch_baseline(runs):
    """The 64-job list through the oracle chain (tests/pipelines.py batch_job)
    on this host: jobs/s from the median of single-thread passes over the
    whole list, and job-parallel over the usable cores (median of `runs`)."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import pipelines as P
    from zimg_amd import workloads as W
    jobs = W.batch_jobs(64)
    cache, planes = {}, []
    for j in jobs:
        key = (tuple(j["src"]), j["bits"])
        if key not in cache:
            cache[key] = W.batch_job_planes(j, seed=900 + len(cache))
        planes.append(cache[key])

    def one_pass():
        for j, pl in zip(jobs, planes):
            P.batch_job(j, pl)
    one_pass()  # warm-up
    n1 = max(3, runs // 3)
    single = []
    for _ in range(n1):
        t0 = time.perf_counter()
        one_pass()
        single.append(time.perf_counter() - t0)
    med1 = statistics.median(single)
    out = {"value": round(len(jobs) / med1, 2), "unit": "jobs/s", "cores": 1, "kind": "port",
           "sample": f"batch64: median of {n1} single-thread passes over the 64-job list through the oracle (tests/pipelines.py)",
           "cpu": cpu_model()}
    from concurrent.futures import ThreadPoolExecutor
    n = usable_cores()
    par = []
    with ThreadPoolExecutor(max_workers=n) as ex:
        list(ex.map(lambda a: P.batch_job(*a), zip(jobs, planes)))
        for _ in range(runs):
            t0 = time.perf_counter()
            list(ex.map(lambda a: P.batch_job(*a), zip(jobs, planes)))
            par.append(time.perf_counter() - t0)
    medn = statistics.median(par)
    out["job_parallel"] = {"value": round(len(jobs) / medn, 2), "cores": n,
                           "sample": f"median of {runs} runs of the 64 jobs, one job per thread at a time"}
    return out


def variant_applies(name, variant):
    """Whether a workload has the form `variant` (the exact path always)."""
    return variant in ("all", "exact") or (variant in ("fp16", "bf16") and name in ("fp32_down", "fp32_up")) or \
        (variant == "approx" and name == "fhd_709_to_2020") or (variant == "learned" and name == "uhd_pq_to_hlg")


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; multi-GPU sharding is design-only (frames are independent)",
                          "n_gpus": args.gpus}))
        sys.exit(2)
    import torch
    if not torch.cuda.is_available():
        print(json.dumps({"error": "no GPU visible"}))
        sys.exit(1)
    frames = args.frames or 64

    def nstreams(name):
        return args.streams or DEFAULT_STREAMS.get(name, 1)

    todo = [args.workload] if args.workload else WORKLOADS
    results = {}
    only = args.variant if args.workload else "all"
    if only != "all" and not variant_applies(args.workload, only):
        print(json.dumps({"error": f"workload {args.workload} has no {only} form"}))
        sys.exit(2)
    for name in todo:
        if only != "all" and only != "exact":
            kw = {"fp16": {"precision": 1}, "bf16": {"precision": 2}, "approx": {"approx": True}}.get(only, {})
            if only == "learned":
                kw = {"surrogates": train_chain_surrogates()}
            results[name] = measure(torch, name, frames, args.steps, args.warmup, streams=nstreams(name), **kw)
        elif name == "batch64":
            results[name] = measure_batch(torch, args.steps, args.warmup)
        else:
            results[name] = measure(torch, name, frames, args.steps, args.warmup, streams=nstreams(name))
        if only == "all" and name != "batch64":
            if name in ("fp32_down", "fp32_up"):
                results[name]["mfma_fp16"] = measure(torch, name, frames, args.steps, args.warmup, precision=1)
            if name == "fhd_709_to_2020":
                results[name]["approximate_gamma"] = measure(torch, name, frames, args.steps, args.warmup, approx=True)
            if name == "uhd_pq_to_hlg":
                learned = measure(torch, name, frames, args.steps, args.warmup, surrogates=train_chain_surrogates(), streams=nstreams(name))
                learned["gain_vs_exact"] = round(learned["fps"] / results[name]["fps"], 3)
                results[name]["learned"] = learned
        if not args.no_cpu:
            results[name]["cpu_baseline"] = cpu_batch_baseline(args.cpu_runs) if name == "batch64" else \
                cpu_baseline(name, args.cpu_runs, full=name in ("fhd_709_to_2020", "uhd_lanczos"))
        print(f"# {name}: {results[name]['fps']} {results[name].get('unit', 'frames/s')}", file=sys.stderr, flush=True)
    primary = todo[0]
    m = results[primary]
    line = {
        "metric": METRIC,
        "value": m["fps"],
        "unit": m.get("unit", "frames/s"),
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": m["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": m["dtype"],
        "data": "synthetic (seeded, zimg_amd/workloads.py), inputs resident in HBM",
        "config": {"workload": primary, "frames_per_step": m["frames"], "graph_launches": m["launches"],
                   "form": "exact" if only == "all" else only, "streams": m.get("streams", 1)},
        "roofline": m["roofline"],
        "cpu_baseline": m.get("cpu_baseline"),
        "one_stream_fps": m.get("one_stream_fps"),
        "variants": {k: m[k] for k in ("approximate_gamma", "mfma_fp16", "learned") if k in m},
        "configs": {k: v for k, v in results.items() if k != primary},
    }
    print(json.dumps(line))


if __name__ == "__main__":
    main()
