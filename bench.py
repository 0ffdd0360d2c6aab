#!/usr/bin/env python3
"""Benchmark of the render chain on one MI355X.

Workload (the single-frame chain): one 6000 x 4000 float32 linear ProPhoto frame of the
synthetic generator (seed 0) on the synthetic 3-layer stock; Philox Poisson grain (film format
35 mm); halation = one mixture of 3 Gaussians sigma 2/8/32 px, weights 0.6/0.3/0.1, scatter off;
FIR for sigma <= 16 (crossover 16.5), recursive above; print, scan, CAM16 gamut, unsharp, RGBA16.
A step is one spk_render of the frame (develop + print + encode) with the input already resident
in HBM and the output written to HBM (spkx_render_device); K steps are timed between two device
synchronisations.

The JSON line reports Mpx/s, the live HIP-event duration and algorithmic-byte roofline of the
dominant kernel, the same kernel against the vector-ALU issue rate that binds it ("valu_roofline":
its VALU instructions per wavefront from the hardware counters in profiles/, the FMA issue rate
measured live by spkx_valu_roofs), and a CPU baseline: the oracle (oracle/chain.cpp, binary32) timed on a bounded
sample (a 1536 x 1024 crop of the same workload) on this host's cores.  It also carries the
batched workload ("batched": 64 frames of mixed sizes and per-frame parameters,
one hipGraph launch, against 64 sequential spk_render and spkx_render_device calls).
`--study precision` runs a reprint sweep instead (one developed negative reprinted, FP16 vs FP32 spectral
integrals against the float64 oracle).
"""
from __future__ import annotations

import argparse
import json
import subprocess
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes per pixel of each stage: a pointwise run reads its input planes once and
# writes its output once; each separable pass of a Gaussian reads one 3-channel plane and writes
# one, plus one read for an accumulating pass (12 B per 3-channel float plane, 8 B per RGBA16 px).
# The whole chain of the workload: upsample 24 + halation mixture of 3 components 168 + develop 36
# + grain blur 48 + print 24 + unsharp blur 48 + unsharp combine and encode 32 = 380 B/px.  Stages
# the engine runs as one kernel keep the bytes of the stages they fuse (unsharp_encode = 48 + 32).
B_ALG = 380
STAGE_BYTES = {"upsample": 24, "develop": 36, "grain_blur": 48, "print": 24, "unsharp_blur": 48, "encode": 32,
               "unsharp_encode": 80, "halation": 168}
STAGE_KERNEL = {"upsample": "k_upsample", "develop": "k_develop", "print": "k_print", "encode": "k_unsharp_encode",
                "unsharp_encode": "k_unsharp_encode2d", "halation": "k_fir2d/k_fir_v/k_fir_h/k_iir_row/k_iir_col",
                "grain_blur": "k_fir2d", "unsharp_blur": "k_fir2d"}
METRIC = "Mpx/s full chain @24MP (frac of HBM peak); batched frames/s; fwd+bwd Mpx/s"


# MI355X vector-ALU issue ceiling: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD per 4 cycles at 2.4 GHz.
VALU_PEAK = 256 * 4 * 2.4e9 / 4

def log(msg):
    print(msg, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--width", type=int, default=6000)
    ap.add_argument("--height", type=int, default=4000)
    ap.add_argument("--policy", choices=["cfg", "ref"], default="cfg",
                    help="cfg: FIR for sigma <= 16; ref: the reference's FIR for sigma < 3")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--cpu-sample", type=str, default="1536x1024")
    ap.add_argument("--spectral", choices=["fp32", "fp16"], default="fp32",
                    help="precision of the print's 81-bin spectral integrals (fp16: matrix cores, FP32 accumulation)")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="first run one plain 1 GiB device-to-device copy (a reference dispatch of known "
                         "traffic for hardware byte counters)")
    ap.add_argument("--study", choices=["precision", "diff"], default=None,
                    help="precision: one developed negative, a reprint sweep of print_exposure 0.5..2.0 in FP32 and "
                         "FP16, each against the float64 oracle (16-bit LSB and CAM16-UCS dE); diff: only the "
                         "differentiable-render measurement (8 frames of 2048x1365), one JSON line")
    ap.add_argument("--sweep", type=int, default=32, help="reprints in the precision study")
    ap.add_argument("--no-batch", action="store_true", help="skip the batched-rendering measurement")
    ap.add_argument("--batch-frames", type=int, default=64)
    ap.add_argument("--no-diff", action="store_true", help="skip the differentiable-render measurement")
    ap.add_argument("--diff-frames", type=int, default=8)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the render chain runs on a single "
              "MI355X and multi-GPU scaling is not part of this build", file=sys.stderr)
        sys.exit(2)
    if args.study == "precision":
        precision_study(args)
        return
    if args.study == "diff":
        import spektralab_engine_amd as spk
        from spektralab_engine_amd import synth
        eng = spk.Engine(synth.default_resources_dir(192))
        print(json.dumps(diff_measure(eng, None, args)), flush=True)
        eng.close()
        return

    import torch
    import spektralab_engine_amd as spk
    from spektralab_engine_amd import synth

    crossover = 16.5 if args.policy == "cfg" else 3.0
    delta = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                 dir_couplers_active=False, glare_active=False, grain_active=True, halation_active=True,
                 film_format_mm=35.0)
    overrides = {"halation_mixture": [[2.0, 0.6], [8.0, 0.3], [32.0, 0.1]], "halation_scatter": False,
                 "fir_crossover": crossover, "spectral_precision": args.spectral}
    res = synth.default_resources_dir(192)
    W, H = args.width, args.height
    log(f"[bench] generating the {W}x{H} frame")
    frame = synth.make_frame(W, H, 0)
    eng = spk.Engine(res)
    sess = eng.open(frame, delta)
    sess.set_overrides(overrides)
    out = torch.empty((H, W, 4), dtype=torch.int16, device="cuda")
    if args.calibrate_copy:
        src = torch.ones(1 << 28, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log("[bench] calibration copy: 1 GiB read, 1 GiB written")
        del src, dst
    log("[bench] warm-up")
    for _ in range(args.warmup):
        sess.render_device(out.data_ptr(), W)
    torch.cuda.synchronize()
    sess.set_profiling(True)
    log(f"[bench] timing {args.steps} steps")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        sess.render_device(out.data_ptr(), W)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    stages = sess.stage_times(reset=True)
    sess.set_profiling(False)
    ms = (t1 - t0) * 1000.0 / args.steps
    px = W * H
    mpxs = px / (ms / 1000.0) / 1e6
    b_alg = B_ALG
    chain_gbs = b_alg * px / (ms / 1000.0) / 1e9
    per_stage = {k: v["ms"] / v["count"] for k, v in stages.items()}
    # the dominant single-kernel stage
    dom = max((k for k in per_stage if k in STAGE_BYTES and "/" not in STAGE_KERNEL[k]), key=lambda k: per_stage[k])
    dom_ms = per_stage[dom]
    dom_bytes = STAGE_BYTES[dom] * px
    achieved = dom_bytes / (dom_ms / 1000.0) / 1e9
    traffic = None
    tf = ROOT / "profiles" / "traffic_chain_24mp.json"
    if tf.exists() and (W, H) == (6000, 4000):
        entry = json.loads(tf.read_text()).get("kernels", {}).get(STAGE_KERNEL[dom])
        if entry:
            traffic = round(entry["hbm_bytes_per_launch"])
    # The compute-bound kernels' roof: vector-ALU issue.  Peak: one wave64 VALU instruction per
    # SIMD every 4 cycles (16 lanes a cycle) at MI355X's peak engine clock, 256 CUs x 4 SIMDs x
    # 2.4 GHz / 4 = 614.4 G wave-instructions/s; the probes' measured rates (spkx_valu_roofs: FMA
    # about 90 % of it, its loop's branch and waits included; exp2 at half of it) are reported
    # beside it.  Achieved: the dominant kernel's VALU instructions per wave from the hardware
    # counters (rocprofv3 SQ_INSTS_VALU / SQ_WAVES of the same workload, in the traffic profile)
    # over the live launch time.
    roofs = eng.valu_roofs()
    valu = None
    if tf.exists() and (W, H) == (6000, 4000):
        entry = json.loads(tf.read_text()).get("kernels", {}).get(STAGE_KERNEL[dom])
        if entry and entry.get("waves_per_launch"):
            peak = VALU_PEAK
            got = entry["valu_insts_per_wave"] * entry["waves_per_launch"] / (dom_ms / 1000.0)
            valu = {"bound": "valu", "kernel": STAGE_KERNEL[dom], "achieved": round(got / 1e9, 1),
                    "peak": round(peak / 1e9, 1), "unit": "G wave-instructions/s", "frac": round(got / peak, 4),
                    "valu_insts_per_wave": round(entry["valu_insts_per_wave"], 1),
                    "waves_per_launch": entry["waves_per_launch"],
                    "peak_source": "one wave64 VALU instruction per SIMD per 4 cycles, 1024 SIMDs at 2.4 GHz",
                    "fma_probe": round(roofs["rates"]["v_fma_f32"]["wave_instructions_per_s"] / 1e9, 1)}
    log(f"[bench] {ms:.3f} ms/frame, {mpxs:.1f} Mpx/s; stages (ms): " +
        ", ".join(f"{k} {v:.3f}" for k, v in per_stage.items()))

    # the reference ABI's own call: spk_render into a pinned host result (the device-to-host copy of
    # the RGBA16 image over PCIe included); reported beside the device-resident value
    import ctypes
    lib = eng.lib
    r = spk.Result()
    lib.spk_render(sess.h, b"full", ctypes.byref(r))
    lib.spk_result_free(ctypes.byref(r))
    a0 = time.perf_counter()
    for _ in range(3):
        r = spk.Result()
        st = lib.spk_render(sess.h, b"full", ctypes.byref(r))
        if st != 0:
            raise spk.SpkError(st, eng.error())
        lib.spk_result_free(ctypes.byref(r))
    abi_ms = (time.perf_counter() - a0) * 1000.0 / 3
    log(f"[bench] spk_render (host result): {abi_ms:.3f} ms/frame")

    batched = None
    if not args.no_batch:
        sess.close()
        batched = batch_measure(eng, res, args)

    differentiable = None
    if not args.no_diff:
        differentiable = diff_measure(eng, res, args)

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(res, delta, overrides, frame, args)

    line = {
        "metric": METRIC,
        "value": round(mpxs, 2),
        "unit": "Mpx/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"plain chain {W}x{H}, Poisson grain + 3-Gaussian halation (2/8/32 px) "
                               f"+ print/scan + CAM16 + unsharp -> RGBA16, policy {args.policy}",
                   "width": W, "height": H, "fir_crossover": crossover, "film_format_mm": 35.0,
                   "stock": "synthetic 3-layer negative / synthetic paper"},
        "chain_hbm_frac": round(chain_gbs / HBM_PEAK_GBS, 4),
        "chain_algorithmic_bytes_per_px": b_alg,
        "stage_ms": {k: round(v, 4) for k, v in per_stage.items()},
        "roofline": {"bound": "hbm", "kernel": STAGE_KERNEL[dom], "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "launch_ms": round(dom_ms, 4), "bytes_per_px": STAGE_BYTES[dom]},
        "valu_roofline": valu,
        "valu_roofs": {k: {"per_cu_per_clock": round(v["per_cu_per_clock"], 3),
                           "g_wave_instructions_per_s": round(v["wave_instructions_per_s"] / 1e9, 1)}
                       for k, v in roofs["rates"].items()},
        "spk_render_host_result": {"ms_per_frame": round(abi_ms, 3), "mpx_s": round(px / (abi_ms / 1000.0) / 1e6, 1),
                                   "note": "PCIe-inclusive: RGBA16 copied into pinned host memory; not the value"},
        "cpu_baseline": cpu,
        "batched": batched,
        "differentiable": differentiable,
    }
    print(json.dumps(line), flush=True)
    sess.close()
    eng.close()


def host_cpu():
    """The host's CPU as lscpu and the affinity mask describe it: model, sockets, cores per socket,
    threads per core, the CPUs this process may use and the physical cores among them."""
    info = {"nproc": len(os.sched_getaffinity(0))}
    try:
        out = subprocess.run(["lscpu"], capture_output=True, text=True, timeout=10).stdout
        keys = {"Model name": "model", "Socket(s)": "sockets", "Core(s) per socket": "cores_per_socket",
                "Thread(s) per core": "threads_per_core"}
        for line in out.splitlines():
            k, _, v = line.partition(":")
            if k.strip() in keys:
                v = v.strip()
                info[keys[k.strip()]] = int(v) if v.isdigit() else v
    except (OSError, subprocess.SubprocessError):
        pass
    info["physical_cores"] = max(1, info["nproc"] // int(info.get("threads_per_core", 1) or 1))
    return info


def cpu_baseline(res, delta, overrides, frame, args):
    """The CPU oracle (the same chain in binary32 and in binary64) on a crop of the benchmark frame,
    one OpenMP thread per physical core: one warm-up render, then the median of five."""
    import spektralab_engine_amd as spk
    from oracle import oracle as orc
    host = host_cpu()
    orc.lib().oracle_set_threads(host["physical_cores"])
    threads = orc.lib().oracle_threads()
    cw, ch = (int(v) for v in args.cpu_sample.split("x"))
    plan = spk.host_plan(res, delta, cw, ch, film_seed=0x4ABEA221, overrides=overrides)
    sample = np.ascontiguousarray(frame[:ch, :cw])
    legs = {}
    for name, f64 in (("f32", False), ("f64", True)):
        log(f"[bench] CPU baseline {name}: oracle on a {cw}x{ch} sample, {threads} threads")
        orc.render(plan, sample, f64=f64)  # warm-up (page-in, thread pool)
        times = []
        for _ in range(5):
            c0 = time.perf_counter()
            orc.render(plan, sample, f64=f64)
            times.append((time.perf_counter() - c0) * 1000.0)
        med = float(np.median(times))
        legs[name] = {"mpx_s": round(cw * ch / (med / 1000.0) / 1e6, 3), "median_ms": round(med, 1),
                      "min_ms": round(min(times), 1), "max_ms": round(max(times), 1)}
    return {"value": legs["f32"]["mpx_s"], "unit": "Mpx/s", "cores": threads, "kind": "port",
            "sample": f"{cw}x{ch} crop of the same frame and stock, same chain; binary32 oracle (value) and "
                      f"binary64 oracle; median of 5 renders after 1 warm-up; {threads} OpenMP threads, "
                      f"one per physical core",
            "legs": legs, "host": host}


def batch_measure(eng, res, args):
    """The batched workload: `batch_frames` frames of log-uniform area 256x256 ..
    4096x2730 with mixed aspect ratios and per-frame print exposure, filter shifts and halation
    sigmas (synth.batch_workload), each in its own open session.  Times three ways of rendering
    every frame once, inputs resident on the device:
      - sequential spk_render (the reference ABI: host-readable RGBA16 per frame, one call each);
      - sequential spkx_render_device (device outputs, one call each);
      - one batch run (spkx_batch_run: one hipGraph launch; autotuned at creation, untimed).
    Returns frames/s and Mpx/s of each and the batch's speedups."""
    import ctypes
    import torch
    import spektralab_engine_amd as spk
    from spektralab_engine_amd import synth

    wl = synth.batch_workload(args.batch_frames, seed=0)
    base = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                dir_couplers_active=False, glare_active=False, grain_active=True, halation_active=True,
                film_format_mm=35.0)
    log(f"[bench] batch workload: {len(wl)} frames, {sum(f['width'] * f['height'] for f in wl) / 1e6:.1f} MP")
    sessions = []
    for f in wl:
        d = dict(base)
        d.update(f["params"])
        s = eng.open(synth.make_frame(f["width"], f["height"], f["frame_seed"]), d)
        s.set_overrides({"halation_mixture": f["mixture"], "halation_scatter": False, "fir_crossover": 16.5})
        sessions.append(s)
    px = sum(f["width"] * f["height"] for f in wl)
    outs = [torch.empty((f["height"], f["width"], 4), dtype=torch.int16, device="cuda") for f in wl]
    lib = eng.lib

    def seq_render():
        for s in sessions:
            r = spk.Result()
            st = lib.spk_render(s.h, b"full", ctypes.byref(r))
            if st != 0:
                raise spk.SpkError(st, eng.error())
            lib.spk_result_free(ctypes.byref(r))

    def seq_device():
        for s, o, f in zip(sessions, outs, wl):
            s.render_device(o.data_ptr(), f["width"])
        torch.cuda.synchronize()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / reps

    t_render = timed(seq_render, 2)
    t_device = timed(seq_device, 3)
    t0 = time.perf_counter()
    b = spk.Batch(eng, sessions, [o.data_ptr() for o in outs], [f["width"] for f in wl])
    t_create = time.perf_counter() - t0

    def batch_run():
        b.run()
        b.sync()

    t_batch = timed(batch_run, 5)
    info = b.info()
    b.close()
    for s in sessions:
        s.close()
    log(f"[bench] batch: spk_render x{len(wl)} {t_render:.1f} ms, render_device x{len(wl)} {t_device:.1f} ms, "
        f"batch {t_batch:.1f} ms (create + tune {t_create:.2f} s)")
    n = len(wl)
    return {"workload": f"batch of {n} frames, {px / 1e6:.1f} MP total, log-uniform area 256x256..4096x2730, "
                        "per-frame print_exposure/filters/halation sigmas",
            "frames": n, "pixels": px,
            "frames_per_s": round(n / (t_batch / 1000.0), 1), "mpx_s": round(px / (t_batch / 1000.0) / 1e6, 1),
            "ms_per_batch": round(t_batch, 3),
            "sequential_spk_render": {"ms": round(t_render, 3), "frames_per_s": round(n / (t_render / 1000.0), 1)},
            "sequential_render_device": {"ms": round(t_device, 3),
                                         "frames_per_s": round(n / (t_device / 1000.0), 1)},
            "speedup_vs_spk_render": round(t_render / t_batch, 3),
            "speedup_vs_render_device": round(t_device / t_batch, 3),
            "tuned_config": info["config"], "create_and_tune_s": round(t_create, 2)}


def diff_measure(eng, res, args, W=2048, H=1365):
    """The differentiable-render workload: `diff_frames` frames of 2048x1365 through the differentiable
    render (grain frozen to each forward pass's counts, the chain workload's halation), an L2 loss
    against targets rendered with the 40 scalar quantities and the enlarger weights perturbed by
    N(0, 0.05) relative.  Times, inputs resident on the device: the differentiable forward pass of
    every frame; forward + loss gradient + reverse pass of every frame; the plain render
    (spkx_render_device) of the same frames.  Returns Mpx/s and the ratios."""
    import torch
    import spektralab_engine_amd as spk
    from spektralab_engine_amd import synth

    n = args.diff_frames
    delta = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                 dir_couplers_active=False, glare_active=False, grain_active=True, halation_active=True,
                 film_format_mm=35.0)
    sess = eng.open(synth.make_frame(W, H, 100), delta)
    sess.set_overrides({"halation_mixture": [[2.0, 0.6], [8.0, 0.3], [32.0, 0.1]], "halation_scatter": False,
                        "fir_crossover": 16.5})
    d = spk.Diff(sess, slots=n)
    xs = [torch.from_numpy(synth.make_frame(W, H, 100 + i)).cuda() for i in range(n)]
    rng = np.random.default_rng(0)
    th0 = d.defaults()
    th_t = th0 * (1.0 + 0.05 * rng.normal(size=th0.size))
    th_t[39] = 0.0
    targets = []
    for i in range(n):
        t = torch.empty_like(xs[i])
        d.forward(i, xs[i].data_ptr(), t.data_ptr(), th_t)
        targets.append(t)
    encs = [torch.empty_like(x) for x in xs]
    gxs = [torch.empty_like(x) for x in xs]
    gth = torch.empty((n, d.QUANTITIES), dtype=torch.float64, device="cuda")
    out = torch.empty((H, W, 4), dtype=torch.int16, device="cuda")
    npx = 3 * W * H

    def fwd():
        for i in range(n):
            d.forward(i, xs[i].data_ptr(), encs[i].data_ptr(), th0)

    def fwd_bwd():
        for i in range(n):
            d.forward(i, xs[i].data_ptr(), encs[i].data_ptr(), th0)
            g = (encs[i] - targets[i]) * (2.0 / npx)
            d.backward(i, xs[i].data_ptr(), g.data_ptr(), gxs[i].data_ptr(), gth[i].data_ptr())

    def plain():
        for i in range(n):
            sess.render_device(out.data_ptr(), W)
        torch.cuda.synchronize()

    def timed(fn, reps):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) * 1000.0 / reps

    t_f = timed(fwd, 3)
    t_fb = timed(fwd_bwd, 3)
    t_p = timed(plain, 3)
    info = d.info()
    px = n * W * H
    log(f"[bench] differentiable x{n} {W}x{H}: forward {t_f:.1f} ms, forward+backward {t_fb:.1f} ms, "
        f"plain render {t_p:.1f} ms")
    d.close()
    sess.close()
    return {"workload": f"differentiable render of {n} frames {W}x{H}, grain frozen, 3-Gaussian halation, L2 loss vs targets at "
                        "quantities perturbed N(0, 0.05) relative; gradients w.r.t. 39 scalars, 243 enlarger weights "
                        "and the input pixels",
            "frames": n, "pixels": px, "quantities": d.QUANTITIES,
            "fwd_bwd_mpx_s": round(px / (t_fb / 1000.0) / 1e6, 1), "fwd_mpx_s": round(px / (t_f / 1000.0) / 1e6, 1),
            "ms_forward": round(t_f, 3), "ms_forward_backward": round(t_fb, 3), "ms_plain_render": round(t_p, 3),
            "fwd_bwd_over_fwd": round(t_fb / t_f, 3), "fwd_bwd_over_plain_render": round(t_fb / t_p, 3),
            "checkpoint_bytes_per_frame": info["checkpoint_bytes_per_slot"]}


def precision_study(args):
    """The reprint study: one developed W x H negative of the chain workload, then
    `sweep` spk_reprint calls at print_exposure 0.5..2.0, each in FP32 and in FP16 spectral
    precision.  Every reprint is compared with the float64 oracle's print of the same negative on a
    central crop (the print is pointwise; the unsharp blur's 2-pixel reach is excluded at the crop
    border).  Prints one JSON line with the errors in 16-bit LSB and CAM16-UCS dE and the reprint
    rates."""
    import torch
    import spektralab_engine_amd as spk
    from spektralab_engine_amd import colour, synth
    from oracle import oracle as orc

    delta = dict(film_stock=synth.FILM_STOCK, print_stock=synth.PAPER_STOCK, auto_exposure=False,
                 dir_couplers_active=False, glare_active=False, grain_active=True, halation_active=True,
                 film_format_mm=35.0)
    overrides = {"halation_mixture": [[2.0, 0.6], [8.0, 0.3], [32.0, 0.1]], "halation_scatter": False,
                 "fir_crossover": 16.5}
    W, H = args.width, args.height
    cw, ch = min(2048, W), min(1536, H)
    x0, y0 = (W - cw) // 2, (H - ch) // 2
    m = 4  # interior margin of the crop
    res = synth.default_resources_dir(192)
    frame = synth.make_frame(W, H, 0)
    eng = spk.Engine(res)
    sess = eng.open(frame, delta)
    sess.set_overrides(overrides)
    out = torch.empty((H, W, 4), dtype=torch.int16, device="cuda")
    sess.render_device(out.data_ptr(), W)
    torch.cuda.synchronize()
    neg = sess.negative("full")[:, y0:y0 + ch, x0:x0 + cw].copy()
    modes = ("fp32", "fp16")
    hist = {k: np.zeros(65536, np.int64) for k in modes}
    de_edges = np.linspace(0.0, 2.0, 20001)
    de_hist = {k: np.zeros(len(de_edges) - 1, np.int64) for k in modes}
    de_max = {k: 0.0 for k in modes}
    times = {k: [] for k in modes}
    exposures = np.linspace(0.5, 2.0, args.sweep)
    for n_e, ev in enumerate(exposures):
        sess.set_params({"print_exposure": float(ev)})
        codes = {}
        for mode in modes:
            sess.set_overrides({"spectral_precision": mode})
            sess.render_device(out.data_ptr(), W, reprint=True)  # first reprint after a change uploads tables
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(3):
                sess.render_device(out.data_ptr(), W, reprint=True)
            torch.cuda.synchronize()
            times[mode].append((time.perf_counter() - t0) * 1000.0 / 3)
            codes[mode] = out[y0:y0 + ch, x0:x0 + cw].cpu().numpy().view(np.uint16)
        plan = bytearray(bytes(sess.last_plan()))
        plan[8:16] = np.array([cw, ch], np.uint32).tobytes()  # the print constants do not depend on the size
        ref = orc.print_((ctypes_buffer(plan)), neg, f64=True)[m:-m, m:-m]
        for mode in modes:
            got = codes[mode][m:-m, m:-m]
            d = np.abs(got[..., :3].astype(np.int64) - ref[..., :3].astype(np.int64)).ravel()
            hist[mode] += np.bincount(d, minlength=65536)[:65536]
            de = colour.delta_e_ucs(got, ref).ravel()
            de_max[mode] = max(de_max[mode], float(de.max()))
            de_hist[mode] += np.histogram(np.minimum(de, de_edges[-1] - 1e-12), bins=de_edges)[0]
        log(f"[precision] print_exposure {ev:.3f}: fp32 {times['fp32'][-1]:.3f} ms, fp16 {times['fp16'][-1]:.3f} ms "
            f"({n_e + 1}/{len(exposures)})")

    def pct(h, q, edges=None):
        c = np.cumsum(h)
        i = int(np.searchsorted(c, q * c[-1]))
        return float(edges[i + 1]) if edges is not None else float(i)

    summary = {}
    for mode in modes:
        h = hist[mode]
        n = h.sum()
        summary[mode] = {
            "lsb_max": int(np.nonzero(h)[0].max()), "lsb_p99": pct(h, 0.99), "lsb_p999": pct(h, 0.999),
            "lsb_mean": float((h * np.arange(65536)).sum() / n),
            "de_max": round(de_max[mode], 5), "de_p99": pct(de_hist[mode], 0.99, de_edges),
            "de_p999": pct(de_hist[mode], 0.999, de_edges),
            "reprint_ms": round(float(np.median(times[mode])), 4),
            "reprint_mpx_s": round(W * H / (np.median(times[mode]) / 1000.0) / 1e6, 1)}
    line = {"study": "precision", "config": {
        "workload": f"one developed {W}x{H} negative of the chain workload, {len(exposures)} reprints "
                    f"at print_exposure 0.5..2.0, FP32 vs FP16 (matrix cores, FP32 accumulation) spectral integrals",
        "reference": f"float64 oracle print of the same negative on the central {cw}x{ch} crop, {m}-px border "
                     "excluded", "pixels_compared": int(hist["fp32"].sum() // 3)},
        "fp32": summary["fp32"], "fp16": summary["fp16"]}
    print(json.dumps(line), flush=True)
    sess.close()
    eng.close()


def ctypes_buffer(b):
    import ctypes
    return ctypes.create_string_buffer(bytes(b), len(b))


if __name__ == "__main__":
    main()
