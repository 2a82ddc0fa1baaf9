#!/usr/bin/env python3
"""Throughput of the FFT3D denoise hot path on one MI355X.

Workload: 4K 10-bit luma (3840x2160), sigma=3, bt=4, bw=bh=32, ow=oh=16,
sharpen=0.4, degrid=1 (mode 1 default), FP32 spectra ring: the configuration
the 4K10 frames/s metric is quoted on. A step is one output frame of the
sequential filter (analysis of the newly needed frame into the ring, temporal
filter over the 4-frame window, inverse FFT, overlap-add, 10-bit pack), with
the input planes already resident in HBM and the output written to HBM. In
steady state one step is one fused launch (the new frame's analysis inside
the filter/synthesis sweep) plus the seam combine.

Prints one JSON line. `value` and `ms_per_step` come from K timed steps with
no events inside (HIP events on the launch stream before and after them).
Per-kernel device times come from a second pass of K steps with HIP events
around every kernel, on the launch stream; `roofline` prices the dominant
kernel's algorithmic bytes per launch (kernel_bytes: per-unit figures times
the units of one launch) against the 8 TB/s HBM peak. `config.forms` runs the
workload in each precision form side by side, each with its error against
the reference CPU build and the FP64 oracle. `cpu_baseline` times the
reference's CPU path (its own sources built into oracle/_ref when present,
else the oracle restatement) single-threaded, in steady state, on configs
A-D (value: the 4K workload).

    python bench.py [--steps K] [--warmup W] [--no-cpu] [--no-host] [--no-served]
                    [--no-variants] [--form NAME]

--form runs the main loop in another precision form of the same workload
(a name from FORMS; default the plain FP32 form, which `value` is quoted
on); the result line then names the form in `config.workload`.
"""
import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "4K10 frames/s (1 GPU), % HBM roofline; p50/p99 frame latency at 32 streams"
W, H, BITS = 3840, 2160, 10
PARAMS = dict(sigma=3.0, bt=4, bw=32, bh=32, ow=16, oh=16, sharpen=0.4)
HBM_PEAK_GBS = 8000.0
SEED = 1002
NUNIQ = 16  # distinct synthetic frames, cycled through the clip


def kernel_bytes(info, bps_in, bps_out, bin_bytes=8):
    """Algorithmic HBM bytes per launch of each kernel. A spectrum is
    bins * bin_bytes with bins = nb * (bw/2+1) * bh (8: complex FP32; 4: the
    FP16 ring)."""
    bins = info.nb * info.fbw * info.bh
    spec = bins * bin_bytes
    return {
        "analysis": W * H * bps_in + spec,          # read the plane, write one spectrum
        "filter_inverse": spec * info.bt,           # read the bt-frame window
        "synthesize": W * H * bps_out,              # write the packed plane
        # fused filter/inverse/overlap-add: read the window, write the plane
        # (the seam tiles' share of the plane is written by combine_seams)
        "filter_synth_sweep": spec * info.bt + W * H * bps_out,
        # the same with the new frame's analysis inside: read its plane and
        # bt-1 stored spectra, write its spectrum and the plane
        "forward_filter_synth_sweep": spec * info.bt + W * H * (bps_in + bps_out),
        "combine_seams": 0,
        "step": W * H * (bps_in + bps_out) + spec * info.bt,
    }


# Configurations of the CPU baseline (the reference's own CPU backend, or the
# oracle restatement when that build is absent): (width, height, kind, filter,
# colour, n1, n2). Each is timed over output frames 0..n1-1 and 0..n2-1 from a
# cold start; (n2 - n1) / (T2 - T1) is the steady-state rate, in which every
# new output frame costs one analysis and one filter/synthesis pass.
CPU_CONFIGS = {
    "A_640x360_u8_bt3": (640, 360, "u8", dict(sigma=2.0, beta=1.0, bt=3, precision=2),
                         "gray", 4, 16),
    "B_1080p10_yuv_plane4_bt4": (1920, 1080, "u10",
                                 dict(sigma=3.0, bt=4, sharpen=0.4, scutoff=0.3, svr=1.0,
                                      smin=4.0, smax=20.0, plane=4), "yuv", 3, 7),
    "C_4k10_bt4": (W, H, "u10", PARAMS, "gray", 3, 6),
    "D_1440p_f32_kalman_bw64": (2560, 1440, "f32",
                                dict(bt=0, kratio=2.0, bw=64, bh=64, sigma=2.0, sigma2=3.0,
                                     sigma3=2.0, sigma4=1.0), "gray", 2, 6),
}


def cpu_info():
    model = "unknown"
    try:
        for ln in open("/proc/cpuinfo"):
            if ln.startswith("model name"):
                model = ln.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0))}


def cpu_baseline():
    """Steady-state frames/s of the reference CPU path, single-threaded (the
    reference holds one instance mutex around every plane of every frame), on
    configs A-D; `value` is config C's (the metric's workload)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, Ref, script, kind_info
    from fft3dcl_amd.synth import make_clip
    ref = Ref.load()
    kind = "reference" if ref is not None else "port"
    lib = ref if ref is not None else Oracle()
    per, total = {}, 0.0
    for name, (w, h, k, kw, color, n1, n2) in CPU_CONFIGS.items():
        bits, st = kind_info(k)
        planes = make_clip(SEED + len(name), w, h, n2 + 2, k, color=color)
        cf = 3 if color == "yuv" else 1
        t = []
        for n in (n1, n2):
            t0 = time.perf_counter()
            lib.run_clip(script(**kw), planes, bits, st, color_family=cf, order=list(range(n)))
            t.append(time.perf_counter() - t0)
        total += sum(t)
        per[name] = {"frames_per_s": round((n2 - n1) / (t[1] - t[0]), 4),
                     "timed": f"frames 0-{n1 - 1} ({t[0]:.2f} s) and 0-{n2 - 1} ({t[1]:.2f} s)"}
    c = per["C_4k10_bt4"]
    return {"value": c["frames_per_s"], "unit": "frames/s", "cores": 1, "kind": kind,
            "sample": "steady state per output frame (one analysis + one filter/synthesis), "
                      "from the difference of two cold-start runs of n1 and n2 output frames "
                      f"per config; {total:.1f} s of single-threaded CPU work in all",
            **cpu_info(), "configs": per}


def gpu_configs(steps=40, warmup=6):
    """Device-resident frames/s of configs A, B and D (the CPU baseline's other
    configs) on the GPU: planes in HBM, every plane of each output frame,
    `steps` timed frames cycled over 8 distinct seeded frames."""
    import torch
    from fft3dcl_amd.filter import GRAY, YUV, Instance
    from fft3dcl_amd.synth import make_clip
    res = {}
    for name, (w, h, k, kw, color, _, _) in CPU_CONFIGS.items():
        if name.startswith("C_"):
            continue
        bits = {"u8": 8, "u10": 10, "f32": 32}[k]
        st = 1 if k == "f32" else 0
        planes = [torch.from_numpy(p).cuda() for p in make_clip(SEED + len(name), w, h, 8, k,
                                                                  color=color)]
        cf = YUV if color == "yuv" else GRAY
        inst = Instance(cf, w, h, bits, st, warmup + steps + 8, **kw)
        stream = torch.cuda.Stream()
        outs = [torch.empty_like(p[0]) for p in planes]
        srcs = [lambda fn, p=p: p[fn % 8] for p in planes]
        for n in range(warmup):
            inst.get_frame(n, srcs, outs, on_device=True, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for n in range(warmup, warmup + steps):
            inst.get_frame(n, srcs, outs, on_device=True, stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        inst.close()
        ms = e0.elapsed_time(e1) / steps
        res[name] = {"frames_per_s": round(1000.0 / ms, 1), "ms_per_frame": round(ms, 4)}
    return res


def host_buffer_rates(frames, steps=60, repeats=3):
    """Frames/s of the same workload when the planes enter and leave through
    host buffers (the VapourSynth plugin's boundary): PCIe included, both
    directions, as the C ABI does it (host-to-device copy before the analysis,
    device-to-host copy after the synthesis, the call returns when the host
    side is done). Pageable numpy buffers through the library's pinned chunks
    (its default), the same buffers handed to hipMemcpyAsync as they are, and
    pinned buffers; then pageable and pinned buffers with each request
    prefetching the next one's new frame (its upload overlaps the request).
    Each rate is the median of `repeats` runs of `steps` frames (the spread,
    min and max, is reported beside it: host copies share the CPU with
    everything else on the machine)."""
    import numpy as np
    import torch
    from fft3dcl_amd.filter import GRAY, Instance
    res = {}
    pinned = torch.from_numpy(frames).pin_memory()
    out_pinned = torch.empty((H, W), dtype=torch.int16).pin_memory()
    for kind in ("pageable", "pageable_unstaged", "pinned", "pageable_prefetch",
                 "pinned_prefetch"):
        inst = Instance(GRAY, W, H, BITS, 0, repeats * steps + 16,
                        prefetch=kind.endswith("prefetch"),
                        **PARAMS)
        if kind == "pageable_unstaged":
            inst.objs[0].set_option(inst.objs[0].OPT_HOST_STAGING, 0)
        if kind.startswith("pageable"):
            out = np.zeros((H, W), np.uint16)
            srcs = [lambda fn: frames[fn % NUNIQ]]
            dst = out
        else:
            srcs = [lambda fn: pinned[fn % NUNIQ]]
            dst = out_pinned
        for n in range(4):
            inst.get_frame(n, srcs, [dst])
        rates = []
        for rep in range(repeats):
            base = 4 + rep * steps
            t0 = time.perf_counter()
            for n in range(base, base + steps):
                inst.get_frame(n, srcs, [dst])
            rates.append(steps / (time.perf_counter() - t0))
        inst.close()
        rates.sort()
        res[kind + "_fps"] = round(rates[len(rates) // 2], 1)
        res[kind + "_fps_range"] = [round(rates[0], 1), round(rates[-1], 1)]
    res["frames"] = steps
    res["repeats"] = repeats
    return res


# Forms of the 4K workload measured side by side: (name, Instance options,
# plane options). "fp32_fft": the block DFTs as FP32 register FFTs; "mfma":
# as matrix products on the matrix cores with FP16 or BF16 operands split
# hi + lo and FP32 accumulation; the filter is FP32 in every form. "half"
# stores the spectra ring as FP16 pairs.
FORMS = [
    ("fp32_fft_fp32_ring", dict(), {}),
    ("fp32_fft_fp16_ring", dict(half=True), {}),
    ("mfma_fp16_split_fp32_ring", dict(dft=1), {}),
    ("mfma_bf16_split_fp32_ring", dict(dft=2), {}),
    ("mfma_fp16_split_fp16_ring", dict(dft=1, half=True), {}),
]


def reference_outputs(frames, outs=(1, 2)):
    """Output frames `outs` of the 4K workload from the reference CPU build
    (or the oracle when it is absent) and from the FP64 oracle, on the first
    four frames of the clip (enough for both windows)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, Ref, script
    clip = [frames[:4]]
    ref = Ref.load()
    lib = ref if ref is not None else Oracle()
    r, _ = lib.run_clip(script(**PARAMS), clip, BITS, 0, order=list(outs))
    o64, _ = Oracle(fp64=True).run_clip(script(**PARAMS), clip, BITS, 0, order=list(outs))
    return ("reference" if ref is not None else "oracle"), r[0][list(outs)], o64[0][list(outs)]


def form_errors(got, ref, o64):
    import numpy as np
    g = got.astype(np.int64)
    dr, d64 = np.abs(g - ref.astype(np.int64)), np.abs(g - o64.astype(np.int64))
    return {"max_lsb_vs_cpu_reference": int(dr.max()),
            "identical_pct_vs_cpu_reference": round(100.0 * float((dr == 0).mean()), 3),
            "max_lsb_vs_fp64_oracle": int(d64.max()),
            "mean_abs_lsb_vs_fp64_oracle": round(float(d64.mean()), 5)}


def forms_section(dev, frames, steps=100, warmup=10, kernel_steps=30):
    """Each form of FORMS on the 4K workload: frames/s (no events in the timed
    steps), per-kernel times (a separate pass with events), and the error of
    output frames 1-2 against the CPU reference and the FP64 oracle, in LSB."""
    import numpy as np
    import torch
    from fft3dcl_amd.filter import GRAY, Instance
    kind, ref, o64 = reference_outputs(frames)
    res = {"errors_on": f"output frames 1-2 of the clip, against the {kind} CPU build "
                        "and the FP64 oracle"}
    for name, inst_kw, plane_opts in FORMS:
        total = steps + warmup + kernel_steps
        inst = Instance(GRAY, W, H, BITS, 0, total + 16, **inst_kw, **PARAMS)
        plane = inst.objs[0]
        for opt, val in plane_opts.items():
            plane.set_option(opt, val)
        stream = torch.cuda.Stream()
        out = torch.empty((H, W), dtype=torch.int16, device="cuda")
        # error frames: a fresh clip of the same frames 0..3
        src4 = [lambda fn: dev[min(fn, 3)]]
        got = []
        for n in range(3):
            inst.get_frame(n, src4, [out], on_device=True, stream=stream.cuda_stream)
            if n >= 1:
                with torch.cuda.stream(stream):
                    got.append(out.clone())
        inst.close()
        inst = Instance(GRAY, W, H, BITS, 0, total + 16, **inst_kw, **PARAMS)
        plane = inst.objs[0]
        for opt, val in plane_opts.items():
            plane.set_option(opt, val)
        srcs = [lambda fn: dev[fn % NUNIQ]]
        for n in range(warmup):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for n in range(warmup, warmup + steps):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        plane.profile()
        plane.set_profiling(True)
        for n in range(warmup + steps, total):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
        plane.set_profiling(False)
        prof = plane.profile()
        info = plane.info
        inst.close()
        ms = e0.elapsed_time(e1) / steps
        g = torch.stack(got).cpu().numpy().view(np.uint16)
        res[name] = {"frames_per_s": round(1000.0 / ms, 1), "ms_per_step": round(ms, 4),
                     "kernels_ms": {k: round(t / c, 4) for k, (c, t) in prof.items()},
                     "ring_bytes_per_spectrum": info.spectrum_bytes,
                     **form_errors(g, ref, o64)}
    return res


def prefetch_variant(dev, steps=100, warmup=10):
    """The same device-resident workload with each request prefetching the
    next one's new frame (fft3d_plane_prefetch with a device source): that
    frame's analysis runs on the plane's upload stream after the current
    sweep, beside its seam combine. Frames/s of both forms, and the largest
    difference from plain requests over 8 frames, in LSB (0: the same
    kernels on the same data)."""
    import numpy as np
    import torch
    from fft3dcl_amd.filter import GRAY, Instance
    outs, res = {}, {}
    for pf in (False, True):
        inst = Instance(GRAY, W, H, BITS, 0, steps + warmup + 16, prefetch=pf, **PARAMS)
        stream = torch.cuda.Stream()
        out = torch.empty((H, W), dtype=torch.int16, device="cuda")
        srcs = [lambda fn: dev[fn % NUNIQ]]
        frames = []
        for n in range(8):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
            with torch.cuda.stream(stream):
                frames.append(out.clone())
        for n in range(8, 8 + warmup):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for n in range(8 + warmup, 8 + warmup + steps):
            inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
        e1.record(stream)
        torch.cuda.synchronize()
        inst.close()
        outs[pf] = torch.stack(frames).cpu().numpy().view(np.uint16).astype(np.int32)
        ms = e0.elapsed_time(e1) / steps
        res["prefetch" if pf else "plain"] = {"frames_per_s": round(1000.0 / ms, 1),
                                              "ms_per_step": round(ms, 4)}
    res["max_lsb_vs_plain"] = int(np.abs(outs[True] - outs[False]).max())
    return res


SERVED_LOOPS = {"one_queue": dict(one_queue=True), "streams": {},
                "batched": dict(batched=True),
                "batched_fp16_mfma": dict(batched=True, form="fp16_mfma")}


def served_setup(seconds=0.5, nstreams=32, uniq=3):
    """The served section's streams, their seeded device frames and outputs."""
    import torch
    from fft3dcl_amd.serve import make_specs
    specs = make_specs(SEED, nstreams, duration=seconds)
    g = torch.Generator(device="cuda").manual_seed(SEED)
    src, out = {}, {}
    for sp in specs:
        dt = torch.uint8 if sp.bits == 8 else torch.int16
        src[sp.sid] = torch.randint(0, 1 << sp.bits, (uniq, sp.height, sp.width), generator=g,
                                    device="cuda", dtype=torch.int32).to(dt)
        out[sp.sid] = torch.empty((sp.height, sp.width), dtype=dt, device="cuda")
    # each stream's frames as ready device views (a server holds its frame
    # buffers; the callback only picks one)
    views = {sid: [t[i] for i in range(uniq)] for sid, t in src.items()}
    sources = {sid: (lambda fn, v=v: v[fn % uniq]) for sid, v in views.items()}
    return specs, sources, out


def served_pass(loop):
    """One capacity pass (every request issued at once) of one served loop
    after a warm-up pass of it: the unit that `rocprofv3 --kernel-trace`
    profiles. One JSON line."""
    from fft3dcl_amd.serve import Server, summarize
    specs, sources, out = served_setup()
    def run():
        return Server(specs, sources, lambda sid, n: out[sid], 64 << 30,
                      **SERVED_LOOPS[loop]).run(paced=False)
    run()
    res = summarize(specs, run())
    print(json.dumps({"served_pass": loop, "frames_per_s": res["frames_per_s"],
                      "requests": res["requests"]}), flush=True)


def served_sample(seconds=0.5, nstreams=32, uniq=3):
    """The served form (32 seeded streams of mixed resolution, bit depth, bt
    and block size, Poisson arrivals at 24-60 frames/s each over `seconds`):
    p50/p99 latency at that offered load, and the capacity of the same
    requests issued all at once, one call per request and batched through
    fft3d_step. Frames are seeded noise made on the GPU (`uniq` per stream,
    cycled): the filter's cost does not depend on them."""
    from fft3dcl_amd.serve import Server, summarize
    specs, sources, out = served_setup(seconds, nstreams, uniq)
    budget = 64 << 30
    Server(specs, sources, lambda sid, n: out[sid], budget).run(paced=False)  # warm-up
    def run(paced, **kw):
        return Server(specs, sources, lambda sid, n: out[sid], budget, **kw).run(paced=paced)

    run(False)  # warm-up of the per-request loop
    run(False, batched=True, form="fp16_mfma")  # and of the batched one (graph captures)
    cap1 = run(False, one_queue=True)
    cap = run(False)
    capb = run(False, batched=True)
    capbm = run(False, batched=True, form="fp16_mfma")
    paced = run(True)
    pacedb = run(True, batched=True)
    pacedbm = run(True, batched=True, form="fp16_mfma")
    offered = sum(sp.nframes for sp in specs) / seconds
    res = {"offered_frames_per_s": round(offered, 1),
           "at_offered_load": summarize(specs, paced),
           "at_offered_load_batched": summarize(specs, pacedb),
           "at_offered_load_batched_fp16_mfma": summarize(specs, pacedbm),
           "capacity_plain_one_queue": summarize(specs, cap1),
           "capacity": summarize(specs, cap),
           "capacity_batched": summarize(specs, capb),
           "capacity_batched_fp16_mfma": summarize(specs, capbm),
           "mix": "seeded: resolutions 640x360..3840x2160, 8/10-bit, bt 0/2/3/4, bw 16/32/64",
           "forms": "capacity_plain_one_queue: the plain per-clip FP32 form, one call per "
                    "request, every clip on one HIP stream in issue order (the reference's one "
                    "in-order queue per process); capacity: one call per request, FP32, a HIP "
                    "stream per clip; batched: one fft3d_step call per pass of the loop "
                    "(batched launches per kernel class, hipGraph replays), FP32 or FP16 ring "
                    "+ MFMA DFT on 32x32 planes"}
    res["batched_fp16_mfma_over_plain_one_queue"] = round(
        res["capacity_batched_fp16_mfma"]["frames_per_s"] /
        res["capacity_plain_one_queue"]["frames_per_s"], 3)
    res["batched_fp16_mfma_over_per_clip_streams"] = round(
        res["capacity_batched_fp16_mfma"]["frames_per_s"] / res["capacity"]["frames_per_s"], 3)
    for k in ("capacity_plain_one_queue", "capacity", "capacity_batched",
              "capacity_batched_fp16_mfma"):
        res[k].pop("p50_ms")
        res[k].pop("p99_ms")
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-host", action="store_true",
                    help="skip the host-buffer (PCIe-inclusive) rate measurement")
    ap.add_argument("--no-served", action="store_true",
                    help="skip the 32-stream served-form sample")
    ap.add_argument("--no-variants", action="store_true",
                    help="skip the precision forms and the device-prefetch variant")
    ap.add_argument("--host-only", action="store_true",
                    help="only measure the host-buffer rates (one JSON line of them)")
    ap.add_argument("--form", default=FORMS[0][0], choices=[f[0] for f in FORMS],
                    help="precision form of the main loop (default: the plain FP32 form)")
    ap.add_argument("--served-pass", choices=sorted(SERVED_LOOPS), default=None,
                    help="only one capacity pass of that served loop (for kernel traces)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a plain device copy of MIB MiB, --steps times (the "
                         "reference point for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        sys.stderr.write("bench.py measures one GPU only: there is no multi-GPU path "
                         "(whole streams would go to separate GPUs, with no communication)\n")
        sys.exit(2)

    import numpy as np
    import torch

    from fft3dcl_amd import capi
    from fft3dcl_amd.filter import GRAY, Instance
    from fft3dcl_amd.synth import make_plane_sequence

    if args.copy_calibration:
        n = args.copy_calibration << 20
        a = torch.empty(n, dtype=torch.uint8, device="cuda")
        b = torch.empty_like(a)
        a.fill_(1)
        torch.cuda.synchronize()
        for _ in range(args.steps):
            b.copy_(a)
        torch.cuda.synchronize()
        print(json.dumps({"copy_calibration_bytes": n, "launches": args.steps}), flush=True)
        return

    capi.dev_init(0)
    if args.served_pass:
        served_pass(args.served_pass)
        return
    frames = make_plane_sequence(SEED, W, H, NUNIQ, "u10")
    if args.host_only:
        print(json.dumps({"host_buffers_pcie_inclusive": host_buffer_rates(frames)}), flush=True)
        return
    dev = torch.from_numpy(frames).to("cuda")
    out = torch.empty((H, W), dtype=torch.int16, device="cuda")
    nframes = max(120, args.warmup + 2 * args.steps + 8)
    form_kw = {f[0]: f[1] for f in FORMS}[args.form]
    inst = Instance(GRAY, W, H, BITS, 0, nframes, **form_kw, **PARAMS)
    plane = inst.objs[0]
    info = plane.info
    stream = torch.cuda.Stream()
    handle = stream.cuda_stream
    srcs = [lambda fn: dev[fn % NUNIQ]]

    def step(n):
        inst.get_frame(n, srcs, [out], on_device=True, stream=handle)

    torch.cuda.synchronize()
    for n in range(args.warmup):
        step(n)
    # timed region: K steps, no events inside
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record(stream)
    for n in range(args.warmup, args.warmup + args.steps):
        step(n)
    ev1.record(stream)
    torch.cuda.synchronize()
    ms_total = ev0.elapsed_time(ev1)
    # kernel pass: the same K steps again with HIP events around every kernel
    # on the launch stream (they add a few microseconds per step, so they stay
    # out of the timed region above)
    base = args.warmup + args.steps
    plane.profile()
    plane.set_profiling(True)
    for n in range(base, base + args.steps):
        step(n)
    plane.set_profiling(False)
    torch.cuda.synchronize()
    prof = plane.profile()
    inst.close()

    ms_step = ms_total / args.steps
    fps = 1000.0 / ms_step
    kb = kernel_bytes(info, 2, 2, 4 if form_kw.get("half") else 8)
    kernels = {}
    for name, (launches, tot) in prof.items():
        avg = tot / launches
        gbs = kb.get(name, 0) / (avg * 1e-3) / 1e9
        kernels[name] = {"launches": launches, "avg_ms": round(avg, 4),
                         "algorithmic_bytes": kb.get(name, 0), "GBps": round(gbs, 1)}
    dom = max(prof, key=lambda k: prof[k][1])
    traffic = None
    tf = ROOT / "profiles" / "traffic.json"
    if tf.exists():
        tj = json.loads(tf.read_text())
        if args.form != FORMS[0][0]:
            tj = tj.get("forms", {}).get(args.form, {})
        t = tj.get(dom)
        traffic = t.get("bytes_per_launch") if isinstance(t, dict) else None
    ach = kernels[dom]["GBps"]
    step_gbs = kb["step"] / (ms_step * 1e-3) / 1e9
    line = {
        "metric": METRIC, "value": round(fps, 2), "unit": "frames/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms_step, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "4K 10-bit luma 3840x2160, sigma=3 bt=4 bw=bh=32 ow=oh=16 "
                               "sharpen=0.4 degrid=1, " +
                               ("FP32 spectra ring" if args.form == FORMS[0][0]
                                else f"form {args.form}"),
                   "frames": nframes, "distinct_frames": NUNIQ, "seed": SEED,
                   "mpix_per_s": round(fps * W * H / 1e6, 1),
                   "step_algorithmic_bytes": kb["step"],
                   "step_GBps": round(step_gbs, 1),
                   "step_hbm_frac": round(step_gbs / HBM_PEAK_GBS, 4),
                   "kernel_events": "a second pass of the same number of steps after the "
                                    "timed region, events around every kernel",
                   "kernels": kernels},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4),
                     "traffic": traffic},
    }
    if not args.no_variants:
        line["config"]["forms"] = forms_section(dev, frames)
        line["config"]["device_prefetch"] = prefetch_variant(dev)
        line["config"]["gpu_configs_A_B_D"] = gpu_configs()
    if not args.no_host:
        line["config"]["host_buffers_pcie_inclusive"] = host_buffer_rates(frames)
    if not args.no_served:
        line["config"]["served_32_streams"] = served_sample()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
