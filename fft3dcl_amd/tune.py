"""Autotuner of the fused sweep's launch parameters (needs the GPU).

For each workload (block size, plane size, ring precision, DFT form) it
times the sweep and seam-combine kernels of the steady-state frame step for
every candidate - block rows per band (the sweep grid: workgroups = strips x
bands) and the sweep kernel built for 2, 3 or 4 resident workgroups per CU
(where the block size has one) - and keeps the fastest. The result is a JSON
table that the library reads at plane creation (tuning.json beside
libfft3d_hip.so, or $FFT3D_TUNING). The reference sizes its work-groups by a
formula instead (src/cl/gpu.cpp:70-81).

    python -m fft3dcl_amd.tune [--out PATH] [--frames N]

Timing uses HIP events around every kernel on the launch stream (the
library's profiling); the candidates only change the summation order of the
partial sums at band seams, so every choice filters the same frame.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

PKG = Path(__file__).resolve().parent

# (width, height, bits, block size, filter parameters, ring/DFT forms)
FORMS_ALL = [(0, 0), (1, 0), (1, 1)]  # (half, dft)
WORKLOADS = []
for w, h in ((640, 360), (1280, 720), (1920, 1080), (960, 540), (3840, 2160), (2560, 1440)):
    for bw in (16, 32, 64):
        forms = FORMS_ALL if bw == 32 else [(0, 0), (1, 0)]
        WORKLOADS.append((w, h, 10, bw, dict(sigma=3.0, bt=4, sharpen=0.4), forms))


def candidates(plane, info):
    """(band rows, occupancy) pairs around the heuristic's band rows."""
    occs = [2]
    for o in (3, 4):
        try:
            plane.set_option(plane.OPT_SWEEP_OCC, o)
            occs.append(o)
        except Exception:
            break
    plane.set_option(plane.OPT_SWEEP_OCC, 0)
    nrows = info.nrows
    rows = sorted({r for r in (1, 2, 3, 4, 5, 6, 8, 10, 12, 16) if r <= nrows} |
                  {min(nrows, info.band_rows)})
    return [(r, o) for o in occs for r in rows]


def time_candidate(inst, plane, srcs, out, stream, first, frames):
    """Sweep + combine milliseconds per frame over output frames
    first..first+frames-1 (each one steady-state step: one new frame)."""
    import torch
    plane.profile()
    plane.set_profiling(True)
    for n in range(first, first + frames):
        inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
    plane.set_profiling(False)
    prof = plane.profile()
    torch.cuda.synchronize()
    t = 0.0
    for name in ("filter_synth_sweep", "combine_seams"):
        c, ms = prof.get(name, (0, 0.0))
        t += ms / max(c, 1)
    return t


def tune(frames: int = 24, log=print) -> dict:
    import torch
    from fft3dcl_amd import capi
    from fft3dcl_amd.filter import GRAY, Instance
    capi.dev_init(0)
    entries = []
    g = torch.Generator(device="cuda").manual_seed(7)
    for w, h, bits, bw, kw, forms in WORKLOADS:
        src = torch.randint(0, 1 << bits, (4, h, w), generator=g, device="cuda",
                            dtype=torch.int32).to(torch.int16)
        out = torch.empty((h, w), dtype=torch.int16, device="cuda")
        srcs = [lambda fn, s=src: s[fn % 4]]
        for half, dft in forms:
            # a clip long enough that every timed frame is a steady-state step
            inst = Instance(GRAY, w, h, bits, 0, 1 << 20, half=bool(half), dft=dft, bw=bw,
                            bh=bw, ow=bw // 2, oh=bw // 2, **kw)
            plane = inst.objs[0]
            info = plane.info
            stream = torch.cuda.Stream()
            best = None
            base = None
            nxt = 0  # next output frame: the clip runs on across candidates
            for R, occ in candidates(plane, info):
                plane.set_option(plane.OPT_BAND_ROWS, R)
                plane.set_option(plane.OPT_SWEEP_OCC, occ)
                for n in range(nxt, nxt + 3):  # warm-up of this launch shape
                    inst.get_frame(n, srcs, [out], on_device=True, stream=stream.cuda_stream)
                t = time_candidate(inst, plane, srcs, out, stream, nxt + 3, frames)
                nxt += 3 + frames
                if R == min(info.nrows, info.band_rows) and occ == 2:
                    base = t
                if best is None or t < best[0]:
                    best = (t, R, occ)
            inst.close()
            t, R, occ = best
            entries.append(dict(bw=bw, bh=bw, width=w, height=h, half=half, dft=dft,
                                band_rows=R, occupancy=occ if not dft else 2,
                                sweep_combine_ms=round(t, 5),
                                heuristic_ms=round(base, 5) if base else None))
            log(json.dumps(entries[-1]))
            sys.stdout.flush()
    return {"version": 1, "device": "gfx950",
            "note": "fused sweep + seam combine per frame, HIP events; written by "
                    "python -m fft3dcl_amd.tune",
            "entries": entries}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", default=str(PKG / "tuning.json"))
    ap.add_argument("--frames", type=int, default=24)
    a = ap.parse_args()
    os.environ["FFT3D_TUNING"] = os.devnull  # measure from the heuristic, not an old table
    t0 = time.time()
    res = tune(a.frames)
    res["seconds"] = round(time.time() - t0, 1)
    Path(a.out).write_text(json.dumps(res, indent=1) + "\n")
    print(f"wrote {a.out} ({len(res['entries'])} entries, {res['seconds']} s)")


if __name__ == "__main__":
    main()
