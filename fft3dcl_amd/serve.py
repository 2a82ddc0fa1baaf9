"""Serving many clips at once: the request loop of the served form.

Each stream is one FFT3DCL(clip, ...) instance over its own clip (its own
resolution, bit depth, temporal mode and block size), with its spectra ring
and Kalman state resident in device memory for as long as it is admitted.
Frame requests arrive per stream as a Poisson process. The loop issues every
arrived request at once on its stream's own HIP stream, so the work of
different streams overlaps on the device while the requests of one stream stay
in order; nothing waits for the GPU except the final collection. Streams are
admitted while their estimated device footprint fits a byte budget; a stream
that does not fit waits until an admitted one has finished (a finished stream
is released only when another waits, since releasing waits for its GPU work).

Latency of a request = completion time on the GPU (an event recorded on the
stream right after the frame's kernels) - its arrival time, both measured from
the start of the run.

Without a GPU (tests against the CPU emulation) the same loop runs with host
buffers and no streams; the results must then equal those of each stream
processed alone.

With batched=True, the requests that one pass of the loop issues, over all
admitted streams, go to the library in one fft3d_step call (each item on its
stream's HIP stream); their completion events are recorded per stream after
the call.
"""
from __future__ import annotations

import math
import time
from dataclasses import dataclass, field

import numpy as np

from . import capi
from .filter import GRAY, Instance

__all__ = ["StreamSpec", "make_specs", "footprint", "Server"]


@dataclass
class StreamSpec:
    sid: int
    width: int
    height: int
    bits: int
    params: dict
    rate: float                  # requests per second (Poisson)
    nframes: int                 # frames of the clip (and requests of the stream)
    arrivals: list = field(default_factory=list)  # seconds from the start


RESOLUTIONS = [(640, 360), (1280, 720), (1920, 1080), (3840, 2160)]


def make_specs(seed: int, n: int = 32, duration: float = 0.5, resolutions=None,
               max_frames: int | None = None) -> list[StreamSpec]:
    """n seeded streams: resolution from `resolutions`, 8 or 10 bit, bt in
    {0, 2, 3, 4}, bw = bh in {16, 32, 64} with ow = bw/2, sigma ~ U[1, 4],
    Poisson arrivals at a rate ~ U[24, 60] per second over `duration` s."""
    rng = np.random.default_rng(seed)
    res = resolutions or RESOLUTIONS
    specs = []
    for sid in range(n):
        w, h = res[int(rng.integers(len(res)))]
        bits = int(rng.choice([8, 10]))
        bt = int(rng.choice([0, 2, 3, 4]))
        bw = int(rng.choice([16, 32, 64]))
        sigma = float(rng.uniform(1.0, 4.0))
        rate = float(rng.uniform(24.0, 60.0))
        t, arr = 0.0, []
        while True:
            t += float(rng.exponential(1.0 / rate))
            if t > duration or (max_frames and len(arr) >= max_frames):
                break
            arr.append(t)
        if not arr:
            arr = [float(rng.uniform(0.0, duration))]
        params = dict(sigma=sigma, bt=bt, bw=bw, bh=bw, ow=bw // 2, oh=bw // 2)
        specs.append(StreamSpec(sid, w, h, bits, params, rate, len(arr), arr))
    return specs


def footprint(spec: StreamSpec) -> int:
    """Device bytes one stream holds while admitted (upper estimate): the ring
    of bt+2 spectra plus the reuse pool, the Kalman state, the staging and
    partial planes, and the unfused blocks buffer."""
    bw = spec.params["bw"]
    bt = spec.params["bt"]
    nx = math.ceil((spec.width + bw // 2) / bw) + 1
    ny = math.ceil((spec.height + bw // 2) / bw) + 1
    nb = 4 * nx * ny
    spec_bytes = nb * bw * bw * 4
    slots = (max(bt, 1) + 2) * 2
    kalman = nb * (bw // 2 + 1) * bw * 16 if bt == 0 else 0
    planes = spec.width * spec.height * (2 + 2 + 16)
    blocks = nb * bw * bw * 4
    return spec_bytes * slots + kalman + planes + blocks


def form_options(form: str, spec: StreamSpec) -> dict:
    """Instance options of a served form for one stream."""
    if form == "fp32":
        return {"precision": 2}  # FP32 ring also for 8-bit streams
    if form != "fp16_mfma":
        raise ValueError(f"unknown served form {form!r}")
    bw, bh = spec.params["bw"], spec.params["bh"]
    opts = {"precision": 0}
    if bw == bh and bw in (16, 32, 64):
        opts["half"] = True
    if bw == bh == 32:
        opts["dft"] = 1
    return opts


class Server:
    """Runs the requests of `specs` with at most `budget` estimated device
    bytes of admitted streams. `sources[sid](fn)` gives frame fn of stream
    sid (a device tensor on a GPU run, a numpy array otherwise); `outputs`
    (sid, n) -> buffer receives the filtered frame."""

    def __init__(self, specs, sources, outputs, budget: int, lib=None, on_device=True,
                 batched: bool = False, form: str = "fp32", one_queue: bool = False):
        """form: "fp32" (FP32 register FFTs and spectra ring, also for 8-bit
        streams, whose default precision=0 would store FP16) or "fp16_mfma"
        (FP16 spectra ring on every square 16/32/64 plane, and the MFMA DFT
        with FP16 hi+lo operands on the 32x32 ones)."""
        self.batched = batched
        self.form = form
        # one_queue: every stream's work on one HIP stream, in issue order -
        # the plain per-clip form (the reference serves all of a process's
        # clips through one in-order queue, src/cl/cl.cpp:175)
        self.one_queue = one_queue
        self.specs = {s.sid: s for s in specs}
        self.sources = sources
        self.outputs = outputs
        self.budget = budget
        self.lib = lib
        self.on_device = on_device
        self.max_admitted = 0

    def run(self, paced: bool = True) -> dict:
        """Issues every request (at its arrival time when `paced`, else all at
        once) and returns per-request latencies (ms) and the run's span."""
        import torch
        reqs = sorted((t, s.sid, n) for s in self.specs.values()
                      for n, t in enumerate(s.arrivals))
        waiting = sorted(self.specs, key=lambda sid: self.specs[sid].arrivals[0])
        admitted: dict[int, dict] = {}
        used = 0
        pending = {sid: [] for sid in self.specs}   # arrived, not yet issued (in order)
        done_events = []                            # (sid, n, arrival, event)
        dev = self.on_device
        start = torch.cuda.Event(enable_timing=True) if dev else None
        shared = torch.cuda.Stream() if dev and self.one_queue else None

        def admit():
            nonlocal used
            while waiting:
                sid = waiting[0]
                need = footprint(self.specs[sid])
                if admitted and used + need > self.budget:
                    return
                waiting.pop(0)
                sp = self.specs[sid]
                inst = Instance(GRAY, sp.width, sp.height, sp.bits, 0, sp.nframes,
                                lib=self.lib, **form_options(self.form, sp), **sp.params)
                # the ring's spectra are allocated now, not inside the loop
                inst.objs[0].set_option(inst.objs[0].OPT_RESERVE_SPECTRA,
                                        max(sp.params["bt"], 1) + 2)
                st = (shared if self.one_queue else torch.cuda.Stream()) if dev else None
                admitted[sid] = dict(inst=inst, stream=st, next=0, need=need)
                used += need
                self.max_admitted = max(self.max_admitted, len(admitted))

        def retire(sid):
            nonlocal used
            a = admitted.pop(sid)
            if dev:
                a["stream"].synchronize()
            a["inst"].close()
            used -= a["need"]

        admit()  # the streams that fit from the start are set up before the clock runs
        if dev:
            torch.cuda.synchronize()
        t0 = time.perf_counter()
        if dev:
            start.record(torch.cuda.current_stream())
        i = 0
        while i < len(reqs) or any(pending.values()):
            now = time.perf_counter() - t0
            while i < len(reqs) and (not paced or reqs[i][0] <= now):
                pending[reqs[i][1]].append((reqs[i][2], reqs[i][0]))
                i += 1
            issued = False
            batch = []  # (sid, n, arrival, request) of this pass when batched
            for sid in list(admitted):
                a = admitted[sid]
                while pending[sid] and pending[sid][0][0] == a["next"]:
                    n, arr = pending[sid].pop(0)
                    h = a["stream"].cuda_stream if dev else None
                    if self.batched:
                        batch.append((sid, n, arr, capi.StepRequest(
                            a["inst"].objs[0], n, self.specs[sid].nframes,
                            lambda fn, s=sid: self.sources[s](fn), self.outputs(sid, n),
                            on_device=dev, stream=h)))
                    else:
                        srcs = [lambda fn, s=sid: self.sources[s](fn)]
                        a["inst"].get_frame(n, srcs, [self.outputs(sid, n)], on_device=dev,
                                            stream=h)
                        ev = None
                        if dev:
                            ev = torch.cuda.Event(enable_timing=True)
                            ev.record(a["stream"])
                        done_events.append((sid, n, arr, ev))
                    a["next"] += 1
                    issued = True
                if self.batched and batch and a["next"] == self.specs[sid].nframes and waiting:
                    self._flush(batch, admitted, done_events, dev)
                    batch = []
                if a["next"] == self.specs[sid].nframes and waiting:
                    # free its memory only when another stream waits for it:
                    # retiring waits for the stream's GPU work
                    retire(sid)
                    admit()
            if batch:
                self._flush(batch, admitted, done_events, dev)
            if not issued and i < len(reqs) and paced:
                time.sleep(max(0.0, min(0.0005, reqs[i][0] - (time.perf_counter() - t0))))
        for sid in list(admitted):
            retire(sid)
        lat, span = [], 0.0
        if dev:
            torch.cuda.synchronize()
            for sid, n, arr, ev in done_events:
                t_done = start.elapsed_time(ev)
                span = max(span, t_done)
                lat.append(t_done - 1000.0 * arr if paced else t_done)
        else:
            span = 1000.0 * (time.perf_counter() - t0)
        return dict(latencies_ms=lat, span_ms=span, requests=len(done_events),
                    max_admitted=self.max_admitted)


    def _flush(self, batch, admitted, done_events, dev):
        """One fft3d_step call for the batched requests, then a completion
        event per stream that had items in it."""
        import torch
        st = capi.step([r for *_, r in batch], lib=self.lib)
        bad = [(b[0], b[1], c) for b, c in zip(batch, st) if c != 0]
        if bad:
            raise capi.FFT3DError(bad[0][2], f"fft3d_step failed for (stream, frame) {bad[:3]}")
        evs = {}
        for sid, n, arr, _ in batch:
            if dev and sid not in evs:
                evs[sid] = torch.cuda.Event(enable_timing=True)
                evs[sid].record(admitted[sid]["stream"])
            done_events.append((sid, n, arr, evs.get(sid)))


def summarize(specs, res) -> dict:
    lat = np.array(res["latencies_ms"]) if res["latencies_ms"] else np.zeros(1)
    px = sum(s.width * s.height * s.nframes for s in specs)
    span_s = res["span_ms"] / 1000.0
    return dict(streams=len(specs), requests=res["requests"],
                frames_per_s=round(res["requests"] / span_s, 1),
                mpix_per_s=round(px / span_s / 1e6, 1),
                p50_ms=round(float(np.percentile(lat, 50)), 3),
                p99_ms=round(float(np.percentile(lat, 99)), 3),
                max_admitted=res["max_admitted"])
