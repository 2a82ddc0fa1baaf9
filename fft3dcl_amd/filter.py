"""Clip-level FFT3DCL filter over the HIP backend (Python host mirror).

Restates the frame scheduling of the reference's VapourSynth filter for the GPU
path (src/fft3dcl.cpp): plane routing (:200-225), backend selection by
`device` (:233-283), the clamped temporal window (:303-331, :424-430), the
forward-if-missing spectra ring (:396-423), process (:431-448) and ring
pruning (:456-467). The same logic in C++ is the VapourSynth plugin
(csrc/fft3dcl_vs.cpp); this module drives it for numpy/torch clips.
"""
from __future__ import annotations

from typing import Sequence

import numpy as np

from . import capi

GRAY, RGB, YUV = 1, 2, 3


def clamp_frame(n: int, num_frames: int) -> int:
    """src/fft3dcl.cpp:303-309"""
    return 0 if n < 0 else (num_frames - 1 if n > num_frames - 1 else n)


def route_planes(color_family: int, plane: int) -> list[int]:
    """Processed plane indices (src/fft3dcl.cpp:200-225)."""
    if color_family == RGB:
        plane = 4
    if color_family == GRAY:
        if plane == 1:
            raise ValueError("FFT3DCL: cannot set chroma processing for a greyscale clip")
        if plane == 4:
            plane = 0
    if plane in (2, 3):
        plane = 1
    if plane not in (0, 1, 4):
        raise ValueError("FFT3DCL: plane must be 0 (luma), 1 (chroma) or 4 (all)")
    if color_family == GRAY or plane == 0:
        return [0]
    if plane == 1:
        return [1, 2]
    return [0, 1, 2]


class Instance:
    """One FFT3DCL(clip, ...) instance: per processed plane a capi.Plane."""

    def __init__(self, color_family: int, width: int, height: int, bits: int,
                 sample_type: int, num_frames: int, ssw: int = 1, ssh: int = 1,
                 lib=None, fuse: bool = False, half: bool = False, prefetch: bool = False,
                 dft: int = 0, **kw):
        """fuse: let the filter pass transform the last missing window frame
        itself (FFT3D_OPT_FUSED_ANALYSIS) instead of a separate analysis.
        half: store the spectra ring as FP16 (FFT3D_OPT_SPECTRA_FP16).
        dft: 0 register FFTs; 1 / 2 the block DFTs on the matrix cores with
        FP16 / BF16 operands split hi + lo (FFT3D_OPT_DFT, 32x32 blocks).
        prefetch: with bt >= 2, each request also transforms the frame that
        the next sequential request adds to the window (fft3d_plane_prefetch),
        so its upload and analysis overlap this request's work."""
        device = int(kw.pop("device", 2))
        if device == 0:
            raise ValueError("FFT3DCL: device=0 selects the CPU backend, which this HIP "
                             "build does not ship (use device=1 or 2)")
        if sample_type == 1 and bits != 32:
            raise ValueError("FFT3DCL: only 32-bit float clips supported (16-bit float not "
                             "ported yet)")
        if sample_type == 0 and not 8 <= bits <= 16:
            raise ValueError("FFT3DCL: integer bit depths 8-16 only")
        self.params = capi.make_params(**kw)
        self.planes = route_planes(color_family, self.params.plane)
        self.nplanes = 1 if color_family == GRAY else 3
        self.num_frames = num_frames
        self.prefetch = prefetch
        self.objs: list[capi.Plane] = []
        for q in self.planes:
            pw = width >> (ssw if q else 0)
            ph = height >> (ssh if q else 0)
            self.objs.append(capi.Plane(self.params, pw, ph, bits, sample_type,
                                        chroma=color_family == YUV and q > 0, lib=lib))
        for o in self.objs:
            o.set_option(capi.Plane.OPT_FUSED_ANALYSIS, int(fuse))
            if half:
                o.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
            if dft:
                o.set_option(capi.Plane.OPT_DFT, dft)
        info = self.objs[0].info
        self.bm, self.bt, self.current = info.bm, info.bt, info.current
        self._conc = None  # per-plane streams and events of device requests

    def window(self, n: int) -> list[int]:
        if self.bm >= 2:
            return [clamp_frame(n - self.current + i, self.num_frames) for i in range(self.bt)]
        return [clamp_frame(n, self.num_frames)]

    def _plane_streams(self, stream):
        """With device buffers and several processed planes, each plane's
        calls go on a stream of its own, after the caller's stream (where the
        sources were written) and before the caller's later work, so the
        planes of a frame overlap on the GPU; the reference runs them one
        after another (src/fft3dcl.cpp:431-448)."""
        import torch
        if self._conc is None:
            self._conc = dict(fork=torch.cuda.Event(),
                              planes=[(torch.cuda.Stream(), torch.cuda.Event())
                                      for _ in self.planes])
        caller = (torch.cuda.ExternalStream(stream) if stream
                  else torch.cuda.current_stream())
        self._conc["fork"].record(caller)
        for ps, _ in self._conc["planes"]:
            ps.wait_event(self._conc["fork"])
        return caller

    def get_frame(self, n: int, src_planes: Sequence, dst_planes: list,
                  on_device: bool = False, stream=None) -> None:
        """src_planes[q](fn) -> the source plane q of frame fn (numpy or device
        tensor); dst_planes[q] receives plane q of output frame n. With device
        buffers the processed planes run concurrently, each on a stream of its
        own between the caller's earlier and later work on `stream`."""
        for q in range(self.nplanes):
            if q not in self.planes:
                dst_planes[q][...] = src_planes[q](n)
        caller = self._plane_streams(stream) if on_device and len(self.planes) > 1 else None
        for k, q in enumerate(self.planes):
            obj = self.objs[k]
            s = self._conc["planes"][k][0].cuda_stream if caller is not None else stream
            win = self.window(n)
            if self.prefetch and self.bm >= 2:
                # the next sequential request's new frame (kept by the prune below)
                nxt = clamp_frame(n - self.current + self.bt, self.num_frames)
                if nxt not in win and not obj.has(nxt):
                    obj.prefetch(nxt, src_planes[q](nxt), on_device)
            missing = [fn for fn in dict.fromkeys(win) if not obj.has(fn)]
            for fn in missing[:-1]:
                obj.forward(fn, src_planes[q](fn), on_device, s)
            if missing:
                fn = missing[-1]
                obj.forward_process(fn, src_planes[q](fn), win, dst_planes[q], on_device,
                                    on_device, s)
            else:
                obj.process(win, dst_planes[q], on_device, s)
            if self.bm >= 2:
                obj.prune(n - self.current - 1, n - self.current + self.bt)
            else:
                # one-frame modes and Kalman (whose state is not in the ring)
                # reuse no spectrum; the reference keeps bt=0 spectra forever
                obj.clear()
            if caller is not None:
                ps, done = self._conc["planes"][k]
                done.record(ps)
                caller.wait_event(done)

    def close(self):
        for o in self.objs:
            o.close()
        self.objs = []


def FFT3DCL(planes: Sequence[np.ndarray], color_family: int = GRAY, bits: int = 8,
            sample_type: int = 0, ssw: int = 1, ssh: int = 1, order=None, lib=None,
            **params) -> list[np.ndarray]:
    """Filters a whole clip given as host arrays, one (frames, h, w) array per
    plane, requesting output frames in `order` (default 0..n-1)."""
    nframes = planes[0].shape[0]
    H, W = planes[0].shape[1:]
    inst = Instance(color_family, W, H, bits, sample_type, nframes, ssw, ssh, lib=lib,
                    **params)
    outs = [np.zeros_like(p) for p in planes]
    srcs = [(lambda fn, a=p: np.ascontiguousarray(a[fn])) for p in planes]
    try:
        for n in (range(nframes) if order is None else order):
            dst = [o[n] for o in outs]
            inst.get_frame(n, srcs, dst)
    finally:
        inst.close()
    return outs
