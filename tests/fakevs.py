"""ctypes driver of the fake VapourSynth host (tests/vs_host/fake_vs.cpp)."""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

BUILD = Path(__file__).resolve().parent / "_build"
FAKE = BUILD / "libfakevs.so"


class Clip(C.Structure):
    _fields_ = [("colorFamily", C.c_int), ("sampleType", C.c_int), ("bits", C.c_int),
                ("ssw", C.c_int), ("ssh", C.c_int), ("width", C.c_int), ("height", C.c_int),
                ("numFrames", C.c_int), ("pad", C.c_int)]


INT_KEYS = {"bw", "bh", "bt", "plane", "mode", "bordersize", "precision", "ow", "oh",
            "wintype", "interlaced", "oldfft", "device"}


class Host:
    def __init__(self):
        self.lib = C.CDLL(str(FAKE))
        self.lib.fv_plugin_info.argtypes = [C.c_char_p, C.c_char_p, C.c_int]
        self.lib.fv_run.argtypes = [C.c_char_p, C.POINTER(Clip), C.c_void_p, C.c_void_p,
                                    C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p,
                                    C.c_int, C.c_void_p, C.POINTER(C.c_int), C.c_char_p, C.c_int,
                                    C.c_char_p, C.c_int]

    def info(self, plugin: str) -> list[str]:
        buf = C.create_string_buffer(4096)
        assert self.lib.fv_plugin_info(plugin.encode(), buf, 4096) == 0, buf.value
        return buf.value.decode().split("|")

    def run(self, plugin: str, planes: list, color_family=1, bits=8, sample_type=0, ssw=1,
            ssh=1, order=None, pad=0, **args):
        """Returns (rc, outputs, requests per output frame, filter mode, error, log)."""
        nplanes = len(planes)
        nframes = planes[0].shape[0]
        H, W = planes[0].shape[1:]
        clip = Clip(color_family, sample_type, bits, ssw, ssh, W, H, nframes, pad)
        order = list(range(nframes)) if order is None else list(order)
        keep = [np.ascontiguousarray(p[f]) for f in range(nframes) for p in planes]
        ins = (C.c_void_p * len(keep))(*[a.ctypes.data for a in keep])
        outs = [np.zeros_like(p) for p in planes]
        optr = (C.c_void_p * (nframes * nplanes))(
            *[outs[q][f].ctypes.data for f in range(nframes) for q in range(nplanes)])
        keys = list(args)
        ck = (C.c_char_p * max(1, len(keys)))(*[k.encode() for k in keys])
        ct = (C.c_int * max(1, len(keys)))(*[0 if k in INT_KEYS else 1 for k in keys])
        cv = (C.c_double * max(1, len(keys)))(*[float(args[k]) for k in keys])
        co = (C.c_int * max(1, len(order)))(*order)
        req = (C.c_int * (8 * max(1, len(order))))()
        fm = C.c_int(-1)
        err = C.create_string_buffer(1024)
        log = C.create_string_buffer(8192)
        rc = self.lib.fv_run(plugin.encode(), C.byref(clip), ins, ck, ct, cv, len(keys), optr,
                             co, len(order), req, C.byref(fm), err, 1024, log, 8192)
        reqs = [[x for x in req[8 * k:8 * k + 8] if x >= 0] for k in range(len(order))]
        return rc, outs, reqs, fm.value, err.value.decode(), log.value.decode()
