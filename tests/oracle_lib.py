"""ctypes bindings of the oracle (oracle/liboracle*.so) and of the reference's
own CPU backend built by oracle/Makefile (oracle/_ref/libfft3dref.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's CPU
baseline use this module.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent.parent
ORACLE_DIR = ROOT / "oracle"

SCRIPT_FIELDS = [
    ("sigma", C.c_float), ("beta", C.c_float), ("bw", C.c_int), ("bh", C.c_int),
    ("bt", C.c_int), ("sharpen", C.c_float), ("plane", C.c_int), ("mode", C.c_int),
    ("bordersize", C.c_int), ("precision", C.c_int), ("degrid", C.c_float),
    ("scutoff", C.c_float), ("svr", C.c_float), ("smin", C.c_float), ("smax", C.c_float),
    ("kratio", C.c_float), ("ow", C.c_int), ("oh", C.c_int), ("wintype", C.c_int),
    ("interlaced", C.c_int), ("sigma2", C.c_float), ("sigma3", C.c_float),
    ("sigma4", C.c_float), ("oldfft", C.c_int),
]


class Script(C.Structure):
    _fields_ = SCRIPT_FIELDS


class Internal(C.Structure):
    _fields_ = [
        ("sigma", C.c_float), ("beta", C.c_float), ("bw", C.c_int), ("bh", C.c_int),
        ("bm", C.c_int), ("bt", C.c_int), ("sharpen", C.c_float), ("sharpenOn", C.c_int),
        ("mode", C.c_int), ("bordersize", C.c_int), ("precision", C.c_int),
        ("degrid", C.c_float), ("degridOn", C.c_int), ("scutoff", C.c_float),
        ("svr", C.c_float), ("smin", C.c_float), ("smax", C.c_float), ("kratio", C.c_float),
        ("ow", C.c_int), ("oh", C.c_int), ("wintype", C.c_int), ("interlaced", C.c_int),
        ("sigma2", C.c_float), ("sigma3", C.c_float), ("sigma4", C.c_float),
        ("usePattern", C.c_int), ("current", C.c_int), ("nx", C.c_int), ("ny", C.c_int),
        ("totw", C.c_int), ("toth", C.c_int), ("oldfft", C.c_int), ("wienerS", C.c_float),
        ("betaX", C.c_float), ("kalmanCN", C.c_float), ("ssmin", C.c_float),
        ("ssmax", C.c_float),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


def script(**kw) -> Script:
    """Script parameters with the reference's defaults (src/fft3dcl.cpp:171-195):
    degrid defaults by mode, ow/oh by the input bw/bh, sigma2/3/4 by sigma."""
    bw = kw.get("bw", 32)
    bh = kw.get("bh", 32)
    sigma = kw.get("sigma", 2.0)
    mode = kw.get("mode", 1)
    d = dict(sigma=sigma, beta=1.0, bw=bw, bh=bh, bt=3, sharpen=0.0, plane=0, mode=mode,
             bordersize=1, precision=0, degrid=1.0 if mode == 1 else 0.0, scutoff=0.3,
             svr=1.0, smin=4.0, smax=20.0, kratio=2.0, ow=bw // 2, oh=bh // 2, wintype=0,
             interlaced=0, sigma2=sigma, sigma3=sigma, sigma4=sigma, oldfft=-1)
    for k, v in kw.items():
        if k not in d:
            raise KeyError(k)
        d[k] = v
    s = Script()
    for k, v in d.items():
        setattr(s, k, v)
    return s


def _f(a):
    return a.ctypes.data_as(C.c_void_p)


class Oracle:
    """The plain-C restatement (FP32 or FP64 twin)."""

    def __init__(self, fp64: bool = False):
        name = "liboracle64.so" if fp64 else "liboracle.so"
        self.lib = C.CDLL(str(ORACLE_DIR / name))
        L = self.lib
        self.real = np.float64 if L.orc_real_size() == 8 else np.float32
        L.orc_validate.argtypes = [C.POINTER(Script), C.c_int, C.c_int, C.c_int,
                                   C.POINTER(Internal), C.POINTER(C.c_char_p)]
        L.orc_plane_new.argtypes = [C.POINTER(Internal), C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int]
        L.orc_plane_new.restype = C.c_void_p
        L.orc_plane_free.argtypes = [C.c_void_p]
        for fn in ("orc_plane_nblocks", "orc_plane_fbw", "orc_plane_ncols"):
            getattr(L, fn).argtypes = [C.c_void_p]
        L.orc_plane_chroma_norm.argtypes = [C.c_void_p]
        L.orc_plane_chroma_norm.restype = C.c_float
        for fn in ("orc_plane_bx", "orc_plane_by", "orc_plane_par", "orc_plane_pattern",
                   "orc_plane_sharpen", "orc_plane_degrid"):
            getattr(L, fn).argtypes = [C.c_void_p, C.c_void_p]
        L.orc_plane_win.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.orc_upload.argtypes = [C.c_void_p, C.c_void_p, C.c_ssize_t, C.c_void_p]
        L.orc_download.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_ssize_t]
        L.orc_compute_spectra.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.orc_build_degrid.argtypes = [C.c_void_p]
        L.orc_filter_synth.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.orc_fft1d.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.orc_fft2d_real_forward.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_fft2d_real_inverse_half.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.orc_round_pow2_clamp.argtypes = [C.c_int]
        L.orc_run_clip.argtypes = [C.POINTER(Script), C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_char_p)]

    # -- parameters
    def validate(self, sp: Script, w: int, h: int, bits: int):
        out = Internal()
        err = C.c_char_p()
        ok = self.lib.orc_validate(C.byref(sp), w, h, bits, C.byref(out), C.byref(err))
        return (out, None) if ok else (None, err.value.decode())

    def round_pow2_clamp(self, v: int) -> int:
        return self.lib.orc_round_pow2_clamp(v)

    # -- FFT
    def fft1d(self, x: np.ndarray, inverse=False) -> np.ndarray:
        a = np.ascontiguousarray(x, dtype=np.complex128 if self.real == np.float64 else np.complex64)
        o = np.empty_like(a)
        self.lib.orc_fft1d(len(a), _f(a), _f(o), int(inverse))
        return o

    def fft2d_forward(self, block: np.ndarray) -> np.ndarray:
        b = np.ascontiguousarray(block, dtype=self.real)
        bh, bw = b.shape
        o = np.empty((bh, bw), dtype=np.complex128 if self.real == np.float64 else np.complex64)
        self.lib.orc_fft2d_real_forward(_f(b), bw, bh, _f(o))
        return o

    def fft2d_inverse_half(self, half: np.ndarray, bw: int) -> np.ndarray:
        h = np.ascontiguousarray(half, dtype=np.complex128 if self.real == np.float64 else np.complex64)
        bh = h.shape[0]
        o = np.empty((bh, bw), dtype=self.real)
        self.lib.orc_fft2d_real_inverse_half(_f(h), bw, bh, _f(o))
        return o

    # -- whole clip
    def run_clip(self, sp: Script, planes: list, bits: int, sample_type: int = 0,
                 color_family: int = 1, ssw: int = 1, ssh: int = 1, order=None,
                 want_float: bool = False):
        """planes: list of (nframes, h, w) arrays (1 for gray, 3 for yuv/rgb).
        Returns (list of output arrays, list of float arrays or None)."""
        nplanes = len(planes)
        nframes = planes[0].shape[0]
        H, W = planes[0].shape[1:]
        if order is None:
            order = list(range(nframes))
        outs = [np.zeros_like(p) for p in planes]
        fl = [np.zeros(p.shape, dtype=self.real) for p in planes] if want_float else None
        ins = (C.c_void_p * (nframes * nplanes))()
        ous = (C.c_void_p * (nframes * nplanes))()
        fls = (C.c_void_p * (nframes * nplanes))()
        keep = []
        for f in range(nframes):
            for q in range(nplanes):
                a = np.ascontiguousarray(planes[q][f])
                keep.append(a)
                ins[f * nplanes + q] = a.ctypes.data
                ous[f * nplanes + q] = outs[q][f].ctypes.data
                fls[f * nplanes + q] = fl[q][f].ctypes.data if want_float else None
        ordr = (C.c_int * len(order))(*order)
        err = C.c_char_p()
        rc = self.lib.orc_run_clip(C.byref(sp), color_family, W, H, ssw, ssh, bits,
                                   sample_type, nframes, ins, ous,
                                   fls if want_float else None, ordr, len(order),
                                   C.byref(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return outs, fl


class Plane:
    """Plane-level oracle object (PlaneRef restatement)."""

    def __init__(self, orc: Oracle, sp: Script, w: int, h: int, bits: int, sample_type=0,
                 chroma=0):
        self.o = orc
        ip, err = orc.validate(sp, w, h, bits)
        if ip is None:
            raise ValueError(err)
        self.ip = ip
        self.w, self.h = w, h
        self.h_ = orc.lib.orc_plane_new(C.byref(ip), w, h, bits, sample_type, chroma)
        L = orc.lib
        self.nb = L.orc_plane_nblocks(self.h_)
        self.fbw = L.orc_plane_fbw(self.h_)
        self.ncols = L.orc_plane_ncols(self.h_)
        self.chroma_norm = L.orc_plane_chroma_norm(self.h_)

    def __del__(self):
        try:
            self.o.lib.orc_plane_free(self.h_)
        except Exception:
            pass

    def tables(self):
        L, bw, bh = self.o.lib, self.ip.bw, self.ip.bh
        t = {}
        for name, which in (("winAna", 0), ("winSynA", 1), ("winSynB", 2)):
            a = np.zeros(bw * bh, np.float32)
            L.orc_plane_win(self.h_, which, _f(a))
            t[name] = a.reshape(bh, bw)
        ps = self.fbw * bh
        a = np.zeros(ps, np.float32)
        t["pattern"] = a.reshape(bh, self.fbw) if L.orc_plane_pattern(self.h_, _f(a)) else None
        a = np.zeros(ps, np.float32)
        t["sharpen"] = a.reshape(bh, self.fbw) if L.orc_plane_sharpen(self.h_, _f(a)) else None
        a = np.zeros(2 * ps, np.float32)
        t["degrid"] = a.view(np.complex64).reshape(bh, self.fbw) \
            if L.orc_plane_degrid(self.h_, _f(a)) else None
        bx = np.zeros(self.nb, np.int32); L.orc_plane_bx(self.h_, _f(bx)); t["bx"] = bx
        by = np.zeros(self.nb, np.int32); L.orc_plane_by(self.h_, _f(by)); t["by"] = by
        par = np.zeros(self.nb, np.uint8); L.orc_plane_par(self.h_, _f(par)); t["par"] = par
        return t

    def upload(self, raw: np.ndarray) -> np.ndarray:
        raw = np.ascontiguousarray(raw)
        out = np.zeros((self.h, self.w), np.float32)
        self.o.lib.orc_upload(self.h_, _f(raw), raw.strides[0], _f(out))
        return out

    def download(self, fl: np.ndarray, dtype) -> np.ndarray:
        fl = np.ascontiguousarray(fl, dtype=np.float32)
        out = np.zeros((self.h, self.w), dtype)
        self.o.lib.orc_download(self.h_, _f(fl), _f(out), out.strides[0])
        return out

    def spectra(self, plane: np.ndarray):
        plane = np.ascontiguousarray(plane, dtype=np.float32)
        n = self.nb * self.fbw * self.ip.bh
        re = np.zeros(n, self.o.real)
        im = np.zeros(n, self.o.real)
        self.o.lib.orc_compute_spectra(self.h_, _f(plane), _f(re), _f(im))
        return (re + 1j * im).reshape(self.nb, self.ip.bh, self.fbw)

    def build_degrid(self):
        self.o.lib.orc_build_degrid(self.h_)

    def filter_synth(self, window) -> np.ndarray:
        res = [np.ascontiguousarray(w.real.ravel(), dtype=self.o.real) for w in window]
        ims = [np.ascontiguousarray(w.imag.ravel(), dtype=self.o.real) for w in window]
        pr = (C.c_void_p * len(window))(*[r.ctypes.data for r in res])
        pi = (C.c_void_p * len(window))(*[i.ctypes.data for i in ims])
        out = np.zeros((self.h, self.w), self.o.real)
        self.o.lib.orc_filter_synth(self.h_, pr, pi, len(window), _f(out))
        return out


class Ref:
    """The reference's own CPU backend (oracle/_ref/libfft3dref.so), or None."""

    PATH = ORACLE_DIR / "_ref" / "libfft3dref.so"

    @classmethod
    def load(cls):
        if not cls.PATH.exists():
            return None
        return cls()

    def __init__(self):
        self.lib = C.CDLL(str(self.PATH))
        L = self.lib
        L.ref_validate.argtypes = [C.POINTER(Script), C.c_int, C.c_int, C.c_int,
                                   C.POINTER(Internal), C.POINTER(C.c_char_p)]
        L.ref_plane_new.argtypes = [C.POINTER(Script), C.c_int, C.c_int, C.c_int, C.c_int,
                                    C.c_int]
        L.ref_plane_new.restype = C.c_void_p
        L.ref_plane_free.argtypes = [C.c_void_p]
        L.ref_plane_nblocks.argtypes = [C.c_void_p]
        L.ref_plane_fbw.argtypes = [C.c_void_p]
        L.ref_plane_table.argtypes = [C.c_void_p, C.c_int, C.c_void_p]
        L.ref_build_degrid.argtypes = [C.c_void_p]
        L.ref_compute_spectra.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p]
        L.ref_filter_synth.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p]
        L.ref_fft1d.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_int]
        L.ref_fft2d_real_forward.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.ref_fft2d_real_inverse_half.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_void_p]
        L.ref_round_pow2_clamp.argtypes = [C.c_int]
        L.ref_run_clip.argtypes = [C.POINTER(Script), C.c_int, C.c_int, C.c_int, C.c_int,
                                   C.c_int, C.c_int, C.c_int, C.c_int, C.c_void_p, C.c_void_p,
                                   C.c_void_p, C.c_void_p, C.c_int, C.POINTER(C.c_char_p)]

    def validate(self, sp: Script, w, h, bits):
        out = Internal()
        err = C.c_char_p()
        ok = self.lib.ref_validate(C.byref(sp), w, h, bits, C.byref(out), C.byref(err))
        return (out, None) if ok else (None, err.value.decode())

    def fft1d(self, x, inverse=False):
        a = np.ascontiguousarray(x, dtype=np.complex64)
        o = np.empty_like(a)
        self.lib.ref_fft1d(len(a), _f(a), _f(o), int(inverse))
        return o

    def fft2d_forward(self, block):
        b = np.ascontiguousarray(block, dtype=np.float32)
        bh, bw = b.shape
        o = np.empty((bh, bw), np.complex64)
        self.lib.ref_fft2d_real_forward(_f(b), bw, bh, _f(o))
        return o

    def fft2d_inverse_half(self, half, bw):
        h = np.ascontiguousarray(half, dtype=np.complex64)
        bh = h.shape[0]
        o = np.empty((bh, bw), np.float32)
        self.lib.ref_fft2d_real_inverse_half(_f(h), bw, bh, _f(o))
        return o

    def run_clip(self, sp, planes, bits, sample_type=0, color_family=1, ssw=1, ssh=1,
                 order=None, want_float=False):
        nplanes = len(planes)
        nframes = planes[0].shape[0]
        H, W = planes[0].shape[1:]
        if order is None:
            order = list(range(nframes))
        outs = [np.zeros_like(p) for p in planes]
        fl = [np.zeros(p.shape, np.float32) for p in planes] if want_float else None
        ins = (C.c_void_p * (nframes * nplanes))()
        ous = (C.c_void_p * (nframes * nplanes))()
        fls = (C.c_void_p * (nframes * nplanes))()
        keep = []
        for f in range(nframes):
            for q in range(nplanes):
                a = np.ascontiguousarray(planes[q][f])
                keep.append(a)
                ins[f * nplanes + q] = a.ctypes.data
                ous[f * nplanes + q] = outs[q][f].ctypes.data
                fls[f * nplanes + q] = fl[q][f].ctypes.data if want_float else None
        ordr = (C.c_int * len(order))(*order)
        err = C.c_char_p()
        rc = self.lib.ref_run_clip(C.byref(sp), color_family, W, H, ssw, ssh, bits,
                                   sample_type, nframes, ins, ous,
                                   fls if want_float else None, ordr, len(order),
                                   C.byref(err))
        if rc != 0:
            raise RuntimeError(err.value.decode())
        return outs, fl


def kind_info(kind: str):
    """(bits, sample_type) for a synth kind."""
    return {"u8": (8, 0), "u10": (10, 0), "u16": (16, 0), "f32": (32, 1)}[kind]
