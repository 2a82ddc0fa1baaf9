"""ctypes binding of the C ABI in include/fft3d_hip.h (libfft3d_hip.so).

Mirrors the reference's GPU backend object (GpuRef, src/cl/gpu.h:16-88) as the
class `Plane`; the clip-level scheduling of its VapourSynth filter
(src/fft3dcl.cpp:311-526) is `fft3dcl_amd.filter.FFT3DCL`.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
DEFAULT_LIB = PKG / "libfft3d_hip.so"

FFT3D_OK = 0
ERRORS = {-1: "EINVAL", -2: "EDEVICE", -3: "ENOMEM", -4: "ELAUNCH", -5: "EMISSING",
          -6: "EUNSUPPORTED"}

PARAM_FIELDS = [
    ("sigma", C.c_float), ("beta", C.c_float), ("bw", C.c_int), ("bh", C.c_int),
    ("bt", C.c_int), ("sharpen", C.c_float), ("plane", C.c_int), ("mode", C.c_int),
    ("bordersize", C.c_int), ("precision", C.c_int), ("degrid", C.c_float),
    ("scutoff", C.c_float), ("svr", C.c_float), ("smin", C.c_float), ("smax", C.c_float),
    ("kratio", C.c_float), ("ow", C.c_int), ("oh", C.c_int), ("wintype", C.c_int),
    ("interlaced", C.c_int), ("sigma2", C.c_float), ("sigma3", C.c_float),
    ("sigma4", C.c_float), ("oldfft", C.c_int),
]
PARAM_NAMES = [n for n, _ in PARAM_FIELDS]


class Params(C.Structure):
    _fields_ = PARAM_FIELDS


class PlaneDesc(C.Structure):
    _fields_ = [("width", C.c_int), ("height", C.c_int), ("bits", C.c_int),
                ("sample_type", C.c_int), ("chroma", C.c_int)]


class PlaneInfo(C.Structure):
    _fields_ = [("nb", C.c_int), ("bw", C.c_int), ("bh", C.c_int), ("ow", C.c_int),
                ("oh", C.c_int), ("mode", C.c_int), ("fbw", C.c_int), ("bm", C.c_int),
                ("bt", C.c_int), ("current", C.c_int), ("ncols", C.c_int),
                ("nrows", C.c_int), ("blocks_per_wg", C.c_int),
                ("spectrum_bytes", C.c_size_t), ("device_bytes", C.c_size_t),
                ("cached_frames", C.c_int), ("dft", C.c_int), ("spectra_fp16", C.c_int),
                ("band_rows", C.c_int), ("sweep_occupancy", C.c_int), ("tuned", C.c_int)]


class KernelTime(C.Structure):
    _fields_ = [("name", C.c_char * 32), ("launches", C.c_int), ("total_ms", C.c_double)]


SOURCE_FN = C.CFUNCTYPE(C.c_void_p, C.c_void_p, C.c_int)


class WorkItem(C.Structure):
    _fields_ = [("plane", C.c_void_p), ("frame", C.c_int), ("num_frames", C.c_int),
                ("source", SOURCE_FN), ("user", C.c_void_p), ("src_on_device", C.c_int),
                ("dst", C.c_void_p), ("bytes", C.c_size_t), ("dst_on_device", C.c_int),
                ("stream", C.c_void_p), ("status", C.c_int)]


EXPORTS = [
    "fft3d_plane_set_profiling", "fft3d_plane_get_profile", "fft3d_plane_set_option",
    "fft3d_params_default", "fft3d_dev_init", "fft3d_plane_create", "fft3d_plane_destroy",
    "fft3d_plane_error", "fft3d_plane_has", "fft3d_plane_forward", "fft3d_plane_process",
    "fft3d_plane_prune", "fft3d_plane_clear", "fft3d_live_instances", "fft3d_live_bytes",
    "fft3d_plane_get_info", "fft3d_plane_get_table", "fft3d_plane_read_spectrum",
    "fft3d_plane_forward_process", "fft3d_plane_prefetch", "fft3d_step",
    "fft3d_plane_process_noalloc",
]


class FFT3DError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(msg)
        self.code = code


_LIBS: dict[str, C.CDLL] = {}


def load(path: str | Path | None = None) -> C.CDLL:
    """Loads (once) and types the C ABI library. Raises if it is missing: the
    GPU path has no fallback."""
    path = str(path or DEFAULT_LIB)
    if path in _LIBS:
        return _LIBS[path]
    if not Path(path).exists():
        raise FileNotFoundError(
            f"{path} is not built; run `python -m fft3dcl_amd.build` (the GPU path has "
            "no CPU fallback)")
    L = C.CDLL(path)
    vp, i, sz = C.c_void_p, C.c_int, C.c_size_t
    L.fft3d_params_default.argtypes = [C.POINTER(Params)]
    L.fft3d_params_default.restype = None
    L.fft3d_dev_init.argtypes = [i, C.c_char_p, sz]
    L.fft3d_plane_create.argtypes = [C.POINTER(Params), C.POINTER(PlaneDesc),
                                     C.POINTER(vp), C.c_char_p, sz]
    L.fft3d_plane_destroy.argtypes = [vp]
    L.fft3d_plane_destroy.restype = None
    L.fft3d_plane_error.argtypes = [vp]
    L.fft3d_plane_error.restype = C.c_char_p
    L.fft3d_plane_has.argtypes = [vp, i]
    L.fft3d_plane_forward.argtypes = [vp, i, vp, sz, i, vp]
    L.fft3d_plane_process.argtypes = [vp, C.POINTER(C.c_int), i, vp, sz, i, vp]
    L.fft3d_plane_forward_process.argtypes = [vp, i, vp, sz, i, C.POINTER(C.c_int), i, vp, sz,
                                              i, vp]
    L.fft3d_plane_prefetch.argtypes = [vp, i, vp, sz, i]
    L.fft3d_plane_process_noalloc.argtypes = [vp, C.POINTER(C.c_int), i, vp, sz, vp]
    L.fft3d_step.argtypes = [C.POINTER(WorkItem), i]
    L.fft3d_plane_prune.argtypes = [vp, i, i]
    L.fft3d_plane_prune.restype = None
    L.fft3d_plane_clear.argtypes = [vp]
    L.fft3d_plane_clear.restype = None
    L.fft3d_live_instances.argtypes = []
    L.fft3d_live_bytes.argtypes = []
    L.fft3d_live_bytes.restype = sz
    L.fft3d_plane_get_info.argtypes = [vp, C.POINTER(PlaneInfo)]
    L.fft3d_plane_get_table.argtypes = [vp, i, vp]
    L.fft3d_plane_read_spectrum.argtypes = [vp, i, vp]
    L.fft3d_plane_set_profiling.argtypes = [vp, i]
    L.fft3d_plane_set_option.argtypes = [vp, i, i]
    L.fft3d_plane_get_profile.argtypes = [vp, C.POINTER(KernelTime), i]
    _LIBS[path] = L
    return L


def make_params(lib: C.CDLL | None = None, **kw) -> Params:
    """Script parameters with the reference's key-dependent defaults
    (src/fft3dcl.cpp:171-195): degrid by mode, ow/oh by bw/bh, sigma2/3/4 by
    sigma."""
    p = Params()
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
            raise TypeError(f"FFT3DCL: unknown parameter '{k}'")
        d[k] = v
    for k, v in d.items():
        setattr(p, k, v)
    return p


def _ptr(x) -> int:
    """Address of a numpy array or a torch tensor (device or host)."""
    if isinstance(x, np.ndarray):
        return x.ctypes.data
    return int(x.data_ptr())


def _buf(x, need: int, itemsize: int, on_device: bool, what: str) -> tuple[int, int]:
    """(address, byte length) of a plane buffer handed to the C ABI, after
    checking it: a C-contiguous numpy array or torch tensor whose elements
    have the plane's sample size, at least `need` bytes long, on the device
    exactly when `on_device`. Raises ValueError otherwise, so that a short or
    strided buffer never reaches a copy or a kernel."""
    if isinstance(x, np.ndarray):
        if on_device:
            raise ValueError(f"{what}: a numpy array is host memory, but on_device is set")
        if not x.flags["C_CONTIGUOUS"]:
            raise ValueError(f"{what}: the array is not C-contiguous")
        size, nbytes = x.dtype.itemsize, x.nbytes
    elif hasattr(x, "data_ptr") and hasattr(x, "element_size"):
        if bool(getattr(x, "is_cuda", False)) != bool(on_device):
            raise ValueError(f"{what}: on_device does not match the tensor's device")
        if not x.is_contiguous():
            raise ValueError(f"{what}: the tensor is not contiguous")
        size, nbytes = x.element_size(), x.element_size() * x.numel()
    else:
        raise ValueError(f"{what}: expected a numpy array or a torch tensor, got "
                         f"{type(x).__name__}")
    if size != itemsize:
        raise ValueError(f"{what}: {size}-byte elements, the plane has {itemsize}-byte samples")
    if nbytes < need:
        raise ValueError(f"{what}: {nbytes} bytes, the plane needs {need}")
    return _ptr(x), nbytes


class Plane:
    """One processed plane of one filter instance (the GpuRef counterpart)."""

    def __init__(self, params: Params, width: int, height: int, bits: int,
                 sample_type: int = 0, chroma: bool = False, lib=None):
        self.lib = load(lib)
        self.desc = PlaneDesc(width, height, bits, sample_type, int(bool(chroma)))
        self.bps = 4 if sample_type == 1 else (1 if bits <= 8 else 2)
        self.dtype = np.float32 if sample_type == 1 else (np.uint8 if bits <= 8 else np.uint16)
        h = C.c_void_p()
        err = C.create_string_buffer(512)
        rc = self.lib.fft3d_plane_create(C.byref(params), C.byref(self.desc), C.byref(h), err,
                                         512)
        if rc != FFT3D_OK:
            raise FFT3DError(rc, err.value.decode())
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.fft3d_plane_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def error(self) -> str:
        return self.lib.fft3d_plane_error(self.h).decode()

    def _check(self, rc: int):
        if rc != FFT3D_OK:
            raise FFT3DError(rc, f"{ERRORS.get(rc, rc)}: {self.error}")

    @property
    def info(self) -> PlaneInfo:
        i = PlaneInfo()
        self._check(self.lib.fft3d_plane_get_info(self.h, C.byref(i)))
        return i

    @property
    def nbytes(self) -> int:
        return self.desc.width * self.desc.height * self.bps

    def has(self, fn: int) -> bool:
        return bool(self.lib.fft3d_plane_has(self.h, fn))

    def buffer(self, x, on_device: bool, what: str) -> tuple[int, int]:
        """(address, bytes) of a checked plane buffer (see _buf)."""
        return _buf(x, self.nbytes, self.bps, on_device, what)

    def forward(self, fn: int, src, on_device: bool = False, stream: int | None = None):
        ptr, n = self.buffer(src, on_device, "forward source")
        self._check(self.lib.fft3d_plane_forward(self.h, fn, ptr, n, int(on_device), stream))

    def process(self, window, dst, on_device: bool = False, stream: int | None = None):
        w = (C.c_int * len(window))(*window)
        ptr, n = self.buffer(dst, on_device, "process output")
        try:
            self._check(self.lib.fft3d_plane_process(self.h, w, len(window), ptr, n,
                                                     int(on_device), stream))
        finally:
            self._pending = None

    def process_noalloc(self, window, dst, stream: int | None = None):
        """process() into a device buffer through fft3d_plane_process_noalloc:
        no device allocation; the kernels' scratch is the library's static
        array, shared by every such call on the GPU (overlapping calls race)."""
        w = (C.c_int * len(window))(*window)
        ptr, n = self.buffer(dst, True, "no-alloc process output")
        self._check(self.lib.fft3d_plane_process_noalloc(self.h, w, len(window), ptr, n, stream))

    def forward_process(self, fn: int, src, window, dst, src_on_device: bool = False,
                        dst_on_device: bool = False, stream: int | None = None):
        """forward(fn, src) then process(window, dst) in one call; with
        OPT_FUSED_ANALYSIS set, frame fn is transformed inside the sweep."""
        w = (C.c_int * len(window))(*window)
        sp, sn = self.buffer(src, src_on_device, "forward_process source")
        dp, dn = self.buffer(dst, dst_on_device, "forward_process output")
        try:
            self._check(self.lib.fft3d_plane_forward_process(
                self.h, fn, sp, sn, int(src_on_device), w, len(window), dp, dn,
                int(dst_on_device), stream))
        finally:
            self._pending = None

    def prefetch(self, fn: int, src, on_device: bool = False):
        """Requests plane fn in the ring; the next process or forward_process
        call carries it out alongside its own work. A host src is kept alive
        until that call returns; a device src must stay unchanged until the
        frame is used."""
        if bool(getattr(src, "is_cuda", False)) != bool(on_device):
            raise ValueError("prefetch: on_device does not match the buffer")
        ptr, n = self.buffer(src, on_device, "prefetch source")
        self._check(self.lib.fft3d_plane_prefetch(self.h, fn, ptr, n, int(on_device)))
        self._pending = src

    def prune(self, lo: int, hi: int):
        self.lib.fft3d_plane_prune(self.h, lo, hi)

    def clear(self):
        self.lib.fft3d_plane_clear(self.h)

    def table(self, which: int) -> np.ndarray | None:
        n = self.lib.fft3d_plane_get_table(self.h, which, None)
        if n <= 0:
            return None
        a = np.zeros(n, np.float32)
        self.lib.fft3d_plane_get_table(self.h, which, a.ctypes.data)
        return a

    OPT_SWEEP, OPT_BAND_ROWS, OPT_FUSED_ANALYSIS, OPT_RESERVE_SPECTRA = 1, 2, 3, 4
    OPT_SPECTRA_FP16 = 5
    OPT_HOST_STAGING, OPT_STAGING_CHUNK = 6, 7
    OPT_DFT = 8  # 0 register FFTs, 1 MFMA with FP16 hi+lo operands, 2 with BF16
    OPT_SWEEP_OCC = 9  # 0 tuned, 2..4 sweep workgroups per CU

    def set_option(self, option: int, value: int):
        self._check(self.lib.fft3d_plane_set_option(self.h, option, int(value)))

    def set_profiling(self, on: bool):
        self._check(self.lib.fft3d_plane_set_profiling(self.h, int(on)))

    def profile(self) -> dict:
        """{kernel: (launches, total_ms)} since the last call (waits for the GPU)."""
        buf = (KernelTime * 8)()
        n = self.lib.fft3d_plane_get_profile(self.h, buf, 8)
        if n < 0:
            self._check(n)
        return {buf[i].name.decode(): (buf[i].launches, buf[i].total_ms) for i in range(n)}

    def read_spectrum(self, fn: int) -> np.ndarray:
        i = self.info
        a = np.zeros((i.nb, i.bh, i.fbw, 2), np.float32)
        self._check(self.lib.fft3d_plane_read_spectrum(self.h, fn, a.ctypes.data))
        return a[..., 0] + 1j * a[..., 1]


def live_instances(lib=None) -> int:
    return load(lib).fft3d_live_instances()


def live_bytes(lib=None) -> int:
    return load(lib).fft3d_live_bytes()


def dev_init(device: int = 0, lib=None):
    err = C.create_string_buffer(256)
    rc = load(lib).fft3d_dev_init(device, err, 256)
    if rc != FFT3D_OK:
        raise FFT3DError(rc, err.value.decode())


class StepRequest:
    """One work item of `step`: output frame `frame` of `plane` into `dst`;
    `source(fn)` gives the plane of any frame the window needs."""

    def __init__(self, plane: Plane, frame: int, num_frames: int, source, dst,
                 on_device: bool = False, stream: int | None = None):
        self.plane, self.frame, self.num_frames = plane, frame, num_frames
        self.source, self.dst, self.on_device, self.stream = source, dst, on_device, stream


def step(requests: list, lib=None) -> list[int]:
    """Runs the requests through one fft3d_step call (the serving path's
    batched host step) and returns each item's status (0 = FFT3D_OK). The
    buffers the sources hand out are kept alive until the call returns."""
    L = load(lib)
    keep = []

    def cb(user, fn):
        # one callback for every item; `user` is the item's index + 1
        r = requests[user - 1]
        buf = r.source(fn)
        if buf is None:
            return None
        try:
            ptr, _ = r.plane.buffer(buf, r.on_device, "step source")
        except ValueError:
            return None  # the item fails with "source gave no plane"
        keep.append(buf)
        return ptr

    c = SOURCE_FN(cb)
    items = (WorkItem * max(1, len(requests)))()
    for k, r in enumerate(requests):
        dptr, dn = r.plane.buffer(r.dst, r.on_device, "step output")
        items[k] = WorkItem(r.plane.h, r.frame, r.num_frames, c, k + 1, int(r.on_device),
                            dptr, dn, int(r.on_device), r.stream, 0)
    L.fft3d_step(items, len(requests))
    return [items[k].status for k in range(len(requests))]
