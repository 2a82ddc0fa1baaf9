"""CPU tests of the C ABI (include/fft3d_hip.h).

* The gfx950 library loads and exports every entry point the header declares
  (no compute call is made without a GPU).
* The same C ABI and kernels, compiled against the CPU emulation of HIP
  (tests/emu), agree with the oracle: this checks the kernels' index
  arithmetic, LDS layouts, packing and the host scheduling on any machine.
* Host logic: parameter validation messages, plane routing, tables.
"""
import ctypes as C
import re
from pathlib import Path

import numpy as np
import pytest

from fft3dcl_amd import capi
from fft3dcl_amd.filter import FFT3DCL, route_planes
from fft3dcl_amd.synth import make_clip
from helpers import lsb_diff, max_abs, mt19937_planes
from oracle_lib import Plane as OPlane, script, kind_info

ROOT = Path(__file__).resolve().parent.parent
HEADER = ROOT / "include" / "fft3d_hip.h"


def declared_symbols():
    txt = HEADER.read_text()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fft3d_\w+)\s*\(", txt)))


def test_header_declares_the_binding_exports():
    assert declared_symbols() == sorted(capi.EXPORTS)


def test_hip_library_exports_every_declared_symbol(hip_lib):
    lib = C.CDLL(hip_lib)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_emu_library_exports_every_declared_symbol(emu_lib):
    lib = C.CDLL(emu_lib)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_params_default_matches_reference_defaults(emu_lib):
    lib = capi.load(emu_lib)
    p = capi.Params()
    lib.fft3d_params_default(C.byref(p))
    want = capi.make_params()
    for n in capi.PARAM_NAMES:
        assert getattr(p, n) == getattr(want, n), n


@pytest.mark.parametrize("kw,msg", [
    (dict(bt=7), "FFT3DCL: valid modes for bt are -1,0,1,2,3,4"),
    (dict(mode=2), "FFT3DCL: mode 2 is not implemented yet in this port (use mode 0 or 1)"),
    (dict(interlaced=1), "FFT3DCL: interlaced is not implemented yet in this port"),
    (dict(bw=256), "FFT3DCL: GPU setup failed: block too large for device local memory (max 64x64)"),
])
def test_create_errors(emu_lib, kw, msg):
    with pytest.raises(capi.FFT3DError) as e:
        capi.Plane(capi.make_params(**kw), 64, 48, 8, lib=emu_lib)
    assert str(e.value) == msg


def test_format_errors(emu_lib):
    with pytest.raises(capi.FFT3DError, match="integer bit depths 8-16 only"):
        capi.Plane(capi.make_params(), 64, 48, 7, lib=emu_lib)
    with pytest.raises(capi.FFT3DError, match="only 32-bit float clips"):
        capi.Plane(capi.make_params(), 64, 48, 16, sample_type=1, lib=emu_lib)


def test_plane_routing():
    assert route_planes(1, 0) == [0] and route_planes(1, 4) == [0]
    assert route_planes(3, 1) == [1, 2] and route_planes(3, 3) == [1, 2]
    assert route_planes(3, 4) == [0, 1, 2] and route_planes(2, 0) == [0, 1, 2]
    with pytest.raises(ValueError, match="cannot set chroma processing for a greyscale clip"):
        route_planes(1, 1)
    with pytest.raises(ValueError, match=r"plane must be 0 \(luma\), 1 \(chroma\) or 4 \(all\)"):
        route_planes(3, 5)


def test_device0_is_rejected(emu_lib):
    planes = make_clip(1, 64, 48, 2, "u8")
    with pytest.raises(ValueError, match="device=0"):
        FFT3DCL(planes, bits=8, device=0, lib=emu_lib)


@pytest.mark.parametrize("kw", [dict(), dict(sharpen=0.5, sigma2=3.0, sigma3=1.0, sigma4=2.0),
                                dict(mode=0, bw=16, bh=32), dict(wintype=1, ow=8, oh=12)])
def test_product_tables_match_oracle(oracle, emu_lib, kw):
    sp = script(**kw)
    pl = capi.Plane(capi.make_params(**kw), 120, 90, 8, lib=emu_lib)
    op = OPlane(oracle, sp, 120, 90, 8)
    op.build_degrid()
    t = op.tables()
    assert np.array_equal(pl.table(0), t["winAna"].ravel())
    if sp.mode == 1:
        assert np.array_equal(pl.table(1), t["winSynA"].ravel())
        assert np.array_equal(pl.table(2), t["winSynB"].ravel())
    if t["pattern"] is not None:
        assert np.array_equal(pl.table(3), t["pattern"].ravel())
    if t["sharpen"] is not None:
        assert np.array_equal(pl.table(4), t["sharpen"].ravel())
    if sp.degrid != 0.0:
        g = pl.table(5).view(np.complex64)
        assert np.abs(g - t["degrid"].ravel()).max() < 1e-6


@pytest.mark.parametrize("kw", [dict(sharpen=0.4), dict(sharpen=0.7, svr=0.5, scutoff=0.2),
                                dict(sharpen=1.0, bw=16, bh=64, degrid=0.5),
                                dict(sharpen=0.3, wintype=1, ow=8, oh=12),
                                dict(sharpen=0.3, wintype=2, bw=8, bh=8)])
def test_separable_factors_rebuild_oracle_tables(oracle, emu_lib, kw):
    """The fused sweep builds each bin's sharpen gain as strength - shY[r]*shX[c]
    and its degrid sample as gx[c]*gy[r] (tables 6 and 7). Both must rebuild
    the oracle's full tables (pinned to the reference build) to float
    rounding."""
    sp = script(**kw)
    pl = capi.Plane(capi.make_params(**kw), 150, 110, 8, lib=emu_lib)
    op = OPlane(oracle, sp, 150, 110, 8)
    op.build_degrid()
    t = op.tables()
    rows = pl.table(6).reshape(-1, 4).astype(np.float64)
    cols = pl.table(7).reshape(-1, 4).astype(np.float64)
    assert rows.shape[0] == sp.bh and cols.shape[0] == sp.bw // 2 + 1
    sharpen = rows[:, 3:4] - rows[:, 0:1] * cols[None, :, 0]
    want = t["sharpen"].reshape(sp.bh, -1)
    assert np.abs(sharpen - want).max() < 1e-6
    gs = (cols[None, :, 1] + 1j * cols[None, :, 2]) * (rows[:, 1:2] + 1j * rows[:, 2:3])
    assert np.abs(gs - t["degrid"].reshape(sp.bh, -1)).max() < 1e-6
    pl.close()


@pytest.mark.parametrize("bw,bh", [(4, 4), (8, 16), (32, 32), (64, 64), (32, 8)])
def test_emu_spectrum_matches_oracle(oracle, emu_lib, bw, bh):
    planes = make_clip(2, 90, 70, 1, "f32")
    pl = capi.Plane(capi.make_params(bw=bw, bh=bh), 90, 70, 32, sample_type=1, lib=emu_lib)
    pl.forward(0, np.ascontiguousarray(planes[0][0]))
    got = pl.read_spectrum(0)
    op = OPlane(oracle, script(bw=bw, bh=bh), 90, 70, 32, sample_type=1)
    want = op.spectra(planes[0][0])
    assert np.abs(got - want).max() < 2e-5 * np.abs(want).max()


EMU_CASES = [
    ("u8", dict(bt=3)), ("u8", dict(bt=-1, sharpen=0.8)), ("u10", dict(bt=4, sharpen=0.4)),
    ("u8", dict(bt=2)), ("u8", dict(bt=1, sigma2=3.0, sigma3=1.5, sigma4=1.0)),
    ("f32", dict(bt=0, kratio=1.5)), ("u8", dict(bt=3, mode=0)),
    ("u10", dict(bt=3, bw=16, bh=16)), ("u16", dict(bt=3, bw=64, bh=32, ow=24, oh=8)),
]


@pytest.mark.parametrize("kind,kw", EMU_CASES)
def test_emu_clip_matches_oracle(oracle, emu_lib, kind, kw):
    bits, st = kind_info(kind)
    planes = make_clip(3, 80, 56, 5, kind)
    order = [0, 1, 2, 4, 3]
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, order=order)
    got = FFT3DCL(planes, bits=bits, sample_type=st, order=order, lib=emu_lib, **kw)
    if kind == "f32":
        assert max_abs(got[0], want[0]) < 1e-5
    else:
        assert lsb_diff(got[0], want[0]) <= 1


def test_emu_yuv_plane4_matches_oracle(oracle, emu_lib):
    planes = make_clip(4, 96, 64, 4, "u8", color="yuv")
    want, _ = oracle.run_clip(script(plane=4, bt=3), planes, 8, 0, color_family=3)
    got = FFT3DCL(planes, color_family=3, bits=8, plane=4, bt=3, lib=emu_lib)
    for g, w in zip(got, want):
        assert lsb_diff(g, w) <= 1


def test_emu_gpu_test_inputs(oracle, emu_lib):
    # tools/gpu_test.cpp: 256x192 mt19937(42) planes, float domain, PASS < 2e-4
    for bt in (1, 2, 3, 4, 0):
        pl = mt19937_planes(42, 256, 192, max(bt, 1))
        planes = [np.stack(pl)]
        _, want = oracle.run_clip(script(bt=bt), planes, 32, 1, want_float=True)
        got = FFT3DCL(planes, bits=32, sample_type=1, bt=bt, lib=emu_lib)
        n = 0 if bt < 2 else (2 if bt == 4 else 1)
        assert max_abs(got[0][n], want[0][n]) < 2e-4


def test_emu_ring_prune_and_counts(emu_lib):
    planes = make_clip(5, 64, 48, 8, "u8")
    inst_before = capi.live_instances(emu_lib)
    from fft3dcl_amd.filter import Instance
    inst = Instance(1, 64, 48, 8, 0, 8, bt=4, lib=emu_lib)
    srcs = [lambda fn: np.ascontiguousarray(planes[0][fn])]
    out = np.zeros((48, 64), np.uint8)
    for n in range(8):
        inst.get_frame(n, srcs, [out])
        # keeps [n-current-1, n-current+bt] (src/fft3dcl.cpp:462), i.e. <= bt+2
        assert inst.objs[0].info.cached_frames <= 6
    assert capi.live_instances(emu_lib) == inst_before + 1
    inst.close()
    assert capi.live_instances(emu_lib) == inst_before


@pytest.mark.parametrize("bt", [0, 1, -1, 2, 4])
def test_emu_ring_stays_bounded_over_long_runs(emu_lib, bt):
    """Over 40 frames the ring never holds more than the window needs: bt >= 2
    keeps at most bt + 2 spectra; bt 1, -1 and Kalman (bt=0, whose state is
    not in the ring) hold none after a frame, and device memory stays flat.
    The reference's bt=0 ring grows by one spectrum per frame
    (src/fft3dcl.cpp:458-466)."""
    from fft3dcl_amd.filter import Instance
    n = 40
    planes = make_clip(6, 64, 48, n, "u8", scene_cut=20)
    inst = Instance(1, 64, 48, 8, 0, n, bt=bt, sigma2=3.0, lib=emu_lib)
    srcs = [lambda fn: np.ascontiguousarray(planes[0][fn])]
    out = np.zeros((48, 64), np.uint8)
    dev = []
    for k in range(n):
        inst.get_frame(k, srcs, [out])
        info = inst.objs[0].info
        assert info.cached_frames <= (bt + 2 if bt >= 2 else 0), (k, info.cached_frames)
        dev.append(info.device_bytes)
    assert max(dev[4:]) == min(dev[4:]), dev
    inst.close()


def test_emu_step_kalman_ring_bounded(emu_lib):
    """fft3d_step on a Kalman plane: no spectrum stays in the ring after an
    item, over a long run, and the output equals the Python host step's."""
    from fft3dcl_amd.filter import Instance
    W, H, n = 96, 64, 30
    clip = make_clip(23, W, H, n, "u8", scene_cut=15)[0]
    src = lambda fn: np.ascontiguousarray(clip[fn])
    inst = Instance(1, W, H, 8, 0, n, bt=0, lib=emu_lib)
    want = np.zeros_like(clip)
    for f in range(n):
        inst.get_frame(f, [src], [want[f]])
    inst.close()
    pl = capi.Plane(capi.make_params(bt=0), W, H, 8, lib=emu_lib)
    got = np.zeros_like(clip)
    for f in range(n):
        assert capi.step([capi.StepRequest(pl, f, n, src, got[f])], lib=emu_lib) == [0]
        assert pl.info.cached_frames == 0
    assert np.array_equal(got, want)
    pl.close()


@pytest.mark.parametrize("kind,kw,W,H,R", [
    ("u8", dict(bt=3), 600, 100, 2),                      # 3 strips x 4 bands
    ("u10", dict(bt=4, sharpen=0.4), 530, 90, 3),         # ragged last strip and band
    ("f32", dict(bt=0, sigma2=3.0, sigma3=2.0, sigma4=1.0), 300, 70, 0),
    ("u8", dict(bt=2, bw=16, bh=16), 700, 60, 5),
    ("u8", dict(bt=3, bw=8, bh=8), 530, 40, 0),
    ("u16", dict(bt=1, bw=64, bh=64), 600, 200, 2),
    ("u8", dict(bt=3, bw=64, bh=32), 700, 100, 1),
    ("u8", dict(bt=-1, bw=4, bh=4), 600, 24, 3),
])
def test_emu_sweep_seams_match_oracle(oracle, emu_lib, kind, kw, W, H, R):
    from fft3dcl_amd.filter import Instance
    bits, st = kind_info(kind)
    planes = make_clip(6, W, H, 4, kind)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    outs = {}
    for sweep in (1, 0):
        inst = Instance(1, W, H, bits, st, 4, lib=emu_lib, **kw)
        inst.objs[0].set_option(capi.Plane.OPT_SWEEP, sweep)
        inst.objs[0].set_option(capi.Plane.OPT_BAND_ROWS, R)
        out = np.zeros_like(planes[0])
        for n in range(4):
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
        inst.close()
        outs[sweep] = out
        if kind == "f32":
            assert max_abs(out, want[0]) < 1e-5
        else:
            assert lsb_diff(out, want[0]) <= 1


@pytest.mark.parametrize("kind,kw,W,H", [
    ("u8", dict(bt=3, precision=2), 300, 90),  # FP32 ring: the fused form runs
    ("u10", dict(bt=4, sharpen=0.4), 270, 70),
    ("f32", dict(bt=0, kratio=1.5), 200, 60),
    ("u8", dict(bt=1, bw=16, bh=64), 300, 140),           # lanes sharing a column
    ("u16", dict(bt=2, bw=64, bh=16, degrid=0.0), 330, 50),
    ("u8", dict(bt=-1, sharpen=0.5, bw=8, bh=8), 120, 40),
])
def test_emu_forward_process_equals_forward_then_process(emu_lib, kind, kw, W, H):
    """fft3d_plane_forward_process (new frame transformed inside the sweep)
    gives the same output frames and ring spectra as forward + process, bit
    for bit, including out-of-order requests."""
    from fft3dcl_amd.filter import Instance
    bits, st = kind_info(kind)
    planes = make_clip(14, W, H, 5, kind)
    order = [0, 1, 2, 4, 3] if kw["bt"] != 0 else list(range(5))
    res = []
    for fuse in (True, False):
        inst = Instance(1, W, H, bits, st, 5, lib=emu_lib, fuse=fuse, **kw)
        out = np.zeros_like(planes[0])
        specs = {}
        for n in order:
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
            specs.update({fn: inst.objs[0].read_spectrum(fn) for fn in range(5)
                          if inst.objs[0].has(fn) and fn not in specs})
        inst.close()
        res.append((out, specs))
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1].keys() == res[1][1].keys()
    for fn in res[0][1]:
        assert np.array_equal(res[0][1][fn], res[1][1][fn])


@pytest.mark.parametrize("chunk,fuse", [(2112, False), (4160, True), (0, False)])
def test_emu_staged_host_transfers(emu_lib, chunk, fuse):
    """Pageable host buffers through the pinned chunks (many chunks with a
    ragged last one, or one chunk split over the copy threads: the plane is
    above 256 KiB) give the same frames as direct copies."""
    from fft3dcl_amd.filter import Instance
    # 264,000 bytes of u16 for the one-chunk case (split over the copy
    # threads above 256 KiB); smaller planes for the many-chunk cases
    W, H = (400, 330) if chunk == 0 else (200, 120)
    planes = make_clip(16, W, H, 3, "u10")
    outs = []
    for staging, ch in ((0, 0), (1, chunk)):
        inst = Instance(1, W, H, 10, 0, 3, lib=emu_lib, fuse=fuse, bt=3, sharpen=0.3)
        pl = inst.objs[0]
        pl.set_option(capi.Plane.OPT_HOST_STAGING, staging)
        pl.set_option(capi.Plane.OPT_STAGING_CHUNK, ch)
        out = np.zeros_like(planes[0])
        for n in range(3):
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
        inst.close()
        outs.append(out)
    assert outs[0].any()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("kw,order,fuse", [
    (dict(bt=4, sharpen=0.4), list(range(6)), False),
    (dict(bt=3), [0, 1, 2, 4, 3, 5], True),                # out of order: prefetch wasted
    (dict(bt=2, bw=16, bh=16), list(range(6)), False),
    (dict(bt=1), list(range(4)), False),                   # one-frame window: no prefetch
])
def test_emu_prefetch_equals_plain(emu_lib, kw, order, fuse):
    """Requests with the next frame prefetched give the same frames and ring
    spectra as plain requests."""
    from fft3dcl_amd.filter import Instance
    W, H = 200, 120
    planes = make_clip(17, W, H, 6, "u10")
    res = []
    for prefetch in (True, False):
        inst = Instance(1, W, H, 10, 0, 6, lib=emu_lib, fuse=fuse, prefetch=prefetch, **kw)
        out = np.zeros_like(planes[0])
        specs = {}
        for n in order:
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
            specs.update({fn: inst.objs[0].read_spectrum(fn) for fn in range(6)
                          if inst.objs[0].has(fn) and fn not in specs})
        inst.close()
        res.append((out, specs))
    assert res[0][0].any()
    assert np.array_equal(res[0][0], res[1][0])
    for fn in res[0][1].keys() & res[1][1].keys():
        assert np.array_equal(res[0][1][fn], res[1][1][fn])


def test_emu_prefetch_calls(emu_lib):
    """A prefetch request is carried out by the next process call, lands the
    same spectrum as forward, and is dropped by a prune that excludes it."""
    planes = make_clip(18, 96, 64, 6, "u8")
    params = capi.make_params(bt=3)
    pl = capi.Plane(params, 96, 64, 8, lib=emu_lib)
    ref = capi.Plane(params, 96, 64, 8, lib=emu_lib)
    out = np.zeros((64, 96), np.uint8)
    for k in range(3):
        pl.forward(k, planes[0][k])
    pl.prefetch(3, planes[0][3])
    assert not pl.has(3)
    pl.process([0, 1, 2], out)
    assert pl.has(3)
    ref.forward(3, planes[0][3])
    want = ref.read_spectrum(3)
    assert np.array_equal(pl.read_spectrum(3), want)
    pl.prefetch(3, planes[0][1])  # already in the ring: no-op
    pl.process([1, 2, 3], out)
    assert np.array_equal(pl.read_spectrum(3), want)
    pl.prefetch(4, planes[0][4])
    pl.prune(1, 3)                # excludes 4: the request goes
    pl.process([1, 2, 3], out)
    assert not pl.has(4)
    pl.prefetch(5, planes[0][5])
    pl.forward(5, planes[0][5])   # forwarded in between: the request finds it there
    pl.process([2, 3, 5], out)
    ref.forward(5, planes[0][5])
    assert np.array_equal(pl.read_spectrum(5), ref.read_spectrum(5))
    # a device source (here the emulation's memory): the analysis reads it
    src4 = np.ascontiguousarray(planes[0][4])
    assert pl.lib.fft3d_plane_prefetch(pl.h, 4, src4.ctypes.data, pl.nbytes, 1) == 0
    pl.process([2, 3, 5], out)
    ref.forward(4, planes[0][4])
    assert np.array_equal(pl.read_spectrum(4), ref.read_spectrum(4))
    rc = pl.lib.fft3d_plane_prefetch(pl.h, 1, planes[0][1].ctypes.data, 10, 0)
    assert rc == -1 and "too small" in pl.error
    with pytest.raises(ValueError):
        class Dev:
            is_cuda = True
        pl.prefetch(1, Dev())
    pl.close()
    ref.close()


@pytest.mark.parametrize("kw,order", [
    (dict(bt=4, sharpen=0.4), [0, 1, 2, 3, 4, 5]),
    (dict(bt=3), [0, 1, 5, 2, 3, 4]),                     # a seek: the window is rebuilt
    (dict(bt=0, sigma2=3.0), [0, 1, 2, 3, 4, 5]),         # Kalman, in order
    (dict(bt=1), [3, 1, 2]),
    (dict(bt=-1, sharpen=0.5), [2, 0]),
])
def test_emu_step_equals_host_step(emu_lib, kw, order):
    """fft3d_step (window, forwards of missing frames, process, prune in the
    library) gives the frames of the Python host step, with the items of two
    planes interleaved in each call."""
    from fft3dcl_amd.filter import Instance
    W, H, n = 120, 80, 6
    clips = [make_clip(20 + c, W, H, n, "u8")[0] for c in range(2)]
    want = []
    for c in range(2):
        inst = Instance(1, W, H, 8, 0, n, lib=emu_lib, **kw)
        out = np.zeros_like(clips[c])
        for f in order:
            inst.get_frame(f, [lambda fn, a=clips[c]: np.ascontiguousarray(a[fn])], [out[f]])
        inst.close()
        want.append(out)
    params = capi.make_params(**kw)
    planes = [capi.Plane(params, W, H, 8, lib=emu_lib) for _ in range(2)]
    got = [np.zeros_like(a) for a in clips]
    for i in range(0, len(order), 2):
        reqs = [capi.StepRequest(planes[c], f, n, lambda fn, a=clips[c]: np.ascontiguousarray(a[fn]),
                                 got[c][f])
                for f in order[i:i + 2] for c in range(2)]
        assert capi.step(reqs, lib=emu_lib) == [0] * len(reqs)
    for c in range(2):
        assert np.array_equal(got[c][order], want[c][order])
        if kw["bt"] >= 2:
            assert planes[c].info.cached_frames <= kw["bt"] + 2
        else:
            assert planes[c].info.cached_frames == 0
        planes[c].close()


def test_emu_step_errors(emu_lib):
    clip = make_clip(22, 96, 64, 3, "u8")[0]
    pl = capi.Plane(capi.make_params(bt=3), 96, 64, 8, lib=emu_lib)
    out = np.zeros_like(clip)
    src = lambda fn: np.ascontiguousarray(clip[fn])
    reqs = [capi.StepRequest(pl, 0, 3, lambda fn: None, out[0]),   # no plane from the source
            capi.StepRequest(pl, 5, 3, src, out[1]),               # frame past the clip
            capi.StepRequest(pl, 1, 3, src, out[1])]               # fine
    st = capi.step(reqs, lib=emu_lib)
    assert st[0] == -1 and st[1] == -1 and st[2] == 0
    assert out[1].any()
    pl.close()


def test_emu_staging_options(emu_lib):
    pl = capi.Plane(capi.make_params(bt=3), 96, 64, 8, lib=emu_lib)
    for opt, bad in ((capi.Plane.OPT_HOST_STAGING, 2), (capi.Plane.OPT_STAGING_CHUNK, 300),
                     (capi.Plane.OPT_STAGING_CHUNK, 128), (capi.Plane.OPT_STAGING_CHUNK, -64)):
        with pytest.raises(capi.FFT3DError) as e:
            pl.set_option(opt, bad)
        assert e.value.code == -1
    pl.set_option(capi.Plane.OPT_STAGING_CHUNK, 1 << 20)
    pl.set_option(capi.Plane.OPT_STAGING_CHUNK, 0)
    pl.close()


def test_emu_forward_process_errors(emu_lib):
    planes = make_clip(15, 96, 64, 4, "u8")
    # the fused analysis runs on the FP32 ring only (8-bit default precision 0: FP16)
    pl = capi.Plane(capi.make_params(bt=3, precision=2), 96, 64, 8, lib=emu_lib)
    pl.set_option(capi.Plane.OPT_FUSED_ANALYSIS, 1)
    with pytest.raises(capi.FFT3DError, match="must be 0 or 1"):
        pl.set_option(capi.Plane.OPT_FUSED_ANALYSIS, 2)
    src = np.ascontiguousarray(planes[0][1])
    out = np.zeros((64, 96), np.uint8)
    with pytest.raises(capi.FFT3DError, match="missing spectra"):
        pl.forward_process(1, src, [0, 1, 2], out)
    assert not pl.has(1)  # nothing was left half-built in the ring
    with pytest.raises(capi.FFT3DError, match="window length"):
        pl.forward_process(1, src, [0, 1], out)
    pl.forward(0, np.ascontiguousarray(planes[0][0]))
    pl.forward(2, np.ascontiguousarray(planes[0][2]))
    pl.forward_process(1, src, [0, 1, 2], out)
    assert pl.has(1)
    # fn outside the window: plain forward + process
    pl.forward_process(3, np.ascontiguousarray(planes[0][3]), [0, 1, 2], out)
    assert pl.has(3)
    pl.close()


def _f16_round(x):
    """Complex array with real and imaginary parts rounded to binary16."""
    return (x.real.astype(np.float16).astype(np.float32)
            + 1j * x.imag.astype(np.float16).astype(np.float32)).astype(np.complex64)


@pytest.mark.parametrize("kind,kw", [("u8", dict(bt=3)), ("u10", dict(bt=4, sharpen=0.4)),
                                     ("u8", dict(bt=2, bw=16, bh=16)),
                                     ("f32", dict(bt=0, bw=64, bh=64))])
def test_emu_fp16_ring_within_parity(oracle, emu_lib, kind, kw):
    """FP16 spectra ring (FFT3D_OPT_SPECTRA_FP16): arithmetic stays FP32 and
    only the stored spectra are rounded, so the output stays within the path's
    parity tolerance of the oracle (which keeps FP32 spectra)."""
    bits, st = kind_info(kind)
    planes = make_clip(17, 120, 88, 5, kind)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    got = FFT3DCL(planes, bits=bits, sample_type=st, lib=emu_lib, half=True, **kw)
    if kind == "f32":
        assert max_abs(got[0], want[0]) < 1e-3
    else:
        assert lsb_diff(got[0], want[0]) <= (2 if bits == 8 else 4)


@pytest.mark.parametrize("kind,kw", [("u8", dict(bt=0, bw=16, bh=16)),
                                     ("u10", dict(bt=0, bw=32, bh=32, kratio=1.5)),
                                     ("u8", dict(bt=0, bw=64, bh=64, sigma2=3.0, sigma3=2.0,
                                                 sigma4=1.0))])
def test_emu_fp16_kalman_state_within_parity(oracle, emu_lib, kind, kw):
    """With the FP16 ring on integer samples the Kalman state is stored as
    FP16 too: every output frame of a 16-frame clip with a scene cut at frame
    8 (the kratio reset) stays within the path's parity tolerance of the
    oracle, whose state is FP32. (Float samples keep FP32 state; the FP16 ring
    alone already reaches 1.9e-3 of range on such a float clip, above the
    float tolerance, so the FP16 ring is a form for integer clips.)"""
    bits, st = kind_info(kind)
    planes = make_clip(29, 96, 80, 16, kind, scene_cut=8)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    got = FFT3DCL(planes, bits=bits, sample_type=st, lib=emu_lib, half=True, **kw)
    for n in range(16):
        if kind == "f32":
            assert max_abs(got[0][n], want[0][n]) < 1e-3, n
        else:
            assert lsb_diff(got[0][n], want[0][n]) <= (2 if bits == 8 else 4), n


def test_emu_fp16_option_restarts_kalman(emu_lib):
    """The Kalman state of integer samples follows the ring: FP16 pairs (8
    bytes per bin) with the FP16 ring, as precision 0 gives 8-bit clips at
    creation, FP32 (16 bytes) otherwise; switching reallocates it. Float
    samples keep FP32 state."""
    pl = capi.Plane(capi.make_params(bt=0, precision=2), 96, 72, 10, lib=emu_lib)
    nbin = pl.info.nb * pl.info.fbw * pl.info.bh
    before = capi.live_bytes(emu_lib)
    pl.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
    assert before - capi.live_bytes(emu_lib) == 8 * nbin
    pl.set_option(capi.Plane.OPT_SPECTRA_FP16, 0)
    assert capi.live_bytes(emu_lib) == before
    pl.close()
    before = capi.live_bytes(emu_lib)
    p8 = capi.Plane(capi.make_params(bt=0), 96, 72, 8, lib=emu_lib)  # precision 0: FP16
    p8.set_option(capi.Plane.OPT_SPECTRA_FP16, 0)
    grown = capi.live_bytes(emu_lib)
    p8.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
    assert grown - capi.live_bytes(emu_lib) == 8 * nbin
    p8.close()
    pf = capi.Plane(capi.make_params(bt=0), 96, 72, 32, sample_type=1, lib=emu_lib)
    before = capi.live_bytes(emu_lib)
    pf.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
    assert capi.live_bytes(emu_lib) == before
    pf.close()


def test_emu_fp16_ring_stores_rounded_spectra(emu_lib):
    frames = make_clip(18, 96, 72, 1, "f32")[0]
    full = capi.Plane(capi.make_params(), 96, 72, 32, sample_type=1, lib=emu_lib)
    half = capi.Plane(capi.make_params(), 96, 72, 32, sample_type=1, lib=emu_lib)
    half.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
    assert half.info.spectrum_bytes * 2 == full.info.spectrum_bytes
    for pl in (full, half):
        pl.forward(0, np.ascontiguousarray(frames[0]))
    a, b = full.read_spectrum(0), half.read_spectrum(0)
    # columns 1..bw/2-1 are stored bins; columns 0 and bw/2 are unpacked from
    # a stored pair, so they are only as close as the rounding of that pair
    assert np.array_equal(b[..., 1:16], _f16_round(a[..., 1:16]))
    assert np.abs(b - a).max() <= 2.0 ** -10 * np.abs(a).max()
    full.close()
    half.close()


def test_emu_fp16_ring_option_errors(emu_lib):
    pl = capi.Plane(capi.make_params(bw=32, bh=16), 96, 72, 8, lib=emu_lib)
    with pytest.raises(capi.FFT3DError, match="square blocks"):
        pl.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
    with pytest.raises(capi.FFT3DError, match="must be 0 or 1"):
        pl.set_option(capi.Plane.OPT_SPECTRA_FP16, 3)
    pl.close()


def test_binding_rejects_bad_buffers(emu_lib):
    """The ctypes binding checks every plane buffer before its address goes
    to the library: short, strided, wrong element size and host/device
    mismatches raise ValueError (a short buffer once corrupted the heap)."""
    W, H = 64, 48
    clip = make_clip(24, W, H, 3, "u8")[0]
    pl = capi.Plane(capi.make_params(bt=1), W, H, 8, lib=emu_lib)
    good = np.ascontiguousarray(clip[0])
    bad = {
        "short": np.zeros(64, np.uint8),
        "strided": np.zeros((H, 2 * W), np.uint8)[:, ::2],
        "element size": np.zeros((H, W), np.uint16),
    }
    for why, buf in bad.items():
        with pytest.raises(ValueError):
            pl.forward(0, buf)
        with pytest.raises(ValueError):
            pl.forward_process(0, good, [0], buf)
    with pytest.raises(ValueError):
        pl.forward(0, good, on_device=True)  # a numpy array is host memory
    pl.forward(0, good)
    with pytest.raises(ValueError):
        pl.process([0], np.zeros(10, np.uint8))
    out = np.zeros((H, W), np.uint8)
    pl.process([0], out)
    assert out.any()
    # step: a bad output fails before the call, a bad source fails its item
    with pytest.raises(ValueError):
        capi.step([capi.StepRequest(pl, 1, 3, lambda fn: good, np.zeros(5, np.uint8))],
                  lib=emu_lib)
    st = capi.step([capi.StepRequest(pl, 1, 3, lambda fn: np.zeros(5, np.uint8), out)],
                   lib=emu_lib)
    assert st == [-1]
    pl.close()


@pytest.mark.parametrize("dft", [1, 2])
@pytest.mark.parametrize("kind,kw", [
    ("u10", dict(bt=4, sharpen=0.4, sigma=3.0)),
    ("u8", dict(bt=3)),
    ("f32", dict(bt=0, sigma2=3.0, sigma3=2.0, sigma4=1.0)),
    ("u8", dict(bt=2, plane=0)),
])
def test_emu_mfma_form_within_parity(oracle, emu_lib, dft, kind, kw):
    """The MFMA DFT form (emulated matrix products with the gfx950 lane
    layout, FP16 / BF16 operands split hi + lo) against the oracle: spectra
    to 2e-6 / 2e-5 of the largest bin, clips within the parity bar; the FP16
    ring combines with it."""
    bits, st = kind_info(kind)
    W, H = 96, 80
    planes = make_clip(51, W, H, 4, kind)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    for half in (False, True):
        got = FFT3DCL(planes, bits=bits, sample_type=st, lib=emu_lib, dft=dft, half=half, **kw)
        if kind == "f32":
            assert max_abs(got[0], want[0]) < 1e-3
        else:
            assert lsb_diff(got[0], want[0]) <= (2 if kind == "u8" else 4)
    pl = capi.Plane(capi.make_params(**kw, precision=2), W, H, bits, st, lib=emu_lib)
    pl.set_option(capi.Plane.OPT_DFT, dft)
    pl.forward(0, np.ascontiguousarray(planes[0][0]))
    got = pl.read_spectrum(0)
    op = OPlane(oracle, script(**kw), W, H, bits, st)
    ref = np.asarray(op.spectra(op.upload(planes[0][0]))).reshape(got.shape)
    assert np.abs(got - ref).max() / np.abs(ref).max() < (2e-6 if dft == 1 else 2e-5)
    pl.close()


def test_emu_mfma_form_option_errors(emu_lib):
    for kw in (dict(bw=16, bh=16), dict(mode=0), dict(bw=64, bh=32)):
        pl = capi.Plane(capi.make_params(bt=3, **kw), 128, 96, 8, lib=emu_lib)
        with pytest.raises(capi.FFT3DError) as e:
            pl.set_option(capi.Plane.OPT_DFT, 1)
        assert e.value.code == -6
        pl.close()
    pl = capi.Plane(capi.make_params(bt=3), 128, 96, 8, lib=emu_lib)
    with pytest.raises(capi.FFT3DError):
        pl.set_option(capi.Plane.OPT_DFT, 3)
    pl.close()


def test_emu_mfma_analysis_runs(emu_lib, monkeypatch):
    """The MFMA analysis on a grid of fewer workgroups than 32-block groups
    (each workgroup then takes a longer run of 4-block chunks, the one-round
    grid of a large plane) gives the spectra of the default grid, bit for
    bit, ragged last chunk included."""
    W, H = 160, 128  # 11 x 9 = 99 blocks of 32x32: 4 groups of 32
    frame = np.ascontiguousarray(make_clip(7, W, H, 1, "u8")[0][0])
    specs = []
    for cus in ("256", "1"):  # one emulated CU holds 2 workgroups
        monkeypatch.setenv("FFT3D_EMU_CUS", cus)
        pl = capi.Plane(capi.make_params(bt=3), W, H, 8, lib=emu_lib)
        pl.set_option(capi.Plane.OPT_DFT, 1)
        pl.forward(0, frame)
        specs.append(pl.read_spectrum(0))
        pl.close()
    assert np.array_equal(specs[0], specs[1])


def _raw_step(lib, reqs):
    """fft3d_step over (plane, frame, nframes, clip, out, on_device) with
    numpy buffers: in the emulation, device memory is host memory, so the
    device flags take the batched path on the same buffers."""
    L = capi.load(lib)
    keep = []
    items = (capi.WorkItem * len(reqs))()
    for k, (pl, f, nf, clip, out, dev) in enumerate(reqs):
        def cb(_user, fn, c=clip):
            a = np.ascontiguousarray(c[fn])
            keep.append(a)
            return a.ctypes.data
        c = capi.SOURCE_FN(cb)
        keep.append(c)
        items[k] = capi.WorkItem(pl.h, f, nf, c, None, int(dev), out.ctypes.data, out.nbytes,
                                 int(dev), None, 0)
    L.fft3d_step(items, len(reqs))
    return [items[k].status for k in range(len(reqs))]


@pytest.mark.parametrize("groups", [((0, 1), (2, 3), (4,)), ((0, 1, 2, 3, 4),)])
def test_emu_batched_step_equals_item_step(emu_lib, groups):
    """The batched step (device buffers: one launch per kernel class over all
    items of a round, each plane's items in rounds) gives the frames of the
    item-by-item step (host buffers), bit for bit, over planes of several
    block sizes, temporal modes, ring precisions and DFT forms, with two
    frames of one plane in the same call, or all five (rounds launched in
    chunks of 1, 2 and 2)."""
    n = 5
    cfgs = [(96, 64, 8, dict(bt=4, sharpen=0.4), {}),
            (80, 72, 10, dict(bt=3, bw=16, bh=16), {}),
            (128, 96, 8, dict(bt=0, sigma2=3.0, bw=64, bh=64), {}),
            (96, 80, 8, dict(bt=2), {"half": 1}),
            (96, 64, 10, dict(bt=-1, sharpen=0.5), {}),
            (96, 64, 8, dict(bt=4), {"dft": 1})]
    res = {}
    for dev in (False, True):
        planes, clips, outs = [], [], []
        for c, (W, H, bits, kw, opt) in enumerate(cfgs):
            pl = capi.Plane(capi.make_params(**kw), W, H, bits, lib=emu_lib)
            if opt.get("half"):
                pl.set_option(capi.Plane.OPT_SPECTRA_FP16, 1)
            if opt.get("dft"):
                pl.set_option(capi.Plane.OPT_DFT, opt["dft"])
            planes.append(pl)
            clips.append(make_clip(60 + c, W, H, n, "u8" if bits == 8 else "u10")[0])
            outs.append(np.zeros_like(clips[-1]))
        for frames in groups:
            reqs = [(planes[c], f, n, clips[c], outs[c][f], dev) for f in frames
                    for c in range(len(cfgs))]
            assert _raw_step(emu_lib, reqs) == [0] * len(reqs)
        for c, pl in enumerate(planes):
            bt = cfgs[c][2 + 1]["bt"]
            assert pl.info.cached_frames <= (bt + 2 if bt >= 2 else 0)
            pl.close()
        res[dev] = outs
    for c in range(len(cfgs)):
        assert np.array_equal(res[True][c], res[False][c]), cfgs[c]
        assert res[True][c].any()


def test_emu_tuning_table_is_read(emu_lib, tmp_path):
    """The library reads the tuning table named by $FFT3D_TUNING once per
    process and applies an entry to the plane that matches it exactly (block
    size, plane size, ring precision, DFT form); an occupancy the block size
    has no kernel for falls back to 2; other planes keep the heuristic."""
    import json
    import subprocess
    import sys
    tab = {"entries": [
        {"bw": 16, "bh": 16, "width": 320, "height": 200, "half": 0, "dft": 0,
         "band_rows": 3, "occupancy": 3},
        {"bw": 32, "bh": 32, "width": 320, "height": 200, "half": 0, "dft": 0,
         "band_rows": 2, "occupancy": 4},
    ]}
    path = tmp_path / "tuning.json"
    path.write_text(json.dumps(tab))
    code = f"""
import sys
sys.path.insert(0, {str(ROOT)!r})
from fft3dcl_amd import capi
lib = {emu_lib!r}
def info(**kw):
    p = capi.Plane(capi.make_params(bt=3, precision=2, **kw), kw.pop('w', 320), 200, 8, lib=lib)
    i = p.info
    r = (i.tuned, i.band_rows, i.sweep_occupancy)
    p.close()
    return r
print(info(bw=16, bh=16), info(bw=32, bh=32), info(bw=8, bh=8))
"""
    env = dict(__import__("os").environ, FFT3D_TUNING=str(path))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, env=env,
                       timeout=120)
    assert r.returncode == 0, r.stderr
    a, b, c = eval(r.stdout.strip().replace(") (", "),("))
    assert a == (1, 3, 3)
    assert b == (1, 2, 2)      # no occupancy-4 sweep for 32x32 blocks
    assert c[0] == 0 and c[2] == 2


@pytest.mark.parametrize("bw,bh,kind,kw", [
    (128, 32, "u8", dict(bt=3)),
    (16, 256, "u10", dict(bt=4, sharpen=0.4)),
    (512, 8, "f32", dict(bt=0, sigma2=3.0)),
    (8, 512, "u8", dict(bt=2, mode=0)),
    (256, 16, "u16", dict(bt=-1, sharpen=0.5, wintype=1)),
])
def test_emu_large_blocks_match_oracle(oracle, emu_lib, bw, bh, kind, kw):
    """Blocks with a side above 64 (LDS FFTs, one workgroup per block), as
    the reference GPU accepts up to bw*bh*8 = 32 KB: spectra to 1e-6 of the
    largest bin and clips within the parity bar."""
    bits, st = kind_info(kind)
    W, H = 250, 130  # ragged against every block size here
    planes = make_clip(3, W, H, 3, kind)
    kw = dict(kw, bw=bw, bh=bh)
    got = FFT3DCL(planes, bits=bits, sample_type=st, lib=emu_lib, **kw)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    if kind == "f32":
        assert max_abs(got[0], want[0]) < 1e-3
    else:
        assert lsb_diff(got[0], want[0]) <= (2 if kind == "u8" else 4)
    pl = capi.Plane(capi.make_params(**kw), W, H, bits, st, lib=emu_lib)
    assert pl.info.blocks_per_wg == 1 and pl.info.band_rows == 0  # unfused path
    pl.forward(0, np.ascontiguousarray(planes[0][0]))
    g = pl.read_spectrum(0)
    op = OPlane(oracle, script(**kw), W, H, bits, st)
    ref = np.asarray(op.spectra(op.upload(planes[0][0]))).reshape(g.shape)
    assert np.abs(g - ref).max() < 1e-6 * np.abs(ref).max()
    pl.close()
