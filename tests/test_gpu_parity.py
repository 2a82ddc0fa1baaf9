"""GPU parity of the HIP path (through the C ABI) against the oracle.

Tolerances (written here, from the parity bar of the path): +/-2 LSB at 8 bit,
+/-4 LSB at 10/16 bit, 1e-3 absolute for float clips (values in [0,1] scale),
and 2e-4 in the normalised float domain for the reference's own GPU-vs-CPU
check (tools/gpu_test.cpp:262-272). Output must be bit-identical run to run.
"""
import numpy as np
import pytest
import torch

from fft3dcl_amd import capi
from fft3dcl_amd.filter import FFT3DCL, Instance
from fft3dcl_amd.synth import make_clip, make_plane_sequence
from helpers import lsb_diff, max_abs, mt19937_planes, tolerance
from oracle_lib import script, kind_info

pytestmark = pytest.mark.gpu


def check(kind, got, want):
    if kind == "f32":
        d = max_abs(got, want)
    else:
        d = lsb_diff(got, want)
    assert d <= tolerance(kind), d
    return d


@pytest.mark.parametrize("bt", [1, 2, 3, 4, 0])
@pytest.mark.parametrize("variant", [dict(), dict(sharpen=0.7), dict(degrid=0.0),
                                     dict(sigma2=3.0)])
def test_reference_gpu_test(gpu, oracle, bt, variant):
    """tools/gpu_test.cpp: 256x192 mt19937(42) float planes, default params with
    bt / sharpen / degrid / sigma2 from argv; PASS if max |GPU-CPU| < 2e-4."""
    planes = [np.stack(mt19937_planes(42, 256, 192, max(bt, 1)))]
    _, want = oracle.run_clip(script(bt=bt, **variant), planes, 32, 1, want_float=True)
    got = FFT3DCL(planes, bits=32, sample_type=1, bt=bt, lib=gpu, **variant)
    n = 0 if bt < 2 else (2 if bt == 4 else 1)
    assert max_abs(got[0][n], want[0][n]) < 2e-4


CONFIGS = {
    # plain method, CPU-runnable: 640x360 u8, 16 frames, sigma 2, bt 3, bw 32
    "A_640x360_u8_bt3": ("u8", 640, 360, 16, "gray", dict(sigma=2.0, beta=1.0, bt=3),
                         None),
    # 1080p YUV420P10 plane=4, bt 4, sharpen (8 of its 64 frames)
    "B_1080p10_yuv_plane4_bt4": ("u10", 1920, 1080, 6, "yuv",
                                 dict(sigma=3.0, bt=4, sharpen=0.4, scutoff=0.3, svr=1.0,
                                      smin=4.0, smax=20.0, plane=4), None),
    # 4K 10-bit luma bt 4 sharpen (5 of its 120 frames)
    "C_4k10_bt4": ("u10", 3840, 2160, 5, "gray", dict(sigma=3.0, bt=4, sharpen=0.4), None),
    # 1440p float Kalman + pattern, 64x64 blocks, scene cut (8 frames, cut at 4)
    "D_1440p_f32_kalman_bw64": ("f32", 2560, 1440, 8, "gray",
                                dict(bt=0, kratio=2.0, bw=64, bh=64, sigma=2.0, sigma2=3.0,
                                     sigma3=2.0, sigma4=1.0), 4),
}


@pytest.mark.parametrize("name", list(CONFIGS))
def test_config_vs_oracle(gpu, oracle, name):
    kind, W, H, n, color, kw, cut = CONFIGS[name]
    bits, st = kind_info(kind)
    planes = make_clip(1000 + len(name), W, H, n, kind, color=color, scene_cut=cut)
    cf = 3 if color == "yuv" else 1
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, color_family=cf)
    got = FFT3DCL(planes, color_family=cf, bits=bits, sample_type=st, lib=gpu, **kw)
    for g, w in zip(got, want):
        d = check(kind, g, w)
        print(f"{name}: max diff {d}")


def test_config_d_full_kalman_run(gpu, oracle):
    """Config D as stated: 2560x1440 float, 96 frames, Kalman (bt=0, kratio 2)
    with the sigma pattern, 64x64 blocks, a scene cut at frame 48. Every frame
    is within 1e-3 of the oracle, no spectrum stays in the ring after a frame
    and the plane's device memory stays flat (the reference's bt=0 ring would
    grow by one spectrum per frame)."""
    W, H, n = 2560, 1440, 96
    kw = dict(bt=0, kratio=2.0, bw=64, bh=64, sigma=2.0, sigma2=3.0, sigma3=2.0, sigma4=1.0)
    planes = make_clip(1040, W, H, n, "f32", scene_cut=48)
    want, _ = oracle.run_clip(script(**kw), planes, 32, 1)
    inst = Instance(1, W, H, 32, 1, n, lib=gpu, **kw)
    src = torch.from_numpy(planes[0]).cuda()
    out = torch.zeros_like(src)
    dev = []
    for k in range(n):
        inst.get_frame(k, [lambda fn: src[fn]], [out[k]], on_device=True)
        info = inst.objs[0].info
        assert info.cached_frames == 0, (k, info.cached_frames)
        dev.append(info.device_bytes)
    torch.cuda.synchronize()
    inst.close()
    assert max(dev[1:]) == min(dev[1:]), dev
    got = out.cpu().numpy()
    d = [max_abs(got[k], want[0][k]) for k in range(n)]
    print(f"config D full run: max diff {max(d):.2e} (frame {int(np.argmax(d))})")
    assert max(d) <= tolerance("f32"), d


@pytest.mark.parametrize("kind,kw,half", [
    ("u8", dict(bt=0, bw=16, bh=16, sigma=2.5), False),       # precision 0: FP16 ring
    ("u10", dict(bt=0, bw=64, bh=64, sigma=3.0, sigma2=4.0, sigma3=2.0, sigma4=1.0), True)])
def test_fp16_kalman_long_run(gpu, oracle, kind, kw, half):
    """The FP16 ring on integer samples with Kalman (bt=0): the state is FP16
    pairs too. 1280x720, 48 frames with a scene cut at frame 24: every frame
    within the path's tolerance of the oracle (FP32 state), device memory
    flat over the clip."""
    W, H, n = 1280, 720, 48
    bits, st = kind_info(kind)
    planes = make_clip(1050 + bits, W, H, n, kind, scene_cut=24)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    inst = Instance(1, W, H, bits, st, n, lib=gpu, half=half, **kw)
    src = torch.from_numpy(planes[0]).cuda()
    out = torch.zeros_like(src)
    dev = []
    for k in range(n):
        inst.get_frame(k, [lambda fn: src[fn]], [out[k]], on_device=True)
        dev.append(inst.objs[0].info.device_bytes)
    torch.cuda.synchronize()
    inst.close()
    assert max(dev[1:]) == min(dev[1:]), dev
    got = out.cpu().numpy()
    d = [lsb_diff(got[k], want[0][k]) for k in range(n)]
    print(f"FP16 Kalman {kind} bw{kw['bw']}: max {max(d)} LSB (frame {int(np.argmax(d))})")
    assert max(d) <= tolerance(kind), d


EDGE = [
    ("u8", 64, 48, 1, dict(bt=3)),            # one-frame clip: window clamps to [0,0,0]
    ("u8", 64, 48, 2, dict(bt=4)),            # two frames, bt 4
    ("u8", 64, 48, 3, dict(bt=3)),
    ("u8", 20, 12, 3, dict(bt=3)),            # plane smaller than one block
    ("u10", 77, 45, 4, dict(bt=2, sharpen=0.5)),  # ragged size
    ("u16", 96, 72, 4, dict(bt=3, beta=2.0)),
    ("u8", 96, 72, 4, dict(bt=3, mode=0)),
    ("u8", 96, 72, 4, dict(bt=3, bw=32, bh=16, ow=10, oh=4, wintype=1)),
    ("u8", 96, 72, 4, dict(bt=1, bw=8, bh=8, wintype=2)),
    ("u8", 96, 72, 4, dict(bt=3, bw=4, bh=4)),
    ("u8", 160, 128, 4, dict(bt=3, bw=64, bh=16)),
    ("f32", 96, 72, 4, dict(bt=-1, sharpen=1.0)),
    ("u8", 96, 72, 5, dict(bt=0, sigma2=4.0, sigma3=2.0, sigma4=1.0)),
    # blocks with a side above 64 (LDS FFTs), up to bw*bh = 4096
    ("u8", 1920, 1080, 4, dict(bt=3, bw=128, bh=32)),
    ("u10", 700, 500, 5, dict(bt=4, sharpen=0.4, bw=16, bh=256)),
    ("f32", 600, 300, 4, dict(bt=0, sigma2=3.0, bw=512, bh=8)),
    ("u8", 500, 600, 4, dict(bt=2, mode=0, bw=8, bh=512)),
    ("u16", 640, 360, 4, dict(bt=-1, sharpen=0.5, bw=256, bh=16)),
    ("u8", 333, 201, 4, dict(bt=1, bw=128, bh=8, wintype=2)),
]


@pytest.mark.parametrize("kind,W,H,n,kw", EDGE)
def test_edge_cases_vs_oracle(gpu, oracle, kind, W, H, n, kw):
    bits, st = kind_info(kind)
    planes = make_clip(7, W, H, n, kind)
    order = list(range(n))[::-1] if n > 2 else None  # backwards seek too
    if kw.get("bt") == 0:
        order = None  # Kalman runs frames in order (fmFrameState)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, order=order)
    got = FFT3DCL(planes, bits=bits, sample_type=st, order=order, lib=gpu, **kw)
    check(kind, got[0], want[0])


def test_yuv_chroma_only(gpu, oracle):
    planes = make_clip(8, 128, 96, 4, "u8", color="yuv")
    want, _ = oracle.run_clip(script(plane=1, bt=3), planes, 8, 0, color_family=3)
    got = FFT3DCL(planes, color_family=3, bits=8, plane=1, bt=3, lib=gpu)
    for g, w in zip(got, want):
        check("u8", g, w)
    assert np.array_equal(got[0], planes[0])  # untouched luma copied


def test_run_to_run_bit_identical(gpu):
    planes = make_clip(9, 640, 360, 6, "u10")
    a = FFT3DCL(planes, bits=10, bt=4, sharpen=0.4, lib=gpu)
    b = FFT3DCL(planes, bits=10, bt=4, sharpen=0.4, lib=gpu)
    assert np.array_equal(a[0], b[0])
    f = make_clip(9, 256, 192, 6, "f32")
    a = FFT3DCL(f, bits=32, sample_type=1, bt=0, lib=gpu)
    b = FFT3DCL(f, bits=32, sample_type=1, bt=0, lib=gpu)
    assert np.array_equal(a[0], b[0])


def test_device_resident_path_equals_host_path(gpu):
    W, H, n = 1280, 720, 6
    frames = make_plane_sequence(10, W, H, n, "u10")
    host = FFT3DCL([frames], bits=10, bt=4, lib=gpu)[0]
    dev_in = torch.from_numpy(frames).cuda()
    dev_out = torch.zeros((n, H, W), dtype=torch.int16, device="cuda")
    stream = torch.cuda.current_stream().cuda_stream
    inst = Instance(1, W, H, 10, 0, n, bt=4, lib=gpu)
    srcs = [lambda fn: dev_in[fn]]
    for k in range(n):
        inst.get_frame(k, srcs, [dev_out[k]], on_device=True, stream=stream)
    torch.cuda.synchronize()
    inst.close()
    assert np.array_equal(dev_out.cpu().numpy().view(np.uint16), host)


@pytest.mark.parametrize("prefetch", [False, True])
def test_device_yuv_planes_on_their_streams_equal_host_path(gpu, prefetch):
    """With device buffers the three planes of a 4:2:0 frame run on streams
    of their own, between the caller's earlier and later work; the frames
    equal the host-buffer path's (planes one after another) bit for bit."""
    W, H, n = 640, 360, 6
    planes = make_clip(77, W, H, n, "u10", color="yuv")
    kw = dict(bt=4, sharpen=0.4, plane=4)
    host = FFT3DCL(planes, color_family=3, bits=10, lib=gpu, **kw)
    dev_in = [torch.from_numpy(p).cuda() for p in planes]
    dev_out = [torch.zeros(p.shape, dtype=torch.int16, device="cuda") for p in planes]
    stream = torch.cuda.Stream()
    inst = Instance(3, W, H, 10, 0, n, lib=gpu, prefetch=prefetch, **kw)
    srcs = [lambda fn, d=d: d[fn] for d in dev_in]
    for k in range(n):
        inst.get_frame(k, srcs, [o[k] for o in dev_out], on_device=True,
                       stream=stream.cuda_stream)
    stream.synchronize()
    inst.close()
    for q in range(3):
        assert np.array_equal(dev_out[q].cpu().numpy().view(np.uint16), host[q]), q


@pytest.mark.parametrize("where,bt,half", [("pageable", 4, False), ("pinned", 4, False),
                                           ("device", 4, False), ("pageable", 3, False),
                                           ("pinned", 2, True)])
def test_prefetch_equals_plain(gpu, where, bt, half):
    """Sequential requests with the next new frame prefetched on the upload
    stream give the same frames as plain requests (4K; host buffers, or
    device buffers on a caller stream; FP32 or FP16 ring)."""
    W, H, n = 3840, 2160, 8
    frames = make_plane_sequence(12, W, H, n, "u10")
    dev = where == "device"
    if where == "pinned":
        src = torch.from_numpy(frames).pin_memory()
    elif dev:
        src = torch.from_numpy(frames).cuda()
    else:
        src = frames
    ts = torch.cuda.Stream()
    outs = []
    for prefetch in (True, False):
        inst = Instance(1, W, H, 10, 0, n, bt=bt, sharpen=0.4, lib=gpu, prefetch=prefetch,
                        half=half)
        if where == "pinned":
            out = torch.zeros((n, H, W), dtype=torch.int16).pin_memory()
        elif dev:
            out = torch.zeros((n, H, W), dtype=torch.int16, device="cuda")
        else:
            out = np.zeros_like(frames)
        torch.cuda.synchronize()
        for k in range(n):
            inst.get_frame(k, [lambda fn: src[fn]], [out[k]], on_device=dev,
                           stream=ts.cuda_stream if dev else None)
        torch.cuda.synchronize()
        inst.close()
        if dev:
            outs.append(out.cpu().numpy().view(np.uint16))
        else:
            outs.append(out.numpy().view(np.uint16) if where == "pinned" else out)
    assert outs[1].any()
    assert np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("dst_on_device", [False, True])
def test_prefetched_frame_used_at_once(gpu, dst_on_device):
    """A process call whose window holds a frame prefetched by the call just
    before waits for that prefetch on its own stream (host or device output;
    with device output nothing blocks in between)."""
    W, H = 3840, 2160
    frames = make_plane_sequence(13, W, H, 6, "u10")
    params = capi.make_params(bt=4)
    res = []
    for pre in (True, False):
        pl = capi.Plane(params, W, H, 10, lib=gpu)
        # a stream of its own (a null handle would select the plane's stream)
        ts = torch.cuda.Stream() if dst_on_device else None
        stream = ts.cuda_stream if dst_on_device else None
        outs = [torch.zeros((H, W), dtype=torch.int16, device="cuda") if dst_on_device
                else np.zeros((H, W), np.uint16) for _ in range(3)]
        torch.cuda.synchronize()
        keep = []
        for k in range(4):
            pl.forward(k, frames[k], False, stream)
        for i, k in enumerate((4, 5)):
            if pre:
                if dst_on_device:
                    # a device source written on the caller stream just before
                    src = torch.empty((H, W), dtype=torch.int16, device="cuda")
                    host = torch.from_numpy(frames[k].view(np.int16))
                    with torch.cuda.stream(ts):
                        src.copy_(host)
                    keep.append(src)
                    pl.prefetch(k, src, True)
                else:
                    pl.prefetch(k, frames[k])   # carried out by this process call
            else:
                pl.forward(k, frames[k], False, stream)
            pl.process([k - 4, k - 3, k - 2, k - 1], outs[i], dst_on_device, stream)
        pl.process([2, 3, 4, 5], outs[2], dst_on_device, stream)
        if dst_on_device:
            ts.synchronize()
            outs = [o.cpu().numpy().view(np.uint16) for o in outs]
        pl.close()
        res.append(outs)
    for a, b in zip(*res):
        assert a.any()
        assert np.array_equal(a, b)


@pytest.mark.parametrize("chunk,fuse", [(0, False), (1 << 20, False), (4160, True)])
def test_host_transfers_staged_direct_pinned(gpu, chunk, fuse):
    """The same frames whichever way host buffers travel: pageable numpy
    buffers through the pinned chunks (4 MiB, 1 MiB, or 4160-byte chunks with
    a ragged last one), pageable buffers handed to hipMemcpyAsync as they
    are, and page-locked buffers straight to the DMA engine."""
    W, H, n = 1920, 1080, 5
    frames = make_plane_sequence(11, W, H, n, "u10")

    def run(staging, src, out):
        inst = Instance(1, W, H, 10, 0, n, bt=4, sharpen=0.4, lib=gpu, fuse=fuse)
        inst.objs[0].set_option(capi.Plane.OPT_HOST_STAGING, staging)
        inst.objs[0].set_option(capi.Plane.OPT_STAGING_CHUNK, chunk)
        for k in range(n):
            inst.get_frame(k, [lambda fn: src[fn]], [out[k]])
        inst.close()
        return out

    staged = run(1, frames, np.zeros_like(frames))
    direct = run(0, frames, np.zeros_like(frames))
    pinned_in = torch.from_numpy(frames).pin_memory()
    pinned_out = run(1, pinned_in, torch.zeros((n, H, W), dtype=torch.int16).pin_memory())
    assert direct.any()
    assert np.array_equal(staged, direct)
    assert np.array_equal(pinned_out.numpy().view(np.uint16), direct)


def test_spectrum_vs_oracle(gpu, oracle):
    from oracle_lib import Plane as OPlane
    frames = make_plane_sequence(11, 333, 201, 1, "f32")
    for bw, bh in ((32, 32), (16, 64), (64, 64), (8, 8)):
        pl = capi.Plane(capi.make_params(bw=bw, bh=bh), 333, 201, 32, sample_type=1, lib=gpu)
        pl.forward(0, np.ascontiguousarray(frames[0]))
        got = pl.read_spectrum(0)
        want = OPlane(oracle, script(bw=bw, bh=bh), 333, 201, 32, sample_type=1).spectra(frames[0])
        assert np.abs(got - want).max() < 1e-5 * np.abs(want).max()
        pl.close()


def test_profiling_counts_launches(gpu):
    frames = make_plane_sequence(12, 640, 360, 5, "u8")
    src = [lambda fn: np.ascontiguousarray(frames[fn])]
    out = np.zeros((360, 640), np.uint8)
    # fused: the last missing frame of each window is analysed inside the sweep
    inst = Instance(1, 640, 360, 8, 0, 5, bt=3, lib=gpu, fuse=True, precision=2)
    inst.objs[0].set_profiling(True)
    for k in range(5):
        inst.get_frame(k, src, [out])
    prof = inst.objs[0].profile()
    fused = prof.get("forward_filter_synth_sweep", (0, 0))[0]
    assert fused == 4 and prof["analysis"][0] == 1  # frame 0 alone, then one new frame a step
    assert fused + prof.get("filter_synth_sweep", (0, 0))[0] == 5
    assert prof["combine_seams"][0] == 5
    assert all(ms > 0 for _, ms in prof.values())
    inst.close()
    inst = Instance(1, 640, 360, 8, 0, 5, bt=3, lib=gpu, fuse=False)
    inst.objs[0].set_profiling(True)
    for k in range(5):
        inst.get_frame(k, src, [out])
    prof = inst.objs[0].profile()
    assert prof["filter_synth_sweep"][0] == 5 and prof["analysis"][0] == 5
    inst.objs[0].set_option(capi.Plane.OPT_SWEEP, 0)
    for k in range(5):
        inst.get_frame(k, src, [out])
    prof = inst.objs[0].profile()
    assert prof["filter_inverse"][0] == 5 and prof["synthesize"][0] == 5
    inst.close()


@pytest.mark.parametrize("kind,kw,W,H", [
    ("u10", dict(bt=4, sharpen=0.4, sigma=3.0), 1920, 1080),
    ("u8", dict(bt=3, precision=2), 1000, 700),  # FP32 ring: the fused form runs
    ("f32", dict(bt=0, kratio=1.5, bw=64, bh=64), 1300, 700),
    ("u8", dict(bt=1, bw=16, bh=64), 900, 500),
    ("u16", dict(bt=2, bw=64, bh=16, degrid=0.0), 700, 300),
])
def test_forward_process_equals_forward_then_process(gpu, kind, kw, W, H):
    """The new frame transformed inside the sweep against forward + process,
    out-of-order requests too. The two launch forms are separate kernels whose
    FFTs the compiler may contract into FMAs differently, so spectra may differ
    by float rounding (1e-6 of the block's magnitude) and outputs by 1 LSB
    (1e-6 for float clips); the CPU emulation test holds them bit-identical."""
    bits, st = kind_info(kind)
    planes = make_clip(16, W, H, 5, kind)
    order = [0, 1, 2, 4, 3] if kw["bt"] != 0 else list(range(5))
    res = []
    for fuse in (True, False):
        inst = Instance(1, W, H, bits, st, 5, lib=gpu, fuse=fuse, **kw)
        out = np.zeros_like(planes[0])
        specs = {}
        for n in order:
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
            specs.update({fn: inst.objs[0].read_spectrum(fn) for fn in range(5)
                          if inst.objs[0].has(fn) and fn not in specs})
        inst.close()
        res.append((out, specs))
    if kind == "f32":
        assert max_abs(res[0][0], res[1][0]) <= 1e-6
    else:
        assert lsb_diff(res[0][0], res[1][0]) <= 1
    assert res[0][1].keys() == res[1][1].keys()
    for fn in res[0][1]:
        a, b = res[0][1][fn], res[1][1][fn]
        assert np.abs(a - b).max() <= 1e-6 * np.abs(b).max()


def test_forward_process_device_buffers(gpu):
    """Fused path reading the new plane in place from device memory and
    writing the output to device memory, on a torch stream."""
    W, H, n = 1280, 720, 6
    frames = make_plane_sequence(17, W, H, n, "u10")
    host = FFT3DCL([frames], bits=10, bt=4, lib=gpu, fuse=False)[0]
    dev_in = torch.from_numpy(frames).cuda()
    dev_out = torch.zeros((n, H, W), dtype=torch.int16, device="cuda")
    stream = torch.cuda.Stream()
    inst = Instance(1, W, H, 10, 0, n, bt=4, lib=gpu, fuse=True)
    for k in range(n):
        inst.get_frame(k, [lambda fn: dev_in[fn]], [dev_out[k]], on_device=True,
                       stream=stream.cuda_stream)
    stream.synchronize()
    inst.close()
    assert lsb_diff(dev_out.cpu().numpy().view(np.uint16), host) <= 1


@pytest.mark.parametrize("kind,kw,W,H,R", [
    ("u10", dict(bt=4, sharpen=0.4, sigma=3.0), 1920, 1080, 0),
    ("u8", dict(bt=3), 1000, 700, 3),
    ("f32", dict(bt=0, bw=64, bh=64), 1300, 700, 2),
    ("u8", dict(bt=2, bw=16, bh=16), 900, 500, 0),
    ("u8", dict(bt=3, bw=8, bh=8), 700, 300, 4),
])
def test_sweep_equals_unfused_path(gpu, oracle, kind, kw, W, H, R):
    """The fused sweep (in-LDS overlap-add, border tiles through partial
    planes) against the unfused filter/inverse + synthesis kernels and the
    oracle, on frames with several strips and bands."""
    bits, st = kind_info(kind)
    planes = make_clip(13, W, H, 4, kind)
    want, _ = oracle.run_clip(script(**kw), planes, bits, st)
    outs = []
    for sweep in (1, 0):
        inst = Instance(1, W, H, bits, st, 4, lib=gpu, **kw)
        inst.objs[0].set_option(capi.Plane.OPT_SWEEP, sweep)
        inst.objs[0].set_option(capi.Plane.OPT_BAND_ROWS, R)
        out = np.zeros_like(planes[0])
        for n in range(4):
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
        inst.close()
        check(kind, out, want[0])
        outs.append(out)
    check(kind, outs[0], outs[1])


@pytest.mark.parametrize("name", ["A_640x360_u8_bt3", "B_1080p10_yuv_plane4_bt4", "C_4k10_bt4"])
def test_fp16_ring_vs_oracle(gpu, oracle, oracle64, name):
    """FP16 spectra ring (the reference's precision=0 storage, here an
    option): within the path's tolerance of the FP32 oracle. Prints the max
    error against the FP32 oracle and its FP64 twin, in LSB."""
    kind, W, H, n, color, kw, cut = CONFIGS[name]
    n = min(n, 5)
    bits, st = kind_info(kind)
    planes = make_clip(1000 + len(name), W, H, n, kind, color=color, scene_cut=cut)
    cf = 3 if color == "yuv" else 1
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, color_family=cf)
    want64, _ = oracle64.run_clip(script(**kw), planes, bits, st, color_family=cf)
    got = FFT3DCL(planes, color_family=cf, bits=bits, sample_type=st, lib=gpu, half=True, **kw)
    for g, w, w64 in zip(got, want, want64):
        d = check(kind, g, w)
        print(f"{name} FP16 ring: max |GPU - oracle| {d} LSB, "
              f"max |GPU - FP64 oracle| {lsb_diff(g, w64)} LSB")


def test_step_host_buffers_equals_host_step(gpu):
    """fft3d_step over two planes (bt=4 and Kalman) with pageable host buffers,
    their items interleaved in each call, gives the frames of the Python host
    step."""
    W, H, n = 1280, 720, 6
    clips = [make_plane_sequence(14 + c, W, H, n, "u10") for c in range(2)]
    kws = [dict(bt=4, sharpen=0.4), dict(bt=0, sigma2=3.0)]
    want = []
    for c in range(2):
        inst = Instance(1, W, H, 10, 0, n, lib=gpu, **kws[c])
        out = np.zeros_like(clips[c])
        for f in range(n):
            inst.get_frame(f, [lambda fn, a=clips[c]: a[fn]], [out[f]])
        inst.close()
        want.append(out)
    planes = [capi.Plane(capi.make_params(**kws[c]), W, H, 10, lib=gpu) for c in range(2)]
    got = [np.zeros_like(a) for a in clips]
    for f in range(0, n, 2):
        reqs = [capi.StepRequest(planes[c], g, n, lambda fn, a=clips[c]: a[fn], got[c][g])
                for g in (f, f + 1) for c in range(2)]
        assert capi.step(reqs, lib=gpu) == [0] * len(reqs)
    for c in range(2):
        planes[c].close()
        assert want[c].any()
        assert np.array_equal(got[c], want[c])


@pytest.mark.parametrize("W,H", [(640, 360), (3840, 2160)])
@pytest.mark.parametrize("dft", [1, 2])
def test_mfma_spectrum_vs_oracle(gpu, oracle, dft, W, H):
    """The MFMA analysis (FP16 or BF16 operands split hi + lo, FP32
    accumulation) against the oracle's FP32 spectra: relative error below
    2e-6 (FP16) and 2e-5 (BF16) of the largest bin. At 4K the plane has
    1,029 groups of 32 blocks, more than one round of resident workgroups:
    the grid is that round, each workgroup with a longer run of blocks."""
    planes = make_clip(41, W, H, 1, "u10")
    from oracle_lib import Plane as OPlane
    kw = dict(bt=3, sigma=3.0)
    inst = Instance(1, W, H, 10, 0, 1, lib=gpu, dft=dft, **kw)
    pl = inst.objs[0]
    assert pl.info.dft == dft
    pl.forward(0, np.ascontiguousarray(planes[0][0]))
    got = pl.read_spectrum(0)
    inst.close()
    op = OPlane(oracle, script(**kw), W, H, 10)
    want = np.asarray(op.spectra(op.upload(planes[0][0]))).reshape(got.shape)
    rel = float(np.abs(got - want).max() / np.abs(want).max())
    print(f"MFMA dft={dft}: spectrum max error {rel:.2e} of the largest bin")
    assert rel < (2e-6 if dft == 1 else 2e-5)


MFMA_CONFIGS = ["A_640x360_u8_bt3", "B_1080p10_yuv_plane4_bt4", "C_4k10_bt4"]


@pytest.mark.parametrize("half", [False, True])
@pytest.mark.parametrize("dft", [1, 2])
@pytest.mark.parametrize("name", MFMA_CONFIGS)
def test_mfma_form_configs_vs_oracle(gpu, oracle, oracle64, name, dft, half):
    """Configs A-C in the MFMA DFT form (FP16 / BF16 hi+lo operands), with the
    FP32 or the FP16 spectra ring, within the parity bar of the oracle; the
    error against the FP64 oracle is printed per configuration."""
    kind, W, H, n, color, kw, cut = CONFIGS[name]
    bits, st = kind_info(kind)
    n = min(n, 4)
    planes = make_clip(1000 + len(name), W, H, n, kind, color=color, scene_cut=cut)
    cf = 3 if color == "yuv" else 1
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, color_family=cf)
    w64, _ = oracle64.run_clip(script(**kw), planes, bits, st, color_family=cf)
    got = FFT3DCL(planes, color_family=cf, bits=bits, sample_type=st, lib=gpu, dft=dft,
                  half=half, **kw)
    d = max(lsb_diff(g, w) for g, w in zip(got, want))
    d64 = max(lsb_diff(g, w) for g, w in zip(got, w64))
    print(f"{name} dft={dft} half={half}: max {d} LSB vs oracle, {d64} LSB vs FP64 oracle")
    for g, w in zip(got, want):
        check(kind, g, w)


def test_mfma_form_option_errors(gpu):
    for kw in (dict(bw=16, bh=16), dict(mode=0), dict(bw=32, bh=32, ow=10, oh=10)):
        pl = capi.Plane(capi.make_params(bt=3, **kw), 128, 96, 8, lib=gpu)
        with pytest.raises(capi.FFT3DError) as e:
            pl.set_option(capi.Plane.OPT_DFT, 1)
        assert e.value.code == -6
        pl.close()


def _tuned_oracle(args):
    """Oracle frames 1-3 of a tuning workload (worker process)."""
    w, h, bw = args
    from oracle_lib import Oracle, script
    clip = make_clip(300 + bw, w, h, 5, "u10")
    out, _ = Oracle().run_clip(script(sigma=3.0, bt=4, sharpen=0.4, bw=bw, bh=bw, ow=bw // 2,
                                      oh=bw // 2), clip, 10, 0, order=[1, 2, 3])
    return (w, h, bw), out[0][1:4]


def test_tuning_table_choices_within_parity(gpu):
    """Every entry of the committed tuning table (fft3dcl_amd/tuning.json): a
    plane of that block size, plane size, ring precision and DFT form picks
    the entry's band rows and sweep occupancy, and its output stays within
    the parity bar of the oracle (+/-4 LSB at 10 bit)."""
    import json
    import multiprocessing as mp
    from pathlib import Path
    tab = json.loads((Path(capi.PKG) / "tuning.json").read_text())["entries"]
    keys = sorted({(e["width"], e["height"], e["bw"]) for e in tab})
    with mp.get_context("spawn").Pool(8) as pool:
        want = dict(pool.map(_tuned_oracle, keys))
    for e in tab:
        w, h, bw = e["width"], e["height"], e["bw"]
        clip = make_clip(300 + bw, w, h, 5, "u10")
        inst = Instance(1, w, h, 10, 0, 5, lib=gpu, half=bool(e["half"]), dft=e["dft"],
                        sigma=3.0, bt=4, sharpen=0.4, bw=bw, bh=bw, ow=bw // 2, oh=bw // 2)
        info = inst.objs[0].info
        assert info.tuned == 1 and info.band_rows == e["band_rows"], e
        assert info.sweep_occupancy == e["occupancy"], e
        out = np.zeros_like(clip[0])
        for n in range(1, 4):
            inst.get_frame(n, [lambda fn: np.ascontiguousarray(clip[0][fn])], [out[n]])
        inst.close()
        d = lsb_diff(out[1:4], want[(w, h, bw)])
        assert d <= 4, (e, d)
