"""fft3d_plane_process_noalloc: the process call that allocates nothing (its
kernels' scratch is the library's static device array) against
fft3d_plane_process on twin planes fed the same frames. The outputs must be
bit-identical: the same kernels add the same partial sums in the same order,
only the partial planes live elsewhere. Two planes in a row, the second
smaller, so the second call runs over the first call's leftovers in the
shared array; a Kalman plane checks that the state update is the same; a
plane whose scratch exceeds the array is rejected with a clear error.
"""
import ctypes as C

import numpy as np
import pytest
import torch

from fft3dcl_amd import capi
from fft3dcl_amd.filter import Instance
from fft3dcl_amd.synth import make_plane_sequence

# (width, height, kind, params): the second is smaller than the first; the
# third is off the fused sweep (mode 0, blocks buffer in the scratch)
CASES = [
    (1280, 720, "u10", dict(bt=4, sharpen=0.4, sigma=3.0)),
    (640, 360, "u8", dict(bt=3)),
    (320, 200, "u8", dict(bt=3, mode=0)),
    (400, 240, "f32", dict(bt=0, sigma2=3.0)),
]


def windows(inst, n):
    return [inst.window(k) for k in range(n)]


def emu_noalloc(pl, window, out):
    w = (C.c_int * len(window))(*window)
    return pl.lib.fft3d_plane_process_noalloc(pl.h, w, len(window), out.ctypes.data,
                                              out.nbytes, None)


def test_emu_noalloc_equals_process(emu_lib):
    """CPU emulation: the same comparison at a quarter of the sizes, with host
    memory standing in for device memory (the emulated 'device address' of
    the array is its host address)."""
    for W, H, kind, kw in CASES:
        W, H, n = W // 4, H // 4, 5
        frames = make_plane_sequence(31, W, H, n, kind)
        bits = {"u8": 8, "u10": 10, "f32": 32}[kind]
        st = 1 if kind == "f32" else 0
        a = Instance(1, W, H, bits, st, n, lib=emu_lib, **kw)
        b = Instance(1, W, H, bits, st, n, lib=emu_lib, **kw)
        pa, pb = a.objs[0], b.objs[0]
        for k, win in enumerate(windows(a, n)):
            for fn in dict.fromkeys(win):
                for pl in (pa, pb):
                    if not pl.has(fn):
                        pl.forward(fn, np.ascontiguousarray(frames[fn]))
            want = np.zeros_like(frames[0])
            got = np.zeros_like(frames[0])
            pa.process(win, want)
            assert emu_noalloc(pb, win, got) == 0, pb.error
            assert np.array_equal(got, want), (W, H, kind, k)
        a.close()
        b.close()


@pytest.mark.gpu
def test_gpu_noalloc_equals_process(gpu):
    capi.dev_init(0, gpu)
    stream = torch.cuda.Stream()
    h = stream.cuda_stream
    for W, H, kind, kw in CASES:
        n = 6
        print(f"no-alloc case {W}x{H} {kind} {kw}", flush=True)
        frames = torch.from_numpy(make_plane_sequence(32, W, H, n, kind)).cuda()
        bits = {"u8": 8, "u10": 10, "f32": 32}[kind]
        st = 1 if kind == "f32" else 0
        a = Instance(1, W, H, bits, st, n, lib=gpu, **kw)
        b = Instance(1, W, H, bits, st, n, lib=gpu, **kw)
        pa, pb = a.objs[0], b.objs[0]
        want = torch.zeros((n,) + tuple(frames.shape[1:]), dtype=frames.dtype, device="cuda")
        got = torch.zeros_like(want)
        for k, win in enumerate(windows(a, n)):
            for fn in dict.fromkeys(win):
                for pl in (pa, pb):
                    if not pl.has(fn):
                        pl.forward(fn, frames[fn], True, h)
            pa.process(win, want[k], True, h)
            pb.process_noalloc(win, got[k], h)
        torch.cuda.synchronize()
        assert torch.equal(got, want), (W, H, kind)
        assert got.float().abs().sum() > 0
        a.close()
        b.close()


@pytest.mark.gpu
def test_gpu_noalloc_rejects_planes_over_the_limit(gpu):
    capi.dev_init(0, gpu)
    W, H = 4096, 4352  # 4 partial planes of W x H floats: 285 MB > 128 MiB
    inst = Instance(1, W, H, 8, 0, 1, bt=1, lib=gpu)
    pl = inst.objs[0]
    src = torch.zeros((H, W), dtype=torch.uint8, device="cuda")
    pl.forward(0, src, True)
    out = torch.zeros_like(src)
    with pytest.raises(capi.FFT3DError, match="128 MiB"):
        pl.process_noalloc([0], out)
    pl.process([0], out, True)  # the allocating call still serves it
    torch.cuda.synchronize()
    inst.close()
