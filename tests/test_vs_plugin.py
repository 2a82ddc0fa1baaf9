"""The VapourSynth API 4 plugin (fft3dcl_amd/csrc/fft3dcl_vs.cpp) driven by a
fake single-threaded host: registration record, argument defaults, error
strings, the arInitial request window, filter mode, reference counting, and
output parity with the oracle's whole-clip path (src/fft3dcl.cpp semantics).

CPU: the plugin built against the C-ABI emulation (tests/_build). GPU: the
shipped plugin over libfft3d_hip.so.
"""
from pathlib import Path

import numpy as np
import pytest

from fakevs import Host, FAKE
from fft3dcl_amd.synth import make_clip
from helpers import lsb_diff, max_abs
from oracle_lib import script, kind_info

ROOT = Path(__file__).resolve().parent.parent
EMU_PLUGIN = ROOT / "tests" / "_build" / "libfft3dcl_vs_emu.so"
HIP_PLUGIN = ROOT / "fft3dcl_amd" / "libfft3dcl_vs.so"

ARGS = ("clip:vnode;"
        "sigma:float:opt;beta:float:opt;bw:int:opt;bh:int:opt;bt:int:opt;sharpen:float:opt;"
        "plane:int:opt;mode:int:opt;bordersize:int:opt;precision:int:opt;degrid:float:opt;"
        "scutoff:float:opt;svr:float:opt;smin:float:opt;smax:float:opt;kratio:float:opt;"
        "ow:int:opt;oh:int:opt;wintype:int:opt;interlaced:int:opt;sigma2:float:opt;"
        "sigma3:float:opt;sigma4:float:opt;oldfft:int:opt;device:int:opt")


@pytest.fixture(scope="module")
def host():
    if not FAKE.exists() or not EMU_PLUGIN.exists():
        pytest.skip("fake host / emulation plugin not built (make -C tests/vs_host)")
    return Host()


def test_registration_record(host):
    # src/fft3dcl.cpp:540-552
    ident, ns, name, func, args = host.info(str(EMU_PLUGIN))
    assert (ident, ns, name, func) == ("com.github.humbleuser33.fft3dcl", "fft3dcl",
                                      "FFT3DCL denoiser (VS port)", "FFT3DCL")
    assert args == ARGS


def test_shipped_plugin_registration(host):
    if not HIP_PLUGIN.exists():
        pytest.skip("libfft3dcl_vs.so not built")
    assert host.info(str(HIP_PLUGIN))[4] == ARGS


@pytest.mark.parametrize("kw,msg", [
    (dict(bt=9), "FFT3DCL: valid modes for bt are -1,0,1,2,3,4"),
    (dict(mode=2), "FFT3DCL: mode 2 is not implemented yet in this port (use mode 0 or 1)"),
    (dict(plane=1), "FFT3DCL: cannot set chroma processing for a greyscale clip"),
    (dict(plane=7), "FFT3DCL: plane must be 0 (luma), 1 (chroma) or 4 (all)"),
    (dict(device=0), "FFT3DCL: device=0 selects the CPU backend, which this HIP build does "
                     "not include (use device=1 or 2)"),
    (dict(bw=200), "FFT3DCL: GPU setup failed: block too large for device local memory (max 64x64)"),
])
def test_create_errors(host, kw, msg):
    planes = make_clip(1, 64, 48, 3, "u8")
    rc, _, _, _, err, _ = host.run(str(EMU_PLUGIN), planes, **kw)
    assert rc == 1 and err == msg  # rc 3 would mean the node reference leaked


def test_format_errors(host):
    planes = make_clip(1, 64, 48, 3, "u16")
    rc, *_, err, _ = host.run(str(EMU_PLUGIN), planes, bits=7)
    assert rc == 1 and err == "FFT3DCL: integer bit depths 8-16 only"
    rc, *_, err, _ = host.run(str(EMU_PLUGIN), planes, bits=16, sample_type=1)
    assert rc == 1 and err.startswith("FFT3DCL: only 32-bit float clips supported")


@pytest.mark.parametrize("bt,current", [(-1, None), (0, None), (1, 0), (2, 1), (3, 1), (4, 2)])
def test_request_window_and_filter_mode(host, bt, current):
    n = 6
    planes = make_clip(2, 64, 48, n, "u8")
    order = [0, 1, 5, 2, 3, 4] if bt != 0 else list(range(n))
    rc, _, reqs, fm, err, _ = host.run(str(EMU_PLUGIN), planes, bt=bt, order=order)
    assert rc == 0, err
    assert fm == (3 if bt == 0 else 1)  # fmFrameState for Kalman, else fmParallelRequests
    for k, f in enumerate(order):
        if bt >= 2:
            want = sorted({min(max(f - current + i, 0), n - 1) for i in range(bt)})
        else:
            want = [f]
        assert reqs[k] == want


@pytest.mark.parametrize("bt,current,pad", [(3, 1, 0), (4, 2, 16)])
def test_prefetch_env(host, monkeypatch, bt, current, pad):
    """FFT3DCL_PREFETCH=1: each request also asks for the frame the next
    sequential request adds, the plugin prefetches it (through its own copy
    when rows are padded), and the output is unchanged."""
    n = 7
    planes = make_clip(5, 96, 64, n, "u8")
    rc0, outs0, _, _, err0, _ = host.run(str(EMU_PLUGIN), planes, bt=bt, pad=pad)
    monkeypatch.setenv("FFT3DCL_PREFETCH", "1")
    rc1, outs1, reqs, _, err1, _ = host.run(str(EMU_PLUGIN), planes, bt=bt, pad=pad)
    assert rc0 == 0 and rc1 == 0, err0 + err1
    for f in range(n):
        want = {min(max(f - current + i, 0), n - 1) for i in range(bt)}
        if f - current + bt < n:
            want.add(f - current + bt)
        assert reqs[f] == sorted(want)
    assert outs0[0].any()
    assert np.array_equal(outs0[0], outs1[0])


CASES = [("u8", dict(bt=3), 1, 0), ("u10", dict(bt=4, sharpen=0.4, sigma=3.0), 1, 16),
         ("u8", dict(bt=0), 1, 0), ("f32", dict(bt=2, sigma2=3.0), 1, 8),
         ("u8", dict(plane=4, bt=3), 3, 0), ("u10", dict(plane=1, bt=4), 3, 24),
         ("u8", dict(bt=1), 2, 0), ("u8", dict(bt=3, bw=128, bh=32), 1, 0)]


def run_vs_vs_oracle(host, oracle, plugin, kind, kw, cf, pad, W=96, H=64, n=5):
    bits, st = kind_info(kind)
    color = "yuv" if cf == 3 else "gray"
    planes = make_clip(3, W, H, n, kind, color=color)
    if cf == 2:  # planar RGB: three full-size planes
        planes = make_clip(3, W, H, n, kind, color="yuv", ssw=0, ssh=0)
    rc, outs, _, _, err, log = host.run(str(plugin), planes, color_family=cf, bits=bits,
                                         sample_type=st, ssw=0 if cf == 2 else 1,
                                         ssh=0 if cf == 2 else 1, pad=pad, **kw)
    assert rc == 0, err
    assert "GPU instance(s) alive" in log
    want, _ = oracle.run_clip(script(**kw), planes, bits, st, color_family=cf,
                              ssw=0 if cf == 2 else 1, ssh=0 if cf == 2 else 1)
    for g, w in zip(outs, want):
        if kind == "f32":
            assert max_abs(g, w) < 1e-3
        else:
            assert lsb_diff(g, w) <= (2 if kind == "u8" else 4)


@pytest.mark.parametrize("kind,kw,cf,pad", CASES)
def test_plugin_output_matches_oracle_emu(host, oracle, kind, kw, cf, pad):
    run_vs_vs_oracle(host, oracle, EMU_PLUGIN, kind, kw, cf, pad)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,kw,cf,pad", CASES)
def test_plugin_output_matches_oracle_gpu(gpu, host, oracle, kind, kw, cf, pad):
    run_vs_vs_oracle(host, oracle, HIP_PLUGIN, kind, kw, cf, pad, W=640, H=360, n=6)


@pytest.mark.gpu
@pytest.mark.parametrize("kind,kw,cf,pad", [CASES[1], CASES[5]])
def test_plugin_prefetch_matches_oracle_gpu(gpu, host, oracle, monkeypatch, kind, kw, cf, pad):
    """The shipped plugin with FFT3DCL_PREFETCH=1 on the GPU (upload stream,
    events, copy-kernel readback) against the oracle."""
    monkeypatch.setenv("FFT3DCL_PREFETCH", "1")
    run_vs_vs_oracle(host, oracle, HIP_PLUGIN, kind, kw, cf, pad, W=640, H=360, n=8)
