"""Pins the oracle (oracle/fft3d_oracle.c) before it is trusted as the checker.

Each check is one a plausible restatement mistake would fail:
  * known answers of the reference's own tests (tools/fft_test.cpp:22-52);
  * closed forms: numpy FFTs, DC = sum, impulse -> flat, Parseval, the
    parameter formulas of src/params.h, the table formulas of
    src/cpu/tables.cpp evaluated independently in numpy;
  * invariants of the whole filter: identity reconstruction (bt=-1, no
    degrid/sharpen), COLA of the analysis x synthesis windows, sigma=0
    identity, a flat frame staying flat under degrid, a static clip under
    Kalman, the Kalman reset at a scene cut;
  * bit-exact agreement with the reference's own CPU backend compiled from its
    sources (oracle/_ref, when present), and with the committed golden
    fixtures that build produced (tests/golden).
"""
import json
from pathlib import Path

import numpy as np
import pytest

from fft3dcl_amd.synth import make_clip
from helpers import mt19937_planes
from oracle_lib import Oracle, Plane, script, kind_info

GOLDEN = Path(__file__).resolve().parent / "golden"


# ------------------------------------------------------------------ FFT

def test_fft_known_answer(oracle):
    # tools/fft_test.cpp:42-47: 8-point FFT of the ramp 0..7
    x = oracle.fft1d(np.arange(8, dtype=np.complex64))
    assert round(float(x[1].real), 4) == -4.0
    assert round(float(x[1].imag), 4) == 9.6569
    back = oracle.fft1d(x, inverse=True)
    assert np.abs(back - np.arange(8)).max() < 1e-5


@pytest.mark.parametrize("n", [2, 4, 8, 16, 32, 64, 128, 256, 512])
def test_fft1d_matches_numpy(oracle, n):
    rng = np.random.default_rng(n)
    x = (rng.standard_normal(n) + 1j * rng.standard_normal(n)).astype(np.complex64)
    got = oracle.fft1d(x)
    want = np.fft.fft(x.astype(np.complex128))
    assert np.abs(got - want).max() <= 1e-5 * np.abs(want).max()
    inv = oracle.fft1d(x, inverse=True)
    assert np.abs(inv - np.fft.ifft(x.astype(np.complex128))).max() < 1e-5


@pytest.mark.parametrize("bw,bh", [(4, 4), (8, 16), (32, 32), (64, 32), (16, 64)])
def test_fft2d_forward_closed_forms(oracle, bw, bh):
    rng = np.random.default_rng(bw * 100 + bh)
    blk = rng.uniform(-1, 1, (bh, bw)).astype(np.float32)
    full = oracle.fft2d_forward(blk)
    want = np.fft.fft2(blk.astype(np.float64))
    assert np.abs(full - want).max() <= 2e-5 * np.abs(want).max()
    assert abs(full[0, 0] - blk.astype(np.float64).sum()) < 1e-4  # DC = sum
    # Parseval
    assert np.isclose((np.abs(full) ** 2).sum() / (bw * bh), (blk.astype(np.float64) ** 2).sum(),
                      rtol=1e-5)
    imp = np.zeros((bh, bw), np.float32)
    imp[0, 0] = 1.0
    assert np.abs(oracle.fft2d_forward(imp) - 1.0).max() == 0.0  # impulse -> flat


def test_fft2d_half_roundtrip(oracle):
    # tools/fft_test.cpp:12-39: mt19937(12345) block, half-spectrum round trip
    bw = bh = 32
    rs = np.random.RandomState(12345)
    u = rs.randint(0, 2 ** 32, size=bw * bh, dtype=np.uint64).astype(np.float32)
    blk = (np.float32(-1) + np.float32(2) * (u / np.float32(4294967296.0))).reshape(bh, bw)
    full = oracle.fft2d_forward(blk)
    assert abs(full[0, 0].imag) < 1e-6
    back = oracle.fft2d_inverse_half(full[:, : bw // 2 + 1], bw)
    assert np.abs(back - blk).max() < 1e-5


# ----------------------------------------------------------- parameters

def test_round_pow2_clamp_closed_form(oracle):
    # src/params.h:62-72 rounds v up to the next power of two, then clamps.
    for v in range(1, 700):
        want = 1 << (v - 1).bit_length() if v > 1 else 1
        want = min(max(want, 4), 512)
        assert oracle.round_pow2_clamp(v) == want, v


@pytest.mark.parametrize("kw,w,h,bits", [
    (dict(), 640, 360, 8),
    (dict(bw=24, bh=40, ow=10, oh=9), 333, 201, 10),
    (dict(bw=64, bh=64), 2560, 1440, 32),
    (dict(bw=32, bh=32, ow=8, oh=12), 100, 80, 8),
    (dict(mode=0), 96, 72, 8),
    (dict(bt=0, sigma=3, kratio=1.5), 1920, 1080, 10),
])
def test_validate_closed_forms(oracle, kw, w, h, bits):
    sp = script(**kw)
    ip, err = oracle.validate(sp, w, h, bits)
    assert err is None
    rp2 = lambda v: min(max(1 << (v - 1).bit_length(), 4), 512)
    bw, bh = rp2(sp.bw), rp2(sp.bh)
    ow = int(np.float32(sp.ow) * np.float32(bw) / np.float32(sp.bw)) & ~1
    oh = int(np.float32(sp.oh) * np.float32(bh) / np.float32(sp.bh))
    oh = min(oh, bh // 2)
    ow = min(ow, bw // 2)
    assert (ip.bw, ip.bh, ip.ow, ip.oh) == (bw, bh, ow, oh)
    if ip.mode == 0 or (bw == 2 * ow and bh == 2 * oh):
        nx, ny = int((w + bw * 1.5 - 1) / bw), int((h + bh * 1.5 - 1) / bh)
    else:
        nx = ((w + ow - 1) // (bw - ow) + 2) // 2
        ny = ((h + oh - 1) // (bh - oh) + 2) // 2
    assert (ip.nx, ip.ny) == (nx, ny)
    f = np.float32
    s = f(sp.sigma) / f(255)
    bt = max(sp.bt, 1)
    assert ip.wienerS == f(f(f(f(s * s) * f(bt)) * f(bw)) * f(bh))
    assert ip.kalmanCN == f(f(f(s * s) * f(bw)) * f(bh))
    assert ip.current == (0 if sp.bt == 1 else 2 if sp.bt == 4 else 1)
    assert ip.precision == (2 if bits > 8 else 0)


def test_validate_errors(oracle):
    msgs = [
        (dict(bt=5), "FFT3DCL: valid modes for bt are -1,0,1,2,3,4"),
        (dict(bt=-2), "FFT3DCL: valid modes for bt are -1,0,1,2,3,4"),
        (dict(mode=2), "FFT3DCL: mode 2 is not implemented yet in this port (use mode 0 or 1)"),
        (dict(interlaced=1), "FFT3DCL: interlaced is not implemented yet in this port"),
    ]
    for kw, msg in msgs:
        ip, err = oracle.validate(script(**kw), 64, 48, 8)
        assert ip is None and err == msg


def test_validate_matches_reference_build(oracle, ref):
    rng = np.random.default_rng(7)
    for _ in range(300):
        kw = dict(bw=int(rng.integers(2, 600)), bh=int(rng.integers(2, 600)),
                  bt=int(rng.integers(-1, 5)), sigma=float(rng.uniform(0, 8)),
                  beta=float(rng.uniform(0.5, 3)), mode=int(rng.integers(0, 2)),
                  sharpen=float(rng.choice([0.0, 0.4])), smin=float(rng.uniform(0, 8)),
                  smax=float(rng.uniform(8, 40)))
        kw["ow"] = int(rng.integers(0, kw["bw"] + 1))
        kw["oh"] = int(rng.integers(0, kw["bh"] + 1))
        sp = script(**kw)
        w, h, bits = int(rng.integers(8, 4000)), int(rng.integers(8, 2200)), int(
            rng.choice([8, 10, 16, 32]))
        a, ea = oracle.validate(sp, w, h, bits)
        b, eb = ref.validate(sp, w, h, bits)
        assert ea == eb
        if a is not None:
            assert a.as_dict() == b.as_dict(), kw


# --------------------------------------------------------------- tables

def _ramps(o, wintype):
    i = np.arange(o)
    c = np.cos(np.pi * (i - o + 0.5) / (o * 2.0))
    if wintype == 1:
        return np.sqrt(c), np.sqrt(c) ** 3
    if wintype == 2:
        return np.ones(o), c * c
    return c, c


@pytest.mark.parametrize("wintype", [0, 1, 2])
@pytest.mark.parametrize("bw,bh,ow,oh", [(32, 32, 16, 16), (16, 32, 4, 10), (64, 64, 32, 32)])
def test_windows_closed_form(oracle, wintype, bw, bh, ow, oh):
    pl = Plane(oracle, script(bw=bw, bh=bh, ow=ow, oh=oh, wintype=wintype), 200, 150, 8)
    t = pl.tables()
    ax, sx = _ramps(pl.ip.ow, wintype)
    ay, sy = _ramps(pl.ip.oh, wintype)

    def prof(n, o, r, rev):
        w = np.ones(n)
        w[:o] = r[::-1] if rev else r
        w[n - o:] = r if rev else r[::-1]
        return w

    A = np.outer(prof(bh, pl.ip.oh, ay, False), prof(bw, pl.ip.ow, ax, False))
    SA = np.outer(prof(bh, pl.ip.oh, sy, False), prof(bw, pl.ip.ow, sx, False))
    SBx = prof(bw, pl.ip.ow, sx, True)
    SB = np.outer(prof(bh, pl.ip.oh, sy, False), SBx[(np.arange(bw) + bw - pl.ip.ow) % bw])
    assert np.array_equal(t["winAna"], A.astype(np.float32))
    assert np.array_equal(t["winSynA"], SA.astype(np.float32))
    assert np.array_equal(t["winSynB"], SB.astype(np.float32))


def test_window_mode0_closed_form(oracle):
    pl = Plane(oracle, script(mode=0, bw=16, bh=32), 64, 64, 8)
    t = pl.tables()
    cy = np.cos((np.arange(32) - 16 + 0.5) * np.pi / 32)
    cx = np.cos((np.arange(16) - 8 + 0.5) * np.pi / 16)
    w = np.sqrt(0.5 * (cx[None, :] ** 2 + cy[:, None] ** 2)).astype(np.float32)
    assert np.array_equal(t["winAna"], w)


@pytest.mark.parametrize("wintype", [0, 1, 2])
def test_window_cola(oracle, wintype):
    # analysis x synthesis windows overlap-add to 1 with 50 % overlap
    W, H = 160, 128
    pl = Plane(oracle, script(wintype=wintype), W, H, 8)
    t = pl.tables()
    acc = np.zeros((H + 64, W + 64))
    for b in range(pl.nb):
        x0, y0 = t["bx"][b] + 32, t["by"][b] + 32
        s = t["winSynB"] if t["par"][b] else t["winSynA"]
        acc[y0:y0 + 32, x0:x0 + 32] += t["winAna"].astype(np.float64) * s
    assert np.abs(acc[32:32 + H, 32:32 + W] - 1.0).max() < 1e-6


def test_pattern_closed_form(oracle):
    sp = script(sigma=2.0, sigma2=3.0, sigma3=1.5, sigma4=1.0, bt=3)
    pl = Plane(oracle, sp, 96, 96, 8)
    pat = pl.tables()["pattern"]
    bw = bh = 32
    fbw = 17
    sc = lambda s: (s / 255.0) ** 2 * 3 * bw * bh
    s, s2, s3, s4 = sc(2.0), sc(3.0), sc(1.5), sc(1.0)
    ft2, ft3 = np.sqrt(0.5) / 2, np.sqrt(0.5) / 4
    for h in range(bh):
        fy = (bh - 2 * abs(h - bh / 2)) / bh
        for w in range(fbw):
            f = np.sqrt(((w / fbw) ** 2 + fy ** 2) * 0.5)
            if f < ft3:
                p = s4 + (s3 - s4) * f / ft3
            elif f < ft2:
                p = s3 + (s2 - s3) * (f - ft3) / (ft2 - ft3)
            else:
                p = s + (s2 - s) * (1 - f) / (1 - ft2)
            assert abs(pat[h, w] - p) <= 1e-5 * abs(p), (h, w)


def test_sharpen_lut_closed_form(oracle):
    pl = Plane(oracle, script(sharpen=0.7, svr=1.3, scutoff=0.25), 96, 96, 8)
    lut = pl.tables()["sharpen"]
    bw = bh = 32
    for j in range(bh):
        dj = j if j < bh // 2 else bh - j
        for i in range(17):
            d2 = dj * dj * 1.3 ** 2 / (bh / 2) ** 2 + i * i / (bw / 2) ** 2
            want = 0.7 * (1 - np.exp(-d2 / (2 * 0.25 ** 2)))
            assert abs(lut[j, i] - want) <= 2e-6, (j, i)


def test_degrid_sample_closed_form(oracle):
    pl = Plane(oracle, script(degrid=0.8), 96, 96, 8)
    pl.build_degrid()
    t = pl.tables()
    spec = np.fft.fft2(t["winAna"].astype(np.float64))
    want = spec[:, :17] / spec[0, 0].real * 0.8
    assert np.abs(t["degrid"] - want).max() < 1e-5


# ------------------------------------------------------------ invariants

@pytest.mark.parametrize("kind", ["u8", "u10", "f32"])
@pytest.mark.parametrize("kw", [dict(), dict(bw=16, bh=16), dict(mode=0), dict(wintype=2),
                                dict(bw=64, bh=32)])
def test_identity_reconstruction(oracle, kind, kw):
    bits, st = kind_info(kind)
    planes = make_clip(11, 96, 72, 2, kind)
    out, fl = oracle.run_clip(script(bt=-1, degrid=0.0, **kw), planes, bits, st, want_float=True)
    if kind == "f32":
        assert np.abs(out[0] - planes[0]).max() < 2e-6
    else:
        assert np.array_equal(out[0], planes[0])


def test_sigma_zero_is_identity(oracle):
    planes = make_clip(12, 96, 72, 1, "f32")
    out, _ = oracle.run_clip(script(bt=1, sigma=0.0, degrid=0.0), planes, 32, 1)
    assert np.abs(out[0] - planes[0]).max() < 2e-6


def test_flat_frame_stays_flat_with_degrid(oracle):
    planes = [np.full((2, 80, 96), 0.5, np.float32)]
    flat, _ = oracle.run_clip(script(bt=1, sigma=4.0), planes, 32, 1)
    grid, _ = oracle.run_clip(script(bt=1, sigma=4.0, degrid=0.0), planes, 32, 1)
    assert np.abs(flat[0] - 0.5).max() < 1e-5
    # without degrid the Wiener filter bites into the window's own spectrum
    assert np.abs(grid[0] - 0.5).max() > 10 * np.abs(flat[0] - 0.5).max()


def test_kalman_static_clip_reproduces_input(oracle):
    one = make_clip(13, 96, 72, 1, "f32")[0]
    planes = [np.repeat(one, 4, axis=0)]
    out, _ = oracle.run_clip(script(bt=0, degrid=0.0), planes, 32, 1)
    assert np.abs(out[0] - planes[0]).max() < 5e-6


def test_kalman_resets_at_scene_cut(oracle):
    planes = make_clip(14, 128, 96, 8, "f32", scene_cut=4)
    out, _ = oracle.run_clip(script(bt=0, kratio=2.0), planes, 32, 1)
    still, _ = oracle.run_clip(script(bt=0, kratio=1e9), planes, 32, 1)
    cut = planes[0][4]
    # with the reset the first frame after the cut follows the new scene;
    # with an unreachable threshold it keeps averaging the old one
    assert np.abs(out[0][4] - cut).mean() < 0.2 * np.abs(still[0][4] - cut).mean()


def test_fp64_twin_close(oracle, oracle64):
    planes = make_clip(15, 96, 72, 5, "f32")
    for bt in (-1, 0, 1, 2, 3, 4):
        a, _ = oracle.run_clip(script(bt=bt, sharpen=0.3), planes, 32, 1)
        b, _ = oracle64.run_clip(script(bt=bt, sharpen=0.3), planes, 32, 1)
        assert np.abs(a[0] - b[0]).max() < 1e-5, bt


# ------------------------------------- against the reference's own build

CASES = [
    ("u8", dict(bt=3)), ("u8", dict(bt=-1)), ("u8", dict(bt=0)), ("u8", dict(bt=1)),
    ("u8", dict(bt=2, sharpen=0.6)), ("u8", dict(bt=4, degrid=0.0)),
    ("u10", dict(bt=4, sharpen=0.4, sigma=3.0)), ("u10", dict(bt=0, kratio=1.5)),
    ("u16", dict(bt=3, beta=1.7)), ("f32", dict(bt=0, sigma2=3.0, sigma3=2.0, sigma4=1.0,
                                                 bw=64, bh=64)),
    ("f32", dict(bt=3, mode=0)), ("u8", dict(bt=3, bw=16, bh=8, ow=6, oh=2)),
    ("u8", dict(bt=2, wintype=1)), ("u10", dict(bt=3, wintype=2)),
    ("u8", dict(bt=3, bw=128, bh=16)),
]


@pytest.mark.parametrize("kind,kw", CASES)
def test_oracle_bit_exact_vs_reference_build(oracle, ref, kind, kw):
    bits, st = kind_info(kind)
    planes = make_clip(21, 150, 90, 6, kind)
    sp = script(**kw)
    order = [0, 1, 2, 5, 4, 3]  # includes a backwards seek
    a, fa = oracle.run_clip(sp, planes, bits, st, order=order, want_float=True)
    b, fb = ref.run_clip(sp, planes, bits, st, order=order, want_float=True)
    assert np.array_equal(a[0], b[0])
    assert np.array_equal(fa[0], fb[0])


def test_oracle_yuv_plane4_vs_reference_build(oracle, ref):
    planes = make_clip(22, 128, 96, 4, "u10", color="yuv")
    sp = script(bt=3, plane=4, sharpen=0.3)
    a, _ = oracle.run_clip(sp, planes, 10, 0, color_family=3)
    b, _ = ref.run_clip(sp, planes, 10, 0, color_family=3)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)


def test_gpu_test_inputs_oracle_vs_reference_build(oracle, ref):
    # tools/gpu_test.cpp:52-58 inputs, default ScriptParams, float domain
    pl = mt19937_planes(42, 256, 192, 3)
    planes = [np.stack(pl)]
    a, fa = oracle.run_clip(script(bt=3), planes, 32, 1, want_float=True)
    b, fb = ref.run_clip(script(bt=3), planes, 32, 1, want_float=True)
    assert np.array_equal(fa[0], fb[0])


# --------------------------------------------------------- golden fixtures

def _golden_cases():
    f = GOLDEN / "cases.json"
    return json.loads(f.read_text()) if f.exists() else []


@pytest.mark.parametrize("case", _golden_cases(), ids=lambda c: c["name"])
def test_oracle_matches_golden(oracle, case):
    d = np.load(GOLDEN / case["file"], allow_pickle=False)
    planes = [d[f"in{q}"] for q in range(case["nplanes"])]
    outs, _ = oracle.run_clip(script(**case["params"]), planes, case["bits"],
                              case["sample_type"], color_family=case["color_family"],
                              order=case["order"])
    for q in range(case["nplanes"]):
        assert np.array_equal(outs[q], d[f"out{q}"]), q
