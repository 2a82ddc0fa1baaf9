"""The served form (fft3dcl_amd/serve.py): many streams of mixed geometry
and temporal mode through one request loop, with admission by device
footprint. Each stream's frames must be the same in the mix as when the
stream is processed alone (bit for bit: every stream has its own plane
objects and the kernels are deterministic)."""
import numpy as np
import pytest

from fft3dcl_amd.filter import GRAY, Instance
from fft3dcl_amd.serve import Server, footprint, form_options, make_specs, summarize
from fft3dcl_amd.synth import make_plane_sequence


def _clips(specs, on_device):
    clips = {}
    for sp in specs:
        kind = "u8" if sp.bits == 8 else "u10"
        a = make_plane_sequence(300 + sp.sid, sp.width, sp.height, sp.nframes, kind)
        if on_device:
            import torch
            a = torch.from_numpy(a).cuda()
        clips[sp.sid] = a
    return clips


def _alone(sp, clip, lib, on_device, form="fp32"):
    inst = Instance(GRAY, sp.width, sp.height, sp.bits, 0, sp.nframes, lib=lib,
                    **form_options(form, sp), **sp.params)
    if on_device:
        import torch
        out = torch.zeros_like(clip)
    else:
        out = np.zeros_like(clip)
    for n in range(sp.nframes):
        inst.get_frame(n, [lambda fn: clip[fn]], [out[n]], on_device=on_device)
    inst.close()
    if on_device:
        import torch
        torch.cuda.synchronize()
        return out.cpu().numpy()
    return out


def _mix(specs, clips, lib, on_device, budget, paced, batched=False):
    if on_device:
        import torch
        outs = {sp.sid: torch.zeros_like(clips[sp.sid]) for sp in specs}
    else:
        outs = {sp.sid: np.zeros_like(clips[sp.sid]) for sp in specs}
    srv = Server(specs, {sid: (lambda fn, c=c: c[fn]) for sid, c in clips.items()},
                 lambda sid, n: outs[sid][n], budget, lib=lib, on_device=on_device,
                 batched=batched)
    res = srv.run(paced=paced)
    if on_device:
        import torch
        torch.cuda.synchronize()
        outs = {k: v.cpu().numpy() for k, v in outs.items()}
    return outs, res


def test_specs_are_seeded_and_mixed():
    a = make_specs(5, 32, duration=0.5)
    b = make_specs(5, 32, duration=0.5)
    assert [(s.width, s.bits, s.params, s.arrivals) for s in a] == \
           [(s.width, s.bits, s.params, s.arrivals) for s in b]
    assert {s.params["bt"] for s in a} == {0, 2, 3, 4}
    assert {s.params["bw"] for s in a} == {16, 32, 64}
    assert all(24.0 <= s.rate <= 60.0 and s.nframes >= 1 for s in a)
    assert all(np.all(np.diff(s.arrivals) > 0) for s in a)


@pytest.mark.parametrize("batched", [False, True])
def test_emu_mix_equals_alone_with_staged_admission(emu_lib, batched):
    specs = make_specs(11, 6, duration=0.2, resolutions=[(96, 64), (128, 80), (160, 96)],
                       max_frames=4)
    clips = _clips(specs, False)
    budget = max(footprint(s) for s in specs) * 2  # about two streams at a time
    outs, res = _mix(specs, clips, emu_lib, False, budget, paced=False, batched=batched)
    assert res["requests"] == sum(s.nframes for s in specs)
    assert 1 <= res["max_admitted"] < len(specs)
    for sp in specs:
        assert np.array_equal(outs[sp.sid], _alone(sp, clips[sp.sid], emu_lib, False)), sp


@pytest.mark.gpu
@pytest.mark.parametrize("batched", [False, True])
def test_gpu_mix_equals_alone(gpu, batched):
    specs = make_specs(12, 8, duration=0.25, resolutions=[(640, 360), (1280, 720)],
                       max_frames=6)
    clips = _clips(specs, True)
    outs, res = _mix(specs, clips, gpu, True, 64 << 30, paced=True, batched=batched)
    s = summarize(specs, res)
    assert res["requests"] == sum(sp.nframes for sp in specs)
    assert s["p99_ms"] >= s["p50_ms"] > 0
    for sp in specs:
        assert np.array_equal(outs[sp.sid], _alone(sp, clips[sp.sid], gpu, True)), sp


def _oracle_stream(args):
    """Oracle output of one served stream (runs in a worker process: the clip
    is regenerated there from its seed)."""
    sid, w, h, bits, params, nframes = args
    from oracle_lib import Oracle, script
    clip = make_plane_sequence(300 + sid, w, h, nframes, "u8" if bits == 8 else "u10")
    out, _ = Oracle().run_clip(script(**params), [clip], bits, 0)
    return sid, out[0]


@pytest.mark.gpu
def test_gpu_config_e_full_shape_vs_oracle(gpu):
    """The served configuration at its full shape: 32 seeded streams over all
    four resolutions (640x360 to 3840x2160), 8/10-bit, bt 0/2/3/4 and
    bw = bh 16/32/64, up to 3 frames each. Every stream's output, per request
    and batched through fft3d_step, is within the parity bar of the oracle
    (+/-2 LSB at 8 bit, +/-4 at 10 bit), the two loops agree bit for bit,
    and each stream alone gives the same frames as in the mix."""
    import multiprocessing as mp
    specs = make_specs(1002, 32, duration=0.5, max_frames=3)
    assert {(s.width, s.height) for s in specs} == {(640, 360), (1280, 720), (1920, 1080),
                                                    (3840, 2160)}
    jobs = [(s.sid, s.width, s.height, s.bits, s.params, s.nframes) for s in specs]
    with mp.get_context("spawn").Pool(8) as pool:
        want = dict(pool.map(_oracle_stream, jobs))
    clips = _clips(specs, True)
    outs, res = _mix(specs, clips, gpu, True, 64 << 30, paced=True)
    outb, _ = _mix(specs, clips, gpu, True, 64 << 30, paced=False, batched=True)
    assert res["requests"] == sum(sp.nframes for sp in specs)
    for sp in specs:
        tol = 2 if sp.bits == 8 else 4
        d = int(np.abs(outs[sp.sid].astype(np.int64) - want[sp.sid].astype(np.int64)).max())
        assert d <= tol, (sp.sid, sp.width, sp.bits, sp.params, d)
        assert np.array_equal(outb[sp.sid], outs[sp.sid]), sp.sid
        assert np.array_equal(_alone(sp, clips[sp.sid], gpu, True), outs[sp.sid]), sp.sid
