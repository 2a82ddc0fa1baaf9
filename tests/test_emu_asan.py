"""Runs a set of emulated kernel launches (tests/emu, thread-per-lane CPU
emulation of the HIP kernels) under AddressSanitizer in a child process, so an
out-of-bounds global or LDS access in the kernels' index arithmetic fails on
the CPU before it could fault a GPU. The host buffers go through the pinned
chunks of the staged transfers (ragged chunks, and the threaded copy of a
plane above 256 KiB), and with the sweep on each request prefetches the next
frame; a batched step call of five frames of two planes covers the step's
host side. LDS arrays are function statics in the
emulation, so their overruns are caught as global-buffer overflows."""
import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
ASAN_LIB = ROOT / "tests" / "_build" / "asan" / "libfft3d_emu.so"

SCRIPT = r"""
import sys, numpy as np
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
from fft3dcl_amd.filter import FFT3DCL, Instance
from fft3dcl_amd.synth import make_clip
from fft3dcl_amd import capi
lib = {lib!r}
cases = [("u8", 8, 0, dict(bt=3)), ("u10", 10, 0, dict(bt=4, sharpen=0.4)),
         ("f32", 32, 1, dict(bt=0, bw=64, bh=64)), ("u8", 8, 0, dict(bt=2, bw=8, bh=16)),
         ("u8", 8, 0, dict(bt=-1, bw=4, bh=4)), ("u8", 8, 0, dict(bt=3, mode=0)),
         ("u16", 16, 0, dict(bt=1, bw=32, bh=16, ow=6, oh=4))]
for kind, bits, st, kw in cases:
    for W, H in ((83, 45), (300, 70)):
        planes = make_clip(1, W, H, 3, kind)
        for sweep in (1, 0):
            inst = Instance(1, W, H, bits, st, 3, lib=lib, prefetch=bool(sweep), **kw)
            inst.objs[0].set_option(1, sweep)
            inst.objs[0].set_option(7, 8192 if sweep else 0)  # pinned chunks of 8 KiB
            out = np.zeros_like(planes[0])
            for n in range(3):
                inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
            inst.close()
# a plane above 256 KiB: one chunk copied by several host threads
planes = make_clip(2, 300, 220, 2, "f32")
inst = Instance(1, 300, 220, 32, 1, 2, lib=lib, bt=2)
out = np.zeros_like(planes[0])
for n in range(2):
    inst.get_frame(n, [lambda fn: np.ascontiguousarray(planes[0][fn])], [out[n]])
inst.close()
# the batched step's host side (descriptors of rounds launched in chunks):
# five frames of two planes in one call, device flags on host buffers (the
# emulation's device memory is host memory)
from test_capi_cpu import _raw_step
pls, reqs = [], []
for c, kw in enumerate((dict(bt=4), dict(bt=3, bw=16, bh=16))):
    pl = capi.Plane(capi.make_params(**kw), 96, 64, 8, lib=lib)
    clip = make_clip(90 + c, 96, 64, 5, "u8")[0]
    out = np.zeros_like(clip)
    pls.append((pl, clip, out))
    reqs += [(pl, f, 5, clip, out[f], True) for f in range(5)]
reqs.sort(key=lambda r: r[1])
assert _raw_step(lib, reqs) == [0] * len(reqs)
for pl, _, out in pls:
    assert out.any()
    pl.close()
print("asan-ok")
"""


def test_emulated_kernels_under_asan():
    if not ASAN_LIB.exists():
        pytest.skip("ASan emulation build missing (make -C tests/emu)")
    libasan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True,
                             text=True).stdout.strip()
    env = dict(os.environ, LD_PRELOAD=libasan, ASAN_OPTIONS="detect_leaks=0")
    code = SCRIPT.format(root=str(ROOT), tests=str(ROOT / "tests"), lib=str(ASAN_LIB))
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True,
                       timeout=900)
    assert r.returncode == 0 and "asan-ok" in r.stdout, (r.stdout[-2000:] + r.stderr[-4000:])
