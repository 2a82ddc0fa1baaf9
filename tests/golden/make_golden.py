"""Writes the golden fixtures in this folder from the reference's own CPU
backend (oracle/_ref/libfft3dref.so, built by oracle/Makefile from the
reference sources in this container).

Each fixture is data only: seeded synthetic input planes and the reference
build's output planes for one parameter set and request order. They pin the
oracle on machines where the reference sources are absent.

    python tests/golden/make_golden.py
"""
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parent))
sys.path.insert(0, str(HERE.parent.parent))

from fft3dcl_amd.synth import make_clip  # noqa: E402
from oracle_lib import Ref, script, kind_info  # noqa: E402

CASES = [
    ("default_u8", "u8", "gray", dict(), [0, 1, 2, 3, 4]),
    ("bt4_sharpen_u10", "u10", "gray", dict(bt=4, sharpen=0.4, sigma=3.0), [0, 1, 2, 3, 4]),
    ("bt2_u8", "u8", "gray", dict(bt=2), [0, 1, 2, 3, 4]),
    ("bt1_pattern_u8", "u8", "gray", dict(bt=1, sigma2=3.0, sigma3=1.5, sigma4=0.5), [0, 1, 2]),
    ("bt0_kalman_f32", "f32", "gray", dict(bt=0, kratio=1.5, sigma2=3.0, sigma3=2.0,
                                             sigma4=1.0, bw=64, bh=64), [0, 1, 2, 3, 4]),
    ("btm1_sharpen_u16", "u16", "gray", dict(bt=-1, sharpen=1.0), [0, 1]),
    ("mode0_u8", "u8", "gray", dict(mode=0, bt=3), [0, 1, 2, 3, 4]),
    ("overlap_u8", "u8", "gray", dict(bw=32, bh=16, ow=10, oh=4, wintype=1), [0, 1, 2, 3, 4]),
    ("wintype2_beta_u10", "u10", "gray", dict(wintype=2, beta=1.6, degrid=0.5), [4, 3, 2, 1, 0]),
    ("yuv_plane4_u8", "u8", "yuv", dict(plane=4, bt=3, sharpen=0.3), [0, 1, 2, 3]),
    ("yuv_chroma_u10", "u10", "yuv", dict(plane=1, bt=4), [0, 1, 2, 3]),
    ("bw8_u8", "u8", "gray", dict(bw=8, bh=8, bt=3), [0, 2, 1, 3]),
]


def main():
    ref = Ref.load()
    if ref is None:
        raise SystemExit("oracle/_ref/libfft3dref.so is not built (make -C oracle)")
    meta = []
    for i, (name, kind, color, kw, order) in enumerate(CASES):
        bits, st = kind_info(kind)
        planes = make_clip(100 + i, 72, 56, 5, kind, color=color)
        cf = 3 if color == "yuv" else 1
        outs, _ = ref.run_clip(script(**kw), planes, bits, st, color_family=cf, order=order)
        arrays = {}
        for q, (p, o) in enumerate(zip(planes, outs)):
            arrays[f"in{q}"] = p
            arrays[f"out{q}"] = o
        fname = f"{name}.npz"
        np.savez_compressed(HERE / fname, **arrays)
        meta.append(dict(name=name, file=fname, params=kw, bits=bits, sample_type=st,
                         color_family=cf, nplanes=len(planes), order=order))
    (HERE / "cases.json").write_text(json.dumps(meta, indent=1) + "\n")
    print(f"wrote {len(meta)} fixtures")


if __name__ == "__main__":
    main()
