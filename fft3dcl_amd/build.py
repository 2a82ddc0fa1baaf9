"""Builds the native parts of fft3dcl_amd in-tree.

    libfft3d_hip.so   the gfx950 kernels + C ABI (include/fft3d_hip.h)
    libfft3dcl_vs.so  the VapourSynth API 4 plugin over that C ABI

Only hipcc / g++ are used; nothing is fetched. Run `python -m fft3dcl_amd.build`.
"""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
INCLUDE = ROOT / "include"
HIP_LIB = PKG / "libfft3d_hip.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# Packed FP32 instructions are disabled for every HIP source of this package.
# The max-ilp machine scheduler orders the latency-bound kernels better than
# the default (sweep 166 -> 162 us at 4K 10-bit bt=4; fewer spilled VGPRs).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
             "-Wno-unused-result", "-Xclang", "-target-feature", "-Xclang",
             "-packed-fp32-ops", "-mllvm", "--amdgpu-sched-strategy=max-ilp"]
# translation units: the C ABI, and the kernels of each block shape
BLOCK_SIZES = (4, 8, 16, 32, 64)
JOBS = min(8, os.cpu_count() or 1)


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")


def build_hip(out: Path = HIP_LIB) -> Path:
    """Compiles the C ABI unit and one kernel unit per block shape, JOBS at a
    time, then links them into libfft3d_hip.so."""
    import tempfile
    from concurrent.futures import ThreadPoolExecutor
    inc = [f"-I{INCLUDE}", f"-I{CSRC}"]
    with tempfile.TemporaryDirectory(prefix="fft3d_build_") as tmpdir:
        units = [(CSRC / "fft3d_capi.hip", [], Path(tmpdir) / "capi.o")]
        units += [(CSRC / "fk_set_bw.hip", [f"-DFK_BW={bw}", f"-DFK_BH={bh}"],
                   Path(tmpdir) / f"s{bw}x{bh}.o")
                  for bw in BLOCK_SIZES for bh in BLOCK_SIZES]
        # the shapes with a side above 64 (LDS FFTs), one unit
        units.append((CSRC / "fk_set_large.hip", ["-DFK_LARGE_SET=1"], Path(tmpdir) / "large.o"))
        # largest shapes first: they take longest
        units.sort(key=lambda u: -sum(int(d.split("=")[1]) for d in u[1]))  # longest first

        def compile_unit(u):
            src, d, o = u
            r = subprocess.run([HIPCC, *HIP_FLAGS, *inc, *d, "-c", str(src), "-o", str(o)],
                               capture_output=True, text=True)
            return None if r.returncode == 0 else f"{src.name} {' '.join(d)}:\n{r.stdout}{r.stderr}"

        with ThreadPoolExecutor(JOBS) as ex:
            failed = [f for f in ex.map(compile_unit, units) if f]
        if failed:
            sys.stderr.write("\n".join(failed))
            raise RuntimeError("build failed: " + "; ".join(f.split(":")[0] for f in failed))
        tmp = out.with_suffix(".so.tmp")
        _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC",
              *[str(o) for _, _, o in units], "-o", str(tmp)])
        os.replace(tmp, out)
    return out


VS_LIB = PKG / "libfft3dcl_vs.so"


def build_plugin(out: Path = VS_LIB) -> Path:
    """The VapourSynth API 4 plugin; loads libfft3d_hip.so from its own folder."""
    src = CSRC / "fft3dcl_vs.cpp"
    tmp = out.with_suffix(".so.tmp")
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-fvisibility=hidden",
          f"-I{INCLUDE}", f"-I{CSRC}", str(src), f"-L{PKG}", "-l:libfft3d_hip.so",
          "-Wl,-rpath,$ORIGIN", "-o", str(tmp)])
    os.replace(tmp, out)
    return out


def build_all() -> list[Path]:
    return [build_hip(), build_plugin()]


if __name__ == "__main__":
    for p in build_all():
        print(p)
