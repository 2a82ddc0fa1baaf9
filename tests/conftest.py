"""Test configuration: the `gpu` marker and shared fixtures.

torch is imported first so that, in a process that also loads
libfft3d_hip.so, the one HIP runtime in use is torch's (both carry the
soname libamdhip64.so.7).
"""
import os
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
sys.path.insert(0, str(ROOT))
sys.path.insert(0, str(Path(__file__).resolve().parent))

import torch  # noqa: E402,F401

HIP_LIB = ROOT / "fft3dcl_amd" / "libfft3d_hip.so"
EMU_LIB = ROOT / "tests" / "_build" / "libfft3d_emu.so"


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU")


@pytest.fixture(scope="session")
def oracle():
    from oracle_lib import Oracle
    return Oracle()


@pytest.fixture(scope="session")
def oracle64():
    from oracle_lib import Oracle
    return Oracle(fp64=True)


@pytest.fixture(scope="session")
def ref():
    from oracle_lib import Ref
    r = Ref.load()
    if r is None:
        pytest.skip("oracle/_ref (the reference's CPU backend build) is not present")
    return r


@pytest.fixture(scope="session")
def emu_lib():
    if not EMU_LIB.exists():
        pytest.skip("CPU emulation library not built (make -C tests/emu)")
    return str(EMU_LIB)


@pytest.fixture(scope="session")
def hip_lib():
    if not HIP_LIB.exists():
        raise RuntimeError("libfft3d_hip.so is not built: run python -m fft3dcl_amd.build")
    return str(HIP_LIB)


@pytest.fixture(scope="session")
def gpu(hip_lib):
    if not torch.cuda.is_available():
        raise RuntimeError("GPU tests need a HIP device")
    from fft3dcl_amd import capi
    capi.dev_init(0, hip_lib)
    return hip_lib
