"""Shared helpers of the test suite (inputs and tolerance checks)."""
from __future__ import annotations

import numpy as np


def mt19937_planes(seed: int, w: int, h: int, n: int) -> list[np.ndarray]:
    """The float planes of the reference's tools/gpu_test.cpp:52-58:
    std::mt19937(seed) fed to uniform_real_distribution<float>(0, 1), one plane
    after the other. numpy's legacy RandomState(seed) is the same generator and
    seeding; generate_canonical<float> divides one 32-bit draw by 2^32."""
    rs = np.random.RandomState(seed)
    u = rs.randint(0, 2 ** 32, size=n * w * h, dtype=np.uint64).astype(np.float32)
    x = u / np.float32(4294967296.0)
    x = np.minimum(x, np.nextafter(np.float32(1), np.float32(0)))
    return [x[i * w * h:(i + 1) * w * h].reshape(h, w).copy() for i in range(n)]


def lsb_diff(a: np.ndarray, b: np.ndarray) -> int:
    return int(np.abs(a.astype(np.int64) - b.astype(np.int64)).max()) if a.size else 0


def tolerance(kind: str) -> float:
    """Per-format parity bar against the CPU reference: +/-2 LSB at 8 bit,
    +/-4 LSB at 10 bit and above, 1e-3 absolute (of the [0,1] range) for float."""
    return {"u8": 2, "u10": 4, "u16": 4, "f32": 1e-3}[kind]


def max_abs(a, b) -> float:
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0
