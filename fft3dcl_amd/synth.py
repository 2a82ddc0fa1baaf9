"""Seeded synthetic clips for the FFT3D denoise path.

Each frame is a background plus three translating sinusoidal gratings and four
translating flat rectangles (painted over the gratings in order), plus i.i.d.
Gaussian noise added last, then clipped and rounded for integer formats:

    v(x, y, t) = base + sum_k A * sin(2*pi*((x - vx_k t) cos(th_k)
                                        + (y - vy_k t) sin(th_k)) / P_k + phi_k)

with P_k ~ U[8, 64] px, th_k ~ U[0, pi), phi_k ~ U[0, 2 pi), A = 40 (8-bit
scale), base = 128, rectangle value ~ U[30, 220], and per-object velocities of
1-3 px/frame in a random direction. 10-bit scales every level by 4 (noise
std 12, clip to [0, 1023]); float divides the 8-bit construction by 255
(noise std 0.012, no clip). A scene cut redraws the grating set.
"""
from __future__ import annotations

import numpy as np

__all__ = ["SceneSpec", "make_plane_sequence", "make_clip"]


class SceneSpec:
    """Random objects of one scene (gratings + rectangles), drawn once."""

    def __init__(self, rng: np.random.Generator, width: int, height: int):
        self.periods = rng.uniform(8.0, 64.0, 3)
        self.theta = rng.uniform(0.0, np.pi, 3)
        self.phase = rng.uniform(0.0, 2 * np.pi, 3)
        sp = rng.uniform(1.0, 3.0, 3)
        ang = rng.uniform(0.0, 2 * np.pi, 3)
        self.gvel = np.stack([sp * np.cos(ang), sp * np.sin(ang)], 1)
        self.rects = []
        for _ in range(4):
            rw = int(rng.integers(max(2, width // 10), max(3, width // 3)))
            rh = int(rng.integers(max(2, height // 10), max(3, height // 3)))
            x0 = float(rng.uniform(0, max(1, width - rw)))
            y0 = float(rng.uniform(0, max(1, height - rh)))
            val = float(rng.uniform(30.0, 220.0))
            s = float(rng.uniform(1.0, 3.0))
            a = float(rng.uniform(0.0, 2 * np.pi))
            self.rects.append((x0, y0, rw, rh, val, s * np.cos(a), s * np.sin(a)))


def _render(scene: SceneSpec, width: int, height: int, t: int, amp: float, base: float,
            scale: float) -> np.ndarray:
    y = np.arange(height, dtype=np.float64)[:, None]
    x = np.arange(width, dtype=np.float64)[None, :]
    img = np.full((height, width), base * scale, dtype=np.float64)
    for k in range(3):
        xs = x - scene.gvel[k, 0] * t
        ys = y - scene.gvel[k, 1] * t
        arg = 2 * np.pi * (xs * np.cos(scene.theta[k]) + ys * np.sin(scene.theta[k])) \
            / scene.periods[k] + scene.phase[k]
        img += amp * scale * np.sin(arg)
    for (x0, y0, rw, rh, val, vx, vy) in scene.rects:
        xa = int(np.floor(x0 + vx * t)) % max(1, width)
        ya = int(np.floor(y0 + vy * t)) % max(1, height)
        img[ya:ya + rh, xa:xa + rw] = val * scale
    return img


def make_plane_sequence(seed: int, width: int, height: int, nframes: int, kind: str,
                        scene_cut: int | None = None) -> np.ndarray:
    """Returns nframes planes of shape (height, width).

    kind: "u8" (8-bit), "u10" (10-bit in uint16), "u16" (16-bit), or "f32"
    (float in [0,1] scale, unclipped).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    scene = SceneSpec(rng, width, height)
    cut_scene = None
    if kind == "u8":
        scale, noise, hi, dtype = 1.0, 3.0, 255, np.uint8
    elif kind == "u10":
        scale, noise, hi, dtype = 4.0, 12.0, 1023, np.uint16
    elif kind == "u16":
        scale, noise, hi, dtype = 257.0, 3.0 * 257.0, 65535, np.uint16
    elif kind == "f32":
        scale, noise, hi, dtype = 1.0 / 255.0, 0.012, None, np.float32
    else:
        raise ValueError(kind)
    out = np.empty((nframes, height, width), dtype=dtype)
    for t in range(nframes):
        sc = scene
        if scene_cut is not None and t >= scene_cut:
            if cut_scene is None:
                cut_scene = SceneSpec(rng, width, height)
            sc = cut_scene
        img = _render(sc, width, height, t, 40.0, 128.0, scale)
        img += rng.normal(0.0, noise, img.shape)
        if hi is None:
            out[t] = img.astype(np.float32)
        else:
            out[t] = np.clip(np.rint(img), 0, hi).astype(dtype)
    return out


def make_clip(seed: int, width: int, height: int, nframes: int, kind: str,
              color: str = "gray", ssw: int = 1, ssh: int = 1,
              scene_cut: int | None = None) -> list[np.ndarray]:
    """Returns a list of planes arrays (one (nframes, h, w) array per plane).

    color: "gray" (1 plane) or "yuv" (3 planes, chroma subsampled by
    2**ssw x 2**ssh, chroma drawn from its own seed).
    """
    planes = [make_plane_sequence(seed, width, height, nframes, kind, scene_cut)]
    if color == "yuv":
        cw, ch = width >> ssw, height >> ssh
        for q in (1, 2):
            planes.append(make_plane_sequence(seed * 7 + q, cw, ch, nframes, kind, scene_cut))
    return planes
