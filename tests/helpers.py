"""Shared scene setup and image metrics for the tests."""
from __future__ import annotations

import math

import numpy as np

# BASELINE workloads: (config, n, seed, width, height)
CONFIG0 = (0, 65536, 0, 1280, 720)
CONFIG1 = (1, 1048576, 1, 1920, 1080)
CONFIG2 = (2, 6000000, 2, 2560, 1440)


def camera_pod(gs, config: int, w: int, h: int, azimuth: float = 0.0):
    """Camera of each synthetic workload (eye, target origin, 60 deg, 0.1..1e4)."""
    if config == 0:
        eye = (0.0, 0.0, 3.0)
    elif config == 1:
        eye = (6.0 * math.sin(azimuth), 0.0, 6.0 * math.cos(azimuth))
    else:
        r, el = 30.0, 0.3
        eye = (r * math.cos(el) * math.sin(azimuth), r * math.sin(el), r * math.cos(el) * math.cos(azimuth))
    return gs.CameraPod.look_at(eye, (0.0, 0.0, 0.0), (w, h), math.pi / 3, 0.1, 1e4)


def oracle_inputs(orc, g):
    """f32-pod inputs for the oracle: positions, colours, SH and the oracle's
    own cov3d of every record."""
    n = g.shape[0]
    pos = np.ascontiguousarray(g["pos"], np.float32)
    color = np.ascontiguousarray(g["color"], np.uint8)
    sh = np.ascontiguousarray(g["sh"].reshape(n, 45), np.float32)
    cov = orc.cov3d_f32(g["scale"], g["rot"]) if n else np.zeros((0, 6), np.float32)
    return pos, color, sh, cov


def oracle_inputs_half(orc, g):
    """f16-pod inputs for the oracle: rotation, scale and SH rounded to f16
    by numpy, cov3d rebuilt by the oracle from the rounded values."""
    n = g.shape[0]
    pos = np.ascontiguousarray(g["pos"], np.float32)
    color = np.ascontiguousarray(g["color"], np.uint8)
    rot = g["rot"].astype(np.float16).astype(np.float32)
    scale = g["scale"].astype(np.float16).astype(np.float32)
    sh = g["sh"].reshape(n, 45).astype(np.float16).astype(np.float32)
    cov = orc.cov3d_f32(scale, rot) if n else np.zeros((0, 6), np.float32)
    return pos, color, np.ascontiguousarray(sh), cov


def max_abs_diff(a, b) -> float:
    return float(np.abs(a.astype(np.float64) - b.astype(np.float64)).max()) if a.size else 0.0


def psnr(a, b, peak: float = 255.0) -> float:
    mse = float(np.mean((a.astype(np.float64) - b.astype(np.float64)) ** 2))
    return float("inf") if mse == 0 else 10.0 * math.log10(peak * peak / mse)
