"""ctypes bindings for the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of ``bench.py``. The product never imports it.

The oracle restates the reference's per-frame splat path in C (see
``oracle/oracle.h`` for the file:line citations and the floating-point
convention). These wrappers take and return numpy arrays.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")

MODE_SPLAT, MODE_ELLIPSE, MODE_POINT = 0, 1, 2
TARGET_RGBA8_STEP, TARGET_RGBA8_F32, TARGET_RGBA32F = 0, 1, 2


class Frame(C.Structure):
    """orc_frame: camera pod, model transform matrices, gaussian transform."""

    _fields_ = [
        ("view", C.c_float * 16),
        ("proj", C.c_float * 16),
        ("size", C.c_float * 2),
        ("model", C.c_float * 16),
        ("sr", C.c_float * 9),
        ("inv_sr", C.c_float * 9),
        ("gsize", C.c_float),
        ("display_mode", C.c_int32),
        ("sh_deg", C.c_int32),
        ("no_sh0", C.c_int32),
        ("max_std_dev", C.c_float),
    ]


SPLAT_F32 = np.dtype(
    [("cx", "<f4"), ("cy", "<f4"), ("ux", "<f4"), ("uy", "<f4"), ("vx", "<f4"), ("vy", "<f4"),
     ("r", "<f4"), ("g", "<f4"), ("b", "<f4"), ("a", "<f4"), ("hx", "<f4"), ("hy", "<f4"),
     ("drawable", "<i4"), ("pad", "<i4"), ("z", "<f4")])
SPLAT_F64 = np.dtype(
    [("cx", "<f8"), ("cy", "<f8"), ("ux", "<f8"), ("uy", "<f8"), ("vx", "<f8"), ("vy", "<f8"),
     ("r", "<f8"), ("g", "<f8"), ("b", "<f8"), ("a", "<f8"), ("hx", "<f8"), ("hy", "<f8"),
     ("drawable", "<i4"), ("pad", "<i4"), ("z", "<f8")])

_lib = None


def build() -> None:
    """Compile liboracle.so from oracle.c with the committed Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE, "liboracle.so"], check=True)


def lib() -> C.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = C.CDLL(_LIB_PATH)
        _declare(_lib)
    return _lib


def _p(a, t):
    return a.ctypes.data_as(C.POINTER(t))


def _declare(L):
    fp, u8p, u32p, dp = C.POINTER(C.c_float), C.POINTER(C.c_uint8), C.POINTER(C.c_uint32), C.POINTER(C.c_double)
    F = C.POINTER(Frame)
    L.orc_camera_forward.argtypes = [C.c_float, C.c_float, fp]
    L.orc_look_to_rh.argtypes = [fp, fp, fp, fp]
    L.orc_perspective_rh.argtypes = [C.c_float, C.c_float, C.c_float, C.c_float, fp]
    L.orc_cov3d_f32.argtypes = [fp, fp, fp]
    L.orc_cov3d_f64.argtypes = [dp, dp, dp]
    L.orc_cov3d_f32_n.argtypes = [C.c_uint64, fp, fp, fp]
    L.orc_cov3d_f64_n.argtypes = [C.c_uint64, dp, dp, dp]
    L.orc_model_mats.argtypes = [fp, fp, fp, fp, fp, fp]
    L.orc_preprocess_f32.argtypes = [F, C.c_uint32, fp, fp, u32p, fp]
    L.orc_preprocess_f32.restype = C.c_uint32
    L.orc_cov2d_axes_f32.argtypes = [F, fp, fp, C.c_float, fp]
    L.orc_cov2d_axes_f64.argtypes = [F, dp, dp, C.c_double, dp]
    L.orc_cov2d_f32.argtypes = [F, fp, fp, fp]
    L.orc_view_color_f32.argtypes = [u8p, fp, fp, C.c_int32, C.c_int32, fp]
    L.orc_view_color_f64.argtypes = [u8p, dp, dp, C.c_int32, C.c_int32, dp]
    L.orc_splats_f32.argtypes = [F, C.c_uint32, u32p, fp, u8p, fp, fp, C.c_void_p]
    L.orc_splats_f64.argtypes = [F, C.c_uint32, u32p, dp, u8p, dp, dp, C.c_void_p]
    L.orc_render_f32.argtypes = [F, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int32, C.c_void_p, C.c_int32]
    L.orc_render_depth_f32.argtypes = [F, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, C.c_int32, C.c_void_p,
                                       C.c_int32, fp, C.c_int32, C.c_int32]
    L.orc_render_f64.argtypes = [F, C.c_uint32, C.c_uint32, C.c_uint32, C.c_void_p, dp, C.c_int32]
    L.orc_stable_sort.argtypes = [C.c_uint32, u32p, u32p]
    L.orc_frame_f32.argtypes = [F, C.c_uint32, fp, u8p, fp, fp, C.c_uint32, C.c_uint32, C.c_int32,
                                u32p, fp, C.c_void_p, C.c_int32]
    L.orc_frame_f32.restype = C.c_uint32


# --------------------------------------------------------------------------
# helpers
# --------------------------------------------------------------------------

def _f32(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float32)
    if shape is not None:
        a = a.reshape(shape)
    return a


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    if shape is not None:
        a = a.reshape(shape)
    return a


IDENTITY4 = np.eye(4, dtype=np.float32)
IDENTITY3 = np.eye(3, dtype=np.float32)


def make_frame(view, proj, size, model=None, sr=None, inv_sr=None, gsize=1.0, display_mode=MODE_SPLAT,
               sh_deg=3, no_sh0=False, max_std_dev=3.0) -> Frame:
    """Build an orc_frame. Matrices are given as numpy (4,4)/(3,3) in
    mathematical [row, col] layout and stored column-major."""
    f = Frame()
    f.view[:] = _f32(view).T.reshape(-1).tolist()
    f.proj[:] = _f32(proj).T.reshape(-1).tolist()
    f.size[:] = [float(size[0]), float(size[1])]
    f.model[:] = _f32(IDENTITY4 if model is None else model).T.reshape(-1).tolist()
    f.sr[:] = _f32(IDENTITY3 if sr is None else sr).T.reshape(-1).tolist()
    f.inv_sr[:] = _f32(IDENTITY3 if inv_sr is None else inv_sr).T.reshape(-1).tolist()
    f.gsize = gsize
    f.display_mode = display_mode
    f.sh_deg = sh_deg
    f.no_sh0 = int(bool(no_sh0))
    f.max_std_dev = max_std_dev
    return f


def frame_from_pod(pod_bytes: bytes, **kw) -> Frame:
    """Build an orc_frame from 144 CameraPod bytes (src/buffer/camera.rs:61-79)."""
    a = np.frombuffer(pod_bytes, dtype=np.float32, count=34)
    view = a[0:16].reshape(4, 4).T
    proj = a[16:32].reshape(4, 4).T
    return make_frame(view, proj, (a[32], a[33]), **kw)


def camera_forward(pitch, yaw):
    out = np.zeros(3, np.float32)
    lib().orc_camera_forward(pitch, yaw, _p(out, C.c_float))
    return out


def look_to_rh(eye, direction, up=(0.0, 1.0, 0.0)):
    """Returns the view matrix as numpy [row, col]."""
    out = np.zeros(16, np.float32)
    e, d, u = _f32(eye), _f32(direction), _f32(up)
    lib().orc_look_to_rh(_p(e, C.c_float), _p(d, C.c_float), _p(u, C.c_float), _p(out, C.c_float))
    return out.reshape(4, 4).T.copy()


def perspective_rh(fov_y, aspect, z_near, z_far):
    out = np.zeros(16, np.float32)
    lib().orc_perspective_rh(fov_y, aspect, z_near, z_far, _p(out, C.c_float))
    return out.reshape(4, 4).T.copy()


def cov3d_f32(scale, rot):
    """(n,3) scale, (n,4) xyzw rotation -> (n,6) cov3d in f32."""
    scale, rot = _f32(scale, (-1, 3)), _f32(rot, (-1, 4))
    out = np.zeros((scale.shape[0], 6), np.float32)
    if scale.shape[0]:
        lib().orc_cov3d_f32_n(scale.shape[0], _p(scale, C.c_float), _p(rot, C.c_float), _p(out, C.c_float))
    return out


def cov3d_f64(scale, rot):
    scale, rot = _f64(scale, (-1, 3)), _f64(rot, (-1, 4))
    out = np.zeros((scale.shape[0], 6), np.float64)
    if scale.shape[0]:
        lib().orc_cov3d_f64_n(scale.shape[0], _p(scale, C.c_double), _p(rot, C.c_double), _p(out, C.c_double))
    return out


def model_mats(pos, rot, scale):
    model = np.zeros(16, np.float32)
    sr = np.zeros(9, np.float32)
    inv = np.zeros(9, np.float32)
    p, r, s = _f32(pos), _f32(rot), _f32(scale)
    lib().orc_model_mats(_p(p, C.c_float), _p(r, C.c_float), _p(s, C.c_float), _p(model, C.c_float),
                         _p(sr, C.c_float), _p(inv, C.c_float))
    return model.reshape(4, 4).T.copy(), sr.reshape(3, 3).T.copy(), inv.reshape(3, 3).T.copy()


def preprocess(frame: Frame, pos, cov3d):
    """Visible indices (ascending) and their f32 depth keys."""
    pos, cov3d = _f32(pos, (-1, 3)), _f32(cov3d, (-1, 6))
    n = pos.shape[0]
    idx = np.zeros(max(n, 1), np.uint32)
    key = np.zeros(max(n, 1), np.float32)
    v = lib().orc_preprocess_f32(C.byref(frame), n, _p(pos, C.c_float), _p(cov3d, C.c_float),
                                 _p(idx, C.c_uint32), _p(key, C.c_float))
    return idx[:v].copy(), key[:v].copy()


def cov2d(frame: Frame, pos, cov3d):
    out = np.zeros(3, np.float32)
    p, c = _f32(pos), _f32(cov3d)
    lib().orc_cov2d_f32(C.byref(frame), _p(p, C.c_float), _p(c, C.c_float), _p(out, C.c_float))
    return out


def cov2d_axes(frame: Frame, pos, cov3d, std_dev, f64=False):
    if f64:
        out = np.zeros(4, np.float64)
        p, c = _f64(pos), _f64(cov3d)
        lib().orc_cov2d_axes_f64(C.byref(frame), _p(p, C.c_double), _p(c, C.c_double), std_dev,
                                 _p(out, C.c_double))
        return out
    out = np.zeros(4, np.float32)
    p, c = _f32(pos), _f32(cov3d)
    lib().orc_cov2d_axes_f32(C.byref(frame), _p(p, C.c_float), _p(c, C.c_float), std_dev, _p(out, C.c_float))
    return out


def view_color(color, sh, direction, sh_deg=3, no_sh0=False, f64=False):
    col = np.ascontiguousarray(color, dtype=np.uint8)
    if f64:
        out = np.zeros(4, np.float64)
        s, d = _f64(sh).reshape(-1), _f64(direction)
        lib().orc_view_color_f64(_p(col, C.c_uint8), _p(s, C.c_double), _p(d, C.c_double), sh_deg,
                                 int(no_sh0), _p(out, C.c_double))
        return out
    out = np.zeros(4, np.float32)
    s, d = _f32(sh).reshape(-1), _f32(direction)
    lib().orc_view_color_f32(_p(col, C.c_uint8), _p(s, C.c_float), _p(d, C.c_float), sh_deg, int(no_sh0),
                             _p(out, C.c_float))
    return out


def splats(frame: Frame, idx, pos, color, sh, cov3d, f64=False):
    """Vertex-stage records for instances idx (draw order)."""
    idx = np.ascontiguousarray(idx, dtype=np.uint32)
    color = np.ascontiguousarray(color, dtype=np.uint8).reshape(-1, 4)
    v = idx.shape[0]
    if f64:
        pos, sh, cov3d = _f64(pos, (-1, 3)), _f64(sh).reshape(-1, 45), _f64(cov3d, (-1, 6))
        out = np.zeros(max(v, 1), SPLAT_F64)
        lib().orc_splats_f64(C.byref(frame), v, _p(idx, C.c_uint32), _p(pos, C.c_double), _p(color, C.c_uint8),
                             _p(sh, C.c_double), _p(cov3d, C.c_double), out.ctypes.data)
        return out[:v]
    pos, sh, cov3d = _f32(pos, (-1, 3)), _f32(sh).reshape(-1, 45), _f32(cov3d, (-1, 6))
    out = np.zeros(max(v, 1), SPLAT_F32)
    lib().orc_splats_f32(C.byref(frame), v, _p(idx, C.c_uint32), _p(pos, C.c_float), _p(color, C.c_uint8),
                         _p(sh, C.c_float), _p(cov3d, C.c_float), out.ctypes.data)
    return out[:v]


def render(frame: Frame, w, h, sp, target=TARGET_RGBA8_STEP, threads=1, f64=False, depth=None, compare=8,
           depth_write=False):
    """Rasterise and blend splat records in draw order; returns (h, w, 4).
    depth: an (h, w) f32 depth attachment tested with the wgpu compare
    function `compare` (1 never .. 8 always) and updated in place when
    depth_write (f32 targets only)."""
    v = sp.shape[0]
    if depth is not None:
        assert not f64 and depth.dtype == np.float32 and depth.shape == (h, w) and depth.flags.c_contiguous
        sp = np.ascontiguousarray(sp, dtype=SPLAT_F32)
        out = np.zeros((h, w, 4), np.float32 if target == TARGET_RGBA32F else np.uint8)
        lib().orc_render_depth_f32(C.byref(frame), w, h, v, sp.ctypes.data, target, out.ctypes.data, threads,
                                   _p(depth, C.c_float), int(compare), int(bool(depth_write)))
        return out
    if f64:
        sp = np.ascontiguousarray(sp, dtype=SPLAT_F64)
        out = np.zeros((h, w, 4), np.float64)
        lib().orc_render_f64(C.byref(frame), w, h, v, sp.ctypes.data, _p(out, C.c_double), threads)
        return out
    sp = np.ascontiguousarray(sp, dtype=SPLAT_F32)
    if target == TARGET_RGBA32F:
        out = np.zeros((h, w, 4), np.float32)
    else:
        out = np.zeros((h, w, 4), np.uint8)
    lib().orc_render_f32(C.byref(frame), w, h, v, sp.ctypes.data, target, out.ctypes.data, threads)
    return out


def stable_sort(keys, vals):
    k = np.array(keys, dtype=np.uint32, copy=True)
    v = np.array(vals, dtype=np.uint32, copy=True)
    lib().orc_stable_sort(k.shape[0], _p(k, C.c_uint32), _p(v, C.c_uint32))
    return k, v


def full_frame(frame: Frame, pos, color, sh, cov3d, w, h, target=TARGET_RGBA8_STEP, threads=1):
    """Whole frame; returns (sorted idx, sorted keys, image)."""
    pos, sh, cov3d = _f32(pos, (-1, 3)), _f32(sh).reshape(-1, 45), _f32(cov3d, (-1, 6))
    color = np.ascontiguousarray(color, dtype=np.uint8).reshape(-1, 4)
    n = pos.shape[0]
    idx = np.zeros(max(n, 1), np.uint32)
    key = np.zeros(max(n, 1), np.float32)
    img = np.zeros((h, w, 4), np.float32 if target == TARGET_RGBA32F else np.uint8)
    v = lib().orc_frame_f32(C.byref(frame), n, _p(pos, C.c_float), _p(color, C.c_uint8), _p(sh, C.c_float),
                            _p(cov3d, C.c_float), w, h, target, _p(idx, C.c_uint32), _p(key, C.c_float),
                            img.ctypes.data, threads)
    return idx[:v].copy(), key[:v].copy(), img
