"""MI355X-native per-frame splat pipeline (Python mirror of the C ABI).

The compute path is libgs3d.so (HIP kernels for gfx950 + C++ host library);
this module binds its C ABI (include/gs3d.h) with ctypes and mirrors the
reference's Rust API shape (wgpu-3dgs-viewer 0.6.1, src/lib.rs):

    dev = Device()
    viewer = Viewer(dev, TextureFormat.RGBA8_UNORM, gaussians)
    viewer.update_camera_with_pod(pod)          # or update_camera(camera, (w, h))
    target = Texture(dev, w, h, TextureFormat.RGBA8_UNORM)
    viewer.render(target)
    dev.poll()
    image = target.read()

There is no CPU fallback: every call goes to libgs3d.so, and a missing or
unloadable library raises.
"""
from __future__ import annotations

import ctypes as C
import enum
import os
from typing import Optional, Sequence

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
# GS3D_LIBRARY names another build of the library (for example one built
# with different tuning defines by build.py --out).
LIB_PATH = os.environ.get("GS3D_LIBRARY") or os.path.join(_HERE, "libgs3d.so")

# gs3d_gaussian (wgpu_3dgs_core::Gaussian): 224 bytes
GAUSSIAN_DTYPE = np.dtype([("rot", "<f4", (4,)), ("pos", "<f4", (3,)), ("color", "u1", (4,)),
                           ("sh", "<f4", (15, 3)), ("scale", "<f4", (3,))])
assert GAUSSIAN_DTYPE.itemsize == 224

STATUS_NAMES = {0: "OK", 1: "INVALID_ARGUMENT", 2: "MODEL_SIZE_EXCEEDS_DEVICE_LIMIT", 3: "OUT_OF_MEMORY",
                4: "DEVICE", 5: "MODEL_NOT_FOUND", 6: "KERNEL"}


class Gs3dError(RuntimeError):
    """A gs3d_status other than GS3D_OK (the reference's error enums, src/error.rs)."""

    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class ModelSizeExceedsDeviceLimit(Gs3dError):
    """PreprocessorCreateError / RendererCreateError::ModelSizeExceedsDeviceLimit."""


class ModelNotFound(Gs3dError):
    """MultiModelViewerAccessError::ModelNotFound."""


class TextureFormat(enum.IntEnum):
    RGBA8_UNORM = 0
    RGBA32_FLOAT = 1
    DEPTH32_FLOAT = 2  # depth attachment


class CompareFunction(enum.IntEnum):
    """wgpu::CompareFunction."""
    NEVER = 1
    LESS = 2
    EQUAL = 3
    LESS_EQUAL = 4
    GREATER = 5
    NOT_EQUAL = 6
    GREATER_EQUAL = 7
    ALWAYS = 8


class GaussianPod(enum.IntEnum):
    SINGLE = 0  # f32 SH + f32 cov3d (the reference's DefaultGaussianPod)
    HALF = 1    # f16 SH + f16 rotation/scale, u8 opacity


class DisplayMode(enum.IntEnum):
    SPLAT = 0
    ELLIPSE = 1
    POINT = 2


class BlendMode(enum.IntEnum):
    """EXACT (the default): back to front, the destination rounded to the
    target format after every splat, as the reference's blender does. FAST:
    front to back in f32, rounded once."""
    FAST = 0
    EXACT = 1


class Stage(enum.IntEnum):
    PREPROCESS = 0
    SORT = 1
    BIN = 2
    TILE_SORT = 3
    COMPOSITE = 4


class CameraPod(C.Structure):
    """CameraPod (src/buffer/camera.rs:61-79), 144 bytes."""

    _fields_ = [("view", C.c_float * 16), ("proj", C.c_float * 16), ("size", C.c_float * 2),
                ("padding", C.c_uint32 * 2)]

    def to_bytes(self) -> bytes:
        return bytes(C.string_at(C.addressof(self), C.sizeof(self)))

    @classmethod
    def from_bytes(cls, b: bytes) -> "CameraPod":
        if len(b) != 144:
            raise ValueError("a CameraPod is 144 bytes")
        return cls.from_buffer_copy(b)

    @classmethod
    def new(cls, camera: "Camera", size) -> "CameraPod":
        """CameraPod::new (src/buffer/camera.rs:72-79)."""
        out = cls()
        _check(lib().gs3d_camera_pod_new(C.byref(camera._c()), int(size[0]), int(size[1]), C.byref(out)))
        return out

    @classmethod
    def look_at(cls, eye, target, size, vertical_fov=np.pi / 3, z_near=0.1, z_far=1e4) -> "CameraPod":
        out = cls()
        e = (C.c_float * 3)(*[float(x) for x in eye])
        t = (C.c_float * 3)(*[float(x) for x in target])
        _check(lib().gs3d_camera_pod_look_at(e, t, float(vertical_fov), float(z_near), float(z_far), int(size[0]),
                                             int(size[1]), C.byref(out)))
        return out


class _CCamera(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("z_start", C.c_float), ("z_end", C.c_float),
                ("vertical_fov", C.c_float), ("pitch", C.c_float), ("yaw", C.c_float)]


class Camera:
    """Camera (src/camera.rs:18-93): position, z range, vertical fov, pitch, yaw."""

    def __init__(self, z=(0.1, 1e4), vertical_fov=np.pi / 3, pos=(0.0, 0.0, 0.0), pitch=0.0, yaw=0.0):
        self.pos = [float(x) for x in pos]
        self.z = (float(z[0]), float(z[1]))
        self.vertical_fov = float(vertical_fov)
        self.pitch = float(pitch)
        self.yaw = float(yaw)

    def _c(self) -> _CCamera:
        c = _CCamera()
        c.pos[:] = self.pos
        c.z_start, c.z_end = self.z
        c.vertical_fov, c.pitch, c.yaw = self.vertical_fov, self.pitch, self.yaw
        return c


_lib = None


def lib() -> C.CDLL:
    """Load libgs3d.so (built by build.py). Raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run __graft_entry__.build() first")
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


def _declare(L):
    vp, u32, i32, f32, u64 = C.c_void_p, C.c_uint32, C.c_int, C.c_float, C.c_uint64
    pp = C.POINTER(C.c_void_p)
    sigs = {
        "gs3d_last_error": ([], C.c_char_p),
        "gs3d_device_create": ([i32, pp], i32),
        "gs3d_device_destroy": ([vp], None),
        "gs3d_device_poll": ([vp], i32),
        "gs3d_device_set_buffer_limit": ([vp, u64], i32),
        "gs3d_device_compute_units": ([vp, C.POINTER(C.c_int)], i32),
        "gs3d_texture_create": ([vp, u32, u32, i32, pp], i32),
        "gs3d_texture_destroy": ([vp], None),
        "gs3d_texture_read": ([vp, vp, C.c_size_t], i32),
        "gs3d_viewer_new": ([vp, i32, vp, u32, i32, pp], i32),
        "gs3d_viewer_destroy": ([vp], None),
        "gs3d_viewer_update_camera": ([vp, vp, u32, u32], i32),
        "gs3d_viewer_update_camera_with_pod": ([vp, vp], i32),
        "gs3d_viewer_update_model_transform": ([vp, vp, vp, vp], i32),
        "gs3d_viewer_update_gaussian_transform": ([vp, f32, i32, u32, i32, f32], i32),
        "gs3d_viewer_render": ([vp, vp], i32),
        "gs3d_viewer_preprocess": ([vp], i32),
        "gs3d_viewer_sort": ([vp], i32),
        "gs3d_viewer_draw": ([vp, vp], i32),
        "gs3d_viewer_draw_standalone": ([vp, vp], i32),
        "gs3d_viewer_set_blend_mode": ([vp, i32], i32),
        "gs3d_viewer_set_tile_size": ([vp, u32, u32], i32),
        "gs3d_viewer_set_sort_items": ([vp, u32], i32),
        "gs3d_viewer_set_pair_capacity": ([vp, u32], i32),
        "gs3d_viewer_set_profiling": ([vp, i32], i32),
        "gs3d_viewer_stage_times": ([vp, vp], i32),
        "gs3d_viewer_update_selection": ([vp, vp, u32], i32),
        "gs3d_viewer_set_depth_stencil": ([vp, i32, i32, i32], i32),
        "gs3d_viewer_render_with_depth": ([vp, vp, vp, i32], i32),
        "gs3d_texture_write": ([vp, vp, C.c_size_t], i32),
        "gs3d_viewer_update_invert_selection": ([vp, i32], i32),
        "gs3d_multi_model_viewer_update_selection": ([vp, C.c_char_p, vp, u32], i32),
        "gs3d_multi_model_viewer_update_invert_selection": ([vp, C.c_char_p, i32], i32),
        "gs3d_viewer_count_fragments": ([vp, vp, vp], i32),
        "gs3d_viewer_gaussian_count": ([vp, vp], i32),
        "gs3d_viewer_read_indirect_args": ([vp, vp], i32),
        "gs3d_viewer_read_radix_sort_indirect_args": ([vp, vp], i32),
        "gs3d_viewer_read_indirect_indices": ([vp, vp, u32], i32),
        "gs3d_viewer_read_gaussians_depth": ([vp, vp, u32], i32),
        "gs3d_depth_buffer_len": ([u32], u32),
        "gs3d_radix_sorter_new": ([vp, u32, pp], i32),
        "gs3d_radix_sorter_destroy": ([vp], None),
        "gs3d_radix_sorter_write": ([vp, vp, vp, u32], i32),
        "gs3d_radix_sorter_sort": ([vp], i32),
        "gs3d_radix_sorter_set_items": ([vp, u32], i32),
        "gs3d_radix_sorter_read": ([vp, vp, vp, u32], i32),
        "gs3d_radix_sort_noalloc": ([vp, vp, vp, vp, vp, u32], i32),
        "gs3d_viewer_pair_count": ([vp, C.POINTER(u32)], i32),
        "gs3d_camera_pod_new": ([vp, u32, u32, vp], i32),
        "gs3d_camera_pod_look_at": ([vp, vp, f32, f32, f32, u32, u32, vp], i32),
        "gs3d_gaussian_cov3d": ([vp, vp], i32),
        "gs3d_pack_half": ([vp, u32, vp, vp], i32),
        "gs3d_synth_gaussians": ([i32, u64, u32, u32, vp], i32),
        "gs3d_ply_read_file": ([C.c_char_p, vp, u64, vp], i32),
        "gs3d_ply_read_memory": ([vp, u64, vp, u64, vp], i32),
        "gs3d_spz_read_file": ([C.c_char_p, vp, u64, vp], i32),
        "gs3d_spz_read_memory": ([vp, u64, vp, u64, vp], i32),
        "gs3d_model_new": ([vp, vp, u32, i32, pp], i32),
        "gs3d_model_destroy": ([vp], None),
        "gs3d_multi_model_viewer_new": ([vp, i32, pp], i32),
        "gs3d_multi_model_viewer_destroy": ([vp], None),
        "gs3d_multi_model_viewer_insert_model": ([vp, C.c_char_p, vp, u32, i32], i32),
        "gs3d_multi_model_viewer_remove_model": ([vp, C.c_char_p], i32),
        "gs3d_multi_model_viewer_update_camera": ([vp, vp, u32, u32], i32),
        "gs3d_multi_model_viewer_update_camera_with_pod": ([vp, vp], i32),
        "gs3d_multi_model_viewer_update_model_transform": ([vp, C.c_char_p, vp, vp, vp], i32),
        "gs3d_multi_model_viewer_update_gaussian_transform": ([vp, f32, i32, u32, i32, f32], i32),
        "gs3d_multi_model_viewer_set_blend_mode": ([vp, i32], i32),
        "gs3d_multi_model_viewer_render": ([vp, vp, C.POINTER(C.c_char_p), u32], i32),
        "gs3d_batch_new": ([vp, C.POINTER(C.c_void_p), u32, i32, pp], i32),
        "gs3d_batch_destroy": ([vp], None),
        "gs3d_batch_set_layout": ([vp, vp, u32], i32),
        "gs3d_batch_autotune": ([vp, vp, i32], i32),
        "gs3d_batch_render": ([vp, vp, i32], i32),
        "gs3d_batch_finish": ([vp], i32),
        "gs3d_batch_read_target": ([vp, u32, vp, C.c_size_t], i32),
        "gs3d_batch_tile_size": ([vp, u32, vp, vp], i32),
        "gs3d_batch_sort_items": ([vp, u32, vp], i32),
        "gs3d_server_new": ([vp, C.POINTER(C.c_void_p), u32, i32, u64, pp], i32),
        "gs3d_server_destroy": ([vp], None),
        "gs3d_server_run": ([vp, vp, u32, C.POINTER(C.c_void_p), vp], i32),
        "gs3d_server_render_alone": ([vp, vp, vp], i32),
    }
    for name, (args, res) in sigs.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res


EXPORTED_SYMBOLS = (
    "gs3d_last_error", "gs3d_device_create", "gs3d_device_destroy", "gs3d_device_poll",
    "gs3d_device_set_buffer_limit", "gs3d_device_compute_units", "gs3d_texture_create", "gs3d_texture_destroy",
    "gs3d_texture_read", "gs3d_viewer_new", "gs3d_viewer_destroy", "gs3d_viewer_update_camera",
    "gs3d_viewer_update_camera_with_pod", "gs3d_viewer_update_model_transform",
    "gs3d_viewer_update_gaussian_transform", "gs3d_viewer_render", "gs3d_viewer_preprocess", "gs3d_viewer_sort",
    "gs3d_viewer_draw", "gs3d_viewer_draw_standalone", "gs3d_viewer_set_blend_mode", "gs3d_viewer_set_tile_size", "gs3d_viewer_set_pair_capacity",
    "gs3d_viewer_set_sort_items", "gs3d_radix_sorter_set_items", "gs3d_batch_sort_items",
    "gs3d_viewer_set_profiling", "gs3d_viewer_stage_times", "gs3d_viewer_count_fragments",
    "gs3d_viewer_update_selection", "gs3d_viewer_update_invert_selection",
    "gs3d_viewer_set_depth_stencil", "gs3d_viewer_render_with_depth", "gs3d_texture_write",
    "gs3d_multi_model_viewer_update_selection", "gs3d_multi_model_viewer_update_invert_selection",
    "gs3d_viewer_gaussian_count", "gs3d_viewer_pair_count",
    "gs3d_viewer_read_indirect_args", "gs3d_viewer_read_radix_sort_indirect_args",
    "gs3d_viewer_read_indirect_indices", "gs3d_viewer_read_gaussians_depth", "gs3d_depth_buffer_len",
    "gs3d_radix_sorter_new", "gs3d_radix_sorter_destroy", "gs3d_radix_sorter_write", "gs3d_radix_sorter_sort",
    "gs3d_radix_sorter_read", "gs3d_radix_sort_noalloc", "gs3d_camera_pod_new", "gs3d_camera_pod_look_at", "gs3d_gaussian_cov3d",
    "gs3d_pack_half", "gs3d_synth_gaussians", "gs3d_ply_read_file", "gs3d_ply_read_memory",
    "gs3d_spz_read_file", "gs3d_spz_read_memory", "gs3d_model_new", "gs3d_model_destroy",
    "gs3d_multi_model_viewer_new", "gs3d_multi_model_viewer_destroy", "gs3d_multi_model_viewer_insert_model",
    "gs3d_multi_model_viewer_remove_model", "gs3d_multi_model_viewer_update_camera",
    "gs3d_multi_model_viewer_update_camera_with_pod", "gs3d_multi_model_viewer_update_model_transform",
    "gs3d_multi_model_viewer_update_gaussian_transform", "gs3d_multi_model_viewer_set_blend_mode",
    "gs3d_multi_model_viewer_render", "gs3d_batch_new", "gs3d_batch_destroy", "gs3d_batch_set_layout",
    "gs3d_batch_autotune", "gs3d_batch_render", "gs3d_batch_finish", "gs3d_batch_read_target",
    "gs3d_batch_tile_size", "gs3d_server_new", "gs3d_server_destroy", "gs3d_server_run",
    "gs3d_server_render_alone",
)

# gs3d_request: model, width, height, padding, arrival_s, CameraPod (144 B)
REQUEST_DTYPE = np.dtype([("model", "<u4"), ("width", "<u4"), ("height", "<u4"), ("padding", "<u4"),
                          ("arrival_s", "<f8"), ("pod", "u1", (144,))])
assert REQUEST_DTYPE.itemsize == 168
VIEW_DTYPE = np.dtype([("model", "<u4"), ("width", "<u4"), ("height", "<u4")])


class ServeStats(C.Structure):
    _fields_ = [("requests", C.c_uint64), ("completed", C.c_uint64), ("seconds", C.c_double),
                ("frames_per_s", C.c_double), ("p50_ms", C.c_double), ("p99_ms", C.c_double),
                ("mean_ms", C.c_double), ("max_ms", C.c_double), ("mean_in_flight", C.c_double),
                ("peak_hbm_bytes", C.c_uint64), ("deferred_for_memory", C.c_uint64)]

    def as_dict(self):
        return {name: getattr(self, name) for name, _ in self._fields_}


def _check(status: int) -> None:
    if status != 0:
        msg = lib().gs3d_last_error().decode(errors="replace")
        if status == 2:
            raise ModelSizeExceedsDeviceLimit(status, msg)
        if status == 5:
            raise ModelNotFound(status, msg)
        raise Gs3dError(status, msg)


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ---------------------------------------------------------------------------
# host helpers
# ---------------------------------------------------------------------------

def synth_gaussians(config: int, n: int, seed: int, first: int = 0) -> np.ndarray:
    """Synthetic scene records (see gs3d_synth_gaussians)."""
    out = np.zeros(n, GAUSSIAN_DTYPE)
    if n:
        _check(lib().gs3d_synth_gaussians(int(config), int(seed), int(first), int(n), _ptr(out)))
    return out


def _read_records(source, read_file, read_memory) -> np.ndarray:
    n = C.c_uint64()
    if isinstance(source, (bytes, bytearray, memoryview, np.ndarray)):
        buf = np.frombuffer(bytes(source), np.uint8)
        _check(read_memory(_ptr(buf), buf.nbytes, None, 0, C.byref(n)))
        out = np.zeros(n.value, GAUSSIAN_DTYPE)
        if n.value:
            _check(read_memory(_ptr(buf), buf.nbytes, _ptr(out), n.value, C.byref(n)))
        return out
    path = os.fsencode(source)
    _check(read_file(path, None, 0, C.byref(n)))
    out = np.zeros(n.value, GAUSSIAN_DTYPE)
    if n.value:
        _check(read_file(path, _ptr(out), n.value, C.byref(n)))
    return out


def read_ply(source) -> np.ndarray:
    """Gaussians::read_from_file(path, GaussiansSource::Ply) (examples/simple.rs:157-160):
    a standard 3DGS PLY file (path) or its bytes, as gs3d_gaussian records."""
    return _read_records(source, lib().gs3d_ply_read_file, lib().gs3d_ply_read_memory)


def read_spz(source) -> np.ndarray:
    """Gaussians::read_from_file(path, GaussiansSource::Spz): an SPZ file (path) or its bytes."""
    return _read_records(source, lib().gs3d_spz_read_file, lib().gs3d_spz_read_memory)


def read_gaussians(path) -> np.ndarray:
    """The examples' loader (examples/simple.rs:157-160): PLY, else SPZ."""
    try:
        return read_ply(path)
    except Gs3dError:
        return read_spz(path)


def gaussians(pos, rot, scale, color, sh=None) -> np.ndarray:
    """Build gs3d_gaussian records from arrays."""
    pos = np.asarray(pos, np.float32).reshape(-1, 3)
    n = pos.shape[0]
    g = np.zeros(n, GAUSSIAN_DTYPE)
    g["pos"] = pos
    g["rot"] = np.asarray(rot, np.float32).reshape(-1, 4)
    g["scale"] = np.asarray(scale, np.float32).reshape(-1, 3)
    g["color"] = np.asarray(color, np.uint8).reshape(-1, 4)
    if sh is not None:
        g["sh"] = np.asarray(sh, np.float32).reshape(-1, 15, 3)
    return g


def cov3d(g: np.ndarray) -> np.ndarray:
    """cov3d of each record as the f32 pod stores it."""
    g = np.ascontiguousarray(g, GAUSSIAN_DTYPE)
    out = np.zeros((g.shape[0], 6), np.float32)
    for i in range(g.shape[0]):
        _check(lib().gs3d_gaussian_cov3d(_ptr(g[i:i + 1]), _ptr(out[i])))
    return out


def pack_half(g: np.ndarray):
    """(rot_scale f16 (n,8), sh f16 (n,48)) as the f16 pod uploads them."""
    g = np.ascontiguousarray(g, GAUSSIAN_DTYPE)
    n = g.shape[0]
    rs = np.zeros((n, 8), np.uint16)
    sh = np.zeros((n, 48), np.uint16)
    if n:
        _check(lib().gs3d_pack_half(_ptr(g), n, _ptr(rs), _ptr(sh)))
    return rs.view(np.float16), sh.view(np.float16)


def depth_buffer_len(n: int) -> int:
    return int(lib().gs3d_depth_buffer_len(int(n)))


# ---------------------------------------------------------------------------
# device objects
# ---------------------------------------------------------------------------

class Device:
    """wgpu::Device + wgpu::Queue: one HIP device and its stream."""

    def __init__(self, hip_device: int = 0):
        h = C.c_void_p()
        _check(lib().gs3d_device_create(int(hip_device), C.byref(h)))
        self._h = h

    @property
    def handle(self):
        return self._h

    def poll(self) -> None:
        """queue.submit + device.poll(wait)."""
        _check(lib().gs3d_device_poll(self._h))

    def set_buffer_limit(self, nbytes: int) -> None:
        _check(lib().gs3d_device_set_buffer_limit(self._h, int(nbytes)))

    def compute_units(self) -> int:
        n = C.c_int()
        _check(lib().gs3d_device_compute_units(self._h, C.byref(n)))
        return n.value

    def close(self) -> None:
        if self._h:
            lib().gs3d_device_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Texture:
    """Render target (wgpu::Texture)."""

    def __init__(self, device: Device, width: int, height: int, fmt: TextureFormat = TextureFormat.RGBA8_UNORM):
        h = C.c_void_p()
        _check(lib().gs3d_texture_create(device.handle, int(width), int(height), int(fmt), C.byref(h)))
        self._h, self.device, self.width, self.height, self.format = h, device, int(width), int(height), fmt

    @property
    def handle(self):
        return self._h

    def read(self) -> np.ndarray:
        """(height, width, 4) uint8 or float32 ((height, width) float32 for a
        depth texture), top row first."""
        if self.format == TextureFormat.DEPTH32_FLOAT:
            out = np.zeros((self.height, self.width), np.float32)
        else:
            dt = np.uint8 if self.format == TextureFormat.RGBA8_UNORM else np.float32
            out = np.zeros((self.height, self.width, 4), dt)
        _check(lib().gs3d_texture_read(self._h, _ptr(out), out.nbytes))
        return out

    def write(self, data: np.ndarray) -> None:
        """queue.write_texture: the whole texture, in read()'s layout."""
        a = np.ascontiguousarray(data)
        _check(lib().gs3d_texture_write(self._h, _ptr(a), a.nbytes))

    def close(self) -> None:
        if self._h:
            lib().gs3d_texture_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Viewer:
    """Viewer (src/lib.rs:65-276): buffers + Preprocessor + RadixSorter + Renderer."""

    def __init__(self, device: Device, texture_format: TextureFormat, gaussians: np.ndarray,
                 pod: GaussianPod = GaussianPod.SINGLE):
        g = np.ascontiguousarray(gaussians, GAUSSIAN_DTYPE)
        h = C.c_void_p()
        _check(lib().gs3d_viewer_new(device.handle, int(texture_format), _ptr(g) if g.shape[0] else None,
                                     int(g.shape[0]), int(pod), C.byref(h)))
        self._h, self.device, self.count, self.format, self.pod = h, device, int(g.shape[0]), texture_format, pod

    @property
    def handle(self):
        return self._h

    # -- Viewer::update_* (src/lib.rs:202-263)
    def update_camera(self, camera: Camera, texture_size) -> None:
        c = camera._c()
        _check(lib().gs3d_viewer_update_camera(self._h, C.byref(c), int(texture_size[0]), int(texture_size[1])))

    def update_camera_with_pod(self, pod: CameraPod) -> None:
        _check(lib().gs3d_viewer_update_camera_with_pod(self._h, C.byref(pod)))

    def update_model_transform(self, pos, rot, scale) -> None:
        p = (C.c_float * 3)(*[float(x) for x in pos])
        r = (C.c_float * 4)(*[float(x) for x in rot])
        s = (C.c_float * 3)(*[float(x) for x in scale])
        _check(lib().gs3d_viewer_update_model_transform(self._h, p, r, s))

    def update_gaussian_transform(self, size: float = 1.0, display_mode: DisplayMode = DisplayMode.SPLAT,
                                  sh_deg: int = 3, no_sh0: bool = False, max_std_dev: float = 3.0) -> None:
        _check(lib().gs3d_viewer_update_gaussian_transform(self._h, float(size), int(display_mode), int(sh_deg),
                                                           int(bool(no_sh0)), float(max_std_dev)))

    # -- recording
    def render(self, target: Texture) -> None:
        """Viewer::render: preprocess -> sort -> render (enqueue only)."""
        _check(lib().gs3d_viewer_render(self._h, target.handle))

    def preprocess(self) -> None:
        _check(lib().gs3d_viewer_preprocess(self._h))

    def sort(self) -> None:
        _check(lib().gs3d_viewer_sort(self._h))

    def draw(self, target: Texture, standalone: bool = False) -> None:
        fn = lib().gs3d_viewer_draw_standalone if standalone else lib().gs3d_viewer_draw
        _check(fn(self._h, target.handle))

    # -- options
    def set_blend_mode(self, mode: BlendMode) -> None:
        _check(lib().gs3d_viewer_set_blend_mode(self._h, int(mode)))

    def set_tile_size(self, tw: int, th: int) -> None:
        _check(lib().gs3d_viewer_set_tile_size(self._h, int(tw), int(th)))

    def set_sort_items(self, items: int) -> None:
        """Keys per thread of the depth and tile sorts' tiles: 8, 12, 16, 20
        (the default) or 24; every choice sorts the same way."""
        _check(lib().gs3d_viewer_set_sort_items(self._h, int(items)))

    def set_pair_capacity(self, pairs: int) -> None:
        """Replaces the (tile, splat) pair lists with ones of `pairs` pairs
        (they grow again when a frame needs more)."""
        _check(lib().gs3d_viewer_set_pair_capacity(self._h, int(pairs)))

    def set_profiling(self, on: bool) -> None:
        _check(lib().gs3d_viewer_set_profiling(self._h, int(bool(on))))

    def stage_times(self) -> np.ndarray:
        out = np.zeros(7, np.float32)
        _check(lib().gs3d_viewer_stage_times(self._h, _ptr(out)))
        return out

    def set_depth_stencil(self, enabled: bool, compare: "CompareFunction" = None, depth_write: bool = False):
        """ViewerCreateOptions::depth_stencil: a depth test for the renderer."""
        cmp = CompareFunction.LESS if compare is None else compare
        _check(lib().gs3d_viewer_set_depth_stencil(self._h, int(bool(enabled)), int(cmp), int(bool(depth_write))))

    def render_with_depth(self, target: "Texture", depth: "Texture", load: bool = False) -> None:
        """Render into target with a Depth32Float attachment (render_with_pass)."""
        _check(lib().gs3d_viewer_render_with_depth(self._h, target.handle, depth.handle, int(bool(load))))

    def update_selection(self, words) -> None:
        """SelectionBuffer contents: one bit per Gaussian, ceil(N / 32) u32 words."""
        w = np.ascontiguousarray(words, np.uint32)
        _check(lib().gs3d_viewer_update_selection(self._h, _ptr(w) if w.size else None, w.size))

    def update_invert_selection(self, invert: bool) -> None:
        """PreprocessorInvertSelectionBuffer::update (initially true)."""
        _check(lib().gs3d_viewer_update_invert_selection(self._h, int(bool(invert))))

    def pair_count(self) -> int:
        """(tile, splat) pairs in the tile lists: the largest count of the frames
        drawn before this call (which waits for them)."""
        n = C.c_uint32()
        _check(lib().gs3d_viewer_pair_count(self._h, C.byref(n)))
        return n.value

    def count_fragments(self, target: Texture) -> int:
        """Fragments of one frame that pass the discard test (waits)."""
        n = C.c_uint64()
        _check(lib().gs3d_viewer_count_fragments(self._h, target.handle, C.byref(n)))
        return n.value

    # -- buffer downloads
    def indirect_args(self) -> np.ndarray:
        out = np.zeros(4, np.uint32)
        _check(lib().gs3d_viewer_read_indirect_args(self._h, _ptr(out)))
        return out

    def radix_sort_indirect_args(self) -> np.ndarray:
        out = np.zeros(3, np.uint32)
        _check(lib().gs3d_viewer_read_radix_sort_indirect_args(self._h, _ptr(out)))
        return out

    def indirect_indices(self, count: Optional[int] = None) -> np.ndarray:
        count = self.count if count is None else int(count)
        out = np.zeros(max(count, 1), np.uint32)
        _check(lib().gs3d_viewer_read_indirect_indices(self._h, _ptr(out), count))
        return out[:count]

    def gaussians_depth(self, count: Optional[int] = None) -> np.ndarray:
        count = depth_buffer_len(self.count) if count is None else int(count)
        out = np.zeros(max(count, 1), np.float32)
        _check(lib().gs3d_viewer_read_gaussians_depth(self._h, _ptr(out), count))
        return out[:count]

    def close(self) -> None:
        if self._h:
            lib().gs3d_viewer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class RadixSorter:
    """RadixSorter (src/radix_sorter.rs:9-96) on its own key/payload buffers."""

    def __init__(self, device: Device, capacity: int):
        h = C.c_void_p()
        _check(lib().gs3d_radix_sorter_new(device.handle, int(capacity), C.byref(h)))
        self._h, self.device, self.capacity, self.n = h, device, int(capacity), 0

    def write(self, keys: np.ndarray, vals: np.ndarray) -> None:
        k = np.ascontiguousarray(keys, np.uint32)
        v = np.ascontiguousarray(vals, np.uint32)
        if k.shape != v.shape:
            raise ValueError("keys and values differ in length")
        _check(lib().gs3d_radix_sorter_write(self._h, _ptr(k), _ptr(v), int(k.shape[0])))
        self.n = int(k.shape[0])

    def sort(self) -> None:
        _check(lib().gs3d_radix_sorter_sort(self._h))

    def set_items(self, items: int) -> None:
        """Keys per thread of the sort's tiles: 8, 12, 16, 20 (the default) or 24."""
        _check(lib().gs3d_radix_sorter_set_items(self._h, int(items)))

    def read(self):
        k = np.zeros(max(self.n, 1), np.uint32)
        v = np.zeros(max(self.n, 1), np.uint32)
        _check(lib().gs3d_radix_sorter_read(self._h, _ptr(k), _ptr(v), self.n))
        return k[:self.n], v[:self.n]

    def close(self) -> None:
        if self._h:
            lib().gs3d_radix_sorter_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


RADIX_SORT_NOALLOC_MAX_KEYS = 4194304


def radix_sort_noalloc(device: Device, keys_in: int, vals_in: int, keys_out: int, vals_out: int, n: int) -> None:
    """The allocation-free sort (no counterpart in the reference): sorts n
    (key, value) pairs at device addresses keys_in/vals_in into
    keys_out/vals_out on the device's queue 0, using a fixed per-GPU scratch
    array instead of allocating. Asynchronous: Device.poll() waits."""
    _check(lib().gs3d_radix_sort_noalloc(device.handle, C.c_void_p(keys_in), C.c_void_p(vals_in),
                                          C.c_void_p(keys_out), C.c_void_p(vals_out), int(n)))


def pods_array(pods) -> np.ndarray:
    """CameraPods as one contiguous (n, 144) byte array."""
    return np.frombuffer(b"".join(p.to_bytes() for p in pods), np.uint8).reshape(-1, 144).copy()


class Model:
    """A resident model (GaussiansBuffer) shared by batched views and served requests."""

    def __init__(self, device: Device, gaussians: np.ndarray, pod: GaussianPod = GaussianPod.HALF):
        g = np.ascontiguousarray(gaussians, GAUSSIAN_DTYPE)
        h = C.c_void_p()
        _check(lib().gs3d_model_new(device.handle, _ptr(g) if g.shape[0] else None, int(g.shape[0]), int(pod),
                                    C.byref(h)))
        self._h, self.device, self.count, self.pod = h, device, int(g.shape[0]), pod

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            lib().gs3d_model_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _handles(objs):
    arr = (C.c_void_p * max(len(objs), 1))(*[o.handle.value if isinstance(o.handle, C.c_void_p) else o.handle
                                              for o in objs])
    return arr


class MultiModelViewer:
    """MultiModelViewer (src/multi_model.rs:289-531)."""

    def __init__(self, device: Device, texture_format: TextureFormat = TextureFormat.RGBA8_UNORM):
        h = C.c_void_p()
        _check(lib().gs3d_multi_model_viewer_new(device.handle, int(texture_format), C.byref(h)))
        self._h, self.device = h, device

    def insert_model(self, key: str, gaussians: np.ndarray, pod: GaussianPod = GaussianPod.SINGLE):
        g = np.ascontiguousarray(gaussians, GAUSSIAN_DTYPE)
        _check(lib().gs3d_multi_model_viewer_insert_model(self._h, key.encode(), _ptr(g) if g.shape[0] else None,
                                                          int(g.shape[0]), int(pod)))

    def remove_model(self, key: str):
        _check(lib().gs3d_multi_model_viewer_remove_model(self._h, key.encode()))

    def update_camera(self, camera: Camera, texture_size):
        c = camera._c()
        _check(lib().gs3d_multi_model_viewer_update_camera(self._h, C.byref(c), int(texture_size[0]),
                                                           int(texture_size[1])))

    def update_camera_with_pod(self, pod: CameraPod):
        _check(lib().gs3d_multi_model_viewer_update_camera_with_pod(self._h, C.byref(pod)))

    def update_model_transform(self, key: str, pos, rot, scale):
        p = (C.c_float * 3)(*[float(x) for x in pos])
        r = (C.c_float * 4)(*[float(x) for x in rot])
        s = (C.c_float * 3)(*[float(x) for x in scale])
        _check(lib().gs3d_multi_model_viewer_update_model_transform(self._h, key.encode(), p, r, s))

    def update_gaussian_transform(self, size=1.0, display_mode=DisplayMode.SPLAT, sh_deg=3, no_sh0=False,
                                  max_std_dev=3.0):
        _check(lib().gs3d_multi_model_viewer_update_gaussian_transform(self._h, float(size), int(display_mode),
                                                                       int(sh_deg), int(bool(no_sh0)),
                                                                       float(max_std_dev)))

    def set_blend_mode(self, mode: BlendMode):
        _check(lib().gs3d_multi_model_viewer_set_blend_mode(self._h, int(mode)))

    def update_selection(self, key: str, words) -> None:
        w = np.ascontiguousarray(words, np.uint32)
        _check(lib().gs3d_multi_model_viewer_update_selection(self._h, key.encode(), _ptr(w) if w.size else None,
                                                              w.size))

    def update_invert_selection(self, key: str, invert: bool) -> None:
        _check(lib().gs3d_multi_model_viewer_update_invert_selection(self._h, key.encode(), int(bool(invert))))

    def render(self, target: Texture, keys: Sequence[str]):
        arr = (C.c_char_p * max(len(keys), 1))(*[k.encode() for k in keys])
        _check(lib().gs3d_multi_model_viewer_render(self._h, target.handle, arr, len(keys)))

    def close(self):
        if self._h:
            lib().gs3d_multi_model_viewer_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """Batched views over resident models, recorded as one HIP graph."""

    def __init__(self, device: Device, models: Sequence[Model], queues: int = 4):
        self._models = list(models)
        h = C.c_void_p()
        _check(lib().gs3d_batch_new(device.handle, _handles(self._models), len(self._models), int(queues),
                                    C.byref(h)))
        self._h, self.device, self.views = h, device, 0
        self._sizes = []

    def set_layout(self, views) -> None:
        """views: sequence of (model index, width, height)."""
        arr = np.array([tuple(v) for v in views], VIEW_DTYPE)
        _check(lib().gs3d_batch_set_layout(self._h, _ptr(arr) if len(arr) else None, len(arr)))
        self.views = len(arr)
        self._sizes = [(int(v[1]), int(v[2])) for v in views]

    def autotune(self, pods, reps: int = 3) -> None:
        a = pods_array(pods)
        _check(lib().gs3d_batch_autotune(self._h, _ptr(a), int(reps)))

    def render(self, pods, use_graph: bool = True) -> None:
        a = pods_array(pods) if not isinstance(pods, np.ndarray) else np.ascontiguousarray(pods)
        _check(lib().gs3d_batch_render(self._h, _ptr(a), int(bool(use_graph))))

    def finish(self) -> None:
        _check(lib().gs3d_batch_finish(self._h))

    def read(self, view: int) -> np.ndarray:
        w, h = self._sizes[view]
        out = np.zeros((h, w, 4), np.uint8)
        _check(lib().gs3d_batch_read_target(self._h, int(view), _ptr(out), out.nbytes))
        return out

    def tile_size(self, view: int):
        tw, th = C.c_uint32(), C.c_uint32()
        _check(lib().gs3d_batch_tile_size(self._h, int(view), C.byref(tw), C.byref(th)))
        return tw.value, th.value

    def sort_items(self, view: int) -> int:
        """Keys per thread of the view's sort tiles (the autotuner's pick)."""
        n = C.c_uint32()
        _check(lib().gs3d_batch_sort_items(self._h, int(view), C.byref(n)))
        return n.value

    def close(self):
        if self._h:
            lib().gs3d_batch_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Server:
    """Continuous-batching request loop over resident models."""

    def __init__(self, device: Device, models: Sequence[Model], queues: int = 8, hbm_budget: int = 64 << 30):
        self._models = list(models)
        h = C.c_void_p()
        _check(lib().gs3d_server_new(device.handle, _handles(self._models), len(self._models), int(queues),
                                     int(hbm_budget), C.byref(h)))
        self._h, self.device = h, device

    def run(self, requests: np.ndarray, capture=None) -> dict:
        """requests: REQUEST_DTYPE array sorted by arrival. capture: request
        indices whose frames to return. Returns (stats dict, {index: image})."""
        r = np.ascontiguousarray(requests, REQUEST_DTYPE)
        images = {}
        ptrs = None
        if capture:
            ptrs = (C.c_void_p * len(r))()
            for i in capture:
                images[int(i)] = np.zeros((int(r[i]["height"]), int(r[i]["width"]), 4), np.uint8)
                ptrs[int(i)] = images[int(i)].ctypes.data
        st = ServeStats()
        _check(lib().gs3d_server_run(self._h, _ptr(r) if len(r) else None, len(r), ptrs, C.byref(st)))
        return st.as_dict(), images

    def render_alone(self, request) -> np.ndarray:
        r = np.ascontiguousarray(np.asarray(request, REQUEST_DTYPE).reshape(1))
        out = np.zeros((int(r[0]["height"]), int(r[0]["width"]), 4), np.uint8)
        _check(lib().gs3d_server_render_alone(self._h, _ptr(r), _ptr(out)))
        return out

    def close(self):
        if self._h:
            lib().gs3d_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
