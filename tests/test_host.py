"""Host-side checks that need no GPU: the C ABI library loads and exports
every symbol include/gs3d.h declares, host packing matches independent
conversions, synthetic scenes are deterministic, and argument errors come
back as gs3d_status codes."""
from __future__ import annotations

import math
import os
import re
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "gs3d.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(gs3d_[a-z0-9_]+)\s*\(", text)))


def test_every_declared_symbol_is_exported(gs):
    lib = gs.lib()
    declared = _declared_symbols()
    assert len(declared) >= 40
    missing = [s for s in declared if not hasattr(lib, s)]
    assert missing == []
    # and the Python binding declares exactly the header's entry points
    assert sorted(gs.EXPORTED_SYMBOLS) == declared


def test_exports_are_plain_c(gs):
    out = subprocess.run(["nm", "-D", "--defined-only", gs.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r" T (gs3d_\w+)$", out, flags=re.M))
    assert set(_declared_symbols()) <= exported


def test_depth_buffer_len(gs):
    # keys_buffer_size (src/radix_sorter.rs:922-939): rounded up to 3840
    for n, want in [(0, 0), (1, 3840), (3840, 3840), (3841, 7680), (65536, 69120), (1048576, 1052160),
                    (6_000_000, 6_001_920)]:
        assert gs.depth_buffer_len(n) == want


def test_half_pod_packing_matches_numpy(gs):
    g = gs.synth_gaussians(2, 2000, seed=5)
    rs, sh = gs.pack_half(g)
    assert np.array_equal(rs[:, :4].view(np.uint16), g["rot"].astype(np.float16).view(np.uint16))
    assert np.array_equal(rs[:, 4:7].view(np.uint16), g["scale"].astype(np.float16).view(np.uint16))
    assert np.all(rs[:, 7].view(np.uint16) == 0)
    assert np.array_equal(sh[:, :45].view(np.uint16), g["sh"].reshape(-1, 45).astype(np.float16).view(np.uint16))
    assert np.all(sh[:, 45:].view(np.uint16) == 0)


def test_synthetic_scenes_are_counter_based(gs):
    a = gs.synth_gaussians(1, 1000, seed=1)
    b = gs.synth_gaussians(1, 300, seed=1, first=500)
    assert a[500:800].tobytes() == b.tobytes()
    c = gs.synth_gaussians(1, 1000, seed=2)
    assert a.tobytes() != c.tobytes()


@pytest.mark.parametrize("config,lo,hi,logmean", [(0, -1, 1, -4.0), (1, -4, 4, -4.0)])
def test_synthetic_distributions(gs, config, lo, hi, logmean):
    g = gs.synth_gaussians(config, 20000, seed=0)
    assert g["pos"].min() >= lo and g["pos"].max() <= hi
    assert abs(g["pos"].mean()) < 0.05 * (hi - lo)
    ls = np.log(g["scale"])
    assert abs(ls.mean() - logmean) < 0.02 and abs(ls.std() - 0.5) < 0.02
    assert np.allclose(np.linalg.norm(g["rot"], axis=1), 1.0, atol=1e-6)
    op = g["color"][:, 3] / 255.0
    assert 0.45 < op.mean() < 0.55  # sigmoid(N(0,1)) has mean 1/2
    rgb = g["color"][:, :3] / 255.0
    assert rgb.min() >= 0.5 - 1e-3 and rgb.max() <= 0.5 + 0.2821 + 1e-2  # 0.5 + C0 * U(0,1)
    assert abs(g["sh"].std() - 0.1) < 0.005


def test_clustered_scene_has_ground_plane(gs):
    g = gs.synth_gaussians(2, 50000, seed=2)
    on_plane = np.abs(g["pos"][:, 1] + 2.0) < 0.3
    frac = on_plane.mean()
    assert 0.18 < frac < 0.3  # 20% plane (+ the clusters that cross it)
    assert abs(np.log(g["scale"]).mean() + 5.0) < 0.02


def test_camera_pod_look_at(gs):
    pod = gs.CameraPod.look_at((0, 0, 3), (0, 0, 0), (1280, 720))
    a = np.frombuffer(pod.to_bytes(), np.float32)
    view = a[:16].reshape(4, 4).T
    assert np.allclose(view @ [0, 0, 0, 1], [0, 0, -3, 1], atol=1e-6)
    proj = a[16:32].reshape(4, 4).T
    assert proj[1, 1] == pytest.approx(1 / math.tan(math.pi / 6), rel=1e-6)
    assert proj[0, 0] == pytest.approx(proj[1, 1] * 720 / 1280, rel=1e-6)


def test_argument_errors_are_status_codes(gs):
    with pytest.raises(gs.Gs3dError) as e:
        gs.synth_gaussians(9, 10, seed=0)
    assert e.value.status == 1
    with pytest.raises(gs.Gs3dError) as e:
        gs.CameraPod.look_at((0, 0, 0), (0, 0, 0), (10, 10))
    assert e.value.status == 1
    with pytest.raises(gs.Gs3dError) as e:
        gs.CameraPod.look_at((0, 0, 1), (0, 0, 0), (0, 10))
    assert e.value.status == 1
    assert "size" in gs.lib().gs3d_last_error().decode()


def test_device_creation_without_gpu_fails_cleanly(gs):
    # on a host without a GPU (this check's home) the device cannot be
    # created; the error must be a status, not a crash
    import torch

    if torch.cuda.device_count() > 0:
        pytest.skip("a GPU is present")
    with pytest.raises(gs.Gs3dError) as e:
        gs.Device(0)
    assert e.value.status in (1, 4)
