"""GPU tests of the rest of the path: display modes, model and Gaussian
transforms, MultiModelViewer, the reference's own end-to-end predicates,
batched views (HIP graph) and the serving loop."""
from __future__ import annotations

import math

import numpy as np
import pytest

from helpers import camera_pod, max_abs_diff, oracle_inputs, psnr

pytestmark = pytest.mark.gpu


def _case(name):
    print(f"[case] {name}", flush=True)


def _small_scene(gs, n=4096, w=320, h=180, seed=0):
    g = gs.synth_gaussians(0, n, seed)
    pod = gs.CameraPod.look_at((0.3, 0.2, 3.0), (0.0, 0.0, 0.0), (w, h))
    return g, pod, w, h


def _render(gs, device, g, pod, w, h, fmt=None, setup=None, pod_kind=None):
    fmt = gs.TextureFormat.RGBA8_UNORM if fmt is None else fmt
    v = gs.Viewer(device, fmt, g, pod_kind if pod_kind is not None else gs.GaussianPod.SINGLE)
    v.update_camera_with_pod(pod)
    if setup:
        setup(v)
    t = gs.Texture(device, w, h, fmt)
    v.render(t)
    device.poll()
    return v, t.read()


@pytest.mark.parametrize("mode", ["ellipse", "point"])
def test_display_modes_match_oracle(gs, orc, device, mode):
    _case(f"display mode {mode}")
    g, pod, w, h = _small_scene(gs)
    dm = gs.DisplayMode.ELLIPSE if mode == "ellipse" else gs.DisplayMode.POINT
    v, img = _render(gs, device, g, pod, w, h, setup=lambda v: v.update_gaussian_transform(
        1.0, dm, 3, False, 3.0))
    frame = orc.frame_from_pod(pod.to_bytes(), display_mode=int(dm))
    pos, color, sh, cov = oracle_inputs(orc, g)
    _, _, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    d = int(np.abs(img.astype(int) - oimg.astype(int)).max())
    close = (np.abs(img.astype(int) - oimg.astype(int)) <= 2).all(axis=2).mean()
    print(f"{mode}: max diff {d}/255, {close * 100:.4f}% of pixels within 2/255, PSNR {psnr(img, oimg):.1f} dB",
          flush=True)
    assert d <= 2
    assert psnr(img, oimg) >= 45.0


def test_gaussian_transform_size_degree_no_sh0(gs, orc, device):
    _case("gaussian transform size 1.7, sh_deg 1, no_sh0")
    g, pod, w, h = _small_scene(gs)
    v, img = _render(gs, device, g, pod, w, h, setup=lambda v: v.update_gaussian_transform(
        1.7, gs.DisplayMode.SPLAT, 1, True, 2.5))
    frame = orc.frame_from_pod(pod.to_bytes(), gsize=1.7, sh_deg=1, no_sh0=True, max_std_dev=2.5)
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, okey, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    vis = int(v.indirect_args()[1])
    assert vis == len(oidx) and np.array_equal(v.indirect_indices(vis), oidx)
    assert max_abs_diff(img, oimg) <= 2


def test_model_transform_bit_exact(gs, orc, device):
    _case("model transform")
    g, pod, w, h = _small_scene(gs, n=20000)
    mp, mr, ms = (0.2, -0.1, 0.3), (0.0, 0.38268343, 0.0, 0.92387953), (1.3, 0.8, 1.1)
    v, img = _render(gs, device, g, pod, w, h, setup=lambda v: v.update_model_transform(mp, mr, ms))
    model, sr, inv_sr = orc.model_mats(mp, mr, ms)
    base = orc.frame_from_pod(pod.to_bytes())
    a = np.frombuffer(pod.to_bytes(), np.float32)
    frame = orc.make_frame(a[:16].reshape(4, 4).T, a[16:32].reshape(4, 4).T, (w, h), model=model, sr=sr,
                           inv_sr=inv_sr)
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, okey, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    vis = int(v.indirect_args()[1])
    assert vis == len(oidx)
    assert np.array_equal(v.indirect_indices(vis), oidx)
    assert np.array_equal(v.gaussians_depth(vis).view(np.uint32), okey.view(np.uint32))
    assert max_abs_diff(img, oimg) <= 2
    del base


def test_empty_model_renders_black(gs, device):
    _case("empty model")
    g = gs.synth_gaussians(0, 0, 0)
    pod = gs.CameraPod.look_at((0, 0, 3), (0, 0, 0), (64, 48))
    v, img = _render(gs, device, g, pod, 64, 48)
    assert v.indirect_args().tolist() == [6, 0, 0, 0]
    assert v.radix_sort_indirect_args().tolist() == [0, 1, 1]
    assert np.all(img[..., :3] == 0) and np.all(img[..., 3] == 255)


def test_stage_times_nested_and_identical_image(gs, device):
    _case("stage times")
    g, pod, w, h = _small_scene(gs)
    v0, ref = _render(gs, device, g, pod, w, h)
    v0.close()
    v, img = _render(gs, device, g, pod, w, h, setup=lambda vv: vv.set_profiling(True))
    ms = v.stage_times()
    v.close()
    assert np.array_equal(img, ref)
    assert ms.shape == (7,) and np.all(ms >= 0.0) and np.all(np.isfinite(ms))
    # the kernel-only spans lie inside their stages
    assert 0.0 < ms[6] <= ms[0]
    assert 0.0 < ms[5] <= ms[4]


def test_model_size_exceeds_device_limit(gs, device):
    _case("model size limit")
    g = gs.synth_gaussians(0, 1000, 0)
    device.set_buffer_limit(1000 * 100)
    try:
        with pytest.raises(gs.ModelSizeExceedsDeviceLimit):
            gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    finally:
        device.set_buffer_limit(1 << 62)


# ---- the reference's e2e predicates (tests/e2e/viewer.rs) on the GPU ------------

def _red_splat(gs, pos=(0.0, 0.0, 1.0)):
    return gs.gaussians(pos=[pos], rot=[(0, 0, 0, 1)], scale=[(1, 1, 1)], color=[(255, 0, 0, 255)])


def _given_camera(gs):
    return gs.Camera(z=(0.1, 1e4), vertical_fov=math.radians(60), pitch=0.1, yaw=0.1)  # given.rs:6-12


def test_reference_update_camera_with_or_without_pod_equal(gs, device):
    _case("update_camera vs update_camera_with_pod")  # tests/e2e/viewer.rs:39-69
    g = _red_splat(gs)
    cam = _given_camera(gs)
    a = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    b = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    a.update_camera_with_pod(gs.CameraPod.new(cam, (1024, 1024)))
    b.update_camera(cam, (1024, 1024))
    ta = gs.Texture(device, 1024, 1024)
    tb = gs.Texture(device, 1024, 1024)
    a.render(ta)
    b.render(tb)
    device.poll()
    assert np.array_equal(ta.read(), tb.read())


@pytest.mark.parametrize("case", ["red", "no_sh0", "behind"])
def test_reference_predicates(gs, device, case):
    _case(f"reference predicate {case}")  # tests/e2e/viewer.rs:71-201
    g = _red_splat(gs, (0.0, 0.0, -1.0) if case == "behind" else (0.0, 0.0, 1.0))
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(gs.CameraPod.new(_given_camera(gs), (1024, 1024)))
    if case == "no_sh0":
        v.update_gaussian_transform(1.0, gs.DisplayMode.SPLAT, 3, True, 3.0)
    t = gs.Texture(device, 1024, 1024)
    v.render(t)
    device.poll()
    img = t.read()
    cnt = (img > 0).reshape(-1, 4).sum(axis=0)
    if case == "red":
        assert cnt[0] > 1 and cnt[1] < 1 and cnt[2] < 1 and cnt[3] > 1
    elif case == "no_sh0":
        assert (cnt[:3] > 1).all() and np.array_equal(img[..., 0], img[..., 1])
    else:
        assert cnt[:3].sum() == 0


# ---- MultiModelViewer (src/multi_model.rs; tests/e2e/multi_model.rs) ---------------

def test_multi_model_viewer_matches_oracle(gs, orc, device):
    _case("multi-model")
    w, h = 320, 180
    ga = gs.synth_gaussians(0, 3000, 1)
    gb = gs.synth_gaussians(0, 3000, 2)
    gb["pos"] += np.float32(0.5)
    pod = gs.CameraPod.look_at((0.0, 0.0, 3.5), (0.0, 0.0, 0.0), (w, h))
    mm = gs.MultiModelViewer(device, gs.TextureFormat.RGBA8_UNORM)
    mm.insert_model("a", ga)
    mm.insert_model("b", gb)
    mm.update_camera_with_pod(pod)
    t = gs.Texture(device, w, h)
    mm.render(t, ["a", "b"])
    device.poll()
    img = t.read()
    frame = orc.frame_from_pod(pod.to_bytes())
    sps = []
    for g in (ga, gb):  # each model sorted on its own, drawn in key order
        pos, color, sh, cov = oracle_inputs(orc, g)
        idx, key = orc.preprocess(frame, pos, cov)
        _, sidx = orc.stable_sort(key.view(np.uint32), idx)
        sps.append(orc.splats(frame, sidx, pos, color, sh, cov))
    oimg = orc.render(frame, w, h, np.concatenate(sps), orc.TARGET_RGBA8_STEP)
    d = max_abs_diff(img, oimg)
    print(f"multi-model: max diff {d}/255 PSNR {psnr(img, oimg):.1f} dB", flush=True)
    assert d <= 2
    with pytest.raises(gs.ModelNotFound):
        mm.render(t, ["a", "missing"])
    with pytest.raises(gs.ModelNotFound):
        mm.update_model_transform("missing", (0, 0, 0), (0, 0, 0, 1), (1, 1, 1))
    mm.remove_model("b")
    with pytest.raises(gs.ModelNotFound):
        mm.remove_model("b")
    mm.render(t, [])
    device.poll()
    assert np.all(t.read()[..., :3] == 0)


# ---- batched views ------------------------------------------------------------------

def test_batch_graph_matches_plain_viewers(gs, device):
    _case("batch graph vs plain viewers")
    models = [gs.Model(device, gs.synth_gaussians(2, 50_000, 3)), gs.Model(device, gs.synth_gaussians(2, 120_000, 4))]
    layout = [(0, 640, 360), (1, 1280, 720), (0, 320, 180), (1, 640, 360), (1, 200, 120)]
    rng = np.random.default_rng(0)

    def pods():
        out = []
        for (m, w, h) in layout:
            r = rng.uniform(8, 30)
            az, el = rng.uniform(0, 2 * math.pi), rng.uniform(-0.3, 0.6)
            eye = (r * math.cos(el) * math.sin(az), r * math.sin(el), r * math.cos(el) * math.cos(az))
            out.append(gs.CameraPod.look_at(eye, (0, 0, 0), (w, h)))
        return out

    b = gs.Batch(device, models, queues=3)
    b.set_layout(layout)
    p1 = pods()
    b.render(p1, use_graph=False)
    b.finish()
    eager = [b.read(i) for i in range(len(layout))]
    b.render(p1, use_graph=True)  # records and replays the graph
    b.finish()
    graph1 = [b.read(i) for i in range(len(layout))]
    p2 = pods()
    for _ in range(3):  # replays with new cameras, eager work in between
        b.render(p2, use_graph=True)
        b.finish()
        again = [b.read(i) for i in range(len(layout))]
    b.render(p2, use_graph=False)
    b.finish()
    eager2 = [b.read(i) for i in range(len(layout))]
    for i in range(len(layout)):
        assert np.array_equal(eager[i], graph1[i])
        assert np.array_equal(eager2[i], again[i])
    # every view equals a plain single-model viewer with the same camera
    for i, (m, w, h) in enumerate(layout):
        g = gs.synth_gaussians(2, [50_000, 120_000][m], [3, 4][m])
        v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g, gs.GaussianPod.HALF)
        v.update_camera_with_pod(p2[i])
        t = gs.Texture(device, w, h)
        v.render(t)
        device.poll()
        assert np.array_equal(t.read(), eager2[i])
    b.autotune(p2, reps=2)
    for i in range(len(layout)):
        assert b.tile_size(i) in ((8, 8), (16, 16), (32, 8))
        assert b.sort_items(i) in (8, 12, 16, 20, 24)
    b.close()


@pytest.mark.parametrize("items", [8, 12, 16, 24])
def test_sort_items_per_thread_change_nothing_in_a_frame(gs, device, items):
    """The depth and tile sorts with 2048- to 6144-key tiles give the frame of
    the default 5120-key tiles: the sorted indices, keys and image."""
    _case(f"viewer sort items {items}")
    g = gs.synth_gaussians(2, 300_000, 5)
    pod = gs.CameraPod.look_at((6.0, 2.0, 14.0), (0.0, 0.0, 0.0), (1280, 720))
    out = []
    for it in (20, items):
        v, img = _render(gs, device, g, pod, 1280, 720, setup=lambda vv: vv.set_sort_items(it),
                         pod_kind=gs.GaussianPod.HALF)
        n = int(v.indirect_args()[1])
        out.append((img, v.indirect_indices(n), v.gaussians_depth(n).view(np.uint32), v.pair_count()))
        v.close()
    assert out[0][1].shape[0] > 20_000 and out[0][3] > 20_000
    assert np.array_equal(out[0][0], out[1][0])
    assert np.array_equal(out[0][1], out[1][1])
    assert np.array_equal(out[0][2], out[1][2])
    assert out[0][3] == out[1][3]


# ---- serving -----------------------------------------------------------------------

def test_server_results_same_alone_and_in_the_mix(gs, device):
    _case("server alone vs mix")
    models = [gs.Model(device, gs.synth_gaussians(2, n, 10 + k)) for k, n in enumerate((30_000, 80_000, 150_000))]
    rng = np.random.default_rng(1)
    n = 60
    reqs = np.zeros(n, gs.REQUEST_DTYPE)
    t = 0.0
    sizes = [(640, 360), (1280, 720), (320, 180)]
    for i in range(n):
        t += rng.exponential(1.0 / 400.0)
        m = int(rng.integers(0, 3))
        w, h = sizes[int(rng.integers(0, 3))]
        az = rng.uniform(0, 2 * math.pi)
        eye = (20 * math.sin(az), 5.0, 20 * math.cos(az))
        reqs[i]["model"], reqs[i]["width"], reqs[i]["height"], reqs[i]["arrival_s"] = m, w, h, t
        reqs[i]["pod"] = np.frombuffer(gs.CameraPod.look_at(eye, (0, 0, 0), (w, h)).to_bytes(), np.uint8)
    srv = gs.Server(device, models, queues=4, hbm_budget=8 << 30)
    capture = [0, 7, 19, 33, 59]
    stats, images = srv.run(reqs, capture=capture)
    print(f"serve: {stats}", flush=True)
    assert stats["completed"] == n and stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    assert stats["mean_in_flight"] > 0
    for i in capture:
        alone = srv.render_alone(reqs[i])
        assert np.array_equal(alone, images[i]), f"request {i} differs alone and in the mix"
    srv.close()


def test_server_request_beyond_the_budget_is_an_error(gs, device):
    _case("server: a request that can never fit the HBM budget")
    model = gs.Model(device, gs.synth_gaussians(2, 10_000, 12))
    reqs = np.zeros(1, gs.REQUEST_DTYPE)
    reqs[0]["model"], reqs[0]["width"], reqs[0]["height"], reqs[0]["arrival_s"] = 0, 320, 180, 0.0
    reqs[0]["pod"] = np.frombuffer(gs.CameraPod.look_at((0, 0, 20), (0, 0, 0), (320, 180)).to_bytes(), np.uint8)
    srv = gs.Server(device, [model], queues=1, hbm_budget=1 << 20)  # less than one view's buffers
    with pytest.raises(gs.Gs3dError, match="HBM budget"):
        srv.run(reqs)
    srv.close()
    model.close()


# ---- PLY ingestion on the device -------------------------------------------------

def test_ply_model_renders_like_oracle(gs, orc, device):
    """The reference's fixture (coverage/model.ply) read by gs3d_ply_read_file
    and rendered: visible set, keys and image against the oracle."""
    import os
    _case("model.ply")
    g = gs.read_ply(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "model.ply"))
    w, h = 640, 480
    pod = gs.CameraPod.look_at((4.0, 4.0, -14.0), (4.0, 4.0, 4.0), (w, h))
    v, img = _render(gs, device, g, pod, w, h)
    frame = orc.frame_from_pod(pod.to_bytes())
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, okey, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    vis = int(v.indirect_args()[1])
    assert vis == len(oidx) == 9
    assert np.array_equal(v.indirect_indices(vis), oidx)
    assert np.array_equal(v.gaussians_depth(vis).view(np.uint32), okey.view(np.uint32))
    print(f"model.ply: max diff {max_abs_diff(img, oimg)}/255 PSNR {psnr(img, oimg):.1f} dB", flush=True)
    assert max_abs_diff(img, oimg) <= 2 and psnr(img, oimg) >= 45.0


# ---- selection early-out (preprocess.wesl:69-78) ---------------------------------

def test_selection_early_out(gs, orc, device):
    """Visible set with a selection bitset: the oracle's visible set with the
    Gaussians whose bit equals the invert flag removed, for both flags, plus
    the all-zero default and the bitset's length check."""
    _case("selection early-out")
    g, pod, w, h = _small_scene(gs, n=5000)
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(pod)
    t = gs.Texture(device, w, h, gs.TextureFormat.RGBA8_UNORM)
    frame = orc.frame_from_pod(pod.to_bytes())
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, okey, _ = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    rng = np.random.default_rng(5)
    bits = rng.integers(0, 2, size=g.shape[0]).astype(bool)
    words = np.zeros((g.shape[0] + 31) // 32, np.uint32)
    for i in np.flatnonzero(bits):
        words[i // 32] |= np.uint32(1 << (i % 32))
    # default: bitset zero, invert true -> nothing hidden
    v.render(t)
    device.poll()
    assert np.array_equal(v.indirect_indices(int(v.indirect_args()[1])), oidx)
    v.update_selection(words)
    for invert in (True, False):
        v.update_invert_selection(invert)
        v.render(t)
        device.poll()
        vis = int(v.indirect_args()[1])
        keep = oidx[bits[oidx] != invert]
        got = v.indirect_indices(vis)
        assert vis == len(keep) and np.array_equal(got, keep), f"invert={invert}"
        # keys of the kept Gaussians are the oracle's
        okeep = okey[bits[oidx] != invert]
        assert np.array_equal(v.gaussians_depth(vis).view(np.uint32), okeep.view(np.uint32))
    with pytest.raises(gs.Gs3dError, match="selection needs"):
        v.update_selection(words[:-1])
    v.close()


# ---- depth attachment (ViewerCreateOptions::depth_stencil) ----------------------

def _depth_plane(sp, w, h):
    """A depth attachment: the left half at a depth in the widest gap between
    the middle half of the splats' fragment depths, the right half at 1."""
    z = np.sort(sp["z"][sp["drawable"] != 0].astype(np.float64))
    lo, hi = len(z) // 4, 3 * len(z) // 4
    gaps = np.diff(z[lo:hi])
    k = int(np.argmax(gaps))
    zt = np.float32(0.5 * (z[lo + k] + z[lo + k + 1]))
    assert gaps[k] > 1e-6
    d = np.ones((h, w), np.float32)
    d[:, : w // 2] = zt
    return d


@pytest.mark.parametrize("blend,write", [("fast", False), ("fast", True), ("exact", False), ("exact", True)])
def test_depth_attachment_matches_oracle(gs, orc, device, blend, write):
    _case(f"depth attachment blend={blend} write={write}")
    g, pod, w, h = _small_scene(gs)
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(pod)
    v.set_blend_mode(gs.BlendMode.FAST if blend == "fast" else gs.BlendMode.EXACT)
    v.set_depth_stencil(True, gs.CompareFunction.LESS, write)
    frame = orc.frame_from_pod(pod.to_bytes())
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, _, _ = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_F32)
    sp = orc.splats(frame, oidx, pos, color, sh, cov)
    d0 = _depth_plane(sp, w, h)
    depth = gs.Texture(device, w, h, gs.TextureFormat.DEPTH32_FLOAT)
    depth.write(d0)
    t = gs.Texture(device, w, h, gs.TextureFormat.RGBA8_UNORM)
    v.render_with_depth(t, depth)
    img = t.read()
    target = orc.TARGET_RGBA8_F32 if blend == "fast" else orc.TARGET_RGBA8_STEP
    # the exact blend tests in draw order against the attachment as updated;
    # the fast blend against the attachment as it was before the draw (the
    # same for Less except where two splats' depth order and draw order
    # disagree by rounding)
    od = d0.copy()
    oimg = orc.render(frame, w, h, sp, target, depth=od, compare=2, depth_write=write and blend == "exact")
    od_seq = d0.copy()
    orc.render(frame, w, h, sp, target, depth=od_seq, compare=2, depth_write=write)
    plain = orc.render(frame, w, h, sp, target)
    print(f"depth {blend} write={write}: max diff {max_abs_diff(img, oimg)}/255, "
          f"{np.mean(np.any(oimg != plain, axis=2)) * 100:.1f}% of pixels changed by the test", flush=True)
    assert np.any(oimg != plain)  # the plane hides something
    assert max_abs_diff(img, oimg) <= 2
    gd = depth.read()
    if write:
        # the nearest kept, passing fragment's depth (clip z / w, ulps apart)
        close = np.isclose(gd, od_seq, rtol=2e-6, atol=0)
        print(f"depth written: {close.mean() * 100:.3f}% of pixels as the oracle's", flush=True)
        assert close.mean() >= 0.999
        assert np.any(gd != d0)
    else:
        assert np.array_equal(gd, d0)
    # a depth-tested renderer needs the attachment, and one of the target's size
    with pytest.raises(gs.Gs3dError, match="depth attachment"):
        v.render(t)
    small = gs.Texture(device, w // 2, h, gs.TextureFormat.DEPTH32_FLOAT)
    with pytest.raises(gs.Gs3dError, match="Depth32Float of the target's size"):
        v.render_with_depth(t, small)
    v.set_depth_stencil(False)
    v.render(t)
    device.poll()
    v.close()


def test_handles_outlive_their_device_handle(gs):
    """Destroying a device handle before the viewers and textures made from
    it only drops a reference: the rest are still usable and destroyable."""
    _case("device destroyed first")
    dev = gs.Device(0)
    g, pod, w, h = _small_scene(gs, n=512)
    v = gs.Viewer(dev, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(pod)
    t = gs.Texture(dev, w, h, gs.TextureFormat.RGBA8_UNORM)
    dev.close()
    v.render(t)
    img = t.read()
    assert img.shape == (h, w, 4)
    v.close()
    t.close()


@pytest.mark.parametrize("compare", ["GREATER", "GREATER_EQUAL", "NOT_EQUAL"])
def test_depth_writes_with_order_dependent_compare(gs, orc, device, compare):
    """With depth writes, Greater / GreaterEqual / NotEqual depend on the
    depths that farther splats drawn before wrote (src/renderer.rs:300-305,
    back-to-front draw order): only the farthest passing splat of a pixel
    draws under Greater. The default (fast) blend sends such draws through
    the in-order path; image and attachment match the oracle's sequential
    render with writes."""
    _case(f"depth writes compare={compare}")
    g, pod, w, h = _small_scene(gs)
    cmp = getattr(gs.CompareFunction, compare)
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(pod)
    v.set_depth_stencil(True, cmp, True)
    frame = orc.frame_from_pod(pod.to_bytes())
    pos, color, sh, cov = oracle_inputs(orc, g)
    oidx, _, _ = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
    sp = orc.splats(frame, oidx, pos, color, sh, cov)
    d0 = _depth_plane(sp, w, h)
    depth = gs.Texture(device, w, h, gs.TextureFormat.DEPTH32_FLOAT)
    depth.write(d0)
    t = gs.Texture(device, w, h, gs.TextureFormat.RGBA8_UNORM)
    v.render_with_depth(t, depth)
    img, gd = t.read(), depth.read()
    od = d0.copy()
    oimg = orc.render(frame, w, h, sp, orc.TARGET_RGBA8_STEP, depth=od, compare=int(cmp), depth_write=True)
    no_write = orc.render(frame, w, h, sp, orc.TARGET_RGBA8_STEP, depth=d0.copy(), compare=int(cmp))
    d, close = max_abs_diff(img, oimg), np.isclose(gd, od, rtol=2e-6, atol=0).mean()
    print(f"depth writes {compare}: max diff {d}/255, depth as the oracle's at {close * 100:.3f}% of pixels, "
          f"{np.mean(np.any(oimg != no_write, axis=2)) * 100:.1f}% of pixels changed by the writes", flush=True)
    if compare != "NOT_EQUAL":
        assert np.any(oimg != no_write)  # the writes matter for this compare
    assert d <= 2
    assert close >= 0.999
    v.close()


# ---- tile-list overflow: grown and recorded again at poll ------------------------------

def test_overflowed_draws_are_recorded_again(gs, orc, device):
    """Lists far too small for the frame: two frames of one viewer into two
    targets (different cameras) before one poll; both draws overflow, the
    poll grows the lists and records both frames again with the cameras they
    were recorded with. Both match the oracle."""
    _case("overflow: two frames before poll")
    n, w, h = 20_000, 320, 180
    g = gs.synth_gaussians(0, n, 5)
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    pods = [gs.CameraPod.look_at((0.3, 0.2, 3.0), (0, 0, 0), (w, h)),
            gs.CameraPod.look_at((-2.5, 0.4, 1.5), (0, 0, 0), (w, h))]
    targets = [gs.Texture(device, w, h) for _ in pods]
    v.set_pair_capacity(64)
    for pod, t in zip(pods, targets):
        v.update_camera_with_pod(pod)
        v.render(t)
    device.poll()
    pos, color, sh, cov = oracle_inputs(orc, g)
    for pod, t in zip(pods, targets):
        frame = orc.frame_from_pod(pod.to_bytes())
        _, _, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP)
        img = t.read()
        assert img[..., :3].max() > 0
        assert max_abs_diff(img, oimg) <= 2
    # the viewer's buffers hold the last recorded frame, as without overflow
    assert np.array_equal(v.indirect_indices(int(v.indirect_args()[1])),
                          orc.full_frame(orc.frame_from_pod(pods[1].to_bytes()), pos, color, sh, cov, w, h)[0])
    v.close()


def test_overflow_of_a_loading_draw_is_an_error(gs, device):
    """A draw that blends over the target's contents (LoadOp::Load) cannot be
    recorded again once it drew nothing: the poll reports it. A texture
    destroyed before the poll drops its draws."""
    _case("overflow: loading draw")
    g, pod, w, h = _small_scene(gs)
    v = gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g)
    v.update_camera_with_pod(pod)
    v.set_depth_stencil(True, gs.CompareFunction.ALWAYS, False)
    depth = gs.Texture(device, w, h, gs.TextureFormat.DEPTH32_FLOAT)
    t = gs.Texture(device, w, h)
    v.set_pair_capacity(16)
    v.render_with_depth(t, depth, load=True)
    with pytest.raises(gs.Gs3dError, match="cannot be recorded again"):
        device.poll()
    # destroyed before the poll: nothing is recorded again, nothing fails
    v.set_pair_capacity(16)
    t2 = gs.Texture(device, w, h)
    v.render_with_depth(t2, depth, load=True)
    t2.close()
    device.poll()
    # a clearing draw that overflows is recorded again by the next poll
    v.render_with_depth(t, depth)
    device.poll()
    assert t.read()[..., :3].max() > 0
    v.close()
