"""Pins the CPU oracle (oracle/) to what the mathematics fixes, independently
of the oracle's own formulas: closed forms, textbook routines from numpy /
scipy in float64, invariants, and the reference's own qualitative test
predicates (tests/e2e/viewer.rs). Each check is chosen so that a plausible
slip in the restatement (a dropped term, a wrong sign or index, a transposed
operand) fails it. CPU only.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.special as sps

F32 = np.float32


def _ulp_diff(a, b):
    ai = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    bi = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    return np.abs(ai - bi)


def _quat_to_mat(q):
    x, y, z, w = [float(v) for v in q]
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - w * z), 2 * (x * z + w * y)],
                     [2 * (x * y + w * z), 1 - 2 * (x * x + z * z), 2 * (y * z - w * x)],
                     [2 * (x * z - w * y), 2 * (y * z + w * x), 1 - 2 * (x * x + y * y)]])


def _camera(orc, eye=(0.0, 0.0, 3.0), target=(0.0, 0.0, 0.0), w=1280, h=720, fov=math.pi / 3, near=0.1, far=1e4,
            **kw):
    d = np.array(target, np.float64) - np.array(eye, np.float64)
    d = (d / np.linalg.norm(d)).astype(np.float32)
    view = orc.look_to_rh(eye, d)
    proj = orc.perspective_rh(fov, w / h, near, far)
    return view, proj, orc.make_frame(view, proj, (w, h), **kw)


# --------------------------------------------------------------------------
# camera (src/camera.rs:71-92)
# --------------------------------------------------------------------------

def test_look_to_rh_is_a_rigid_view(orc):
    rng = np.random.default_rng(0)
    for _ in range(20):
        eye = rng.uniform(-5, 5, 3).astype(np.float32)
        pitch, yaw = rng.uniform(-1.4, 1.4), rng.uniform(0, 2 * math.pi)
        f = orc.camera_forward(pitch, yaw)
        v = orc.look_to_rh(eye, f).astype(np.float64)
        r = v[:3, :3]
        assert np.allclose(r @ r.T, np.eye(3), atol=1e-5)  # orthonormal rotation
        assert np.allclose(v @ np.append(eye, 1.0), [0, 0, 0, 1], atol=1e-4)  # eye -> origin
        assert np.allclose(r @ f, [0, 0, -1], atol=1e-5)  # right-handed: forward -> -Z
        fw = np.array([math.cos(pitch) * math.sin(yaw), math.sin(pitch), math.cos(pitch) * math.cos(yaw)])
        assert np.allclose(f, fw, atol=1e-6)


def test_perspective_rh_closed_form(orc):
    fov, aspect, n, f = 1.0, 16 / 9, 0.1, 1e4
    p = orc.perspective_rh(fov, aspect, n, f).astype(np.float64)
    fy = 1.0 / math.tan(fov / 2)
    assert p[1, 1] == pytest.approx(fy, rel=1e-6)
    assert p[0, 0] == pytest.approx(fy / aspect, rel=1e-6)
    assert p[2, 2] == pytest.approx(f / (n - f), rel=1e-6)
    assert p[3, 2] == -1.0 and p[2, 3] == pytest.approx(n * f / (n - f), rel=1e-6)
    # depth 0..1: the near plane maps to 0, the far plane to 1
    for z, want in ((n, 0.0), (f, 1.0)):
        c = p @ np.array([0, 0, -z, 1.0])
        assert c[2] / c[3] == pytest.approx(want, abs=1e-5)


def test_camera_pod_new_matches_reference_layout(gs, orc):
    # tests/buffer/camera.rs:95-125: CameraPod == {camera.view(), camera.projection(w/h), size}
    cam = gs.Camera(z=(0.1, 1e4), vertical_fov=math.radians(60), pos=(0.3, -0.2, 1.0), pitch=0.1, yaw=0.1)
    pod = gs.CameraPod.new(cam, (1024, 768))
    a = np.frombuffer(pod.to_bytes(), np.float32)
    view = orc.look_to_rh(cam.pos, orc.camera_forward(cam.pitch, cam.yaw))
    proj = orc.perspective_rh(F32(cam.vertical_fov), F32(1024) / F32(768), 0.1, 1e4)
    assert len(pod.to_bytes()) == 144
    assert np.array_equal(a[0:16].view(np.uint32), view.T.reshape(-1).view(np.uint32))
    assert np.array_equal(a[16:32].view(np.uint32), proj.T.reshape(-1).view(np.uint32))
    assert a[32] == 1024 and a[33] == 768
    assert np.frombuffer(pod.to_bytes(), np.uint32)[34:].tolist() == [0, 0]


# --------------------------------------------------------------------------
# cov3d (wgpu-3dgs-core packing; parity unpinned against the reference)
# --------------------------------------------------------------------------

def test_cov3d_matches_float64_rssr(gs, orc):
    g = gs.synth_gaussians(0, 500, seed=11)
    c32 = orc.cov3d_f32(g["scale"], g["rot"])
    for i in range(len(g)):
        r = _quat_to_mat(g["rot"][i])
        m = r @ np.diag(g["scale"][i].astype(np.float64))
        s = m @ m.T
        want = np.array([s[0, 0], s[0, 1], s[0, 2], s[1, 1], s[1, 2], s[2, 2]])
        # f32 error relative to the matrix scale (off-diagonals cancel)
        assert np.allclose(c32[i], want, rtol=0, atol=2e-6 * np.abs(np.diag(s)).max())
        assert np.all(np.linalg.eigvalsh(s) >= -1e-18)
    # the product's host packing (f32 pod) is the same arithmetic, bit for bit
    assert np.array_equal(gs.cov3d(g).view(np.uint32), c32.view(np.uint32))


# --------------------------------------------------------------------------
# preprocess (preprocess.wesl:60-126): depth keys and culling
# --------------------------------------------------------------------------

def test_depth_key_closed_form(orc):
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1))
    zs = np.array([0.2, 0.5, 1.0, 2.0, 3.0, 10.0, 100.0, 5e3], np.float32)
    pos = np.stack([np.zeros_like(zs), np.zeros_like(zs), -zs], 1)
    cov = np.tile(np.array([1e-4, 0, 0, 1e-4, 0, 1e-4], np.float32), (len(zs), 1))
    idx, key = orc.preprocess(frame, pos, cov)
    assert idx.tolist() == list(range(len(zs)))  # all straight ahead: visible
    n, f = 0.1, 1e4
    ndcz = (f / (n - f) * (-zs.astype(np.float64)) + n * f / (n - f)) / zs.astype(np.float64)
    want = 1.0 - ndcz
    assert np.allclose(key, want, rtol=2e-6, atol=1e-7)
    assert np.all(np.diff(key) < 0)  # farther -> smaller key -> drawn first


def test_cull_special_cases(orc):
    view, proj, frame = _camera(orc)
    tiny = [1e-6, 0, 0, 1e-6, 0, 1e-6]
    big = [4.0, 0, 0, 4.0, 0, 4.0]
    cases = [
        ((0.0, 0.0, 0.0), tiny, True),     # centre of the view
        ((0.0, 0.0, 5.0), big, False),     # behind the camera: ndc.z > 1
        ((5.0, 0.0, 0.0), tiny, False),    # off screen, small
        ((3.5, 0.0, 0.0), big, True),      # off screen (ndc.x 1.14), bound reaches in
        ((0.0, 0.0, 3.0 - 0.05), tiny, False),  # nearer than the near plane
    ]
    pos = np.array([c[0] for c in cases], np.float32)
    cov = np.array([c[1] for c in cases], np.float32)
    idx, _ = orc.preprocess(frame, pos, cov)
    want = [i for i, c in enumerate(cases) if c[2]]
    assert idx.tolist() == want


def test_negative_minor_eigenvalue_culls_offscreen(orc):
    # an indefinite cov3d makes lambda_min < 0: cov2d_axes returns 0
    # (utils.wesl:66-68), so an off-screen centre stays culled
    view, proj, frame = _camera(orc)
    bad = np.array([[4.0, 0, 0, -4.0, 0, 4.0]], np.float32)
    ax = orc.cov2d_axes(frame, (3.5, 0, 0), bad[0], 3.0)
    assert np.all(ax == 0)
    idx, _ = orc.preprocess(frame, np.array([[3.5, 0, 0]], np.float32), bad)
    assert idx.size == 0


# --------------------------------------------------------------------------
# cov2d / cov2d_axes (utils.wesl:25-79)
# --------------------------------------------------------------------------

def _cov2d_f64(view, proj, size, pos, c):
    w = view[:3, :3].astype(np.float64)
    t = view.astype(np.float64) @ np.append(np.asarray(pos, np.float64), 1.0)
    fx, fy = proj[0, 0] * size[0] * 0.5, proj[1, 1] * size[1] * 0.5
    j = np.array([[fx / t[2], 0, -fx * t[0] / t[2] ** 2], [0, fy / t[2], -fy * t[1] / t[2] ** 2], [0, 0, 0]])
    s = np.array([[c[0], c[1], c[2]], [c[1], c[3], c[4]], [c[2], c[4], c[5]]], np.float64)
    m = j @ w
    return (m @ s @ m.T)[:2, :2]


def test_cov2d_axes_match_eigendecomposition(gs, orc):
    view, proj, frame = _camera(orc, eye=(0.5, 0.3, 3.0), target=(0, 0, 0))
    g = gs.synth_gaussians(0, 200, seed=3)
    cov = orc.cov3d_f32(g["scale"], g["rot"]) * np.float32(50.0)
    sd = 3.0
    for i in range(len(g)):
        c2 = _cov2d_f64(view, proj, (1280, 720), g["pos"][i], cov[i])
        evals, evecs = np.linalg.eigh(c2)
        ax = orc.cov2d_axes(frame, g["pos"][i], cov[i], sd).astype(np.float64)
        major, minor = ax[:2], ax[2:]
        assert np.linalg.norm(major) == pytest.approx(min(sd * math.sqrt(evals[1]), 1024), rel=2e-4)
        assert np.linalg.norm(minor) == pytest.approx(min(sd * math.sqrt(max(evals[0], 0)), 1024), rel=2e-3, abs=1e-3)
        # the major axis is the eigenvector of the larger eigenvalue
        cosang = abs(major @ evecs[:, 1]) / np.linalg.norm(major)
        assert cosang == pytest.approx(1.0, abs=1e-4)
        assert abs(major @ minor) <= 1e-3 * np.linalg.norm(major) * max(np.linalg.norm(minor), 1e-9)


def test_isotropic_on_axis_is_a_circle_and_axes_clamp(orc):
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=1024, h=1024)
    s, z, sd = 0.01, 2.0, 3.0
    ax = orc.cov2d_axes(frame, (0, 0, -z), [s * s, 0, 0, s * s, 0, s * s], sd)
    focal = proj[1, 1] * 1024 * 0.5
    r = sd * focal * s / z
    assert np.linalg.norm(ax[:2]) == pytest.approx(r, rel=1e-5)
    assert np.linalg.norm(ax[2:]) == pytest.approx(r, rel=1e-5)
    # utils.wesl:73-74: lengths clamp at 1024 px
    big = orc.cov2d_axes(frame, (0, 0, -z), [4.0, 0, 0, 4.0, 0, 4.0], sd)
    assert np.linalg.norm(big[:2]) == pytest.approx(1024.0, rel=1e-6)
    # utils.wesl:71: an axis-aligned Gaussian (cov.xy = 0, lambda_max = cov.xx)
    # takes the (0, 1) direction instead of normalising a zero vector
    aa = orc.cov2d_axes(frame, (0, 0, -z), [s * s, 0, 0, 4 * s * s, 0, s * s], sd)
    assert np.all(np.isfinite(aa))


# --------------------------------------------------------------------------
# view_color (utils.wesl:82-136) against real spherical harmonics
# --------------------------------------------------------------------------

def _real_sh_cs(l, m, d):
    """Real SH built from scipy's complex SH (Condon-Shortley phase included)."""
    x, y, z = d
    theta = math.acos(max(-1.0, min(1.0, z)))  # polar
    phi = math.atan2(y, x)                     # azimuth
    if m == 0:
        return float(np.real(sps.sph_harm_y(l, 0, theta, phi)))
    ylm = sps.sph_harm_y(l, abs(m), theta, phi)
    return float(math.sqrt(2) * (np.real(ylm) if m > 0 else np.imag(ylm)))


@pytest.mark.parametrize("k", range(15))
def test_sh_basis_functions(orc, k):
    l = 1 if k < 3 else (2 if k < 8 else 3)
    m = k - {1: 0, 2: 3, 3: 8}[l] - l
    rng = np.random.default_rng(k)
    for _ in range(8):
        d = rng.normal(size=3)
        d = d / np.linalg.norm(d)
        sh = np.zeros((15, 3))
        sh[k, 0] = 0.1
        out = orc.view_color([0, 0, 0, 255], sh, d, sh_deg=3, no_sh0=True, f64=True)
        basis = (out[0] - 0.5) / 0.1
        # the constants at utils.wesl:83-87 are 7-8 significant digits
        assert basis == pytest.approx(_real_sh_cs(l, m, d), rel=2e-6, abs=2e-7)
        assert out[1] == 0.5 and out[2] == 0.5  # other channels untouched


def test_view_color_dc_degree_and_clamp(orc):
    d = np.array([0.0, 0.0, 1.0])
    sh = np.full((15, 3), 0.3)
    c = orc.view_color([255, 128, 0, 51], sh, d, sh_deg=0)
    assert c[0] == pytest.approx(1.0) and c[1] == pytest.approx(128 / 255) and c[2] == 0.0
    assert c[3] == pytest.approx(0.2)
    neg = orc.view_color([0, 0, 0, 255], -sh, d, sh_deg=3)
    assert np.all(neg[:3] >= 0.0)  # max(result, 0) (utils.wesl:135)
    grey = orc.view_color([255, 0, 0, 255], np.zeros((15, 3)), d, sh_deg=3, no_sh0=True)
    assert np.allclose(grey[:3], 0.5)


# --------------------------------------------------------------------------
# sort (radix_sort.wgsl semantics)
# --------------------------------------------------------------------------

@pytest.mark.parametrize("n", [0, 1, 7, 3840, 100_003])
def test_stable_sort_matches_numpy(orc, n):
    rng = np.random.default_rng(n)
    keys = rng.integers(0, 50, n).astype(np.uint32) * np.uint32(0x01010101)  # heavy ties
    vals = np.arange(n, dtype=np.uint32)
    k, v = orc.stable_sort(keys, vals)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[o]) and np.array_equal(v, vals[o])


# --------------------------------------------------------------------------
# compositing (render.wesl:143-151 + ALPHA_BLENDING)
# --------------------------------------------------------------------------

def _single_splat_scene(orc, pos, s, color, target, w=64, h=64):
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=w, h=h)
    cov = np.array([[s * s, 0, 0, s * s, 0, s * s]], np.float32)
    p = np.array([pos], np.float32)
    idx, key = orc.preprocess(frame, p, cov)
    sp = orc.splats(frame, idx, p, np.array([color], np.uint8), np.zeros((1, 45), np.float32), cov)
    return frame, proj, orc.render(frame, w, h, sp, target)


def test_single_splat_image_closed_form(orc):
    s, z, sd, w = 0.01, 1.0, 3.0, 64
    frame, proj, img = _single_splat_scene(orc, (0, 0, -z), s, (255, 128, 64, 204), orc.TARGET_RGBA32F, w, w)
    focal = proj[1, 1] * w * 0.5
    lam = (focal * s / z) ** 2                     # cov2d eigenvalue (px^2)
    ys, xs = np.mgrid[0:w, 0:w]
    d2 = (xs + 0.5 - w / 2) ** 2 + (ys + 0.5 - w / 2) ** 2
    r2 = 4.0 * d2 / (sd * sd * lam)                # |qo|^2 (render.wesl:116-121)
    alpha = np.where(r2 <= sd * sd, 0.8 * np.exp(-r2), 0.0)
    for ch, c in enumerate((1.0, 128 / 255, 64 / 255)):
        assert np.allclose(img[..., ch], c * alpha, atol=2e-6)
    assert np.allclose(img[..., 3], 1.0, atol=1e-6)  # clear alpha 1 stays 1


def test_two_splats_over_operator(orc):
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=32, h=32)
    pos = np.array([[0, 0, -2.0], [0, 0, -1.0]], np.float32)
    cov = np.array([[1e-4, 0, 0, 1e-4, 0, 1e-4]] * 2, np.float32)
    col = np.array([[255, 0, 0, 255], [0, 0, 255, 128]], np.uint8)
    idx, key = orc.preprocess(frame, pos, cov)
    k, order = orc.stable_sort(key.view(np.uint32), idx)
    assert order.tolist() == [0, 1]  # far first
    sp = orc.splats(frame, order, pos, col, np.zeros((2, 45), np.float32), cov)
    img = orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F)
    a_far, a_near = sp["a"][0], sp["a"][1]
    # each splat alone gives its fragment alpha; together, the nearer one
    # (drawn last) covers the farther: C = c_n a_n + c_f a_f (1 - a_n)
    al_far = orc.render(frame, 32, 32, sp[:1], orc.TARGET_RGBA32F)[16, 16, 0]
    al_near = orc.render(frame, 32, 32, sp[1:], orc.TARGET_RGBA32F)[16, 16, 2]
    assert img[16, 16, 2] == pytest.approx(al_near, rel=1e-6)
    assert img[16, 16, 0] == pytest.approx(al_far * (1.0 - al_near / 1.0), rel=1e-5)
    assert a_far == 1.0 and a_near == pytest.approx(128 / 255)


def test_unorm8_per_step_rounding_is_sticky(orc):
    # 1000 faint white splats over black: in f32 they add up; blending into
    # Rgba8Unorm rounds every step back to 0 (src*a < 0.5/255)
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=16, h=16)
    n = 1000
    pos = np.tile(np.array([[0, 0, -1.0]], np.float32), (n, 1))
    cov = np.tile(np.array([[1.0, 0, 0, 1.0, 0, 1.0]], np.float32), (n, 1))
    col = np.tile(np.array([[255, 255, 255, 0]], np.uint8), (n, 1))
    col[:, 3] = 0
    sp = orc.splats(frame, np.arange(n, dtype=np.uint32), pos, col, np.zeros((n, 45), np.float32), cov)
    sp["a"] = np.float32(0.001)
    step = orc.render(frame, 16, 16, sp, orc.TARGET_RGBA8_STEP)
    f32 = orc.render(frame, 16, 16, sp, orc.TARGET_RGBA32F)
    assert step[8, 8, 0] == 0
    assert f32[8, 8, 0] > 0.5


# --------------------------------------------------------------------------
# the reference's own end-to-end predicates (tests/e2e/viewer.rs)
# --------------------------------------------------------------------------

def _reference_single_red(orc, gs, pos=(0.0, 0.0, 1.0), no_sh0=False, size=256):
    cam = gs.Camera(z=(0.1, 1e4), vertical_fov=math.radians(60), pitch=0.1, yaw=0.1)  # given.rs:6-12
    pod = gs.CameraPod.new(cam, (size, size))
    frame = orc.frame_from_pod(pod.to_bytes(), no_sh0=no_sh0)
    g = gs.gaussians(pos=[pos], rot=[(0, 0, 0, 1)], scale=[(1, 1, 1)], color=[(255, 0, 0, 255)])
    p = g["pos"].astype(np.float32)
    cov = orc.cov3d_f32(g["scale"], g["rot"])
    idx, key, img = orc.full_frame(frame, p, g["color"], np.zeros((1, 45), np.float32), cov, size, size,
                                   orc.TARGET_RGBA8_STEP)
    return idx, img


def test_reference_predicate_red_splat(orc, gs):
    idx, img = _reference_single_red(orc, gs)  # tests/e2e/viewer.rs:71-96
    assert idx.tolist() == [0]
    assert (img[..., 0] > 0).sum() > 1 and (img[..., 1] > 0).sum() < 1 and (img[..., 2] > 0).sum() < 1
    assert (img[..., 3] > 0).sum() > 1


def test_reference_predicate_no_sh0_grey(orc, gs):
    idx, img = _reference_single_red(orc, gs, no_sh0=True)  # tests/e2e/viewer.rs:98-155
    for ch in range(3):
        assert (img[..., ch] > 0).sum() > 1
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 1], img[..., 2])


def test_reference_predicate_behind_camera(orc, gs):
    idx, img = _reference_single_red(orc, gs, pos=(0.0, 0.0, -1.0))  # tests/e2e/viewer.rs:157-201
    assert idx.size == 0
    assert img[..., :3].sum() == 0


# ---- depth attachment (pipeline depth_stencil, src/renderer.rs:300-305) ----------

def _two_splat_scene(orc):
    """A far red splat and a near half-transparent blue one on the axis,
    drawn far first; their fragment depths are the centres' clip z / w."""
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=32, h=32)
    pos = np.array([[0, 0, -2.0], [0, 0, -1.0]], np.float32)
    cov = np.array([[1e-2, 0, 0, 1e-2, 0, 1e-2]] * 2, np.float32)  # 3 sd: ~4 and ~8 px
    col = np.array([[255, 0, 0, 255], [0, 0, 255, 128]], np.uint8)
    sp = orc.splats(frame, np.array([0, 1], np.uint32), pos, col, np.zeros((2, 45), np.float32), cov)
    return frame, proj, sp


def test_fragment_depth_closed_form(orc):
    # perspective_rh (depth 0..1): z_ndc = far / (far - near) * (1 - near / d)
    frame, proj, sp = _two_splat_scene(orc)
    n, f = 0.1, 1e4
    for k, d in enumerate((2.0, 1.0)):
        assert sp["z"][k] == pytest.approx(f / (f - n) * (1.0 - n / d), rel=1e-6)
    assert sp["z"][1] < sp["z"][0]


def test_depth_test_compare_functions(orc):
    frame, proj, sp = _two_splat_scene(orc)
    plain = orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F)
    ones = lambda: np.ones((32, 32), np.float32)  # noqa: E731
    zeros = lambda: np.zeros((32, 32), np.float32)  # noqa: E731
    # every drawn fragment lies in front of the far plane: Less against 1.0,
    # and Always, draw everything; Never and Less against 0.0 draw nothing
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=ones(), compare=2), plain)
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=zeros(), compare=8), plain)
    black = orc.render(frame, 32, 32, sp[:0], orc.TARGET_RGBA32F)
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=ones(), compare=1), black)
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=zeros(), compare=2), black)
    # a plane between the two: only the near splat passes Less, only the far
    # one passes Greater
    mid = np.full((32, 32), 0.5 * (sp["z"][0] + sp["z"][1]), np.float32)
    near_only = orc.render(frame, 32, 32, sp[1:], orc.TARGET_RGBA32F)
    far_only = orc.render(frame, 32, 32, sp[:1], orc.TARGET_RGBA32F)
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=mid.copy(), compare=2), near_only)
    assert np.array_equal(orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=mid.copy(), compare=5), far_only)
    assert far_only[16, 16, 0] > 0 and near_only[16, 16, 2] > 0


def test_depth_write_follows_draw_order(orc):
    frame, proj, sp = _two_splat_scene(orc)
    # drawn near first with depth writes and Less, the far splat then fails
    # wherever the near one kept a fragment
    rev = sp[::-1].copy()
    d = np.ones((32, 32), np.float32)
    img = orc.render(frame, 32, 32, rev, orc.TARGET_RGBA32F, depth=d, compare=2, depth_write=True)
    near_only = orc.render(frame, 32, 32, sp[1:], orc.TARGET_RGBA32F)
    near_mask = near_only[..., 2] > 0
    assert np.array_equal(img[near_mask], near_only[near_mask])
    assert np.all(d[near_mask] == sp["z"][1])
    # without writes the far splat blends over the near one there
    img2 = orc.render(frame, 32, 32, rev, orc.TARGET_RGBA32F, depth=np.ones((32, 32), np.float32), compare=2)
    assert img2[16, 16, 0] > 0 and img[16, 16, 0] == 0
    # back to front (the renderer's order) a write changes nothing for Less
    d3 = np.ones((32, 32), np.float32)
    img3 = orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F, depth=d3, compare=2, depth_write=True)
    assert np.array_equal(img3, orc.render(frame, 32, 32, sp, orc.TARGET_RGBA32F))


# --------------------------------------------------------------------------
# model transform (core's model_transform_mat / model_scale_rot_mat /
# model_transform_inv_sr_mat, called at utils.wesl:37, render.wesl:65,85;
# fed by Viewer::update_model_transform, src/lib.rs:217-225)
# --------------------------------------------------------------------------

def _rand_unit_quat(rng):
    q = rng.normal(size=4)
    return q / np.linalg.norm(q)


def test_model_mats_are_translate_rotate_scale(orc):
    # model (p, 1) = R(q) diag(s) p + t, built from the textbook rotation
    # matrix in float64; inv_sr is the inverse of sr. Non-uniform scales make
    # R S and S R differ, so a swapped or transposed factor fails.
    rng = np.random.default_rng(7)
    for _ in range(25):
        q = _rand_unit_quat(rng)
        s = rng.uniform(0.2, 3.0, 3)
        t = rng.uniform(-5, 5, 3)
        model, sr, inv_sr = (m.astype(np.float64) for m in orc.model_mats(t, q, s))
        r = _quat_to_mat(q)
        want_sr = r @ np.diag(s)
        assert np.allclose(sr, want_sr, atol=2e-6 * s.max())
        for _k in range(4):
            p = rng.uniform(-3, 3, 3)
            assert np.allclose(model @ np.append(p, 1.0), np.append(want_sr @ p + t, 1.0), atol=2e-5 * (1 + s.max()))
        assert np.allclose(inv_sr @ sr, np.eye(3), atol=5e-6 * s.max() / s.min())
        assert np.allclose(inv_sr, np.linalg.inv(want_sr), rtol=0, atol=5e-6 / s.min())


def test_model_mats_known_rotation(orc):
    # 90 degrees about +Z (q = (0, 0, sin 45, cos 45)): x -> y, y -> -x
    h = math.sqrt(0.5)
    model, sr, inv_sr = orc.model_mats((1.0, 2.0, 3.0), (0.0, 0.0, h, h), (2.0, 3.0, 4.0))
    assert np.allclose(model @ [1, 0, 0, 1], [1.0, 4.0, 3.0, 1.0], atol=1e-6)   # 2 x -> (0, 2, 0), + t
    assert np.allclose(model @ [0, 1, 0, 1], [-2.0, 2.0, 3.0, 1.0], atol=1e-6)  # 3 y -> (-3, 0, 0), + t
    assert np.allclose(model @ [0, 0, 1, 1], [1.0, 2.0, 7.0, 1.0], atol=1e-6)   # 4 z
    assert np.allclose(inv_sr @ [0.0, 2.0, 0.0], [1.0, 0.0, 0.0], atol=1e-6)


def _model_frame(orc, view, proj, size, t, q, s, **kw):
    model, sr, inv_sr = orc.model_mats(t, q, s)
    return orc.make_frame(view, proj, size, model=model, sr=sr, inv_sr=inv_sr, **kw)


def test_preprocess_applies_model_transform(orc):
    # model_to_world (preprocess.wesl:82): a point translated and rotated by
    # the model transform has the depth key of its world position
    view, proj, _ = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1))
    h = math.sqrt(0.5)
    frame = _model_frame(orc, view, proj, (1280, 720), (0.0, 0.0, -4.0), (0.0, h, 0.0, h), (1.0, 1.0, 2.0))
    # model-space x = 1 rotates (90 deg about +Y) to world -z, then t
    pos = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0], [0.0, 0.0, 3.0]], np.float32)
    cov = np.tile(np.array([1e-6, 0, 0, 1e-6, 0, 1e-6], np.float32), (3, 1))
    idx, key = orc.preprocess(frame, pos, cov)
    # world: (0, 0, -5), (0, 0, -3), (2 * 3 = 6 along +x ... ) -> (6, 0, -4): off screen
    assert idx.tolist() == [0, 1]
    n, f = 0.1, 1e4
    for k, d in zip(idx, (5.0, 3.0)):
        assert key[k] == pytest.approx(1.0 - f / (f - n) * (1.0 - n / d), rel=1e-5)


# --------------------------------------------------------------------------
# view direction fed to view_color (render.wesl:58-73)
# --------------------------------------------------------------------------

def _one_splat_color(orc, frame, pos, coeff_index, value, color=(128, 128, 128, 255)):
    sh = np.zeros((1, 15, 3), np.float32)
    sh[0, coeff_index, :] = value
    cov = np.array([[1e-4, 0, 0, 1e-4, 0, 1e-4]], np.float32)
    sp = orc.splats(frame, np.array([0], np.uint32), np.array([pos], np.float32), np.array([color], np.uint8),
                    sh.reshape(1, 45), cov, f64=False)
    return np.array([sp["r"][0], sp["g"][0], sp["b"][0]], np.float64)


def test_view_direction_straight_ahead(orc):
    # The direction view_color sees is normalize(world - camera) in model
    # space. Looking down a world axis at a splat straight ahead, only the
    # degree-1 term of that axis survives: c1 (-y sh0 + z sh1 - x sh2)
    # (utils.wesl:100-102), with c1 = 0.4886025.
    c1, base, val = 0.4886025, 128 / 255, 0.2
    cases = [  # (eye, target, splat, coefficient, expected sign of its term)
        ((0, 0, 0), (0, 0, -1), (0, 0, -2), 1, -1.0),   # dir -z: + z sh1 -> -sh1
        ((0, 0, 0), (0, 0, 1), (0, 0, 2), 1, +1.0),     # dir +z
        ((0, 0, 0), (1, 0, 0), (2, 0, 0), 2, -1.0),     # dir +x: - x sh2
        ((0, 0, 0), (-1, 0, 0), (-2, 0, 0), 2, +1.0),   # dir -x
        ((0, 5, 0), (0, 0, 0.001), (0, 0, 0), 0, +1.0),  # dir -y: - y sh0 -> +sh0
    ]
    for eye, target, p, k, sign in cases:
        view, proj, frame = _camera(orc, eye=eye, target=target, sh_deg=1)
        got = _one_splat_color(orc, frame, p, k, val)
        assert np.allclose(got, base + sign * c1 * val, atol=2e-6), (eye, target, k, got)
        # the other two degree-1 coefficients contribute nothing on this axis
        for other in {0, 1, 2} - {k}:
            assert np.allclose(_one_splat_color(orc, frame, p, other, val), base, atol=2e-6)


def test_view_direction_uses_inverse_model_scale_rotation(orc):
    # render.wesl:64-65: dir = -normalize(inv(R S) (campos - world)), against
    # numpy's inverse in float64, for random model transforms and cameras,
    # read through a degree-1-only SH (colour = 0.5 + c1 (-y s0 + z s1 - x s2)).
    rng = np.random.default_rng(5)
    c1 = 0.4886025
    for _ in range(20):
        q, s, t = _rand_unit_quat(rng), rng.uniform(0.5, 2.0, 3), rng.uniform(-1, 1, 3)
        eye = rng.uniform(-6, 6, 3)
        view, proj, _ = _camera(orc, eye=tuple(eye), target=(0, 0, 0))
        frame = _model_frame(orc, view, proj, (1280, 720), t, q, s, sh_deg=1, no_sh0=True)
        p = rng.uniform(-1, 1, 3)
        sr = _quat_to_mat(q) @ np.diag(s)
        world = sr @ p + t
        d = np.linalg.inv(sr) @ (world - eye)
        d /= np.linalg.norm(d)
        sh = np.zeros((1, 15, 3), np.float32)
        sh[0, 0, 0], sh[0, 1, 1], sh[0, 2, 2] = 0.1, 0.1, 0.1  # r reads y, g reads z, b reads x
        cov = np.array([[1e-4, 0, 0, 1e-4, 0, 1e-4]], np.float32)
        sp = orc.splats(frame, np.array([0], np.uint32), np.array([p], np.float32),
                        np.array([[0, 0, 0, 255]], np.uint8), sh.reshape(1, 45), cov)
        got = np.array([sp["r"][0], sp["g"][0], sp["b"][0]], np.float64)
        want = np.maximum(0.5 + 0.1 * c1 * np.array([-d[1], d[2], -d[0]]), 0.0)
        assert np.allclose(got, want, atol=5e-6), (got, want)


# --------------------------------------------------------------------------
# footprint of a splat in pixel space (render.wesl:106-151, camera.wesl:18-20)
# --------------------------------------------------------------------------

def _pixel_of(view, proj, size, world):
    clip = proj.astype(np.float64) @ (view.astype(np.float64) @ np.append(world, 1.0))
    ndc = clip[:2] / clip[3]
    return np.array([(ndc[0] + 1) * 0.5 * size[0], (1 - ndc[1]) * 0.5 * size[1]])


def _expected_footprint(view, proj, size, world, cov3, sd, opacity):
    """EWA footprint restated from first principles in float64: the pixel
    Jacobian of the projection at the centre by central differences (pixel y
    pointing down), Sigma_px = J Sigma J^T, and the shader's falloff: the quad
    corner +-sd reaches sd sqrt(lambda) * sd / 2 px along each axis, so
    |qo|^2 = (4 / sd^2) d^T Sigma_px^-1 d, alpha = a exp(-|qo|^2), discarded
    past |qo|^2 = sd^2."""
    eps = 1e-4
    jac = np.zeros((2, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = eps
        jac[:, k] = (_pixel_of(view, proj, size, world + e) - _pixel_of(view, proj, size, world - e)) / (2 * eps)
    sig = jac @ cov3 @ jac.T
    c = _pixel_of(view, proj, size, world)
    w, h = size
    ys, xs = np.mgrid[0:h, 0:w]
    d = np.stack([xs + 0.5 - c[0], ys + 0.5 - c[1]], -1)
    m = np.einsum("...i,ij,...j->...", d, np.linalg.inv(sig), d)
    r2 = 4.0 / (sd * sd) * m
    return np.where(r2 <= sd * sd, opacity * np.exp(-r2), 0.0), r2, c, sig


@pytest.mark.parametrize("tilt", [45.0, -30.0, 120.0])
def test_anisotropic_footprint_orientation(orc, tilt):
    # A 10:1 needle tilted in the image plane, straight ahead and off axis:
    # every pixel matches the first-principles footprint, so a mirrored
    # (un-flipped y), transposed or rotated ellipse fails.
    w = h = 96
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=w, h=h)
    a = math.radians(tilt)
    q = (0.0, 0.0, math.sin(a / 2), math.cos(a / 2))  # rotation about the view axis
    scale = (0.2, 0.02, 0.02)
    cov = orc.cov3d_f32(np.array([scale]), np.array([q]))
    for p in ((0.0, 0.0, -2.0), (0.3, -0.2, -2.5)):
        sp = orc.splats(frame, np.array([0], np.uint32), np.array([p], np.float32),
                        np.array([[255, 255, 255, 255]], np.uint8), np.zeros((1, 45), np.float32), cov)
        img = orc.render(frame, w, h, sp, orc.TARGET_RGBA32F)
        cov3 = cov[0].astype(np.float64)
        cov3m = np.array([[cov3[0], cov3[1], cov3[2]], [cov3[1], cov3[3], cov3[4]], [cov3[2], cov3[4], cov3[5]]])
        want, r2, c, sig = _expected_footprint(view, proj, (w, h), np.array(p), cov3m, 3.0, 1.0)
        assert np.allclose([sp["cx"][0], sp["cy"][0]], c, atol=1e-3)
        # pixels whose |qo|^2 lies at the discard edge may go either way in f32
        edge = np.abs(r2 - 9.0) < 1e-3
        assert np.allclose(img[..., 0][~edge], want[~edge], atol=2e-4)
        # the major axis runs along the direction the tilt fixes on screen:
        # world (cos a, sin a) appears at pixel (cos a, -sin a)
        evals, evecs = np.linalg.eigh(sig)
        major = evecs[:, 1]
        assert abs(abs(major @ [math.cos(a), -math.sin(a)]) - 1.0) < 2e-3
        lit = img[..., 0] > 0.05
        assert lit.sum() > 10
        ys, xs = np.nonzero(lit)
        dd = np.stack([xs + 0.5 - c[0], ys + 0.5 - c[1]], 1)
        assert np.abs(dd @ [math.sin(a), math.cos(a)]).max() < 3.0  # a thin band across the minor axis
        assert np.abs(dd @ [math.cos(a), -math.sin(a)]).max() > 15.0  # and long along the major one


# --------------------------------------------------------------------------
# ellipse and point display modes (render.wesl:92-104, 153-166)
# --------------------------------------------------------------------------

def test_ellipse_mode_outline_ring(orc):
    # alpha = a inside (sd - 0.1)^2, 1 on the ring (sd - 0.1)^2 < |qo|^2 <= sd^2,
    # discarded outside; with a = 0.2 the three regions give c 0.2, c, 0
    w = h = 128
    view, proj, _ = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=w, h=h)
    sd = 3.0
    frame = orc.make_frame(view, proj, (w, h), display_mode=orc.MODE_ELLIPSE, max_std_dev=sd)
    s, z = 0.15, 2.0
    cov3 = np.array([[s * s, 0, 0, s * s, 0, 2.0 * s * s]], np.float32)
    sp = orc.splats(frame, np.array([0], np.uint32), np.array([[0, 0, -z]], np.float32),
                    np.array([[255, 255, 255, 51]], np.uint8), np.zeros((1, 45), np.float32), cov3)
    img = orc.render(frame, w, h, sp, orc.TARGET_RGBA32F)
    cm = np.diag([s * s, s * s, 2.0 * s * s]).astype(np.float64)
    _, r2, _, _ = _expected_footprint(view, proj, (w, h), np.array([0.0, 0.0, -z]), cm, sd, 1.0)
    e2 = (sd - 0.1) ** 2
    firm = (np.abs(r2 - sd * sd) > 1e-3) & (np.abs(r2 - e2) > 1e-3)
    want = np.where(r2 > sd * sd, 0.0, np.where(r2 > e2, 1.0, 0.2))
    assert np.allclose(img[..., 0][firm], want[firm], atol=1e-6)
    assert (want == 1.0).sum() > 20 and (want == 0.2).sum() > 100  # both regions are present


def test_point_mode_quad_extent(orc):
    # quad offset +-0.01 size, scaled by (aspect, 1) / |view| in clip space
    # (render.wesl:92-96, aspect = H / W, camera.wesl:8-10): a square of
    # half side 0.01 size H / (2 |view|) pixels, colour at alpha 1
    w, h = 200, 100
    view, proj, _ = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=w, h=h)
    size, z = 40.0, 2.0
    frame = orc.make_frame(view, proj, (w, h), display_mode=orc.MODE_POINT, gsize=size)
    sp = orc.splats(frame, np.array([0], np.uint32), np.array([[0, 0, -z]], np.float32),
                    np.array([[255, 0, 0, 10]], np.uint8), np.zeros((1, 45), np.float32),
                    np.array([[1e-4, 0, 0, 1e-4, 0, 1e-4]], np.float32))
    half = 0.01 * size * h / (2.0 * z)  # 10 px
    assert sp["hx"][0] == pytest.approx(half, rel=1e-6) and sp["hy"][0] == pytest.approx(half, rel=1e-6)
    img = orc.render(frame, w, h, sp, orc.TARGET_RGBA32F)
    lit = img[..., 0] > 0
    ys, xs = np.nonzero(lit)
    assert lit.sum() == int(2 * half) ** 2
    assert xs.min() == w // 2 - 10 and xs.max() == w // 2 + 9 and ys.min() == h // 2 - 10 and ys.max() == h // 2 + 9
    assert np.all(img[..., 0][lit] == 1.0)  # alpha 1 whatever the opacity
    # farther away the square shrinks as 1 / |view|
    sp2 = orc.splats(frame, np.array([0], np.uint32), np.array([[0, 0, -2 * z]], np.float32),
                     np.array([[255, 0, 0, 10]], np.uint8), np.zeros((1, 45), np.float32),
                     np.array([[1e-4, 0, 0, 1e-4, 0, 1e-4]], np.float32))
    assert sp2["hx"][0] == pytest.approx(half / 2, rel=1e-6)


# --------------------------------------------------------------------------
# remaining oracle entry points: cov2d, the f64 twins, whole frame
# --------------------------------------------------------------------------

def test_cov2d_matches_float64(gs, orc):
    view, proj, frame = _camera(orc, eye=(0.5, 0.3, 3.0), target=(0, 0, 0))
    g = gs.synth_gaussians(0, 100, seed=4)
    cov = orc.cov3d_f32(g["scale"], g["rot"]) * np.float32(50.0)
    for i in range(len(g)):
        c2 = _cov2d_f64(view, proj, (1280, 720), g["pos"][i], cov[i])
        got = orc.cov2d(frame, g["pos"][i], cov[i]).astype(np.float64)
        scale = np.abs(np.diag(c2)).max()
        assert np.allclose(got, [c2[0, 0], c2[0, 1], c2[1, 1]], rtol=0, atol=2e-5 * scale)


def test_f64_twins_agree_with_f32(gs, orc):
    view, proj, frame = _camera(orc, eye=(0.5, 0.3, 3.0), target=(0, 0, 0))
    g = gs.synth_gaussians(0, 200, seed=9)
    c32 = orc.cov3d_f32(g["scale"], g["rot"])
    c64 = orc.cov3d_f64(g["scale"].astype(np.float64), g["rot"].astype(np.float64))
    dscale = np.abs(c64[:, [0, 3, 5]]).max(1, keepdims=True)  # off-diagonals cancel: error relative to the diagonal
    assert np.all(np.abs(c32 - c64) <= 2e-6 * dscale)
    for i in range(0, 200, 20):
        a32 = orc.cov2d_axes(frame, g["pos"][i], c32[i] * 50, 3.0)
        a64 = orc.cov2d_axes(frame, g["pos"][i], c64[i] * 50, 3.0, f64=True)
        assert np.allclose(a32, a64, rtol=1e-3, atol=1e-3 * np.abs(a64).max())
    # the f64 render equals the closed-form single-splat image as the f32 one does
    s, z, w = 0.01, 1.0, 48
    view, proj, frame = _camera(orc, eye=(0, 0, 0), target=(0, 0, -1), w=w, h=w)
    cov = np.array([[s * s, 0, 0, s * s, 0, s * s]], np.float64)
    sp = orc.splats(frame, np.array([0], np.uint32), np.array([[0, 0, -z]], np.float64),
                    np.array([[255, 255, 255, 255]], np.uint8), np.zeros((1, 45)), cov, f64=True)
    img = orc.render(frame, w, w, sp, f64=True)
    want, _, _, _ = _expected_footprint(view, proj, (w, w), np.array([0, 0, -z], np.float64), np.diag([s * s] * 3),
                                        3.0, 1.0)
    assert np.allclose(img[..., 0], want, atol=1e-6)


def test_full_frame_is_the_composition_of_its_stages(gs, orc):
    # Viewer::render (src/lib.rs:266-275): preprocess -> stable sort by key
    # bits -> vertex stage -> blend, composed by hand from the pinned stages
    n, w, h = 3000, 160, 90
    g = gs.synth_gaussians(0, n, seed=12)
    view, proj, frame = _camera(orc, w=w, h=h)
    pos = np.ascontiguousarray(g["pos"], np.float32)
    sh = np.ascontiguousarray(g["sh"].reshape(n, 45), np.float32)
    cov = orc.cov3d_f32(g["scale"], g["rot"])
    idx, key = orc.preprocess(frame, pos, cov)
    order = np.argsort(key.view(np.uint32), kind="stable")
    sp = orc.splats(frame, idx[order], pos, g["color"], sh, cov)
    img = orc.render(frame, w, h, sp, orc.TARGET_RGBA8_STEP)
    fidx, fkey, fimg = orc.full_frame(frame, pos, g["color"], sh, cov, w, h, orc.TARGET_RGBA8_STEP, threads=2)
    assert np.array_equal(fidx, idx[order]) and np.array_equal(fkey.view(np.uint32), key[order].view(np.uint32))
    assert np.array_equal(fimg, img)
