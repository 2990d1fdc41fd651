"""The batched and serving workloads at their full sizes, on the GPU.

Batched (bench.py's mix section): 4 resident models of 0.5M/1M/3M/8M
Gaussians (clustered scene, f16 pod), 32 views of uneven resolution up to
3840x2160, one HIP graph recorded and replayed. Every view is byte-equal to a
plain Viewer of its model with the same camera and tile size, and the
8M-Gaussian 3840x2160 view and one view of each model match the oracle: the
visible set, depth keys and sorted order bit-exact, the image within 2/255
and at least 45 dB. The reference analogue is MultiModelViewer::render
(src/multi_model.rs:476-530: per model preprocess and sort, then the draws).

Serving (bench.py's serve section): 8 resident models totalling 40M
Gaussians, a Poisson stream of 400 requests/s and a denser one of 1500/s
(several requests on the device at once); one captured request per model is
the same alone and in the mix, and matches the oracle as above.
"""
from __future__ import annotations

import numpy as np
import pytest

import bench
from helpers import max_abs_diff, oracle_inputs_half, psnr

pytestmark = pytest.mark.gpu

ORACLE_THREADS = 16


def _case(msg):
    print(f"[case] {msg}", flush=True)


def _oracle_check(gs, orc, g, pod, img, label, viewer=None):
    """The image against the oracle (and, with a plain viewer that rendered
    the same frame, its visible set, keys and order bit-exact)."""
    frame = orc.frame_from_pod(pod.to_bytes())
    pos, color, sh, cov = oracle_inputs_half(orc, g)
    w, h = int(pod.size[0]), int(pod.size[1])
    oidx, okey, oimg = orc.full_frame(frame, pos, color, sh, cov, w, h, orc.TARGET_RGBA8_STEP,
                                      threads=ORACLE_THREADS)
    d, p = max_abs_diff(img, oimg), psnr(img, oimg)
    print(f"{label}: V={len(oidx)} max diff {d:.0f}/255 PSNR {p:.1f} dB", flush=True)
    if viewer is not None:
        v = int(viewer.indirect_args()[1])
        assert v == len(oidx)
        assert np.array_equal(viewer.indirect_indices(v), oidx)
        assert np.array_equal(viewer.gaussians_depth(v).view(np.uint32), okey.view(np.uint32))
    assert d <= 2 and p >= 45.0, (label, d, p)


# ---- batched: 4 models, 32 views ----------------------------------------------------

@pytest.fixture(scope="module")
def mix(gs, device):
    _case("batched workload: synthesising 12.5M Gaussians")
    gauss = [gs.synth_gaussians(2, n, bench.MIX_SEED + i) for i, n in enumerate(bench.MIX_SIZES)]
    models = [gs.Model(device, g, gs.GaussianPod.HALF) for g in gauss]
    views, pod_sets = bench.mix_workload(gs)
    yield gauss, models, views, pod_sets
    for m in models:
        m.close()


def test_batched_mix_matches_plain_viewers_and_oracle(gs, orc, device, mix):
    gauss, models, views, pod_sets = mix
    assert (3, 3840, 2160) in views
    pods = [gs.CameraPod.from_bytes(bytes(p)) for p in pod_sets[0]]
    b = gs.Batch(device, models, queues=4)
    b.set_layout(views)
    b.autotune(pods, reps=1)
    tiles = [b.tile_size(i) for i in range(len(views))]
    _case(f"batched: tiles {tiles}")
    b.render(pod_sets[0], use_graph=True)  # records the graph and runs it
    b.finish()
    b.render(pod_sets[0], use_graph=True)  # replays it
    b.finish()
    out = [b.read(i) for i in range(len(views))]
    b.close()

    plain = [gs.Viewer(device, gs.TextureFormat.RGBA8_UNORM, g, gs.GaussianPod.HALF) for g in gauss]
    checked = set()
    for i, (m, w, h) in enumerate(views):
        _case(f"batched view {i}: model {m} ({bench.MIX_SIZES[m]}) {w}x{h} tile {tiles[i]}")
        v = plain[m]
        v.set_tile_size(*tiles[i])
        v.update_camera_with_pod(pods[i])
        t = gs.Texture(device, w, h)
        v.render(t)
        device.poll()
        img = t.read()
        t.close()
        assert np.array_equal(img, out[i]), f"view {i} differs from a plain viewer"
        # the 8M-Gaussian 4K view, and the first view of every model
        if (m, w, h) == (3, 3840, 2160) or m not in checked:
            checked.add(m)
            _oracle_check(gs, orc, gauss[m], pods[i], img, f"batched view {i} (model {m}, {w}x{h})", viewer=v)
    assert checked == {0, 1, 2, 3}
    for v in plain:
        v.close()


# ---- serving: 8 models, 40M Gaussians ------------------------------------------------

@pytest.fixture(scope="module")
def serve(gs, device):
    rng = np.random.default_rng(4)
    sizes = bench.serve_model_sizes(rng)
    _case(f"serving workload: models {sizes}")
    models = [gs.Model(device, gs.synth_gaussians(2, n, bench.SERVE_SEED + i), gs.GaussianPod.HALF)
              for i, n in enumerate(sizes)]
    req = bench.serve_requests(gs, rng, sizes, 400.0, 2.0)
    srv = gs.Server(device, models, queues=8)
    # one captured request per model: its first
    capture = [int(np.nonzero(req["model"] == m)[0][0]) for m in range(len(sizes))]
    stats, images = srv.run(req, capture=capture)
    print(f"serve: {stats}", flush=True)
    yield sizes, srv, req, capture, stats, images
    srv.close()
    for m in models:
        m.close()


def test_served_stream_completes_and_overlaps(serve):
    sizes, srv, req, capture, stats, images = serve
    assert abs(sum(sizes) - 40_000_000) <= 8
    assert stats["completed"] == stats["requests"] == len(req) > 600
    assert 0 < stats["p50_ms"] <= stats["p99_ms"] <= stats["max_ms"]
    assert stats["mean_in_flight"] > 0


def test_served_frames_same_alone_and_in_the_mix(serve):
    sizes, srv, req, capture, stats, images = serve
    for i in capture:
        _case(f"served request {i}: model {int(req[i]['model'])} {int(req[i]['width'])}x{int(req[i]['height'])}")
        alone = srv.render_alone(req[i])
        assert np.array_equal(alone, images[i]), f"request {i} differs alone and in the mix"


def test_dense_stream_overlaps_and_is_the_same_alone(gs, serve):
    """1500 requests/s: by Little's law several requests are on the device at
    once (continuous batching, not a queue); each captured request renders the
    same alone as in that mix."""
    sizes, srv, req, capture, stats, images = serve
    rng = np.random.default_rng(9)
    dense = bench.serve_requests(gs, rng, sizes, 1500.0, 1.0)
    cap = [int(np.nonzero(dense["model"] == m)[0][0]) for m in range(len(sizes))]
    _case(f"dense stream: {len(dense)} requests")
    dstats, dimages = srv.run(dense, capture=cap)
    print(f"dense serve: {dstats}", flush=True)
    assert dstats["completed"] == dstats["requests"] == len(dense)
    assert dstats["mean_in_flight"] > 1.0
    for i in cap:
        alone = srv.render_alone(dense[i])
        assert np.array_equal(alone, dimages[i]), f"dense request {i} differs alone and in the mix"


def test_served_frames_match_oracle(gs, orc, serve):
    sizes, srv, req, capture, stats, images = serve
    for i in capture:
        m = int(req[i]["model"])
        _case(f"served request {i} vs oracle: model {m} ({sizes[m]})")
        g = gs.synth_gaussians(2, sizes[m], bench.SERVE_SEED + m)
        pod = gs.CameraPod.from_bytes(bytes(req[i]["pod"]))
        _oracle_check(gs, orc, g, pod, images[i], f"served request {i} (model {m})")
        del g
