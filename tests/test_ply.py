"""PLY ingestion (gs3d_ply_read_file / gs3d_ply_read_memory), host only.

The reference loads models through wgpu-3dgs-core's
Gaussians::read_from_file(path, GaussiansSource::Ply) (examples/simple.rs:
157-160); core is not in the reference tree. Its only PLY fixture,
coverage/model.ply (copied to tests/golden/model.ply), encodes round values:
pure primaries, a 127/255 grey, opacity +inf, scales ln(0.5)/ln(1.5)/...,
unit quaternions with rot_0 the large (w) component. Reading it back to those
values pins the byte layout and the DC-colour / logit-opacity / log-scale /
rot_0 = w conventions. The f_rest channel-major order is the published 3DGS
layout; core's own choice is parity unpinned (the fixture's f_rest are zero).
"""
import math
import os

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
MODEL_PLY = os.path.join(HERE, "golden", "model.ply")
C0 = 0.28209479177387814


def _ply_bytes(fields, rows, fmt="binary_little_endian", types=None, pre=b""):
    """A PLY file with one vertex element of the given float fields."""
    types = types or {}
    head = ["ply", f"format {fmt} 1.0", "comment made by the test"]
    head.append(f"element vertex {len(rows)}")
    for f in fields:
        head.append(f"property {types.get(f, 'float')} {f}")
    head.append("end_header")
    text = ("\n".join(head) + "\n").encode()
    if fmt == "ascii":
        body = "".join(" ".join(repr(float(v)) for v in r) + "\n" for r in rows).encode()
        return text + body
    dt = np.dtype([(f, {"float": "<f4", "double": "<f8", "uchar": "u1"}[types.get(f, "float")]) for f in fields])
    arr = np.zeros(len(rows), dt)
    for i, r in enumerate(rows):
        arr[i] = tuple(r)
    return text + pre + arr.tobytes()


STD_FIELDS = (["x", "y", "z", "nx", "ny", "nz", "f_dc_0", "f_dc_1", "f_dc_2"] +
              [f"f_rest_{i}" for i in range(45)] +
              ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])


def _std_row(pos, dc, rest, opacity, log_scale, q_wxyz):
    return list(pos) + [0, 0, 0] + list(dc) + list(rest) + [opacity] + list(log_scale) + list(q_wxyz)


def test_reference_fixture_reads_to_its_encoded_values(gs):
    g = gs.read_ply(MODEL_PLY)
    assert g.shape == (9,)
    colors = [(255, 0, 0), (0, 255, 0), (0, 0, 255), (255, 255, 0), (0, 255, 255), (255, 0, 255),
              (255, 255, 255), (127, 127, 127), (0, 0, 0)]
    assert [tuple(int(c) for c in r["color"][:3]) for r in g] == colors
    assert np.all(g["color"][:, 3] == 255)  # opacity +inf -> sigmoid 1
    np.testing.assert_allclose(g["scale"][0], [0.5, 1.0, 0.75], rtol=1e-6)
    np.testing.assert_allclose(g["scale"][4], [1.5, 1.0, 0.75], rtol=1e-6)
    np.testing.assert_allclose(np.linalg.norm(g["rot"], axis=1), 1.0, rtol=1e-6)
    # rot_0 is w: the identity rotation of vertex 6 is (x, y, z, w) = (0, 0, 0, 1)
    np.testing.assert_array_equal(g["rot"][6], [0, 0, 0, 1])
    assert np.all(g["rot"][:, 3] > 0.9)
    np.testing.assert_array_equal(g["pos"][1], [0, 8, 4])
    assert np.all(g["sh"] == 0)


def test_memory_and_file_agree(gs):
    data = open(MODEL_PLY, "rb").read()
    a = gs.read_ply(MODEL_PLY)
    b = gs.read_ply(data)
    assert a.tobytes() == b.tobytes()


def test_round_trip_of_known_records(gs):
    rng = np.random.default_rng(7)
    n = 50
    pos = rng.normal(size=(n, 3)).astype(np.float32)
    q = rng.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    u8 = rng.integers(0, 256, size=(n, 4))
    sh = rng.normal(0, 0.1, size=(n, 15, 3)).astype(np.float32)
    log_scale = rng.normal(-3, 0.5, size=(n, 3))
    rows = []
    for i in range(n):
        dc = [(u8[i, c] / 255.0 - 0.5) / C0 for c in range(3)]
        a = u8[i, 3] / 255.0
        opacity = math.log(a / (1 - a)) if 0 < a < 1 else (math.inf if a >= 1 else -math.inf)
        rest = [sh[i, k, c] for c in range(3) for k in range(15)]  # channel-major
        rows.append(_std_row(pos[i], dc, rest, opacity, log_scale[i], [q[i, 3], q[i, 0], q[i, 1], q[i, 2]]))
    g = gs.read_ply(_ply_bytes(STD_FIELDS, rows))
    np.testing.assert_array_equal(g["pos"], pos)
    np.testing.assert_array_equal(g["color"], u8.astype(np.uint8))
    np.testing.assert_allclose(g["rot"], q, atol=1e-6)
    np.testing.assert_allclose(g["scale"], np.exp(log_scale), rtol=1e-6)
    np.testing.assert_array_equal(g["sh"], sh)


def test_ascii_equals_binary_and_property_order_is_free(gs):
    rng = np.random.default_rng(8)
    rows = [_std_row(rng.normal(size=3), rng.normal(size=3), rng.normal(0, 0.1, 45), rng.normal(),
                     rng.normal(-3, 0.5, 3), rng.normal(size=4)) for _ in range(7)]
    a = gs.read_ply(_ply_bytes(STD_FIELDS, rows))
    b = gs.read_ply(_ply_bytes(STD_FIELDS, rows, fmt="ascii"))
    np.testing.assert_allclose(a["pos"], b["pos"], rtol=1e-6)
    np.testing.assert_allclose(a["sh"], b["sh"], rtol=1e-6, atol=1e-7)
    # shuffled property order, extra properties, double columns
    perm = list(np.random.default_rng(9).permutation(len(STD_FIELDS)))
    fields = [STD_FIELDS[i] for i in perm] + ["extra"]
    rows2 = [[r[i] for i in perm] + [42.0] for r in rows]
    c = gs.read_ply(_ply_bytes(fields, rows2, types={"x": "double", "f_rest_7": "double"}))
    np.testing.assert_allclose(c["pos"], a["pos"], rtol=1e-6)
    np.testing.assert_array_equal(c["color"], a["color"])
    np.testing.assert_allclose(c["sh"], a["sh"], rtol=1e-6, atol=1e-7)


def test_lower_sh_degree_fills_zeros(gs):
    fields = (["x", "y", "z", "f_dc_0", "f_dc_1", "f_dc_2"] + [f"f_rest_{i}" for i in range(9)] +
              ["opacity", "scale_0", "scale_1", "scale_2", "rot_0", "rot_1", "rot_2", "rot_3"])
    rest = list(np.arange(9, dtype=np.float32) + 1)
    row = [0, 0, 0, 0, 0, 0] + rest + [0, 0, 0, 0, 1, 0, 0, 0]
    g = gs.read_ply(_ply_bytes(fields, [row]))
    # degree 1: three coefficients per channel, channel-major
    np.testing.assert_array_equal(g["sh"][0, :3], [[1, 4, 7], [2, 5, 8], [3, 6, 9]])
    assert np.all(g["sh"][0, 3:] == 0)
    assert tuple(g["color"][0]) == (128, 128, 128, 128)  # 0.5 rounds up; sigmoid(0) = 0.5


def test_empty_and_error_cases(gs):
    assert gs.read_ply(_ply_bytes(STD_FIELDS, [])).shape == (0,)
    with pytest.raises(gs.Gs3dError, match="missing vertex property opacity"):
        gs.read_ply(_ply_bytes([f for f in STD_FIELDS if f != "opacity"], [[0.0] * (len(STD_FIELDS) - 1)]))
    full = _ply_bytes(STD_FIELDS, [[0.0] * len(STD_FIELDS)] * 3)
    with pytest.raises(gs.Gs3dError, match="shorter"):
        gs.read_ply(full[:-5])
    with pytest.raises(gs.Gs3dError, match="magic"):
        gs.read_ply(b"plx\n" + full[4:])
    with pytest.raises(gs.Gs3dError, match="unsupported format"):
        gs.read_ply(full.replace(b"binary_little_endian", b"binary_big_endian"))
    with pytest.raises(gs.Gs3dError, match="cannot open"):
        gs.read_ply(os.path.join(HERE, "golden", "does_not_exist.ply"))
