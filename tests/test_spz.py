"""SPZ ingestion (gs3d_spz_read_file / gs3d_spz_read_memory), host only.

The reference reads SPZ through wgpu-3dgs-core's GaussiansSource::Spz
(examples/simple.rs:157-160); core is not in the reference tree and the
reference holds no SPZ file, so the reader is "parity unpinned": these tests
pin it against files packed here from the published layout (header, then
positions, alphas, colours, scales, rotations, SH, gzip-compressed) and
check that each field decodes to the value it encodes within its
quantisation step.
"""
import gzip
import math
import os
import struct

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
C0 = 0.28209479177387814


def _pack_spz(pos, alpha_u8, dc, log_scale, quat_xyzw, sh, version=2, frac_bits=12, sh_degree=3):
    n = pos.shape[0]
    dim = {0: 0, 1: 3, 2: 8, 3: 15}[sh_degree]
    out = bytearray(struct.pack("<IIIBBBB", 0x5053474E, version, n, sh_degree, frac_bits, 0, 0))
    fixed = np.round(pos * (1 << frac_bits)).astype(np.int64)
    for i in range(n):
        for c in range(3):
            out += int(fixed[i, c] & 0xFFFFFF).to_bytes(3, "little")
    out += bytes(alpha_u8.astype(np.uint8))
    out += bytes(np.clip(np.round(dc * 0.15 * 255 + 127.5), 0, 255).astype(np.uint8).reshape(-1))
    out += bytes(np.clip(np.round((log_scale + 10) * 16), 0, 255).astype(np.uint8).reshape(-1))
    q = quat_xyzw / np.linalg.norm(quat_xyzw, axis=1, keepdims=True)
    if version == 2:
        q = q * np.where(q[:, 3:4] < 0, -1, 1)
        out += bytes(np.clip(np.round((q[:, :3] + 1) * 127.5), 0, 255).astype(np.uint8).reshape(-1))
    else:
        for i in range(n):
            largest = int(np.argmax(np.abs(q[i])))
            qi = q[i] * (-1 if q[i, largest] < 0 else 1)
            comp = 0
            for c in range(4):
                if c == largest:
                    continue
                mag = int(round(abs(qi[c]) / math.sqrt(0.5) * 511))
                comp = (comp << 10) | (mag | ((1 << 9) if qi[c] < 0 else 0))
            comp |= largest << 30
            out += comp.to_bytes(4, "little")
    out += bytes(np.clip(np.round(sh[:, :dim, :] * 128 + 128), 0, 255).astype(np.uint8).reshape(-1))
    return gzip.compress(bytes(out))


def _records(n, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-50, 50, size=(n, 3))
    alpha = rng.integers(0, 256, size=n)
    dc = rng.uniform(-1.5, 1.5, size=(n, 3))
    log_scale = rng.uniform(-9, 5, size=(n, 3))
    q = rng.normal(size=(n, 4))
    sh = rng.uniform(-0.9, 0.9, size=(n, 15, 3))
    return pos, alpha, dc, log_scale, q, sh


def _same_rotation(a, b, atol):
    # q and -q are one rotation
    d = np.minimum(np.abs(a - b).max(axis=1), np.abs(a + b).max(axis=1))
    assert d.max() <= atol, d.max()


@pytest.mark.parametrize("version", [2, 3])
def test_spz_fields_decode_to_what_they_encode(gs, version):
    n = 200
    pos, alpha, dc, log_scale, q, sh = _records(n, version)
    g = gs.read_spz(_pack_spz(pos, alpha, dc, log_scale, q, sh, version=version))
    assert g.shape == (n,)
    np.testing.assert_allclose(g["pos"], pos, atol=0.5 / 4096 + 1e-5)       # 12 fractional bits
    assert np.array_equal(g["color"][:, 3], alpha.astype(np.uint8))        # alpha stored as sigmoid * 255
    # colour: dc quantised in steps of 1 / (0.15 * 255), then rgb = 0.5 + C0 dc -> u8
    want = np.clip(np.floor((0.5 + C0 * dc) * 255 + 0.5), 0, 255)
    assert np.abs(g["color"][:, :3].astype(int) - want).max() <= 2
    np.testing.assert_allclose(np.log(g["scale"]), log_scale, atol=0.5 / 16 + 1e-5)
    qn = q / np.linalg.norm(q, axis=1, keepdims=True)
    if version == 2:
        # x, y, z quantised in steps of 1 / 127.5 with w >= 0 rebuilt from
        # them (w's error grows as w -> 0, as the format is)
        qs = qn * np.where(qn[:, 3:4] < 0, -1, 1)
        np.testing.assert_allclose(g["rot"][:, :3], qs[:, :3], atol=0.5 / 127.5 + 2e-3)
        assert np.all(g["rot"][:, 3] >= 0)
    else:
        _same_rotation(g["rot"], qn, atol=0.003)
    np.testing.assert_allclose(np.linalg.norm(g["rot"], axis=1), 1.0, atol=1e-6)
    np.testing.assert_allclose(g["sh"], np.round(sh * 128) / 128, atol=1e-6)


def test_spz_lower_degree_and_file_path(gs, tmp_path):
    pos, alpha, dc, log_scale, q, sh = _records(10, 7)
    data = _pack_spz(pos, alpha, dc, log_scale, q, sh, sh_degree=1)
    g = gs.read_spz(data)
    assert np.all(g["sh"][:, 3:] == 0)
    np.testing.assert_allclose(g["sh"][:, :3], np.round(sh[:, :3] * 128) / 128, atol=1e-6)
    f = tmp_path / "m.spz"
    f.write_bytes(data)
    assert gs.read_spz(str(f)).tobytes() == g.tobytes()
    # the examples' loader: PLY first, then SPZ
    assert gs.read_gaussians(str(f)).tobytes() == g.tobytes()
    ply = os.path.join(HERE, "golden", "model.ply")
    assert gs.read_gaussians(ply).tobytes() == gs.read_ply(ply).tobytes()


def test_spz_errors(gs):
    pos, alpha, dc, log_scale, q, sh = _records(5, 8)
    good = _pack_spz(pos, alpha, dc, log_scale, q, sh)
    with pytest.raises(gs.Gs3dError, match="not a gzip stream"):
        gs.read_spz(b"ply\n")
    raw = gzip.decompress(good)
    with pytest.raises(gs.Gs3dError, match="bad magic"):
        gs.read_spz(gzip.compress(b"XXXX" + raw[4:]))
    with pytest.raises(gs.Gs3dError, match="shorter than its point count"):
        gs.read_spz(gzip.compress(raw[:-10]))
    with pytest.raises(gs.Gs3dError, match="unsupported version"):
        gs.read_spz(gzip.compress(raw[:4] + struct.pack("<I", 9) + raw[8:]))
    assert gs.read_spz(_pack_spz(pos[:0], alpha[:0], dc[:0], log_scale[:0], q[:0], sh[:0])).shape == (0,)


def test_spz_rejects_fractional_bits_and_bounds_inflation(gs):
    pos, alpha, dc, log_scale, q, sh = _records(5, 9)
    raw = gzip.decompress(_pack_spz(pos, alpha, dc, log_scale, q, sh))
    for bits in (25, 32, 255):
        bad = raw[:13] + bytes([bits]) + raw[14:]
        with pytest.raises(gs.Gs3dError, match="fractional bits"):
            gs.read_spz(gzip.compress(bad))
    # a stream far longer than its point count needs (a gzip bomb behind a
    # small header) is read up to what the header's count needs
    bomb = gzip.compress(raw + bytes(64 << 20), compresslevel=1)
    assert len(bomb) < (1 << 20)
    g = gs.read_spz(bomb)
    assert g.tobytes() == gs.read_spz(gzip.compress(raw)).tobytes()
    # and a truncated stream is an error, not a partial read
    with pytest.raises(gs.Gs3dError):
        gs.read_spz(gzip.compress(raw)[:-12])
