"""Standalone RadixSorter (src/radix_sorter.rs:9-96) on the GPU vs a stable
sort by key (numpy.argsort(kind="stable")), including ragged sizes around the
default 5120-key tile, heavy ties and extreme keys; and the same for the
2048-, 3072-, 4096- and 6144-key tiles (8, 12, 16 and 24 keys per thread) the
batch autotuner may pick."""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _check(gs, device, keys, vals=None, items=None):
    n = keys.shape[0]
    if vals is None:
        vals = np.arange(n, dtype=np.uint32)
    s = gs.RadixSorter(device, max(n, 1))
    if items is not None:
        s.set_items(items)
    s.write(keys, vals)
    s.sort()
    k, v = s.read()
    order = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[order])
    assert np.array_equal(v, vals[order])
    s.close()


@pytest.mark.parametrize("n", [0, 1, 2, 255, 4095, 4096, 4097, 5119, 5120, 5121, 65537, 1_000_003, 8_000_000,
                               12_000_000])
def test_sort_random_sizes(gs, device, n):
    print(f"[case] sort random n={n}", flush=True)
    rng = np.random.default_rng(n)
    _check(gs, device, rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32))


@pytest.mark.parametrize("kind", ["all_equal", "few_distinct", "sorted", "reversed", "extremes", "depth_like"])
def test_sort_key_patterns(gs, device, kind):
    print(f"[case] sort pattern {kind}", flush=True)
    n = 300_001
    rng = np.random.default_rng(7)
    if kind == "all_equal":
        keys = np.full(n, 0x3D000000, np.uint32)
    elif kind == "few_distinct":
        keys = rng.choice(np.array([0, 7, 0x3F800000, 0xFFFFFFFF], np.uint32), n)
    elif kind == "sorted":
        keys = np.sort(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32))
    elif kind == "reversed":
        keys = np.sort(rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32))[::-1].copy()
    elif kind == "extremes":
        keys = rng.choice(np.array([0, 0xFFFFFFFF], np.uint32), n)
    else:
        # f32 depth keys 1 - ndc.z near 0.03: thousands of equal keys
        z = rng.uniform(2.0, 4.0, n).astype(np.float32)
        keys = (np.float32(1.0) - (np.float32(1.0) - np.float32(0.1) / z)).view(np.uint32)
    vals = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
    _check(gs, device, keys, vals)


def test_sort_repeated_calls_reuse_state(gs, device):
    print("[case] sort repeated", flush=True)
    rng = np.random.default_rng(3)
    s = gs.RadixSorter(device, 200_000)
    for n in (200_000, 17, 123_457, 0, 4096):
        keys = rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        vals = np.arange(n, dtype=np.uint32)
        s.write(keys, vals)
        s.sort()
        k, v = s.read()
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k, keys[order]) and np.array_equal(v, vals[order])
    s.close()


@pytest.mark.parametrize("items", [8, 12, 16, 24])
@pytest.mark.parametrize("n", [0, 1, 2047, 2048, 2049, 3071, 3072, 3073, 4097, 6143, 6144, 6145, 1_000_003,
                               12_000_000])
def test_sort_items_per_thread_random_sizes(gs, device, items, n):
    print(f"[case] sort items={items} random n={n}", flush=True)
    rng = np.random.default_rng(n + items)
    _check(gs, device, rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32), items=items)


@pytest.mark.parametrize("items", [8, 12, 16, 24])
def test_sort_items_per_thread_ties(gs, device, items):
    print(f"[case] sort items={items} ties", flush=True)
    rng = np.random.default_rng(11)
    n = 300_001
    keys = rng.choice(np.array([0, 7, 0x3F800000, 0xFFFFFFFF], np.uint32), n)
    _check(gs, device, keys, rng.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32), items=items)
    z = rng.uniform(2.0, 4.0, n).astype(np.float32)
    keys = (np.float32(1.0) - (np.float32(1.0) - np.float32(0.1) / z)).view(np.uint32)
    _check(gs, device, keys, items=items)


def test_sort_items_per_thread_rejects_other_values(gs, device):
    s = gs.RadixSorter(device, 16)
    for bad in (0, 4, 10, 28, 32):
        with pytest.raises(gs.Gs3dError):
            s.set_items(bad)
    s.close()
