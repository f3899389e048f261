"""Stand-alone radix sort on the GPU (tpd_sort_pairs) against numpy's stable argsort: bit-exact
keys and values, stability under heavy key duplication, ragged sizes around the 4096-key
partition, and the key widths of the render path."""
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def eng():
    from torpedo_amd.engine import GaussianEngine
    e = GaussianEngine()
    yield e
    e.close()


def gpu_sort(eng, keys, vals, key_bits):
    import torch
    from torpedo_amd.engine import sort_pairs_device
    n = keys.size
    kt = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    vt = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    ka = torch.empty_like(kt)
    va = torch.empty_like(vt)
    torch.cuda.synchronize()
    in_alt = sort_pairs_device(eng, kt.data_ptr(), vt.data_ptr(), ka.data_ptr(), va.data_ptr(), n, key_bits,
                               torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    k = (ka if in_alt else kt).cpu().numpy().view(np.uint64)
    v = (va if in_alt else vt).cpu().numpy().view(np.uint32)
    return k, v


CASES = [
    (1, 8, "uniform"), (2, 16, "uniform"), (4095, 45, "uniform"), (4096, 45, "uniform"), (4097, 45, "dup"),
    (100_000, 44, "dup"), (1_000_003, 45, "tiles"), (3_000_000, 64, "uniform"), (5_000_000, 45, "tiles"),
]


@pytest.mark.parametrize("n,bits,kind", CASES)
def test_sort_matches_stable_argsort(eng, n, bits, kind):
    print(f"[sort] n={n} bits={bits} kind={kind}", flush=True)
    rng = np.random.default_rng(n + bits)
    mask = np.uint64((1 << bits) - 1) if bits < 64 else np.uint64(0xFFFFFFFFFFFFFFFF)
    if kind == "uniform":
        keys = rng.integers(0, 2 ** 63, n, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, n, dtype=np.uint64)
        keys &= mask
    elif kind == "dup":
        keys = rng.integers(0, 17, n, dtype=np.uint64) << np.uint64(20)
    else:  # tile|depth keys as the renderer makes them
        tiles = rng.integers(0, 8160, n, dtype=np.uint64)
        depth = rng.uniform(0.01, 100.0, n).astype(np.float32).view(np.uint32).astype(np.uint64)
        keys = (tiles << np.uint64(32)) | depth
    vals = rng.permutation(n).astype(np.uint32)
    k, v = gpu_sort(eng, keys, vals, bits)
    o = np.argsort(keys, kind="stable")
    assert np.array_equal(k, keys[o])
    assert np.array_equal(v, vals[o])


@pytest.mark.parametrize("n,bits", [(70_001, 27), (300_000, 31), (4097, 17)])
def test_sort_carries_bits_above_key_bits(eng, n, bits):
    """Bits at and above key_bits ride along with the key without ordering it (the batched frame's
    item sort carries each item's emission record there): the result equals a stable argsort of
    the low key_bits bits, with the high bits intact."""
    rng = np.random.default_rng(n)
    low = rng.integers(0, 1 << bits, n, dtype=np.uint64) >> np.uint64(rng.integers(0, 6))  # duplicates too
    keys = low | (rng.integers(1, 1 << (64 - bits), n, dtype=np.uint64) << np.uint64(bits))
    vals = rng.permutation(n).astype(np.uint32)
    k, v = gpu_sort(eng, keys, vals, bits)
    o = np.argsort(low, kind="stable")
    assert np.array_equal(k, keys[o])
    assert np.array_equal(v, vals[o])


def test_sort_empty_and_invalid(eng):
    from torpedo_amd._capi import TpdError
    k, v = gpu_sort(eng, np.zeros(0, np.uint64), np.zeros(0, np.uint32), 32)
    assert k.size == 0
    with pytest.raises(TpdError):
        gpu_sort(eng, np.zeros(4, np.uint64), np.zeros(4, np.uint32), 0)


NUM_SHAPES = 12  # 8-bit digits x 7 block shapes, 9-bit digits x 5


def test_shape_count(eng):
    from torpedo_amd._capi import TpdError
    eng.setSortShape(NUM_SHAPES - 1)
    with pytest.raises(TpdError):
        eng.setSortShape(NUM_SHAPES)
    eng.setSortShape(0)


@pytest.mark.parametrize("shape", range(NUM_SHAPES))
def test_every_block_shape(eng, shape):
    print(f"[sort] shape={shape}", flush=True)
    eng.setSortShape(shape)
    try:
        rng = np.random.default_rng(100 + shape)
        for n in (3, 2047, 2049, 8191, 12289, 700_001):
            tiles = rng.integers(0, 3600, n, dtype=np.uint64)
            depth = rng.uniform(0.01, 100.0, n).astype(np.float32).view(np.uint32).astype(np.uint64)
            keys = (tiles << np.uint64(32)) | depth
            keys[: n // 3] = keys[0]  # a long run of one key
            vals = np.arange(n, dtype=np.uint32)
            k, v = gpu_sort(eng, keys, vals, 44)
            o = np.argsort(keys, kind="stable")
            assert np.array_equal(k, keys[o]) and np.array_equal(v, vals[o]), (shape, n)
        # full 64-bit keys: the most passes any digit width needs
        keys = rng.integers(0, 2 ** 63, 100_003, dtype=np.uint64) * np.uint64(2) + rng.integers(0, 2, 100_003,
                                                                                                   dtype=np.uint64)
        keys[::7] = keys[3]
        vals = np.arange(keys.size, dtype=np.uint32)
        k, v = gpu_sort(eng, keys, vals, 64)
        o = np.argsort(keys, kind="stable")
        assert np.array_equal(k, keys[o]) and np.array_equal(v, vals[o]), (shape, 64)
    finally:
        eng.setSortShape(0)
