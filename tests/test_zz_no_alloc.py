"""tpd_sort_pairs_noalloc (the allocation-free sort: fixed per-GPU scratch, no hipMalloc) against
tpd_sort_pairs on the same inputs, bit for bit, through the C ABI: two calls in a row, the second on
fewer pairs than the first (the scratch keeps the last call's contents), under an 8-bit and a 9-bit
radix shape; sizes beyond the scratch limit are refused."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

LIMIT = 1 << 22


@pytest.fixture(scope="module")
def eng():
    from torpedo_amd.engine import GaussianEngine
    e = GaussianEngine()
    yield e
    e.close()


def tile_depth_keys(rng, n, tiles=8160):
    t = rng.integers(0, tiles, n, dtype=np.uint64)
    depth = rng.uniform(0.01, 100.0, n).astype(np.float32).view(np.uint32).astype(np.uint64)
    keys = (t << np.uint64(32)) | depth
    keys[: n // 5] = keys[n // 2]  # a long run of one key: stability shows
    return keys


def run_sort(fn, eng, keys, vals, key_bits):
    import torch
    kt = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    vt = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    ka = torch.empty_like(kt)
    va = torch.empty_like(vt)
    torch.cuda.synchronize()
    in_alt = fn(eng, kt.data_ptr(), vt.data_ptr(), ka.data_ptr(), va.data_ptr(), keys.size, key_bits,
                torch.cuda.current_stream().cuda_stream)
    torch.cuda.synchronize()
    return (ka if in_alt else kt).cpu().numpy().view(np.uint64), (va if in_alt else vt).cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("shape", [4, 6])
def test_noalloc_sort_equals_sort(eng, shape):
    from torpedo_amd.engine import sort_pairs_device, sort_pairs_device_noalloc
    eng.setSortShape(shape)
    try:
        rng = np.random.default_rng(shape)
        for n, bits in ((3_000_001, 45), (123_457, 64), (LIMIT, 45)):
            print(f"[noalloc] shape {shape}: n={n} key bits {bits}", flush=True)
            keys = tile_depth_keys(rng, n) if bits < 64 else rng.integers(0, 2 ** 63, n, dtype=np.uint64)
            vals = rng.permutation(n).astype(np.uint32)
            k0, v0 = run_sort(sort_pairs_device, eng, keys, vals, bits)
            k1, v1 = run_sort(sort_pairs_device_noalloc, eng, keys, vals, bits)
            assert np.array_equal(k1, k0) and np.array_equal(v1, v0), (shape, n)
            if n < LIMIT:
                o = np.argsort(keys, kind="stable")
                assert np.array_equal(k1, keys[o]) and np.array_equal(v1, vals[o]), (shape, n)
    finally:
        eng.setSortShape(0)


def test_noalloc_sort_limits(eng):
    from torpedo_amd._capi import TpdError
    from torpedo_amd.engine import sort_pairs_device_noalloc
    print("[noalloc] limits", flush=True)
    k, v = run_sort(sort_pairs_device_noalloc, eng, np.zeros(0, np.uint64), np.zeros(0, np.uint32), 32)
    assert k.size == 0
    with pytest.raises(TpdError, match="CAPACITY"):
        sort_pairs_device_noalloc(eng, 0, 0, 0, 0, LIMIT + 1, 45)
    with pytest.raises(TpdError, match="INVALID"):
        sort_pairs_device_noalloc(eng, 0, 0, 0, 0, 16, 0)
