"""Measurement entry points (tpd_measure_copy, tpd_measure_bandwidth, tpd_measure_run_scatter): each
runs on small sizes and returns a positive time within what HBM allows, and bad arguments are
rejected before any GPU work."""
import pytest

pytestmark = pytest.mark.gpu

PEAK_GBS = 8000.0


@pytest.fixture(scope="module")
def eng():
    from torpedo_amd.engine import GaussianEngine
    e = GaussianEngine()
    yield e
    e.close()


def test_streams_and_gathers(eng):
    for kind in eng.BANDWIDTH_KINDS:
        gbs = eng.measureBandwidth(kind, 64 << 20, iters=3)
        print(f"[measure] {kind}: {gbs:.0f} GB/s", flush=True)
        # 64 MiB fits the last-level cache, so the rate may exceed the HBM peak, but not tenfold
        assert 0.0 < gbs < 10 * PEAK_GBS
    for kind in eng.GATHER_KINDS:
        rate = eng.measureGather(kind, 32 << 20, iters=2)
        print(f"[measure] {kind}: {rate:.1f} G reads/s", flush=True)
        assert rate > 0.0
    assert eng.measureCopy(64 << 20, iters=3) > 0.0


@pytest.mark.parametrize("key_bytes,bins,keys_per_block,blocks_per_cu", [
    (4, 512, 16384, 0), (4, 512, 8192, 2), (8, 512, 12288, 0), (8, 256, 6144, 1), (4, 256, 4096, 4)])
def test_run_scatter(eng, key_bytes, bins, keys_per_block, blocks_per_cu):
    n = keys_per_block * 300
    print(f"[measure] run scatter: {key_bytes}-byte keys, {bins} digits, {keys_per_block} keys per block, "
          f"{blocks_per_cu} blocks/CU", flush=True)
    gbs = eng.measureRunScatter(key_bytes, n, bins, keys_per_block, blocks_per_cu, iters=3)
    assert 0.0 < gbs < 10 * PEAK_GBS


@pytest.mark.parametrize("args", [
    (2, 512 * 16 * 4, 512, 8192),       # key width other than 4 or 8
    (4, 512 * 16 * 4, 128, 8192),       # digit count other than 256 or 512
    (4, 8192 * 3 + 1, 512, 8192),       # n not a multiple of the block
    (4, 1000 * 3, 512, 1000),           # block not a multiple of 512
    (4, 32768 * 2, 512, 32768),         # more than 32 keys per thread of 512
    (4, 0, 512, 8192),                  # no keys
])
def test_run_scatter_rejects(eng, args):
    from torpedo_amd._capi import TpdError
    with pytest.raises(TpdError):
        eng.measureRunScatter(*args)
