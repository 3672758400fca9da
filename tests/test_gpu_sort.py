"""GPU radix sort of ray keys (hiprt_sort_keys, the bounce-queue sort).

The path tracer's images do not depend on its queue order, so they only show
that the sort is a permutation; these tests pin the order itself: the
permutation equals numpy's stable argsort of the keys' low `bits` bits for
every size class of the tiled sort (empty, below one 64-item round, one tile
of 4096 and one item either side, several tiles, tile counts that are and are
not multiples of the 8 XCDs), every pass count (1..4 eight-bit passes) and
key distributions with many ties (constant, descending, clustered runs as the
bounce queues hold, uniform).
"""
import numpy as np
import pytest

import hiprt

pytestmark = pytest.mark.gpu


def _keys(n, bits, kind, seed):
    rng = np.random.default_rng(seed)
    mask = np.uint64((1 << bits) - 1)
    if kind == "uniform":
        k = rng.integers(0, 1 << 32, n, dtype=np.uint64)
    elif kind == "constant":
        k = np.full(n, rng.integers(0, 1 << 32), dtype=np.uint64)
    elif kind == "descending":
        k = np.arange(n, 0, -1, dtype=np.uint64) * np.uint64(2654435761)
    else:   # clustered: runs of a slowly changing key with low-bit noise, high bits above `bits` random
        run = rng.integers(0, 1 << 32, max(n // 64 + 1, 1), dtype=np.uint64)
        k = run[np.arange(n) // 64] ^ rng.integers(0, 64, n, dtype=np.uint64)
        k ^= rng.integers(0, 1 << 8, n, dtype=np.uint64) << np.uint64(24)
    k &= np.uint64(0xFFFFFFFF)
    return k.astype(np.uint32), (k & mask).astype(np.uint32)


def _sort(torch, keys, bits):
    d = torch.from_numpy(keys.view(np.int32)).cuda()
    perm = hiprt.sort_keys(d, bits)
    torch.cuda.synchronize()
    assert np.array_equal(d.cpu().numpy().view(np.uint32), keys)   # the wrapper leaves its input alone
    return perm.cpu().numpy().view(np.uint32)


@pytest.mark.parametrize("n", [0, 1, 63, 2049, 4095, 4096, 4097, 8 * 4096, 9 * 4096 + 5, 100_003])
@pytest.mark.parametrize("bits,kind", [(16, "clustered"), (8, "uniform"), (19, "descending"), (32, "uniform"),
                                       (1, "uniform"), (13, "constant"), (24, "clustered")])
def test_sort_equals_stable_argsort(torch_cuda, hip_ctx, n, bits, kind):
    print(f"case: sort n={n} bits={bits} keys={kind}", flush=True)
    keys, low = _keys(n, bits, kind, seed=n * 31 + bits)
    got = _sort(torch_cuda, keys, bits)
    want = np.argsort(low, kind="stable").astype(np.uint32)
    assert np.array_equal(got, want)


def test_sort_queue_sized(torch_cuda, hip_ctx):
    """A bounce queue's size (3 M rays, 16-bit keys, two passes): the
    permutation is the stable order and the keys it visits are
    non-decreasing."""
    n = 3_000_000
    print(f"case: sort n={n} bits=16 clustered", flush=True)
    keys, low = _keys(n, 16, "clustered", seed=7)
    got = _sort(torch_cuda, keys, 16)
    assert np.array_equal(got, np.argsort(low, kind="stable").astype(np.uint32))
    assert (np.diff(low[got].astype(np.int64)) >= 0).all()


def test_sort_rejects_bad_bits(torch_cuda, hip_ctx):
    d = torch_cuda.zeros(16, dtype=torch_cuda.int32, device="cuda")
    for bits in (0, 33):
        with pytest.raises(hiprt.HiprtError):
            hiprt.sort_keys(d, bits)
