"""The library's radix sort (skw_sort_pairs_u32) against numpy's stable sort on the same keys:
keys equal on the sorted bits keep their input order, bits outside the range are ignored, and
the inputs are left unchanged."""
import numpy as np
import pytest

import skwarka_amd as sk
from tests.conftest import from_dev

pytestmark = pytest.mark.gpu


def sort_both(torch, keys, vals, begin, end):
    n = len(keys)
    d_ki = torch.from_numpy(keys.view(np.int32).copy()).to("cuda")
    d_vi = torch.from_numpy(vals.view(np.int32).copy()).to("cuda")
    d_ko = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    d_vo = torch.empty(max(n, 1), dtype=torch.int32, device="cuda")
    sk.sort_pairs_u32(d_ki.data_ptr(), d_ko.data_ptr(), d_vi.data_ptr(), d_vo.data_ptr(), n, begin, end)
    torch.cuda.synchronize()
    mask = np.uint32(((1 << (end - begin)) - 1) << begin) if end > begin else np.uint32(0)
    order = np.argsort(keys & mask, kind="stable")
    assert np.array_equal(d_ki.cpu().numpy().view(np.uint32), keys), "input keys were written"
    assert np.array_equal(d_vi.cpu().numpy().view(np.uint32), vals), "input values were written"
    got_k = d_ko.cpu().numpy().view(np.uint32)[:n]
    got_v = d_vo.cpu().numpy().view(np.uint32)[:n]
    assert np.array_equal(got_k, keys[order])
    assert np.array_equal(got_v, vals[order])


@pytest.mark.parametrize("n", [1, 63, 4096, 4097, 70001, 1 << 20])
@pytest.mark.parametrize("bits", [(0, 32), (12, 30), (0, 21), (3, 4), (7, 7)])
def test_sort_matches_stable_argsort(torch_cuda, n, bits):
    print("case: radix sort n=%d bits=%s" % (n, bits), flush=True)
    rng = np.random.default_rng(n + 31 * bits[0] + bits[1])
    keys = rng.integers(0, 1 << 32, n, dtype=np.uint64).astype(np.uint32)
    # many keys equal on the sorted bits, and the all-ones key (a missed lookup's key)
    keys[rng.random(n) < 0.3] = 0xFFFFFFFF
    keys[rng.random(n) < 0.2] = keys[0]
    vals = np.arange(n, dtype=np.uint32) ^ np.uint32(0x5A5A5A5A)
    sort_both(torch_cuda, keys, vals, *bits)


def test_sort_few_distinct_keys_large(torch_cuda):
    # 5M pairs over 7 distinct keys: long equal runs crossing many tiles
    print("case: radix sort n=5000000, 7 distinct keys", flush=True)
    rng = np.random.default_rng(5)
    keys = rng.integers(0, 7, 5_000_000).astype(np.uint32) * np.uint32(0x01010101)
    vals = rng.integers(0, 1 << 32, len(keys), dtype=np.uint64).astype(np.uint32)
    sort_both(torch_cuda, keys, vals, 0, 32)


def test_sort_empty_and_bad_arguments(torch_cuda):
    print("case: radix sort n=0 and argument checks", flush=True)
    sk.sort_pairs_u32(0, 0, 0, 0, 0, 0, 32)
    with pytest.raises(sk.SkwError):
        sk.sort_pairs_u32(0, 0, 0, 0, 10, 0, 32)
    d = torch_cuda.empty(16, dtype=torch_cuda.int32, device="cuda")
    for b, e in ((-1, 8), (0, 33), (9, 8)):
        with pytest.raises(sk.SkwError):
            sk.sort_pairs_u32(d.data_ptr(), d.data_ptr(), d.data_ptr(), d.data_ptr(), 4, b, e)
