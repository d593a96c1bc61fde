"""The allocation-free sort (mray_radix_sort_u64_u32_noalloc) against the
workspace sort (mray_radix_sort_u64_u32) through the C-ABI: bit-exact keys and
values on two calls in a row, the second smaller (the shared scratch keeps the
first call's contents), plus the size limit and the trivial sizes."""

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(n, seed):
    rng = np.random.default_rng(seed)
    keys = rng.integers(0, 1 << 63, size=n, dtype=np.uint64)
    keys[::3] = keys[0]                     # long runs of equal keys: stability
    keys[1::7] = np.uint64((1 << 63) - 1)
    vals = rng.permutation(n).astype(np.uint32)
    return keys, vals


def _sort_both(torch, keys, vals):
    from mray_amd import native
    dk1 = torch.from_numpy(keys.view(np.int64).copy()).cuda()
    dv1 = torch.from_numpy(vals.view(np.int32).copy()).cuda()
    dk2, dv2 = dk1.clone(), dv1.clone()
    native.radix_sort_u64_u32(dk1, dv1)
    native.radix_sort_u64_u32_noalloc(dk2, dv2)
    torch.cuda.synchronize()
    return (dk1.cpu().numpy().view(np.uint64), dv1.cpu().numpy().view(np.uint32),
            dk2.cpu().numpy().view(np.uint64), dv2.cpu().numpy().view(np.uint32))


def test_noalloc_sort_matches_workspace_sort(torch_gpu):
    torch = torch_gpu
    for n, seed in ((1 << 20, 1), (70001, 2)):      # full size, then a smaller call
        print(f"noalloc sort: n={n} seed={seed}", flush=True)
        keys, vals = _inputs(n, seed)
        k1, v1, k2, v2 = _sort_both(torch, keys, vals)
        assert np.array_equal(k2, k1) and np.array_equal(v2, v1)
        order = np.argsort(keys, kind="stable")
        assert np.array_equal(k2, keys[order]) and np.array_equal(v2, vals[order])


def test_noalloc_sort_limits(torch_gpu):
    from mray_amd import native
    torch = torch_gpu
    L = native.lib()
    # beyond the scratch: rejected on the host before any device work
    assert L.mray_radix_sort_u64_u32_noalloc(None, None, native.NOALLOC_SORT_MAX + 1, None) == 1
    assert b"exceed" in L.mray_last_error()
    # 0 and 1 pairs are already sorted
    assert L.mray_radix_sort_u64_u32_noalloc(None, None, 0, None) == 0
    one_k = torch.tensor([5], dtype=torch.int64, device="cuda")
    one_v = torch.tensor([9], dtype=torch.int32, device="cuda")
    native.radix_sort_u64_u32_noalloc(one_k, one_v)
    assert int(one_k.item()) == 5 and int(one_v.item()) == 9
    # a tile-edge size
    keys, vals = _inputs(4097, 3)
    k1, v1, k2, v2 = _sort_both(torch, keys, vals)
    assert np.array_equal(k2, k1) and np.array_equal(v2, v1)
