"""mmo_topk_dev_noalloc, the top-k entry point that allocates nothing (its scratch is a
per-GPU static device array), against mmo_topk_dev through the C ABI: bit-identical
energies and indices on two calls in a row, the second with a smaller input (the
scratch keeps the first call's contents), ties and NaN included; sizes beyond the
scratch limit are rejected with an error before any launch."""
import numpy as np
import pytest

from mmo_amd import ff
from mmo_amd._lib import ERR_BAD_ARG, MMOError

pytestmark = pytest.mark.gpu

MAX_N = 32768 * 4096


def _energies(n, seed):
    import torch
    rng = np.random.default_rng(seed)
    e = rng.normal(0.0, 50.0, n).round(1) + 0.0     # rounding gives many ties; no -0.0
    e[rng.integers(0, n, max(1, n // 1000))] = np.nan
    e[rng.integers(0, n, 3)] = np.inf
    return torch.tensor(e, device="cuda:0")


def _both(ctx, e, k):
    import torch
    v0, i0 = ctx.topk(e, k)
    kk = min(k, e.numel())
    v1 = torch.full((kk,), -1.0, dtype=torch.float64, device=e.device)
    i1 = torch.full((kk,), -1, dtype=torch.int64, device=e.device)
    ctx.topk_noalloc(e, k, v1, i1)
    torch.cuda.synchronize()
    return v0.cpu().numpy(), i0.cpu().numpy(), v1.cpu().numpy(), i1.cpu().numpy()


def test_noalloc_topk_matches_the_allocating_entry_point():
    ctx = ff.Context(0)
    cases = [(36_864_000, 100, 1), (5_000, 2048, 2), (1_000_003, 7, 3), (10, 64, 4)]
    for n, k, seed in cases:       # each call smaller than the one before it
        print("no-alloc top-k n=%d k=%d" % (n, k), flush=True)
        e = _energies(n, seed)
        v0, i0, v1, i1 = _both(ctx, e, k)
        assert np.array_equal(i0, i1), (n, k)
        assert np.array_equal(v0, v1, equal_nan=True), (n, k)
        ref = np.lexsort((np.arange(n), np.nan_to_num(e.cpu().numpy(), nan=np.inf)))
        ref = ref[np.argsort(np.isnan(e.cpu().numpy()[ref]), kind="stable")][:min(k, n)]
        assert np.array_equal(i1, ref), (n, k)


def test_noalloc_topk_rejects_sizes_beyond_its_scratch():
    import torch
    ctx = ff.Context(0)
    e = torch.zeros(4, dtype=torch.float64, device="cuda:0")
    v = torch.empty(4, dtype=torch.float64, device="cuda:0")
    i = torch.empty(4, dtype=torch.int64, device="cuda:0")
    from mmo_amd._lib import lib
    import ctypes as ct
    rc = lib.mmo_topk_dev_noalloc(ctx.h, ct.c_void_p(e.data_ptr()), ct.c_int64(MAX_N + 1), 4,
                                  ct.c_void_p(v.data_ptr()), ct.c_void_p(i.data_ptr()), None)
    assert rc == ERR_BAD_ARG
    assert b"scratch limit" in lib.mmo_last_error(ctx.h)
    with pytest.raises(MMOError):
        ctx.topk_noalloc(e, 0, v, i)
