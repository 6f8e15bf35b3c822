"""ipf_branch_order_noalloc: the branch sort with its histograms in the
library's fixed device scratch array instead of a hipMalloc'd buffer.

Bar: the order equals ipf_branch_order's (an integer permutation: bit for
bit), on two calls in a row through the C ABI, the second on a context with
fewer points (the scratch keeps the first call's histograms: nothing of
them may leak into the second), then once more on the first context; the
device's free memory is the same before and after each no-alloc call.
"""
import ctypes

import numpy as np
import pytest

import intpointsfem_amd as ipf
from intpointsfem_amd import wellbore as wb

pytestmark = pytest.mark.gpu


def _ctx(mesh):
    m = wb.make_mesh(mesh)
    sm = ipf.StructMatrix(m, device=0)
    return m, sm, sm.context()


def _keys(n, seed, gpu):
    import torch
    rng = np.random.default_rng(seed)
    k = rng.integers(0, 5, n).astype(np.int32)
    k[rng.random(n) < 0.01] = 7  # out-of-range keys sort last
    k[: n // 3] = 1  # a long run of one key
    return torch.as_tensor(k, device=gpu)




def test_branch_order_noalloc_matches_branch_order(gpu):
    assert ipf.lib().ipf_branch_order_noalloc_max_points() == 53_687_040
    _, sm_big, big = _ctx("40x40-q2")
    _, sm_small, small = _ctx((12, 10, 2, 1.1))
    n_big, n_small = big.info()["npts"], small.info()["npts"]
    assert n_small < n_big
    keys_big = _keys(n_big, 1, gpu)
    keys_small = _keys(n_small, 2, gpu)
    # the no-alloc entry reads the keys given by pointer: pass them through
    # the C ABI directly (branch_dev, order_dev)
    for ctx, keys, n in ((big, keys_big, n_big), (small, keys_small, n_small), (big, keys_big, n_big)):
        import torch
        order_na = torch.full((n,), -1, dtype=torch.int32, device=gpu)
        order_ref = torch.full((n,), -2, dtype=torch.int32, device=gpu)
        # the allocating entry point first: it loads the sort kernels (the
        # runtime maps a kernel's code on its first launch) and allocates its
        # own histograms, outside the window measured below
        ctx.branch_order(keys, order_ref)
        ctx.synchronize()
        free0 = torch.cuda.mem_get_info()[0]
        rc = ipf.lib().ipf_branch_order_noalloc(ctx.handle, ctypes.c_void_p(keys.data_ptr()),
                                                ctypes.c_void_p(order_na.data_ptr()))
        assert rc == 0, ipf.lib().ipf_last_error(ctx.handle)
        ctx.synchronize()
        free1 = torch.cuda.mem_get_info()[0]
        a, b = order_na.cpu().numpy(), order_ref.cpu().numpy()
        print(f"{n} points: orders equal {np.array_equal(a, b)}, free memory change {free1 - free0}", flush=True)
        assert np.array_equal(a, b)
        assert np.array_equal(np.sort(a), np.arange(n))
        k = keys.cpu().numpy()
        kk = np.where((k >= 0) & (k < 5), k, 4)
        assert np.all(np.diff(kk[a]) >= 0)
        assert free1 == free0
