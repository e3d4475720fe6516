"""The allocation-free order statistics (mx_order_stats_static, needs an
MI355X): called through the C-ABI, its outputs equal those of
mx_order_stats with a workspace of the caller's on the same photons, on
two calls in a row, the second with fewer photons (the static workspace
keeps the first call's contents), and a call beyond its size limit is
refused with an error before anything runs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _photons(n, seed, dev):
    import torch
    rng = np.random.default_rng(seed)
    orders = np.arange(-12, 3)
    order = rng.choice(np.append(orders, -2 ** 31), size=n).astype(np.int32)
    value = rng.normal(0, 1, n) + 0.1 * order
    value[rng.random(n) < 0.01] = np.nan
    prob = rng.random(n)
    return (torch.as_tensor(order, device=dev), torch.as_tensor(value, device=dev),
            torch.as_tensor(prob, device=dev), orders)


def _call(fn, order, value, prob, orders, dev, workspace=True):
    import torch
    from marxs_amd import _lib
    n, no = order.numel(), len(orders)
    ords = torch.as_tensor(orders.astype(np.int32), device=dev)
    out = {k: torch.full((no,), -7.0, dtype=torch.float64, device=dev)
           for k in ('prob_sum', 'mean', 'median', 'std')}
    cnt = torch.full((no,), -7, dtype=torch.int64, device=dev)
    kept = torch.full((no,), -7, dtype=torch.int64, device=dev)
    a = _lib.MxOrderStats()
    a.n, a.order, a.value, a.prob = n, order.data_ptr(), value.data_ptr(), prob.data_ptr()
    a.orders, a.n_orders = ords.data_ptr(), no
    a.flags, a.sigma, a.maxiters = _lib.MX_OS_FILTER, 3.0, 5
    ws = None
    if workspace:
        wsb = int(_lib.lib.mx_order_stats_workspace(n, no, 1))
        ws = torch.empty(wsb, dtype=torch.uint8, device=dev)
        a.workspace, a.workspace_bytes = ws.data_ptr(), wsb
    a.prob_sum, a.count = out['prob_sum'].data_ptr(), cnt.data_ptr()
    a.mean, a.median, a.std = out['mean'].data_ptr(), out['median'].data_ptr(), out['std'].data_ptr()
    a.n_kept = kept.data_ptr()
    rc = fn(ctypes.byref(a), ctypes.c_void_p(torch.cuda.current_stream(dev).cuda_stream))
    torch.cuda.synchronize(dev)
    res = {k: v.cpu().numpy() for k, v in out.items()}
    res['count'], res['n_kept'] = cnt.cpu().numpy(), kept.cpu().numpy()
    return rc, res


def test_static_workspace_equals_caller_workspace(gpu):
    from marxs_amd import _lib
    for n, seed in ((2_000_000, 1), (300_000, 2)):
        order, value, prob, orders = _photons(n, seed, gpu)
        rc1, ref = _call(_lib.lib.mx_order_stats, order, value, prob, orders, gpu)
        rc2, got = _call(_lib.lib.mx_order_stats_static, order, value, prob, orders, gpu,
                         workspace=False)
        assert rc1 == 0 and rc2 == 0, _lib.lib.mx_last_error()
        assert ref['count'].sum() > 0.9 * n * 15 / 16
        for k in ref:
            assert np.array_equal(ref[k], got[k], equal_nan=True), (n, k)


def test_static_workspace_size_limit(gpu):
    from marxs_amd import _lib
    n = 4_000_000
    need = int(_lib.lib.mx_order_stats_workspace(n, 15, 1))
    assert need > _lib.MX_OS_STATIC_BYTES
    order, value, prob, orders = _photons(n, 3, gpu)
    rc, got = _call(_lib.lib.mx_order_stats_static, order, value, prob, orders, gpu, workspace=False)
    assert rc == 1
    assert b'static workspace' in _lib.lib.mx_last_error()
    # nothing ran: the outputs keep their fill value
    assert np.all(got['prob_sum'] == -7.0) and np.all(got['count'] == -7)
