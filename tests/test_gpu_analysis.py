"""Per-order figures of merit on the GPU (mx_order_stats) against the
oracle's restatement (needs an MI355X).  Sums of probability within 1e-12
relative (summation order differs), counts and kept-value counts exact,
clipped mean / std within 1e-9 relative, medians exact (they are elements
or the mean of two elements of the same sorted values)."""
import numpy as np
import pytest

from oracle import analysis as OA

pytestmark = pytest.mark.gpu


def _t(x, gpu, dtype=None):
    import torch
    return torch.as_tensor(np.asarray(x), device=gpu, dtype=dtype)


def _oracle_stats(order, value, prob, orders, res_filter=False):
    out = {k: [] for k in ('prob_sum', 'count', 'mean', 'median', 'std', 'n_kept')}
    for o in orders:
        m = order == o
        out['prob_sum'].append(prob[m].sum())
        c = m & (np.isfinite(value) & (prob > 0) if res_filter else True)
        out['count'].append(int(c.sum()))
        mean, med, std = OA.sigma_clipped_stats(value[c])
        out['mean'].append(mean)
        out['median'].append(med)
        out['std'].append(std)
    return {k: np.array(v) for k, v in out.items()}


@pytest.mark.parametrize('n, res_filter', [(200_000, False), (200_000, True), (1000, False),
                                           (3_000_000, True)])
def test_order_stats_against_oracle(gpu, n, res_filter):
    import torch
    from marxs_amd.analysis import order_stats
    rng = np.random.default_rng(n + res_filter)
    orders = np.arange(-12, 3)
    order = rng.integers(-14, 4, n).astype(np.int32)
    order[rng.random(n) < 0.05] = -2 ** 31                      # no order
    value = 300. * order + rng.normal(0, 1, n)
    out_idx = rng.random(n) < 0.003
    value[out_idx] += rng.choice([-1, 1], out_idx.sum()) * rng.uniform(10, 100, out_idx.sum())
    value[rng.random(n) < 0.01] = np.nan
    value[rng.random(n) < 0.001] = np.inf
    prob = rng.uniform(0, 1, n)
    prob[rng.random(n) < 0.05] = 0.
    got = order_stats(_t(order, gpu), _t(prob, gpu), orders, value=_t(value, gpu),
                      res_filter=res_filter)
    ref = _oracle_stats(order, value, prob, orders, res_filter)
    assert np.allclose(got['prob_sum'], ref['prob_sum'], rtol=1e-12, atol=0)
    assert np.array_equal(got['count'], ref['count'])
    for k in ('mean', 'std'):
        assert np.allclose(got[k], ref[k], rtol=1e-9, atol=1e-12, equal_nan=True), k
    assert np.array_equal(got['median'], ref['median'], equal_nan=True)


def test_clipping_closed_form_on_gpu(gpu):
    from marxs_amd.analysis import sigma_clipped_stats
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, 100_000)
    data = np.concatenate([u, [100., -150.], [np.nan, np.inf]])
    rng.shuffle(data)
    m, med, s = sigma_clipped_stats(_t(data, gpu))
    assert np.isclose(m, np.mean(u), rtol=1e-12) and med == np.median(u)
    assert np.isclose(s, np.std(u), rtol=1e-12)
    # empty and all-NaN input
    assert np.isnan(sigma_clipped_stats(_t(np.full(10, np.nan), gpu))[0])


def test_reference_known_answers_on_gpu(gpu):
    """marxs/analysis/tests/test_grating.py:86-96, 150-238"""
    from marxs_amd import analysis as A
    from marxs_amd.table import PhotonTable
    p = PhotonTable({'a': np.array([0, 1, 2, 2, 1]), 'probability': [.1, 1., 1., .3, .5]})
    aeff = A.effectivearea_from_photonlist(p, [2, 1, 0, -1], 5, A_geom=2., ordercol='a')
    assert aeff[3] == 0 and np.isclose(aeff[2], .04) and np.isclose(aeff[1], .6)
    rng = np.random.default_rng(0)
    order = np.zeros(200, dtype=int)
    order[50:] = 5
    x = np.full(200, -100.)
    x[50:] = rng.normal(scale=1, size=150)
    t = PhotonTable({'order': order, 'xpos': x, 'probability': np.ones(200),
                     'CCD': np.where(np.arange(200) < 100, 0, -1)})
    r = A.CaptureResAeff(A_geom=10., order_col='order', orders=[0, 2, 5], dispersion_coord='xpos')
    out = r(t)
    ref = OA.capture_res_aeff(order, x, np.ones(200), [0, 2, 5], A_geom=10.)
    assert np.allclose(out['Aeff'], [2.5, 0., 7.5]) and np.isclose(out['Aeffgrat'], 7.5)
    assert np.isnan(out['R'][1]) and np.isclose(out['R'][2], ref['R'][2], rtol=1e-9)
    t['probability'] = np.zeros(200)
    out = r(t)
    assert np.all(out['Aeff'] == 0) and np.isnan(out['R'][2])
    t['probability'] = np.ones(200)
    t['xpos'] = np.full(200, np.nan)
    out = r(t)
    assert np.allclose(out['Aeff'], [2.5, 0., 7.5]) and np.isnan(out['R'][2])
    t['xpos'] = x
    gaps = A.CaptureResAeff_CCDgaps(A_geom=10., order_col='order', orders=[0, 2, 5],
                                    dispersion_coord='xpos', aeff_filter_col='CCD')
    out = gaps(t)
    assert np.allclose(out['Aeff'], [2.5, 0., 2.5]) and np.isnan(out['R'][1])
    with pytest.raises(A.AnalysisError):
        A.resolvingpower_from_photonlist(t[:10], [5], col='xpos')


def test_trace_then_capture_on_device(gpu):
    """Trace the 200-facet Rowland instrument on the GPU, then CaptureResAeff
    on the device columns; the oracle's figures of merit on the same
    (downloaded) columns agree."""
    from marxs_amd import configs, analysis as A
    from tests.helpers import make_inputs, gpu_trace
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(200)
    ci = CompiledInstrument(inst)
    ph, u = make_inputs(400_000, 7, ci.n_uniform_slots, 'lines')
    got, st = gpu_trace(ci, ph, u, precision=64)
    orders = np.arange(-12, 3)
    cap = A.CaptureResAeff(A_geom=np.pi * (600 ** 2 - 300 ** 2), orders=orders,
                           dispersion_coord='det_x', zeropos=0.)
    t = {'order': got['order'].astype(np.int32), 'det_x': got['det_x'],
         'probability': got['prob']}
    out = cap(t)
    ref = OA.capture_res_aeff(got['order'], got['det_x'], got['prob'], orders,
                              A_geom=np.pi * (600 ** 2 - 300 ** 2), zeropos=0.)
    assert np.allclose(out['Aeff'], ref['Aeff'], rtol=1e-12)
    assert np.allclose(out['R'], ref['R'], rtol=1e-9, equal_nan=True)
    assert np.isfinite(out['R']).sum() >= 5


def test_order_stats_edge_cases(gpu):
    """No photons, orders absent from the list, one photon, duplicate orders."""
    from marxs_amd.analysis import order_stats
    e = np.zeros(0)
    r = order_stats(_t(e.astype(np.int32), gpu), _t(e, gpu), [0, 1], value=_t(e, gpu))
    assert np.array_equal(r['count'], [0, 0]) and np.all(r['prob_sum'] == 0)
    assert np.all(np.isnan(r['mean'])) and np.array_equal(r['n_kept'], [0, 0])
    r = order_stats(_t(np.array([5], np.int32), gpu), _t([0.5], gpu), [0, 5], value=_t([2.5], gpu))
    assert r['count'].tolist() == [0, 1] and r['prob_sum'].tolist() == [0., 0.5]
    assert r['mean'][1] == 2.5 and r['median'][1] == 2.5 and r['std'][1] == 0.
    with pytest.raises(ValueError, match='distinct'):
        order_stats(_t(np.array([5], np.int32), gpu), _t([0.5], gpu), [5, 5])


def test_order_stats_more_than_64_orders(gpu):
    from marxs_amd.analysis import order_stats
    rng = np.random.default_rng(4)
    order = rng.integers(-50, 50, 100_000).astype(np.int32)
    value = order * 10. + rng.normal(0, 1, order.size)
    prob = rng.random(order.size)
    orders = np.arange(-50, 50)
    got = order_stats(_t(order, gpu), _t(prob, gpu), orders, value=_t(value, gpu))
    ref = _oracle_stats(order, value, prob, orders)
    assert np.array_equal(got['count'], ref['count'])
    assert np.allclose(got['mean'], ref['mean'], rtol=1e-9)


def test_find_best_detector_position_known_answers(gpu):
    """marxs/analysis/tests/test_analysis.py:6-41: rays converging at x = 1
    (or z = 0.5 with the rotated detector) give the best focus there."""
    import marxs_amd as mx
    from marxs_amd import analysis as A
    rng = np.random.default_rng(2)
    p = mx.generate_test_photons(100)
    d = np.asarray(p['dir']).copy()
    d[:, 1] = rng.random(100)
    p['dir'] = d
    pos = np.asarray(p['pos']).copy()
    pos[:, 1] = 0.5
    p['pos'] = pos
    mx.propagate(p, -10)
    out = A.find_best_detector_position(p, objective_func_args={'colname': 'det_x'})
    assert np.isclose(out.x, 1.)

    p = mx.generate_test_photons(100)
    d = np.asarray(p['dir']).copy()
    d[:, 0] = rng.random(100)
    d[:, 2] = -1
    p['dir'] = d
    pos = np.asarray(p['pos']).copy()
    pos[:, 2] = 0.5
    p['pos'] = pos
    mx.propagate(p, -10)
    out = A.find_best_detector_position(p, objective_func_args={'colname': 'det_x'},
                                        orientation=np.array([[0., 1., 0.], [0., 0., 1.],
                                                              [1., 0., 0.]]))
    assert np.isclose(out.x, 0.5)

    p = mx.generate_test_photons(100)
    d = np.asarray(p['dir']).copy()
    d[:, 2] = rng.random(100)
    p['dir'] = d
    pos = np.asarray(p['pos']).copy()
    pos[:, 1] = 0.5
    p['pos'] = pos
    mx.propagate(p, -10)
    out = A.find_best_detector_position(p, A.mean_width_2d, {})
    assert np.isclose(out.x, 1.0)


def test_resolvingpower_from_photonlist_robust_known_answers(gpu):
    """marxs/analysis/tests/test_grating.py:47-84 (statistical answers,
    10 % relative)."""
    from marxs_amd import analysis as A
    rng = np.random.default_rng(3)
    size = 10000
    orders = [1, 2, 3]
    ph = {'col1': 2000 + rng.normal(size=size), 'col2': 1000 + rng.normal(size=size),
          'order': rng.choice(orders, size=size)}
    res, pos, std = A.resolvingpower_from_photonlist_robust([ph, ph], orders, ['col1', 'col2'], [0, 0])
    assert np.allclose(res, 1000 / 2.3548, rtol=0.1) and np.allclose(std, 1, rtol=0.1)
    assert np.allclose(pos, 1000, rtol=0.1)
    res, pos, std = A.resolvingpower_from_photonlist_robust([ph, ph], orders, ['col1', 'col2'],
                                                            [0, -1000])
    assert np.allclose(res, 2000 / 2.3548, rtol=0.1)
    ph = {'col1': 2000 + rng.normal(size=size), 'col2': 1000 + rng.normal(size=size),
          'ord': rng.choice([0, 1, 2], size=size)}
    ph['col1'][ph['ord'] == 0] = 0
    res, pos, std = A.resolvingpower_from_photonlist_robust([ph, ph], [1, 2], ['col1', 'col2'],
                                                            [None, 10000], ordercol='ord')
    assert np.allclose(res, 2000 / 2.3548, rtol=0.1) and np.allclose(pos, 2000, rtol=0.1)
    with pytest.raises(ValueError, match='zeropositions'):
        A.resolvingpower_from_photonlist_robust([ph, ph], [1], ['col1'], [0, 0])


def test_detected_fraction(gpu):
    """marxs/analysis/analysis.py:84-107 on the photon list of the
    reference's effective-area test (test_grating.py:87-96)."""
    from marxs_amd import analysis as A
    from marxs_amd.table import PhotonTable
    p = PhotonTable({'a': np.array([0, 1, 2, 2, 1]), 'probability': [.1, 1., 1., .3, .5]})
    assert np.allclose(A.detected_fraction(p, [2, 1, 0, -1], col='a'), [1.3 / 5, 1.5 / 5, .1 / 5, 0.])


def _pre_grating_photons(parts, n, seed, energy):
    """Photons after aperture and lens (FP64 GPU trace), monochromatic."""
    import marxs_amd as mx
    from tests.helpers import make_inputs, gpu_trace
    from marxs_amd.tracer import CompiledInstrument
    ci = CompiledInstrument(mx.Sequence(elements=[parts['aperture'], parts['lens']]))
    ph, u = make_inputs(n, seed, ci.n_uniform_slots)
    ph['energy'] = np.full(n, energy)
    out, st = gpu_trace(ci, ph, u)
    return {k: out[k] for k in ('pos', 'dir', 'pol', 'energy', 'prob')}


def _table(ph):
    from marxs_amd.table import PhotonTable
    h = lambda a, w: np.hstack([a, np.full((len(a), 1), w)])
    return PhotonTable({'pos': h(ph['pos'], 1.), 'dir': h(ph['dir'], 0.),
                        'polarization': h(ph['pol'], 0.), 'energy': ph['energy'],
                        'probability': ph['prob']})


def test_resolvingpower_per_order_rowland_against_oracle(gpu):
    """gratings.py:26-130 with a RowlandTorus: every photon sent into one
    order at a time, traced to a tube on the Rowland circle; the oracle runs
    the same loop (order selectors with one order, the tube, sigma-clipped
    statistics) on the same photons.  R and FWHM within 1e-9 relative."""
    import marxs_amd as mx
    from marxs_amd import configs, analysis as A
    from marxs_amd.geometry import Cylinder
    from tests.helpers import oracle_trace
    inst, parts = configs.cat_rowland_instrument(200)
    ph = _pre_grating_photons(parts, 20000, 21, 0.8)
    orders = np.array([-7, -4, 0])
    res, fwhm, info = A.resolvingpower_per_order(parts['gratings'], _table(ph), orders,
                                                 detector=parts['rowland'])
    assert info['method'] == 'Circular detector on Rowland circle'
    det = mx.CircularDetector(geometry=Cylinder.from_rowland(
        parts['rowland'], width=1e5, rotation=np.pi, kwargs={'phi_lim': [-np.pi / 2, np.pi / 2]}))
    n = len(ph['prob'])
    z = oracle_trace(mx.Sequence(elements=[det]), ph, np.zeros((0, n)))
    zeropos = OA.sigma_clipped_stats(z['detpix_x'][z['prob'] > 0])[0]
    rng = np.random.default_rng(0)
    for i, o in enumerate(orders):
        sel = mx.OrderSelector([o])
        for e in parts['gratings'].elements:
            e.order_selector = sel
        g = oracle_trace(mx.Sequence(elements=[parts['gratings']]), ph, rng.random((2, n)))
        keep = (g['order'] == o) & (g['prob'] > 0)
        sub = {k: g[k][keep] for k in ('pos', 'dir', 'pol', 'energy', 'prob')}
        t = oracle_trace(mx.Sequence(elements=[det]), sub, np.zeros((0, keep.sum())))
        mean, med, std = OA.sigma_clipped_stats(t['detpix_x'])
        assert np.isclose(fwhm[i], 2.3548 * std, rtol=1e-9, atol=0)
        assert np.isclose(res[i], abs((mean - zeropos) / (2.3548 * std)), rtol=1e-9, atol=0)
    assert res[0] > 100 and res[1] > 100 and res[2] < 1


def test_resolvingpower_per_order_flat_detectors(gpu):
    """The two other detector modes of gratings.py:26-130: a user detector
    (the instrument's own CCD array, column det_x) and None (a flat detector
    at the best focus of each order)."""
    from marxs_amd import configs, analysis as A
    inst, parts = configs.cat_rowland_instrument(200)
    ph = _pre_grating_photons(parts, 20000, 22, 0.8)
    orders = np.array([-6, 0])
    res, fwhm, info = A.resolvingpower_per_order(parts['gratings'], _table(ph), orders, detector=None)
    assert info['method'] == 'Detector position numerically optimized'
    assert len(info['fit_results']) == 2
    # the best focus lies near the Rowland circle, about 6 m behind the gratings
    assert all(-500 < r.x < 500 for r in info['fit_results'])
    assert np.all(np.isfinite(fwhm)) and res[0] > 100
    import marxs_amd as mx
    res2, fwhm2, info2 = A.resolvingpower_per_order(parts['gratings'], _table(ph), orders,
                                                    detector=mx.FlatDetector(zoom=1e5),
                                                    colname='det_x')
    assert info2['method'] == 'User defined detector'
    assert np.all(np.isfinite(fwhm2)) and res2[0] > 100
    # a detector at the focus of order 0: narrower there than the best focus
    # found for the dispersed order, wider for that order
    assert fwhm2[1] <= fwhm[1] * 1.01 and fwhm2[0] >= fwhm[0]
