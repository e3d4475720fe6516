"""The oracle's figures of merit on the CPU against the reference's known
answers (marxs/analysis/tests/test_grating.py) and closed forms."""
import numpy as np
import pytest

import marxs_amd as mx
from oracle import analysis as OA
from oracle import trace as OT
from oracle.describe import describe


def test_aeff_known_answer():
    """test_grating.py:86-96"""
    aeff = OA.effectivearea_from_photonlist([0, 1, 2, 2, 1], [.1, 1., 1., .3, .5],
                                            [2, 1, 0, -1], 5, A_geom=2.)
    assert aeff[3] == 0 and np.isclose(aeff[2], .04) and np.isclose(aeff[1], .6)


def _capture_photons(prob=1., xpos=None, seed=0):
    rng = np.random.default_rng(seed)
    order = np.zeros(200, dtype=int)
    order[50:] = 5
    x = np.full(200, -100.) if xpos is None else xpos
    if xpos is None:
        x[50:] = rng.normal(scale=1, size=150)
    return order, x, np.full(200, prob)


def test_capture_res_aeff_known_answers():
    """test_grating.py:150-211"""
    order, x, p = _capture_photons()
    out = OA.capture_res_aeff(order, x, p, [0, 2, 5], A_geom=10.)
    assert np.allclose(out['Aeff'], [2.5, 0., 7.5]) and np.isclose(out['Aeff0'], 2.5)
    assert np.isclose(out['Aeffgrat'], 7.5) and np.isnan(out['R'][1])
    assert np.isclose(out['R'][2], 50., rtol=.8)
    # probability 0: no area, no resolving power
    order, x, p = _capture_photons(prob=0.)
    out = OA.capture_res_aeff(order, x, p, [0, 2, 5], A_geom=10.)
    assert np.all(out['Aeff'] == 0) and np.isnan(out['R'][2])
    # non-finite coordinates: area counted, resolving power not
    order, x, p = _capture_photons(xpos=np.full(200, np.nan))
    out = OA.capture_res_aeff(order, x, p, [0, 2, 5], A_geom=10.)
    assert np.allclose(out['Aeff'], [2.5, 0., 7.5]) and np.isnan(out['R'][2])


def test_average_R_Aeff_known_answers():
    """test_grating.py:106-147 (scalar cases)"""
    r, a = OA.average_R_Aeff(np.array([np.nan, np.nan]), np.array([1, 2]))
    assert r is np.ma.masked and a == 3
    assert OA.average_R_Aeff(np.array([5., 2.]), np.array([1, 2]))[0] == pytest.approx(3)


def test_clipping_closed_form():
    """Uniform values in [-1, 1] (3 std = 1.73 > 1) plus far outliers: the
    first round drops the outliers, the second keeps everything, so the
    result is numpy's mean / median / std of the uniform values."""
    rng = np.random.default_rng(1)
    u = rng.uniform(-1, 1, 1000)
    data = np.concatenate([u, [100., -150.], [np.nan, np.inf]])
    rng.shuffle(data)
    m, med, s = OA.sigma_clipped_stats(data)
    kept = data[np.isfinite(data) & (np.abs(data) <= 1)]     # u, in the shuffled order
    assert (m, med, s) == (np.mean(kept), np.median(kept), np.std(kept))
    assert np.isclose(m, np.mean(u), rtol=1e-12) and med == np.median(u)


def test_clipping_brute_force():
    """Each round's kept set is what a value-by-value test of the bounds
    keeps; the loop stops at the first round that removes nothing."""
    rng = np.random.default_rng(2)
    x = np.concatenate([rng.normal(0, 1, 300), rng.normal(0, 8, 20)])
    kept = x.copy()
    for _ in range(5):
        cen, std = np.median(kept), np.std(kept)
        new = np.array([v for v in kept if cen - 3 * std <= v <= cen + 3 * std])
        if len(new) == len(kept):
            break
        kept = new
    assert OA.sigma_clipped_stats(x) == (np.mean(kept), np.median(kept), np.std(kept))


def test_resolving_power_thin_lens_grating():
    """test_grating.py:16-44: focused beam, orders -1/0/+1 on a detector:
    R > 10 in the dispersed orders, < 1 in order 0; cross-dispersion R < 1
    and positions 0."""
    t = mx.generate_test_photons(1000)
    rng = np.random.default_rng(3)
    pos = np.asarray(t['pos'], dtype=float)[:, :3].copy()
    pos[:, 1:] = -0.1 + 0.2 * rng.random((1000, 2))
    ph = {'pos': pos, 'dir': np.asarray(t['dir'], dtype=float)[:, :3].copy(),
          'pol': np.asarray(t['polarization'], dtype=float)[:, :3].copy(),
          'energy': np.asarray(t['energy'], dtype=float), 'prob': np.ones(1000)}
    orders = np.array([-1, 0, 1])
    inst = mx.Sequence(elements=[mx.ThinLens(focallength=1),
                                 mx.FlatGrating(position=[-0.01, 0, 0], d=1e-4,
                                                order_selector=mx.OrderSelector(orders)),
                                 mx.FlatDetector(position=[-1, 0, 0])])
    els = describe(inst)
    u = rng.random((OT.uniform_slots(els)[1], 1000))
    out = OT.trace(els, ph, u)
    res, p, s = OA.resolvingpower_from_photonlist(out['order'], out['det_x'], orders)
    assert res[0] > 10 and res[1] < 1 and res[2] > 10
    res, p, s = OA.resolvingpower_from_photonlist(out['order'], out['det_x'], orders, zeropos=0)
    assert res[0] > 10 and res[1] < 1 and res[2] > 10
    resy, py, sy = OA.resolvingpower_from_photonlist(out['order'], out['det_y'], orders, zeropos=0)
    assert np.all(resy < 1.) and np.allclose(py, 0, atol=1e-4)
    with pytest.raises(OA.AnalysisError):
        OA.resolvingpower_from_photonlist(out['order'][:10], out['det_x'][:10], orders)
