"""The oracle's element-major driver, its invariants, the pixel-index
convention, and planted mistakes that the oracle's pins must catch (CPU)."""
import numpy as np
import pytest

from oracle import trace as OT
from oracle.describe import describe

import test_oracle as T


def _oracle_photons(t):
    return {'pos': np.asarray(t['pos'], dtype=float)[:, :3].copy(),
            'dir': np.asarray(t['dir'], dtype=float)[:, :3].copy(),
            'pol': np.asarray(t['polarization'], dtype=float)[:, :3].copy(),
            'energy': np.asarray(t['energy'], dtype=float),
            'prob': np.asarray(t['probability'], dtype=float)}


def test_driver_parallel_numbering_hits_element_3():
    """marxs/simulator/tests/test_parallel.py:11-23: four detectors, the test
    photons all land on the one with id 3."""
    import marxs_amd as mx
    pos = [[0, -10.1, -10.1], [0, .1, -10.1], [0, -10.1, .1], [0, .1, .1]]
    det = mx.Parallel(elem_class=mx.FlatDetector, elem_args={'pixsize': 0.01, 'zoom': 5},
                      elem_pos={'position': pos}, id_col='CCD_ID')
    out = OT.trace(describe(det), _oracle_photons(mx.generate_test_photons(5)),
                   np.zeros((0, 5)))
    assert np.all(out['CCD_ID'] == 3)


def test_driver_invariants_on_rowland_array():
    """marxs/optics/tests/test_all_optics.py:191-214: after every element the
    polarization is a unit vector perpendicular to the direction and the
    probability never increases."""
    from marxs_amd import configs
    from helpers import make_inputs
    inst, parts = configs.cat_rowland_instrument(200)
    els = describe(inst)
    ph, u = make_inputs(3000, 21, 4)
    prev = np.ones(3000)
    for k in range(1, len(els) + 1, 37):
        out = OT.trace(els[:k], ph, u)
        d, p = out['dir'], out['pol']
        assert np.allclose(np.linalg.norm(p, axis=1), 1, atol=1e-12)
        assert np.allclose(np.sum(d * p, axis=1), 0, atol=1e-12)
        assert np.all(out['prob'] <= prev + 1e-15)
        prev = out['prob']
    full = OT.trace(els, ph, u)
    assert (full['CCD'] >= 0).sum() > 100


def test_driver_last_writer_wins():
    """simulator.py:42-49 with optics/base.py:201-209: a photon that passes
    two overlapping facets of one Parallel is diffracted by both; the id and
    local columns hold the second facet, the order column the second order,
    and the direction equals two successive diffractions."""
    import marxs_amd as mx
    sel = mx.OrderSelector([1])
    par = mx.Parallel(elem_class=mx.FlatGrating,
                      elem_args={'zoom': [1, 20, 20], 'd': 1e-3, 'order_selector': sel},
                      elem_pos={'position': [[100., 0, 0], [50., 3., 0]]}, id_col='facet')
    n = 4
    # photon 0/1 pass both facets, photon 2 (y = 22) only the second,
    # photon 3 (y = -18) only the first (it leaves it at y = -17.94, outside
    # the second facet's y range [-17, 23])
    ph = {'pos': np.array([[200., 0., 0.], [200., 0.5, 1.], [200., 22., 0.], [200., -18., 0.]]),
          'dir': np.tile([-1., 0, 0], (n, 1)), 'pol': np.tile([0., 0, 1], (n, 1)),
          'energy': np.ones(n), 'prob': np.ones(n)}
    els = describe(par)
    out = OT.trace(els, ph, np.zeros((2, n)))
    pos, d, ids = ph['pos'].copy(), ph['dir'].copy(), np.full(n, -1)
    for el in els:
        fr = OT.frame(el['pos4d'])
        hit, q, ly, lz = OT.intersect(fr, pos, d)
        new, _ = OT.diffract(el, fr, d[hit], ph['energy'][hit], ly[hit],
                             np.ones(int(hit.sum()), int), None)
        d[hit] = new
        pos[hit] = q[hit]
        ids[hit] = el['id_num']
    assert ids.tolist() == [1, 1, 1, 0]
    assert np.array_equal(out['facet'], ids)
    assert np.array_equal(out['dir'], d) and np.array_equal(out['pos'], pos)
    assert np.all(out['order'] == 1)


def test_pixel_index_convention():
    """Pixel k covers detpix in [k - 0.5, k + 0.5) (pixel centres at integer
    detpix, marxs/optics/detector.py:26-32); outside [0, npix) gives -1."""
    dp = np.array([0., 0.49, 0.5, 1.49999, -0.5, -0.51, 9.49, 9.5, np.nan])
    assert OT.pixel_index(dp, 10).tolist() == [0, 0, 1, 1, 0, -1, 9, -1, -1]


# ------------------------------------------------------ planted mistakes
def _swap_local(orig):
    def f(fr, pos, d):
        hit, q, ly, lz = orig(fr, pos, d)
        return hit, q, lz, ly
    return f


def _reverse_ray(orig):
    return lambda fr, pos, d: orig(fr, pos, -d)


def _flip_order(orig):
    return lambda el, fr, d, e, ly, m, pt: orig(el, fr, d, e, ly, -np.asarray(m), pt)


def _next_energy_row(orig):
    def f(tab, energy, u):
        if tab['energy'] is None:
            return orig(tab, energy, u)
        e = np.asarray(tab['energy'])
        row = np.argmin(np.abs(e[None, :] - energy[:, None]), axis=1)
        shifted = e[np.minimum(row + 1, len(e) - 1)]
        return orig(tab, shifted, u)
    return f


def _inverse_transport(orig):
    return lambda d1, d2, pol: orig(d2, d1, pol)


def _no_half_pixel(orig):
    return lambda dp, n: orig(np.asarray(dp) - 0.5, n)


def _unnormalised_cdf(orig):
    def f(sel):
        tab = orig(sel)
        if sel['type'] == 'orderselector':
            p = np.asarray(sel['p'], dtype=float)
            tab = dict(tab, cdf=np.cumsum(p)[None, :])
        return tab
    return f


MUTANTS = [
    ('ray reversed', 'intersect', _reverse_ray, ['test_intersect_forwards_backwards']),
    ('local axes swapped', 'intersect', _swap_local, ['test_intersect_zoomed']),
    ('order sign flipped', 'diffract', _flip_order, ['test_grating_equation_angles']),
    ('neighbouring energy row', 'select_order', _next_energy_row, ['test_efficiencyfile_known']),
    ('cdf not normalised', 'selector_table', _unnormalised_cdf,
     ['test_orderselector_inverse_cdf_matches_numpy_choice']),
    ('inverse transport', 'parallel_transport', _inverse_transport,
     ['test_paralleltransport_rodrigues']),
    ('pixel edges shifted', 'pixel_index', _no_half_pixel, []),
]


@pytest.mark.parametrize('name,attr,make,checks', MUTANTS, ids=[m[0] for m in MUTANTS])
def test_planted_mistake_is_caught(monkeypatch, name, attr, make, checks):
    monkeypatch.setattr(OT, attr, make(getattr(OT, attr)))
    fns = [getattr(T, c) for c in checks]
    if attr == 'pixel_index':
        fns = [test_pixel_index_convention]
    for fn in fns:
        with pytest.raises(AssertionError):
            fn()


def test_per_photon_selector_loop_equals_vectorised():
    """The reference's one-photon-at-a-time EfficiencyFile loop
    (grating.py:90-96), restated for the CPU baseline, selects the same
    orders and totals as the vectorised selector on the 200-facet workload,
    including draws on the cumulative probabilities themselves."""
    from marxs_amd import configs
    from helpers import make_inputs
    inst, parts = configs.cat_rowland_instrument(200)
    from marxs_amd.compiler import compile_tree
    ns = compile_tree(inst).n_uniform_slots
    ph, u = make_inputs(3000, 21, ns)
    els = describe(inst)
    a = OT.trace(els, ph, u)
    b = OT.trace(els, ph, u, selector_loop=True)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=np.asarray(a[k]).dtype.kind == 'f'), k
    tab = OT.selector_table(els[2]['selector'])
    e = np.repeat(tab['energy'][:4], 15)
    uu = np.tile(tab['cdf'][0], 4)
    va = OT.select_order(tab, e, uu)
    vb = OT.select_order_loop(tab, e, uu)
    for x, y in zip(va, vb):
        assert np.array_equal(x, y)
