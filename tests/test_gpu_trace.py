"""GPU parity of the HIP trace against the oracle (needs an MI355X).

FP64 path: integer outputs (facet, order, CCD, pixel) bit-exact except for
photons the oracle places within 1e-6 px of a pixel edge (counted and
reported); every float column within 1e-9 relative.
FP32 path: detector coordinates within 0.5 um RMS of the FP64 oracle.
"""
import numpy as np
import pytest

from marxs_amd import configs
import marxs_amd as mx
from marxs_amd.mathutils import axangle2mat
from tests.helpers import (gpu_trace, oracle_trace, compare, make_inputs,
                           near_pixel_boundary)

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _ci(inst, **kw):
    from marxs_amd.tracer import CompiledInstrument
    return CompiledInstrument(inst, **kw)


def _parity(inst, n, seed, energies='uniform', grid=True, n_draws=2, exempt=True, max_near_frac=1e-3):
    ci = _ci(inst, grid=grid, n_draws=n_draws)
    ph, u = make_inputs(n, seed, ci.n_uniform_slots, energies)
    ref = oracle_trace(inst, ph, u, n_draws=n_draws)
    got, st = gpu_trace(ci, ph, u, precision=64, grid=grid)
    rep = compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref) if exempt else None,
                  max_near_frac=max_near_frac)
    return ref, got, st, rep


@pytest.mark.parametrize('grid', [True, False])
def test_flat_grid_fp64_parity(gpu, grid):
    """121 coplanar FlatGrating facets (shared edges), one 8192^2 detector.
    The zero order (half the photons) focuses exactly onto the corner of four
    pixels, where the pixel index is decided by the last bit of the detector
    coordinate: those photons' pixel indices may differ (they are counted;
    0.17 % of the photons were measured, the bound here is 0.5 %).  Every
    other integer is exact and every float within 1e-9 relative."""
    ref, got, st, rep = _parity(configs.flat_grid_instrument(), 20000, 1, grid=grid,
                                max_near_frac=5e-3)
    assert np.isfinite(got['det_x']).sum() > 10000
    assert rep['facet'][1:] == (0, 0) and rep['order'][1:] == (0, 0)
    assert rep['pix_x'][1] == 0 and rep['pix_y'][1] == 0


@pytest.mark.parametrize('grid', [True, False])
def test_flat_grid_shifted_detector_exact_pixels(gpu, grid):
    """The same with the detector moved by a third of a pixel in y and z, so
    that the zero order lands inside a pixel: every pixel index must equal
    the oracle's with no exemption (detector.py:59-75 pixel convention)."""
    ref, got, st, rep = _parity(configs.flat_grid_instrument(det_shift=0.024 / 3), 20000, 1,
                                grid=grid, exempt=False)
    assert (got['pix_x'] >= 0).sum() > 10000
    assert rep['pix_x'][1:] == (0, 0) and rep['pix_y'][1:] == (0, 0)


@pytest.mark.parametrize('grid', [True, False])
def test_cat_rowland_200_fp64_parity(gpu, grid):
    inst, parts = configs.cat_rowland_instrument(200)
    ref, got, st, rep = _parity(inst, 40000, 2, grid=grid)
    assert (got['CCD'] >= 0).sum() > 1000
    assert rep['CCD'][1] == 0


def test_cat_rowland_600_fp64_parity_line_spectrum(gpu):
    inst, parts = configs.cat_rowland_instrument(600)
    ref, got, st, rep = _parity(inst, 40000, 3, energies='lines')
    assert (got['facet'] >= 0).sum() > 10000


def test_fp32_detector_error(gpu):
    """FP32 local-frame tracer: <= 0.5 um RMS detector error vs the FP64
    oracle; facet / CCD edge disagreements counted."""
    inst, parts = configs.cat_rowland_instrument(600)
    ci = _ci(inst)
    ph, u = make_inputs(40000, 4, ci.n_uniform_slots, 'lines')
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=32)
    both = (ref['CCD'] >= 0) & (got['CCD'] == ref['CCD']) & (got['facet'] == ref['facet'])
    err = np.hypot(got['det_x'] - ref['det_x'], got['det_y'] - ref['det_y'])[both] * 1e3
    rms = float(np.sqrt(np.mean(err ** 2)))
    edge = int((got['facet'] != ref['facet']).sum() + (got['CCD'] != ref['CCD']).sum())
    print(f'FP32: rms {rms:.4f} um max {err.max():.4f} um edge mismatches {edge}')
    assert both.sum() > 5000
    assert rms <= 0.5
    assert edge <= 0.001 * len(err)
    assert (got['order'] == ref['order'])[got['facet'] == ref['facet']].all()


# ---------------------------------------------------------------- elements
def _single(elements, ph, u, **kw):
    inst = mx.Sequence(elements=elements)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(inst, ph, u, **kw)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    return ref, got, st


def _photons(n, seed, spread=0.05, x0=100.):
    rng = np.random.default_rng(seed)
    d = np.column_stack([-np.ones(n), rng.normal(0, spread, (n, 2))])
    d /= np.linalg.norm(d, axis=1)[:, None]
    pol = np.cross(d, rng.normal(size=(n, 3)))
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    pos = np.column_stack([np.full(n, x0), rng.uniform(-30, 30, (n, 2))])
    return {'pos': pos, 'dir': d, 'pol': pol, 'energy': rng.uniform(0.3, 2, n),
            'prob': np.ones(n)}, rng.random((4, n))


def test_apertures(gpu):
    ph, u = _photons(5000, 1)
    del ph['pos']
    for ap in (mx.CircleAperture(position=[50, 1, 2], zoom=[1, 30, 30], r_inner=10,
                                 phi=[0.3, 2.5]),
               mx.RectangleAperture(position=[50, 0, 0], zoom=[1, 20, 10],
                                    orientation=mx.configs.axangle2mat([0, 1, 1], .4))):
        _single([ap], ph, u[:2])


def test_lenses(gpu):
    ph, u = _photons(5000, 2)
    _single([mx.PerfectLens(focallength=500, position=[10, 0, 5], zoom=[1, 40, 40],
                            d_center_optical_axis=3.)], ph, u[:0])
    _single([mx.ThinLens(focallength=300, position=[10, 1, 0], zoom=[1, 40, 40])], ph, u[:0])


@pytest.mark.parametrize('cls', ['flat', 'cat', 'nonparallel', 'reflection'])
def test_gratings(gpu, cls):
    ph, u = _photons(8000, 3, spread=0.03)
    sel = mx.OrderSelector([-3, -1, 0, 1, 2], [.1, .3, .2, .2, .1])
    kw = dict(d=2e-4, order_selector=sel, zoom=[1, 25, 25], groove_angle=.2,
              orientation=mx.configs.axangle2mat([0, 0, 1], .03))
    if cls == 'flat':
        g = mx.FlatGrating(**kw)
    elif cls == 'cat':
        g = mx.CATGrating(**kw)
    elif cls == 'nonparallel':
        g = mx.NonParallelCATGrating(d_blaze_mm=1e-3, blaze_center=0.01, **kw)
    else:
        g = mx.FlatGrating(transmission=False, **kw)
    det = mx.FlatDetector(pixsize=0.01, position=[-300, 0, 0], zoom=[1, 200, 200])
    _single([g, det], ph, u[:2])


def test_efficiency_file_and_ccd_array(gpu):
    ph, u = _photons(8000, 4, spread=0.02)
    eff = configs.blaze_efficiency()
    grat = mx.Parallel(elem_class=mx.CATGrating,
                       elem_args={'d': 2e-4, 'order_selector': eff, 'zoom': [1, 10, 10]},
                       elem_pos={'position': [[0, y, z] for y in (-20, 0, 20) for z in (-20, 0, 20)]},
                       id_col='facet')
    ccd = mx.Parallel(elem_class=mx.FlatDetector,
                      elem_args={'pixsize': 0.024, 'zoom': [1, 12.288, 12.288]},
                      elem_pos={'position': [[-400, y, 0] for y in np.arange(-200, 201, 25.)]},
                      id_col='CCD')
    ref, got, st = _single([grat, ccd], ph, u[:2])
    assert (got['CCD'] >= 0).sum() > 100


def test_multiple_interactions_and_draw_reuse(gpu):
    """Two stacked grating planes in one stage: every photon interacts twice,
    the second interaction uses the second draw; with one draw per stage the
    first draw is reused (counted)."""
    ph, u = _photons(4000, 5, spread=0.001)
    sel = mx.OrderSelector([-1, 0, 1])
    g = mx.Parallel(elem_class=mx.FlatGrating,
                    elem_args={'d': 1e-3, 'order_selector': sel, 'zoom': [1, 40, 40]},
                    elem_pos={'position': [[0, 0, 0], [-5, 0, 0]]}, id_col='facet')
    inst = mx.Sequence(elements=[g])
    for nd in (2, 1):
        ref = oracle_trace(inst, ph, u[:nd], n_draws=nd)
        got, st = gpu_trace(inst, ph, u[:nd], n_draws=nd)
        compare(ref, got, rtol=RTOL)
        assert st['multi_hit'] > 3000
        assert (st['slot_reuse'] > 0) == (nd == 1)


def test_empty_and_degenerate_inputs(gpu):
    inst = configs.flat_grid_instrument()
    ph, u = make_inputs(0, 1, 4)
    got, st = gpu_trace(inst, ph, u)
    assert got['pos'].shape == (0, 3)
    # NaN directions and photons that miss everything
    ph, u = make_inputs(64, 2, 4)
    ph['dir'][:8] = np.nan
    ph['dir'][8:16] = [1., 0, 0]
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(inst, ph, u)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert np.all(got['facet'][8:16] == -1)


def test_status_errors_raise(gpu):
    ph, u = _photons(100, 6)
    bad = mx.EfficiencyFile.from_arrays([1.], [[0.6, 0.6]], [0, 1])
    g = mx.FlatGrating(d=1e-3, order_selector=bad, zoom=[1, 40, 40])
    with pytest.raises(ValueError, match='probability outside'):
        gpu_trace(mx.Sequence(elements=[g]), ph, u[:2])


def test_table_api_and_existing_columns(gpu):
    """Element __call__ on a PhotonTable: reference known answers through the
    GPU path (marxs/optics/tests/test_detector.py:16-35)."""
    pos = np.array([[0, 0., -0.25, 1.], [0., 9.9, 1., 1.], [0., 10.1, 1., 1.]])
    d = np.ones((3, 4))
    d[:, 3] = 0
    t = mx.PhotonTable({'pos': pos, 'dir': d, 'energy': [1, 2., 3.],
                        'polarization': np.tile([0, 1., -1., 0], (3, 1)),
                        'probability': [1., 1., 1.]})
    det = mx.FlatDetector(zoom=[1., 10., 5.], pixsize=0.5)
    t = det(t)
    nan = np.nan
    assert np.allclose(t['det_x'], [0, 9.9, nan], equal_nan=True)
    assert np.allclose(t['det_y'], [-0.25, 1., nan], equal_nan=True)
    assert np.allclose(t['detpix_x'], [19.5, 39.3, nan], equal_nan=True)
    assert np.allclose(t['detpix_y'], [9, 11.5, nan], equal_nan=True)
    # a second detector writing the same columns keeps values where it misses
    det2 = mx.FlatDetector(zoom=[1., 1., 1.], pixsize=0.5, position=[0, 9.9, 1.])
    t = det2(t)
    # photon 0 misses det2 (keeps its det_x from det); photons 1, 2 hit it
    assert np.allclose(t['det_x'], [0, 0, 0.2])
    assert np.allclose(t['detpix_x'], [19.5, 1.5, 1.9])
    # order column: float with NaN for photons that met no grating
    g = mx.CATGrating(d=1. / 5000, order_selector=mx.OrderSelector([5]), zoom=2)
    p = mx.generate_test_photons(5)
    p['pos'][:, 1] = np.arange(5)
    p = g(p)
    assert np.all(p['order'][:3] == 5) and np.all(np.isnan(p['order'][3:]))
    assert np.all(np.isnan(p['grat_y'][3:]))


def test_parallel_numbering_hits_element_3(gpu):
    """marxs/simulator/tests/test_parallel.py:11-23"""
    pos = [[0, -10.1, -10.1], [0, .1, -10.1], [0, -10.1, .1], [0, .1, .1]]
    det = mx.Parallel(elem_class=mx.FlatDetector, elem_args={'pixsize': 0.01, 'zoom': 5},
                      elem_pos={'position': pos}, id_col='CCD_ID')
    p = det(mx.generate_test_photons(5))
    assert np.all(p['CCD_ID'] == 3)


def test_cat_order_convention_table(gpu):
    """marxs/optics/tests/test_grating.py:200-223 through the GPU path."""
    dirs = np.array([[-1, 0., 0., 0], [-1, 0.01, -0.01, 0], [-1, 0.01, 0.01, 0],
                     [-1, -0.01, 0.01, 0], [-1, -0.01, -0.01, 0]])
    pol = np.tile([0., 0, 1, 0], (5, 1))
    mk = lambda: mx.PhotonTable({'pos': np.ones((5, 4)), 'dir': dirs.copy(), 'energy': np.ones(5),
                                 'polarization': pol, 'probability': np.ones(5)})
    p5 = mx.CATGrating(d=1. / 5000, order_selector=mx.OrderSelector([5]), zoom=2)(mk())
    m5 = mx.CATGrating(d=1. / 5000, order_selector=mx.OrderSelector([-5]), zoom=2)(mk())
    assert np.all(p5['dir'][1:3, 1] > 0) and np.all(p5['dir'][3:, 1] < 0)
    assert np.all(m5['dir'][1:3, 1] < 0) and np.all(m5['dir'][3:, 1] > 0)


@pytest.mark.parametrize('name', ['flat_grid_1000', 'cat_rowland_200_2000', 'cat_stack_200_2000'])
def test_golden_vectors(gpu, name):
    """The FP64 HIP path against the committed oracle vectors (the flat
    grid's zero order on a pixel corner: pixel indices there may differ,
    counted, at most 0.5 % of the photons)."""
    from helpers import load_golden
    from marxs_amd import configs
    meta, els, ph, u, out = load_golden(name)
    inst = getattr(configs, meta['builder'])(**meta['kwargs'])
    inst = inst[0] if isinstance(inst, tuple) else inst
    got, st = gpu_trace(inst, ph, u, precision=64)
    compare(out, got, rtol=1e-9, near=near_pixel_boundary(out),
            max_near_frac=5e-3 if name.startswith('flat_grid') else 1e-3)


@pytest.mark.parametrize('grid', ['even', 'uneven', 'two'])
def test_efficiency_row_search_ties_and_ends(gpu, grid):
    """Nearest energy row (first on ties, grating.py:93-94) for energies on
    grid points, exactly half-way between rows (ties), off both ends, and for
    evenly and unevenly spaced tables (different search paths in the kernel)."""
    energies = {'even': np.round(np.arange(0.3, 2.0001, 0.1), 10),
                'uneven': np.array([0.3, 0.35, 0.5, 0.9, 1.0, 1.6, 2.0]),
                'two': np.array([0.5, 1.5])}[grid]
    orders = np.array([-2, -1, 0, 1])
    rng = np.random.default_rng(9)
    prob = rng.uniform(0.05, 0.2, (len(energies), len(orders)))
    eff = mx.EfficiencyFile.from_arrays(energies, prob, orders)
    mids = 0.5 * (energies[1:] + energies[:-1])
    e = np.concatenate([energies, mids, [0.01, 0.29, 2.5, 10.], rng.uniform(0, 3, 2000)])
    n = len(e)
    ph, u = _photons(n, 10, spread=0.0)
    ph['energy'] = e
    ph['pos'][:, 1:] = rng.uniform(-5, 5, (n, 2))
    grat = mx.FlatGrating(d=1e-3, order_selector=eff, zoom=[1, 20, 20])
    ref, got, st = _single([grat], ph, u[:2])
    assert np.array_equal(ref['order'], got['order'])


def _interp_table():
    import os
    path = os.path.join(os.path.dirname(__file__), 'golden', 'grating_efficiency.csv')
    return mx.InterpolateEfficiencyTable(path)


def test_interpolated_efficiency_table_parity(gpu):
    """InterpolateEfficiencyTable (catgrating.py:63-144, degree 1) through the
    FP64 kernel: orders, probabilities and blaze angles against the oracle,
    for wavelengths and blaze angles inside and beyond the table."""
    ph, u = _photons(6000, 11, spread=0.04)
    ph['energy'] = np.random.default_rng(12).uniform(0.3, 2.0, len(ph['energy']))
    # the table scaled to totals below 1 (its rows summing to exactly 1 can
    # interpolate to 1 + 1 ulp, which the reference rejects as a probability)
    t = _interp_table()
    t = mx.InterpolateEfficiencyTable.from_arrays(t.wave, t.theta, t.orders, 0.9 * t.data)
    grat = mx.CATGrating(d=2e-4, order_selector=t, zoom=[1, 40, 40],
                         orientation=axangle2mat([0, 0, 1], np.deg2rad(1.5)))
    ref, got, st = _single([grat], ph, u[:2])
    assert np.array_equal(ref['order'], got['order'])
    assert len(set(got['order'][got['order'] > -100].tolist())) == 3


def test_interpolated_table_in_use(gpu):
    """test_cat.py:108-117: test photons (blaze 0, clamped to the 1 deg row)
    at 1 keV go to order 0 with probability 0.5."""
    cat = mx.CATGrating(order_selector=_interp_table(), d=0.001)
    p = cat(mx.generate_test_photons(5000))
    assert np.isclose((np.asarray(p['order']) == 0).sum(), 2500, rtol=.05)


def test_cat_rowland_blaze_dependent_table_parity(gpu):
    """200 CAT facets whose order selector is a blaze-dependent
    InterpolateEfficiencyTable (the bench's interp workload at C2 size)."""
    inst, parts = configs.cat_rowland_instrument(200, efficiency=configs.blaze_interp_efficiency())
    ref, got, st, rep = _parity(inst, 40000, 21, energies='lines')
    assert rep['order'][1] == 0 and rep['facet'][1] == 0
    assert len(set(got['order'][got['facet'] >= 0].tolist())) >= 12


@pytest.mark.parametrize('scale', [0.0, 0.5, 2.0])
def test_interpolated_total_estimate_only_guides(gpu, scale):
    """The table summed over the orders only estimates the total for picking
    the order in the first pass; the pick is then checked with the exact
    running sums and total.  With the summed table scaled by 0, 0.5 or 2 the
    check fails for (nearly) every photon and the second pass decides: FP64
    orders, probabilities and positions still equal the oracle's."""
    from marxs_amd.tracer import DeviceTables
    inst, parts = configs.cat_rowland_instrument(200, efficiency=configs.blaze_interp_efficiency())
    ci = _ci(inst)
    t = ci.tables
    k = int(np.nonzero(t.sel_i[:, 8] & 8)[0][0])
    nw, nt, o = int(t.sel_i[k, 0]), int(t.sel_i[k, 9]), int(t.sel_i[k, 11])
    assert o >= 0
    t.sel_d[o:o + nw * nt] *= scale
    ci.dev_tables = DeviceTables(t, ci.device)
    ph, u = make_inputs(20000, 23, ci.n_uniform_slots, 'lines')
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    rep = compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert rep['order'][1] == 0 and rep['facet'][1] == 0
    assert len(set(got['order'][got['facet'] >= 0].tolist())) >= 10


def test_blaze_sharing_kernel_equals_plain(gpu):
    """Tables with a blaze-dependent selector (blaze_stages > 0) run the
    plain kernels that form the blaze angle once for the selector and the
    diffraction.  The kernels that form it in the diffraction, run on the
    same tables with blaze_stages cleared, give FP64 columns identical bit
    for bit, and FP32 orders equal for all but float-rounding cases."""
    inst, parts = configs.cat_rowland_instrument(200, efficiency=configs.blaze_interp_efficiency())
    ci = _ci(inst)
    assert ci.dev_tables.struct.blaze_stages == 1
    ph, u = make_inputs(30000, 24, ci.n_uniform_slots, 'lines')
    out = {}
    for bz in (1, 0):
        ci.dev_tables.struct.blaze_stages = bz
        for prec in (64, 32):
            out[bz, prec], st = gpu_trace(ci, ph, u, precision=prec)
    ci.dev_tables.struct.blaze_stages = 1
    a, b = out[1, 64], out[0, 64]
    assert set(a) == set(b)
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=a[k].dtype.kind == 'f'), k
    assert (a['facet'] >= 0).sum() > 3000
    a, b = out[1, 32], out[0, 32]
    same = (a['facet'] == b['facet']) & (a['facet'] >= 0)
    assert same.sum() > 3000
    assert float((a['order'] == b['order'])[same].mean()) >= 0.999


def test_cat_rowland_blaze_dependent_table_fp32(gpu):
    """The FP32 kernel reads the interpolated table from its LDS copy (values
    rounded to float32): against the oracle, orders agree for all but a few
    photons whose draw falls within float rounding of a cumulative
    probability, and the detector position within 0.5 um RMS."""
    inst, parts = configs.cat_rowland_instrument(200, efficiency=configs.blaze_interp_efficiency())
    ci = _ci(inst)
    assert ci.tables.sel_i[0, 6] >= 0          # staged in the FP32 LDS image
    ph, u = make_inputs(40000, 22, ci.n_uniform_slots, 'lines')
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=32)
    same = (got['facet'] == ref['facet']) & (ref['facet'] >= 0)
    agree = float((got['order'] == ref['order'])[same].mean())
    both = same & (got['order'] == ref['order']) & (got['CCD'] == ref['CCD']) & (ref['CCD'] >= 0)
    err = np.hypot(got['det_x'] - ref['det_x'], got['det_y'] - ref['det_y'])[both] * 1e3
    print(f'FP32 interp: order agreement {agree:.6f}, rms {np.sqrt(np.mean(err ** 2)):.4f} um')
    assert same.sum() > 5000 and agree >= 0.999
    assert both.sum() > 1000 and np.sqrt(np.mean(err ** 2)) <= 0.5


# ------------------------------------------------------- lens reflectivity
def _refl_spline(f, x, y):
    from scipy.interpolate import RectBivariateSpline
    xg, yg = np.meshgrid(x, y, indexing='ij')
    return RectBivariateSpline(x, y, f(xg, yg), kx=1, ky=1)


def test_perfect_lens_reflectivity_known_answer(gpu):
    """marxs/optics/tests/test_mirror.py:62-80 through the GPU: probability
    0.367205 for the reference's test photon and spline."""
    import marxs_amd as mx
    interp = _refl_spline(lambda x, y: np.exp(-np.sqrt((x / 2) ** 2 + y ** 2)),
                          np.linspace(-3, 3, 100), np.linspace(-3, 3, 100))
    lens = mx.PerfectLens(focallength=123.0, reflectivity_interpolator=interp)
    ph = {'pos': np.array([[1., 0, 0]]), 'dir': np.array([[-1., 0, 0]]),
          'pol': np.array([[0., 1, 0]]), 'energy': np.ones(1), 'prob': np.ones(1)}
    for prec in (64, 32):
        got, st = gpu_trace(mx.Sequence(elements=[lens]), ph, np.zeros((0, 1)), precision=prec)
        assert np.allclose(got['energy'], 1)
        assert np.allclose(got['pol'], [0, 1, 0])
        assert np.allclose(got['prob'], 0.367205)


def test_perfect_lens_reflectivity_parity(gpu):
    """A tabulated reflectivity (energy 0.2-2.5 keV x angle 0-0.02 rad) on the
    200-facet workload's lens: FP64 equal to the oracle, which calls scipy's
    RectBivariateSpline itself (indices exact, floats 1e-9 relative); FP32
    probabilities within 1e-5 relative."""
    from marxs_amd.tracer import CompiledInstrument
    rng = np.random.default_rng(12)
    e, a = np.round(np.arange(0.2, 2.5001, 0.1), 10), np.linspace(0., 0.02, 21)
    vals = (0.95 - 0.1 * e[:, None] - 8. * a[None, :]) * (1 - 0.02 * rng.random((len(e), len(a))))
    from scipy.interpolate import RectBivariateSpline
    interp = RectBivariateSpline(e, a, vals, kx=1, ky=1)
    inst, parts = configs.cat_rowland_instrument(200)
    parts['lens'].reflectivity_interpolator = interp
    ci = _ci(inst)
    ph, u = make_inputs(30000, 13, ci.n_uniform_slots)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert ref['prob'].max() < 0.95 and (ref['prob'] > 0).sum() > 1000
    g32, _ = gpu_trace(ci, ph, u, precision=32)
    ok = (ref['facet'] == g32['facet']) & (ref['CCD'] == g32['CCD'])
    rel = np.abs(g32['prob'] - ref['prob'])[ok] / np.maximum(ref['prob'][ok], 1e-300)
    assert float(rel.max()) < 1e-5


def test_perfect_lens_reflectivity_uneven_grid(gpu):
    """Reflectivity on an uneven (geometric) energy grid and an uneven angle
    grid, with photon energies on the grid points, one ulp either side and
    beyond both ends: the kernel's interval search (a guess for an evenly
    spaced grid, checked, else bisection) against scipy's spline in the
    oracle, FP64 floats within 1e-9 relative."""
    from scipy.interpolate import RectBivariateSpline
    rng = np.random.default_rng(31)
    e = np.geomspace(0.2, 2.5, 17)
    a = np.array([0., 0.001, 0.0025, 0.003, 0.006, 0.02])
    vals = (0.95 - 0.1 * e[:, None] - 8. * a[None, :]) * (1 - 0.02 * rng.random((len(e), len(a))))
    interp = RectBivariateSpline(e, a, vals, kx=1, ky=1)
    inst, parts = configs.cat_rowland_instrument(200)
    parts['lens'].reflectivity_interpolator = interp
    ci = _ci(inst)
    n = 30000
    ph, u = make_inputs(n, 32, ci.n_uniform_slots)
    en = np.concatenate([e, np.nextafter(e, 0), np.nextafter(e, 10), [0.1, 3.0]])
    ph['energy'] = en[rng.integers(0, len(en), n)]
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert (ref['prob'] > 0).sum() > 1000


def test_perfect_lens_reflectivity_needs_extended_kernel(gpu):
    """Tables with a reflectivity record select the extended kernel; the
    plain kernel, given such tables with ext_stages cleared, leaves the
    probability alone and reports the lens in the bad-argument counter."""
    import torch
    import marxs_amd as mx
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    interp = _refl_spline(lambda x, y: np.exp(-np.sqrt((x / 2) ** 2 + y ** 2)),
                          np.linspace(-3, 3, 100), np.linspace(-3, 3, 100))
    ci = CompiledInstrument(mx.Sequence(elements=[mx.PerfectLens(focallength=123.0,
                                                                 reflectivity_interpolator=interp)]))
    assert ci.dev_tables.struct.ext_stages == 1
    ci.dev_tables.struct.ext_stages = 0
    for prec in (64, 32):
        dp = DevicePhotons(np.array([[-1., 0, 0]] * 3).T, np.array([[0., 1, 0]] * 3).T, np.ones(3),
                           np.ones(3), device=gpu, pos=np.array([[1., 0, 0]] * 3).T)
        st = ci.launch(dp, torch.zeros((0, 3), dtype=torch.float64, device=gpu), prec).cpu().numpy()
        assert st[5] == 3
        assert np.array_equal(dp.prob.cpu().numpy(), np.ones(3))


def test_perfect_lens_reflectivity_rejected_forms(gpu):
    import marxs_amd as mx
    from marxs_amd.compiler import CompileError
    from scipy.interpolate import RectBivariateSpline
    x = np.linspace(0, 1, 5)
    for bad in (lambda e, a, grid=False: np.ones_like(e),
                RectBivariateSpline(x, x, np.ones((5, 5)), kx=2, ky=2)):
        with pytest.raises(CompileError):
            _ci(mx.Sequence(elements=[mx.PerfectLens(focallength=10., reflectivity_interpolator=bad)]))


def test_plane_test_counter(gpu):
    """MX_FLAG_COUNT_TESTS counts the ray-plane tests: without the culling
    grid every photon tests every element of every array stage once (the
    scan after a hit goes on with the later elements), 1 + 200 + 16 per
    photon on the 200-facet workload; with the grid far fewer; counting
    changes no result."""
    import torch
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    inst, parts = configs.cat_rowland_instrument(200)
    n = 30011
    for grid in (False, True):
        ci = CompiledInstrument(inst, grid=grid)
        ph, u = make_inputs(n, 17, ci.n_uniform_slots)
        outs = []
        for ct in (True, False):
            dp = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu)
            st = ci.launch(dp, torch.as_tensor(u, device=gpu), 64, count_tests=ct).cpu().numpy()
            outs.append((dp, st))
        (a, sa), (b, sb) = outs
        assert torch.equal(a.pos, b.pos) and torch.equal(a.icols['CCD'], b.icols['CCD'])
        assert sb[7] == 0
        if not grid:
            assert sa[7] == n * (1 + 200 + 16)
        else:
            assert n < sa[7] < 0.05 * n * (1 + 200 + 16)


@pytest.mark.parametrize('prec', [64, 32])
def test_every_photon_traced_once_any_grid(gpu, prec):
    """The waves of a workgroup take 64-photon chunks of its contiguous share
    of the photons: for photon counts around the chunk and workgroup sizes
    and grids of 1 to 8 workgroups per CU (and more workgroups than photon
    chunks), every photon is traced once (output columns filled with a
    sentinel first are all overwritten) and the results equal those of one
    workgroup per CU bit for bit (FP64) or exactly (FP32: the same kernel
    code runs for every photon, only the grid changes)."""
    import torch
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    from marxs_amd.batch import outputs_like
    ci = CompiledInstrument(configs.cat_rowland_instrument(200)[0])
    for n in (1, 63, 64, 65, 1023, 1025, 64 * 256 + 1, 200_003):
        ph, u = make_inputs(n, 50 + n % 97, ci.n_uniform_slots, 'lines')
        src = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu)
        uu = torch.as_tensor(u, device=gpu)
        ref = None
        for wpc in (1, 2, 3, 8):
            out = outputs_like(src)
            for t in (out.pos, out.dir, out.pol, out.prob):
                t.fill_(-7.5)
            st = ci.launch(out, uu, prec, src=src, wg_per_cu=wpc).cpu().numpy()
            cols = {'pos': out.pos, 'dir': out.dir, 'pol': out.pol, 'prob': out.prob}
            cols.update(out.fcols)
            cols.update(out.icols)
            cols = {k: v.cpu().numpy() for k, v in cols.items()}
            for k in ('pos', 'dir', 'pol', 'prob'):
                assert not np.any(cols[k] == -7.5), (n, wpc, k)
            if ref is None:
                ref = (cols, st)
                continue
            for k in cols:
                assert np.array_equal(cols[k], ref[0][k], equal_nan=cols[k].dtype.kind == 'f'), (n, wpc, k)
            assert np.array_equal(st, ref[1]), (n, wpc)
