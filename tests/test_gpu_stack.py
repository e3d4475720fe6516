"""GPU parity of stack stages (FlatStack: CAT grating membrane, quality
factor, L1 cross-support, L2 absorption and diffraction; stand-alone
Gaussian scatter) against the oracle (needs an MI355X).

FP64: integer columns (facet, orders, CCD, pixels) bit-exact except photons
within 1e-6 px of a pixel edge; float columns within 1e-9 relative.  The
scatter's normal deviate is the inverse normal CDF of a uniform draw (HIP's
normcdfinv on the GPU, scipy's ndtri in the oracle; both within a few ulp).
"""
import os

import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import configs, catgrating as CG
from marxs_amd.scatter import RandomGaussianScatter
from tests.helpers import (gpu_trace, oracle_trace, compare, make_inputs,
                           near_pixel_boundary)

pytestmark = pytest.mark.gpu

RTOL = 1e-9
SI = os.path.join(os.path.dirname(__file__), 'golden', 'SiTransmission.csv')


def _stack_instrument(n=200, **kw):
    return configs.cat_rowland_instrument(n, facet_class=CG.CATL1L2Stack,
                                          facet_kwargs={'l1_transmission': SI, **kw})


def _ci(inst, **kw):
    from marxs_amd.tracer import CompiledInstrument
    return CompiledInstrument(inst, **kw)


@pytest.mark.parametrize('grid', [True, False])
def test_cat_stack_rowland_fp64_parity(gpu, grid):
    """200 CATL1L2Stack facets on the Rowland torus, 16 CCDs."""
    inst, parts = _stack_instrument()
    ci = _ci(inst, grid=grid)
    ph, u = make_inputs(40000, 11, ci.n_uniform_slots)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64, grid=grid)
    rep = compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert (got['facet'] >= 0).sum() > 5000
    assert (got['CCD'] >= 0).sum() > 1000
    assert rep['facet'][1] == 0 and rep['order'][1] == 0 and rep['order_L1'][1] == 0
    # photons through the L1 bars: order_L1 0 with the bar transmission
    assert ((got['order_L1'] == 0) & (got['facet'] >= 0)).sum() > 100


def test_cat_stack_line_spectrum_600(gpu):
    inst, parts = _stack_instrument(600)
    ci = _ci(inst)
    ph, u = make_inputs(40000, 12, ci.n_uniform_slots, 'lines')
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))


def test_cat_stack_fp32(gpu):
    """FP32 path: plane tests and layers in float; detector
    coordinates within 0.5 um RMS of the FP64 oracle."""
    inst, parts = _stack_instrument(600)
    ci = _ci(inst)
    ph, u = make_inputs(40000, 13, ci.n_uniform_slots, 'lines')
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=32)
    both = ((ref['CCD'] >= 0) & (got['CCD'] == ref['CCD']) & (got['facet'] == ref['facet'])
            & (got['order_L1'] == ref['order_L1']))
    err = np.hypot(got['det_x'] - ref['det_x'], got['det_y'] - ref['det_y'])[both] * 1e3
    rms = float(np.sqrt(np.mean(err ** 2)))
    print(f'FP32 stack: rms {rms:.4f} um over {both.sum()} photons')
    assert both.sum() > 3000
    assert rms <= 0.5
    same = got['facet'] == ref['facet']
    assert (got['order'] == ref['order'])[same].all()
    assert (got['order_L1'] == ref['order_L1'])[same].all()


def _photons(n, seed, spread=0.02):
    rng = np.random.default_rng(seed)
    d = np.column_stack([-np.ones(n), rng.normal(0, spread, (n, 2))])
    d /= np.linalg.norm(d, axis=1)[:, None]
    pol = np.cross(d, rng.normal(size=(n, 3)))
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    pos = np.column_stack([np.full(n, 100.), rng.uniform(-20, 20, (n, 2))])
    return {'pos': pos, 'dir': d, 'pol': pol, 'energy': rng.uniform(0.2, 3., n),
            'prob': np.ones(n)}


def _run(elements, n=20000, seed=1, n_draws=2, precision=64):
    inst = mx.Sequence(elements=elements)
    ci = _ci(inst, n_draws=n_draws)
    ph = _photons(n, seed)
    u = np.random.default_rng(seed + 100).random((ci.n_uniform_slots, n))
    ref = oracle_trace(inst, ph, u, n_draws=n_draws)
    got, st = gpu_trace(ci, ph, u, precision=precision)
    rep = compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    return ref, got, st, rep


def test_standalone_layers(gpu):
    """Each layer as an element of its own, in sequence, then a detector."""
    sel = mx.OrderSelector([-2, -1, 0, 1, 2], [0.1, 0.2, 0.3, 0.2, 0.1])
    els = [CG.L1(order_selector=CG.l1_order_selector, l1_transmission=SI,
                 position=[50, 0, 0], zoom=[1, 30, 30]),
           CG.L2Abs(position=[40, 0, 0], zoom=[1, 30, 30],
                    orientation=mx.mathutils.axangle2mat([0, 1, 0], 0.3)),
           RandomGaussianScatter(scatter=1e-4, position=[30, 0, 0], zoom=[1, 30, 30]),
           CG.L2Diffraction(position=[20, 0, 0], zoom=[1, 30, 30]),
           RandomGaussianScatter(scatter=0., position=[10, 0, 0], zoom=[1, 30, 30]),
           mx.FlatDetector(position=[0, 0, 0], zoom=[1, 60, 60], pixsize=0.01)]
    ref, got, st, rep = _run(els)
    assert 'scatter' in got and 'L2Diffraction' in got and 'order_L1' in got
    assert np.isfinite(got['det_x']).sum() > 10000


def test_stack_with_interpolated_efficiency(gpu):
    """Membrane with an InterpolateEfficiencyTable (blaze-dependent)."""
    wave = np.linspace(0.4, 6.5, 9)
    theta = np.deg2rad(np.linspace(0, 3, 7))
    orders = np.arange(-3, 2)
    data = np.random.default_rng(5).uniform(0.01, 0.18, (len(orders), len(wave), len(theta)))
    tab = mx.InterpolateEfficiencyTable.from_arrays(wave, theta, orders, data)
    st = CG.CATL1L2Stack(order_selector=tab, l1_transmission=SI, zoom=[1, 25, 25],
                         orientation=mx.mathutils.axangle2mat([0, 0, 1], np.deg2rad(1.5)),
                         position=[50, 0, 0], id_col='facet', id_num=3)
    det = mx.FlatDetector(position=[0, 0, 0], zoom=[1, 60, 60], pixsize=0.01)
    ref, got, st_, rep = _run([st, det])
    assert (got['facet'] == 3).sum() > 10000


def test_stack_multiple_hits_and_draw_reuse(gpu):
    """Two stacks in one stage hit one after the other; one draw set per
    interaction (n_draws=1: the second interaction reuses slot 0..4)."""
    sel = mx.OrderSelector([-1, 0, 1], [0.3, 0.4, 0.3])
    mk = lambda x: CG.CATL1L2Stack(order_selector=sel, l1_transmission=SI, zoom=[1, 30, 30],
                                   position=[x, 0, 0], id_col='facet', id_num=int(x))
    det = mx.FlatDetector(position=[0, 0, 0], zoom=[1, 60, 60], pixsize=0.01)
    for n_draws in (1, 2):
        ref, got, st, rep = _run([mk(60.), mk(50.), det], n_draws=n_draws)
        assert (got['facet'] == 50).sum() > 10000
        assert st['multi_hit'] > 10000
        assert (st['slot_reuse'] > 0) == (n_draws == 1)


def test_transmission_out_of_range_raises(gpu):
    """Energies outside the L1 table raise like scipy's interp1d."""
    l1 = CG.L1(order_selector=mx.OrderSelector([1]), l1_transmission=(np.array([0.1, 1.0]),
                                                                      np.array([0.2, 0.6])),
               l1_dims={'bardepth': 0.001, 'period': 0.005, 'barwidth': 0.004},
               position=[50, 0, 0], zoom=[1, 30, 30])
    inst = mx.Sequence(elements=[l1])
    ci = _ci(inst)
    ph = _photons(2000, 3)
    u = np.random.default_rng(7).random((ci.n_uniform_slots, 2000))
    with pytest.raises(ValueError, match='interpolation range'):
        oracle_trace(inst, ph, u)
    with pytest.raises(ValueError, match='interpolation range'):
        gpu_trace(ci, ph, u, precision=64)
    ph['energy'] = np.random.default_rng(8).uniform(0.2, 0.9, 2000)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL)


def test_batched_stack_variants_equal_single(gpu):
    """Stack instruments in the batched kernel: each variant's result equals
    its single trace bit for bit (FP64)."""
    import torch
    from marxs_amd.batch import BatchedTrace, outputs_like
    from marxs_amd.tracer import DevicePhotons
    cis, inputs, unis = [], [], []
    for k, n in enumerate((60, 150)):
        ci = _ci(_stack_instrument(n)[0])
        ph, u = make_inputs(20000, 20 + k, ci.n_uniform_slots)
        cis.append(ci)
        inputs.append(DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu))
        unis.append(torch.as_tensor(u, device=gpu))
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    bt.run()
    for v, ci in enumerate(cis):
        dp = outputs_like(inputs[v])
        ci.launch(dp, unis[v], 64, src=inputs[v])
        a = {'pos': dp.pos, 'dir': dp.dir, 'prob': dp.prob, **dp.fcols, **dp.icols}
        b = bt.outputs[v]
        b = {'pos': b.pos, 'dir': b.dir, 'prob': b.prob, **b.fcols, **b.icols}
        for k in a:
            x, y = a[k].cpu().numpy(), b[k].cpu().numpy()
            assert np.array_equal(x, y, equal_nan=x.dtype.kind == 'f'), k


def test_radial_mirror_scatter_and_catsupportbars(gpu):
    """Aperture, lens, RadialMirrorScatter at the lens (blur in and across
    the plane of reflection), CAT stack facets, catsupportbars, CCDs."""
    from marxs_amd.scatter import RadialMirrorScatter
    inst0, parts = _stack_instrument(200)
    rms = RadialMirrorScatter(inplanescatter=2e-5, perpplanescatter=4e-6,
                              position=[12500., 0, 0], zoom=[1, 700, 700])
    inst = mx.Sequence(elements=[parts['aperture'], parts['lens'], rms, parts['gratings'],
                                 CG.catsupportbars, parts['ccds']])
    ci = _ci(inst)
    ph, u = make_inputs(40000, 31, ci.n_uniform_slots)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
    assert np.all(got['prob'][got['facet'] < 0] == 0)
    assert (got['inplanescatter'] != 0).all() and (got['perpplanescatter'] != 0).all()
    got32, st32 = gpu_trace(ci, ph, u, precision=32)
    both = (ref['CCD'] >= 0) & (got32['CCD'] == ref['CCD']) & (got32['facet'] == ref['facet'])
    err = np.hypot(got32['det_x'] - ref['det_x'], got32['det_y'] - ref['det_y'])[both] * 1e3
    assert both.sum() > 1000 and np.sqrt(np.mean(err ** 2)) <= 0.5


@pytest.mark.parametrize('sig', [(0.1, 0.), (0.1, 0.01), (0., 0.1), (0., 0.)])
def test_radial_scatter_sigmas(gpu, sig):
    """The reference's scatter test set-ups (test_scatter.py:11-75) on the
    GPU: photons at (0, 1, 0) and (0, 0, 1), detector 1 mm behind."""
    from marxs_amd.scatter import RadialMirrorScatter
    det = mx.FlatDetector(position=[-1, 0, 0], zoom=1000)
    for pos in ([0., 1., 0.], [0., 0., 1.]):
        inst = mx.Sequence(elements=[RadialMirrorScatter(inplanescatter=sig[0],
                                                         perpplanescatter=sig[1]), det])
        ci = _ci(inst)
        t = mx.generate_test_photons(2000)
        ph = {'pos': np.tile(np.array(pos), (2000, 1)),
              'dir': np.asarray(t['dir'], dtype=float)[:, :3].copy(),
              'pol': np.asarray(t['polarization'], dtype=float)[:, :3].copy(),
              'energy': np.asarray(t['energy'], dtype=float),
              'prob': np.asarray(t['probability'], dtype=float)}
        u = np.random.default_rng(9).random((ci.n_uniform_slots, 2000))
        ref = oracle_trace(inst, ph, u)
        got, st = gpu_trace(ci, ph, u, precision=64)
        compare(ref, got, rtol=RTOL, near=near_pixel_boundary(ref))
        if pos[1] == 1. and sig[0] == 0.:
            assert np.all(got['det_x'] == 1.0)


def test_quality_factor_on_its_own_reads_the_order_column(gpu):
    """catgrating.py:158-160: probability * factor ** (order ** 2) from the
    photon's 'order' column; NaN order gives NaN; without the column the
    reference's KeyError."""
    qf = CG.QualityFactor()
    n = 1000
    rng = np.random.default_rng(3)
    p = mx.generate_test_photons(n)
    order = rng.integers(-6, 7, n).astype(float)
    order[::97] = np.nan
    p['order'] = order
    p['probability'] = np.full(n, 0.5)
    out = qf(p)
    want = 0.5 * qf.factor ** (order ** 2)
    got = np.asarray(out['probability'])
    assert np.allclose(got, want, rtol=1e-14, equal_nan=True)
    assert np.isnan(got[::97]).all()
    with pytest.raises(KeyError):
        qf(mx.generate_test_photons(5))
