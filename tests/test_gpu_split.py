"""The split trace (mx_trace_split: the plain kernel up to each photon's
first hit of the stack stage, the layer programs and later stages over the
queued photons; opt-in, CompiledInstrument(split=True)) against the
one-kernel trace of the same tables (needs an MI355X).

Photons are independent and the split changes only which kernel runs which
part of a photon's trace.  FP64 (no contraction, the reference's operation
order in both kernels): every output column and every status counter equal
bit for bit.  FP32: the two kernels contract float multiply-adds of the
culling and scan differently, so float columns differ in the last bits
(detector coordinates by up to 0.2 um, a few thousandths of the FP32 path's
own error budget of 0.5 um against FP64); element, order and pixel indices
and the status counters stay equal.  test_split_oracle_parity_fp64
checks the split FP64 trace against the oracle directly."""
import os

import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import configs, catgrating as CG
from tests.helpers import make_inputs

pytestmark = pytest.mark.gpu

SI = os.path.join(os.path.dirname(__file__), 'golden', 'SiTransmission.csv')


def _both(ci, ph, u, prec, gpu, src=True, count_tests=False):
    """Trace the same inputs fused and split; returns both photon sets and
    status vectors."""
    import torch
    from marxs_amd.tracer import DevicePhotons
    uu = torch.as_tensor(np.ascontiguousarray(u), dtype=torch.float64, device=gpu)
    out = []
    for split in (False, True):
        pos = ph['pos'].T if 'pos' in ph else None
        s = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], pos=pos, device=gpu)
        if src:
            d = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], pos=pos, device=gpu)
            st = ci.launch(d, uu, prec, src=s, split=split, count_tests=count_tests)
        else:
            d = s
            st = ci.launch(d, uu, prec, split=split, count_tests=count_tests)
        torch.cuda.synchronize()
        out.append((d, st.cpu().numpy()))
    return out


def _assert_close32(a, b, max_int_frac=1e-4):
    """FP32: indices and counters equal (up to max_int_frac of the photons,
    for ties at an edge), float columns within float rounding: positions and
    detector coordinates 1 um, directions 1e-6, probabilities 1e-6
    relative."""
    import torch
    (da, sa), (db, sb) = a, b
    n = da.n
    assert np.array_equal(sa, sb), (sa, sb)
    for name in da.icols:
        assert int((da.icols[name] != db.icols[name]).sum()) <= max_int_frac * n, name
    same = torch.ones(n, dtype=torch.bool, device=da.pos.device)
    for name in da.icols:
        same &= da.icols[name] == db.icols[name]

    def close(x, y, atol, rtol=0.):
        ok = (torch.abs(x - y) <= atol + rtol * torch.abs(x)) | (torch.isnan(x) & torch.isnan(y))
        return bool(ok[..., same].all())
    assert close(da.pos, db.pos, 1e-3)
    assert close(da.dir, db.dir, 1e-6)
    assert close(da.prob, db.prob, 0., 1e-6)
    for name in ('det_x', 'det_y', 'y', 'z', 'grat_y', 'grat_z'):
        if name in da.fcols:
            assert close(da.fcols[name], db.fcols[name], 1e-3), name


def _assert_equal(a, b):
    import torch
    (da, sa), (db, sb) = a, b
    for name in ('pos', 'dir', 'pol', 'prob'):
        x, y = getattr(da, name), getattr(db, name)
        assert torch.equal(torch.nan_to_num(x, nan=7.25), torch.nan_to_num(y, nan=7.25)), name
        assert torch.equal(torch.isnan(x), torch.isnan(y)), name
    assert set(da.fcols) == set(db.fcols) and set(da.icols) == set(db.icols)
    for name in da.fcols:
        x, y = da.fcols[name], db.fcols[name]
        assert torch.equal(torch.isnan(x), torch.isnan(y)), name
        assert torch.equal(torch.nan_to_num(x, nan=7.25), torch.nan_to_num(y, nan=7.25)), name
    for name in da.icols:
        assert torch.equal(da.icols[name], db.icols[name]), name
    assert np.array_equal(sa, sb), (sa, sb)


@pytest.mark.parametrize('prec', [32, 64])
def test_split_equals_fused_stack_rowland(gpu, prec):
    """200 CATL1L2Stack facets, 16 CCDs, 60 011 photons (not a multiple of a
    workgroup): equal columns and counters, plane-test counts included."""
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(200, facet_class=CG.CATL1L2Stack,
                                                 facet_kwargs={'l1_transmission': SI})
    ci = CompiledInstrument(inst, split=True)
    assert ci.tables.split_stage == 2 and ci.uses_split()
    ph, u = make_inputs(60011, 21, ci.n_uniform_slots, 'lines')
    a, b = _both(ci, ph, u, prec, gpu, count_tests=True)
    (_assert_equal if prec == 64 else _assert_close32)(a, b)
    d = b[0]
    assert int((d.icols['facet'] >= 0).sum()) > 5000 and int((d.icols['CCD'] >= 0).sum()) > 1000
    assert b[1][7] > 0


@pytest.mark.parametrize('prec', [32, 64])
def test_split_multi_hit_and_in_place(gpu, prec):
    """Two stack elements one behind the other (every photon through both:
    the second kernel resumes the scan after the queued hit) and a detector;
    traced in place."""
    from marxs_amd.tracer import CompiledInstrument
    sel = mx.OrderSelector([-1, 0, 1], [0.3, 0.4, 0.3])
    a = CG.CATL1L2Stack(order_selector=sel, l1_transmission=SI, position=[60, 0, 0], zoom=[1, 20, 20])
    b = CG.CATL1L2Stack(order_selector=sel, l1_transmission=SI, position=[50, 0, 0], zoom=[1, 20, 20])
    det = mx.FlatDetector(pixsize=0.01, position=[0, 0, 0], zoom=[1, 30, 30])
    ci = CompiledInstrument(mx.Sequence(elements=[a, b, det]), n_draws=3, split=True)
    assert ci.tables.split_stage == 0 and ci.uses_split()
    n = 5003
    rng = np.random.default_rng(3)
    pos = np.column_stack([np.full(n, 100.), rng.uniform(-25, 25, n), rng.uniform(-25, 25, n)])
    ph = {'pos': pos, 'dir': np.tile([-1., 0, 0], (n, 1)), 'pol': np.tile([0., 1, 0], (n, 1)),
          'energy': rng.uniform(0.3, 2., n), 'prob': np.ones(n)}
    u = rng.random((ci.n_uniform_slots, n))
    ra, rb = _both(ci, ph, u, prec, gpu, src=False)
    (_assert_equal if prec == 64 else _assert_close32)(ra, rb)
    st = rb[1]
    assert st[4] > 1000      # photons that met both stacks


def test_split_without_culling_grid(gpu):
    """Every element tested (no culling grid): the first kernel's scan and
    the second kernel's continuation run over all elements; FP64 equal."""
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(50, facet_class=CG.CATL1L2Stack,
                                                 facet_kwargs={'l1_transmission': SI})
    ci = CompiledInstrument(inst, grid=False, split=True)
    ph, u = make_inputs(20011, 24, ci.n_uniform_slots, 'lines')
    a, b = _both(ci, ph, u, 64, gpu)
    _assert_equal(a, b)
    assert int((b[0].icols['facet'] >= 0).sum()) > 500


def test_split_no_hits_and_empty(gpu):
    """Photons that meet no facet (the queue stays empty) and a zero-length
    trace."""
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(50, facet_class=CG.CATL1L2Stack,
                                                 facet_kwargs={'l1_transmission': SI})
    ci = CompiledInstrument(inst, split=True)
    ph, u = make_inputs(3000, 22, ci.n_uniform_slots)
    ph['dir'][:] = [1., 0, 0]     # away from the instrument: the aperture passes none
    for prec in (32, 64):
        a, b = _both(ci, ph, u, prec, gpu)
        (_assert_equal if prec == 64 else _assert_close32)(a, b)
        assert int((b[0].icols['facet'] >= 0).sum()) == 0
    e = {k: v[:0] for k, v in ph.items()}
    a, b = _both(ci, e, u[:, :0], 32, gpu)
    assert b[0].n == 0


def test_split_not_used_where_tables_do_not_allow(gpu):
    """Radial mirror scatter stacks, tube detectors, several stack stages or
    regrouping keep the one-kernel trace."""
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(50)
    assert CompiledInstrument(inst, split=True).tables.split_stage is None
    assert not CompiledInstrument(inst, split=True).uses_split()
    inst, parts = configs.cat_rowland_instrument(50, facet_class=CG.CATL1L2Stack,
                                                 facet_kwargs={'l1_transmission': SI})
    assert not CompiledInstrument(inst, regroup=True, split=True).uses_split()
    assert not CompiledInstrument(inst, split=False).uses_split()
    # the default is the one-kernel trace
    ci = CompiledInstrument(inst)
    assert not ci.uses_split(precision=64) and not ci.uses_split(precision=32)


def test_split_oracle_parity_fp64(gpu):
    """The split FP64 trace of 200 CATL1L2Stack facets against the oracle:
    indices exact (pixel edges aside), floats within 1e-9 relative."""
    from marxs_amd.tracer import CompiledInstrument
    from tests.helpers import gpu_trace, oracle_trace, compare, near_pixel_boundary
    inst, parts = configs.cat_rowland_instrument(200, facet_class=CG.CATL1L2Stack,
                                                 facet_kwargs={'l1_transmission': SI})
    ci = CompiledInstrument(inst, split=True)
    ph, u = make_inputs(30000, 23, ci.n_uniform_slots)
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    rep = compare(ref, got, rtol=1e-9, near=near_pixel_boundary(ref))
    assert rep['facet'][1] == 0 and rep['order'][1] == 0 and rep['order_L1'][1] == 0
    assert (got['facet'] >= 0).sum() > 4000
