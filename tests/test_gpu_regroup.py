"""Regrouping rays between stages (needs an MI355X).

With regrouping on (off by default), before every culled array stage the
trace kernel exchanges rays between the threads of a workgroup so that waves
are all interacting or all passing, every ray at most max_frac of a
workgroup's rays can interact (exchange or not decided per chunk).
Which thread carries a photon must not change anything: every output
column and every status counter equals the trace without regrouping bit for
bit, in both precisions, for photon counts that leave a partial last chunk,
and the regrouping kernel is the one that runs on these workloads."""
import numpy as np
import pytest

from helpers import make_inputs

pytestmark = pytest.mark.gpu


def _columns(dp):
    out = {'pos': dp.pos, 'dir': dp.dir, 'pol': dp.pol, 'prob': dp.prob}
    out.update(dp.fcols)
    out.update(dp.icols)
    return {k: v.cpu().numpy() for k, v in out.items()}


def _both(ci, n, seed, precision, energies='uniform', gpu=None):
    import torch
    from marxs_amd.tracer import DevicePhotons
    ph, u = make_inputs(n, seed, ci.n_uniform_slots, energies)
    res = {}
    for rg in (True, False, 'half'):
        dp = DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu)
        uni = torch.as_tensor(u, device=gpu)
        kw = {'regroup': True, 'regroup_max_frac': 0.5} if rg == 'half' else {'regroup': rg}
        st = ci.launch(dp, uni, precision, **kw)
        res[rg] = (_columns(dp), st.cpu().numpy())
    return res


def _assert_same(res):
    b = res[False][0]
    for key in (True, 'half'):
        a = res[key][0]
        assert a.keys() == b.keys()
        for k in a:
            assert np.array_equal(a[k], b[k], equal_nan=a[k].dtype.kind == 'f'), (key, k)
        assert np.array_equal(res[key][1], res[False][1])


def _instruments():
    from marxs_amd import configs
    return {'flat_grid': configs.flat_grid_instrument(),
            'cat200': configs.cat_rowland_instrument(200)[0],
            'cat600': configs.cat_rowland_instrument(600)[0]}


@pytest.mark.parametrize('name', ['flat_grid', 'cat200', 'cat600'])
@pytest.mark.parametrize('precision', [64, 32])
def test_regrouped_trace_equals_plain_trace(gpu, name, precision):
    from marxs_amd.tracer import CompiledInstrument
    ci = CompiledInstrument(_instruments()[name])
    assert ci.regroup_pairs(precision, regroup=True) >= 64
    assert ci.regroup_pairs(precision) == 0
    res = _both(ci, 100_003, 7, precision, 'lines' if name == 'cat600' else 'uniform', gpu)
    _assert_same(res)
    cols = res[True][0]
    assert (cols['facet'] >= 0).sum() > 10_000


@pytest.mark.parametrize('precision', [64, 32])
def test_regrouped_stack_trace_equals_plain_trace(gpu, precision):
    """The extended kernel (CATL1L2Stack facets: layer program, two more
    columns, five draws per interaction) regroups before the stack stage."""
    from marxs_amd import configs, catgrating
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(200, facet_class=catgrating.CATL1L2Stack)
    ci = CompiledInstrument(inst)
    assert ci.regroup_pairs(precision, regroup=True) >= 64
    _assert_same(_both(ci, 50_001, 8, precision, gpu=gpu))


def test_small_and_single_chunk(gpu):
    """Fewer photons than one workgroup, and exactly one workgroup."""
    from marxs_amd.tracer import CompiledInstrument
    ci = CompiledInstrument(_instruments()['cat200'])
    for n in (1, 63, 700, 1024, 1025):
        for prec in (64, 32):
            _assert_same(_both(ci, n, 9 + n, prec, gpu=gpu))
