"""Batched trace of many instrument variants in one launch (needs an MI355X).

Each variant's batched result must equal the single-variant trace of the
same variant bit for bit in FP64 (and the oracle, through the single-trace
parity), whatever the tile plan; hipGraph replays must repeat the eager
result; the autotuner must leave a plan that still gives the same result."""
import numpy as np
import pytest

from helpers import make_inputs, oracle_trace, compare, near_pixel_boundary

pytestmark = pytest.mark.gpu

SIZES = [20000, 7001, 30000, 0, 12345]


@pytest.fixture(scope='module')
def study(gpu):
    import torch
    from marxs_amd import configs
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    variants = configs.design_variants(len(SIZES) - 1, np.random.default_rng(11),
                                       facet_range=(50, 300))
    trees = [v[0] for v in variants] + [configs.flat_grid_instrument()]
    cis, inputs, unis, hosts = [], [], [], []
    for k, (tree, n) in enumerate(zip(trees, SIZES)):
        ci = CompiledInstrument(tree)
        ph, u = make_inputs(n, 100 + k, ci.n_uniform_slots)
        cis.append(ci)
        inputs.append(DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu))
        unis.append(torch.as_tensor(u, device=gpu))
        hosts.append((tree, ph, u))
    return cis, inputs, unis, hosts


def _single(ci, src, u, precision):
    from marxs_amd.batch import outputs_like
    dp = outputs_like(src)
    st = ci.launch(dp, u, precision, src=src)
    return dp, st


def _columns(dp):
    out = {'pos': dp.pos, 'dir': dp.dir, 'pol': dp.pol, 'prob': dp.prob}
    out.update(dp.fcols)
    out.update(dp.icols)
    return {k: v.cpu().numpy() for k, v in out.items()}


def _assert_identical(a, b):
    for k in a:
        assert np.array_equal(a[k], b[k], equal_nan=a[k].dtype.kind == 'f'), k


def test_batched_fp64_equals_single_variant_traces(study):
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = study
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    st = bt.run()
    for v, ci in enumerate(cis):
        dp, sst = _single(ci, inputs[v], unis[v], 64)
        _assert_identical(_columns(dp), _columns(bt.outputs[v]))
        assert st[v]['multi_hit'] == int(sst.cpu()[4])
    assert len(bt.outputs[3].prob) == 0


def test_batched_variant_against_oracle(study):
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = study
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    bt.run()
    v = 1
    tree, ph, u = hosts[v]
    ref = oracle_trace(tree, ph, u)
    got = _columns(bt.outputs[v])
    got = {k: (x.T if x.ndim == 2 else x) for k, x in got.items()}
    got = {k: (x.astype(np.int64) if x.dtype.kind == 'i' else x) for k, x in got.items()}
    compare(ref, got, rtol=1e-9, near=near_pixel_boundary(ref))


@pytest.mark.parametrize('plan', [
    dict(tile_photons=512, order='given'),
    dict(tile_photons=1, order='cost'),
    dict(tile_photons=4096, scale_by_cost=True),
    dict(tile_photons=[100, 2000, 512, 7, 999], order=[4, 2, 0, 1, 3], wg_per_cu=2),
    dict(tile_photons=1024, lds_budget=8192),
])
def test_tile_plans_give_identical_results(study, plan):
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = study
    ref = BatchedTrace(cis, inputs, unis, precision=64)
    ref.run()
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    bt.set_plan(**plan)
    bt.run()
    for v in range(len(cis)):
        _assert_identical(_columns(ref.outputs[v]), _columns(bt.outputs[v]))


def test_batched_fp32_matches_single_fp32(study):
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = study
    bt = BatchedTrace(cis, inputs, unis, precision=32)
    bt.run()
    for v, ci in enumerate(cis):
        dp, _ = _single(ci, inputs[v], unis[v], 32)
        a, b = _columns(dp), _columns(bt.outputs[v])
        for k in a:
            if a[k].dtype.kind == 'f':
                ok = ~np.isnan(a[k])
                assert np.array_equal(np.isnan(a[k]), np.isnan(b[k])), k
                assert np.allclose(a[k][ok], b[k][ok], rtol=0, atol=1e-6), k
            else:
                assert np.array_equal(a[k], b[k]), k


def test_graph_replay_repeats_eager_result(study):
    import torch
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = study
    bt = BatchedTrace(cis, inputs, unis, precision=32)
    bt.run()
    eager = [_columns(o) for o in bt.outputs]
    bt.capture()
    for rep in range(3):
        for o in bt.outputs:
            o.prob.fill_(-1.0)
            o.pos.fill_(float('nan'))
        # unrelated eager work between replays
        _single(cis[0], inputs[0], unis[0], 64)
        torch.cuda.synchronize()
        bt.replay()
        torch.cuda.synchronize()
        assert int(bt.raw_status()[:, 5].sum()) == 0
        for v in range(len(cis)):
            _assert_identical(eager[v], _columns(bt.outputs[v]))


def test_autotuner_picks_a_plan_with_the_same_result(study):
    from marxs_amd.batch import BatchedTrace
    from marxs_amd.autotune import tune_batched, tune_trace
    cis, inputs, unis, hosts = study
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    bt.run()
    ref = [_columns(o) for o in bt.outputs]
    best, table = tune_batched(bt, reps=2)
    assert best in [c for c, _ in table] and all(ms > 0 for _, ms in table)
    bt.run()
    for v in range(len(cis)):
        _assert_identical(ref[v], _columns(bt.outputs[v]))
    from marxs_amd.batch import outputs_like
    dp = outputs_like(inputs[0])
    wpc, res = tune_trace(cis[0], dp, unis[0], 64, candidates=(2, 8), reps=2, src=inputs[0])
    assert cis[0].wg_per_cu[64] == wpc and set(res) == {2, 8}


def test_run_tolerances_batched_matches_sequential_loop(gpu):
    """run_tolerances semantics (marxs/design/tolerancing.py:224-283): the
    batched run gives the per-variant analysis of the sequential loop."""
    import marxs_amd as mx
    from marxs_amd import configs, tolerancing
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(120)
    n = 20000
    ph, _ = make_inputs(n, 7, 0)
    table = mx.PhotonTable({'dir': np.c_[ph['dir'], np.zeros(n)],
                            'polarization': np.c_[ph['pol'], np.zeros(n)],
                            'energy': ph['energy'], 'probability': ph['prob']})
    u = np.random.default_rng(8).random((CompiledInstrument(inst).n_uniform_slots, n))

    def analyse(p):
        sel = (np.asarray(p['order']) == -6) & (np.asarray(p['CCD']) >= 0)
        return {'n': int(sel.sum()), 'y': float(np.mean(np.asarray(p['pos'])[sel, 1]))}

    seq = []
    params = [{'dx': s, 'rx': 1e-4 * s, 'seed': k} for k, s in enumerate([0., 0.2, 1.0])]

    def wig(e, dx, rx, seed):
        tolerancing.wiggle(e, dx=dx, rx=rx, rng=np.random.default_rng(seed))

    for pars in params:
        wig(parts['gratings'], **pars)
        out = CompiledInstrument(inst)(table.copy(), uniforms=u)
        seq.append({**pars, **analyse(out)})
    bat = tolerancing.run_tolerances_batched(table, inst, wig, parts['gratings'], params, analyse,
                                             uniforms=u)
    assert seq == bat
    assert seq[0]['n'] > 100 and seq[0]['y'] != seq[2]['y']


def test_run_tolerances_batched_with_device_figures_of_merit(gpu):
    """CaptureResAeff after every variant (tolerancing.py:280) on the device
    columns of the batched trace equals the same analysis of the host tables
    written back from it."""
    import marxs_amd as mx
    from marxs_amd import configs, tolerancing
    from marxs_amd.analysis import CaptureResAeff
    from marxs_amd.tracer import CompiledInstrument
    inst, parts = configs.cat_rowland_instrument(120)
    n = 50000
    ph, _ = make_inputs(n, 9, 0)
    ph['energy'][:] = 0.65
    table = mx.PhotonTable({'dir': np.c_[ph['dir'], np.zeros(n)],
                            'polarization': np.c_[ph['pol'], np.zeros(n)],
                            'energy': ph['energy'], 'probability': ph['prob']})
    u = np.random.default_rng(10).random((CompiledInstrument(inst).n_uniform_slots, n))
    # zeroth order misses the CCD row: its position is given
    cap = CaptureResAeff(A_geom=1., orders=np.arange(-12, 3), dispersion_coord='det_x',
                         zeropos=0.)
    params = [{'dx': s, 'seed': k} for k, s in enumerate([0., 0.5])]

    def wig(e, dx, seed):
        tolerancing.wiggle(e, dx=dx, rng=np.random.default_rng(seed))

    host = tolerancing.run_tolerances_batched(table, inst, wig, parts['gratings'], params, cap,
                                              uniforms=u)
    dev = tolerancing.run_tolerances_batched(table, inst, wig, parts['gratings'], params, cap,
                                             uniforms=u, analyze_on_device=True)
    for a, b in zip(host, dev):
        assert np.array_equal(a['Aeff'], b['Aeff'])
        assert np.array_equal(a['R'], b['R'], equal_nan=True)
    assert np.isfinite(dev[0]['R']).sum() >= 3


# --------------------------------------------------------- 64-variant study
C5_PHOTONS = 20_000


@pytest.fixture(scope='module')
def c5(gpu):
    """The batching workload's shape: 64 design variants (Rowland R
    5000-7000 mm, blaze 1-2 deg, 50-800 facets, each with its own 16-CCD
    row), 2e4 photons each (the workload itself uses 2e5; fewer keep the
    test short, the tables are the workload's)."""
    import torch
    from marxs_amd import configs
    from marxs_amd.tracer import CompiledInstrument, DevicePhotons
    variants = configs.design_variants(64, np.random.default_rng(5), facet_range=(50, 800))
    cis, inputs, unis, hosts = [], [], [], []
    for k, (tree, parts, pars) in enumerate(variants):
        ci = CompiledInstrument(tree)
        ph, u = make_inputs(C5_PHOTONS, 500 + k, ci.n_uniform_slots)
        cis.append(ci)
        inputs.append(DevicePhotons(ph['dir'].T, ph['pol'].T, ph['energy'], ph['prob'], device=gpu))
        unis.append(torch.as_tensor(u, device=gpu))
        hosts.append((tree, ph, u, pars))
    return cis, inputs, unis, hosts


def test_c5_batched_fp64_equals_each_single_trace(c5):
    """One batched FP64 launch over all 64 variants (26 126 facets); every
    variant's columns equal its own single-variant trace bit for bit, and
    its status counters equal the single trace's
    (run_tolerances semantics, marxs/design/tolerancing.py:275-281)."""
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = c5
    nf = [h[3]['n_facets'] for h in hosts]
    assert min(nf) < 100 and max(nf) > 750 and len(cis) == 64
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    st = bt.run()
    for v, ci in enumerate(cis):
        dp, sst = _single(ci, inputs[v], unis[v], 64)
        _assert_identical(_columns(dp), _columns(bt.outputs[v]))
        assert st[v]['multi_hit'] == int(sst.cpu()[4]), v
        assert st[v]['grid_fallback'] == int(sst.cpu()[3]), v


def test_c5_variants_against_oracle(c5):
    """The smallest and the largest variant, and two large variants whose
    loose culling lists do not fit in LDS and are read from global memory,
    against the oracle's trace of the same photons (indices exact, floats
    1e-9 relative)."""
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = c5
    nf = np.array([h[3]['n_facets'] for h in hosts])
    order = np.argsort(nf)
    spill = [int(v) for v in order[::-1] if cis[v].tables.stage_i[2, 39] < 0]
    assert len(spill) >= 3
    pick = sorted({int(order[0]), int(order[-1]), spill[1], spill[2]})
    bt = BatchedTrace(cis, inputs, unis, precision=64)
    bt.run()
    for v in pick:
        tree, ph, u, pars = hosts[v]
        ref = oracle_trace(tree, ph, u)
        got = _columns(bt.outputs[v])
        got = {k: (x.T if x.ndim == 2 else x) for k, x in got.items()}
        got = {k: (x.astype(np.int64) if x.dtype.kind == 'i' else x) for k, x in got.items()}
        print('variant', v, 'facets', nf[v])
        rep = compare(ref, got, rtol=1e-9, near=near_pixel_boundary(ref))
        assert rep['facet'][1:] == (0, 0) and rep['order'][1:] == (0, 0) and rep['CCD'][1:] == (0, 0)
        assert (got['facet'] >= 0).sum() > 500 and (got['CCD'] >= 0).sum() > 100


def test_c5_batched_fp32_close_to_fp64(c5):
    """The batched FP32 launch over the 64 variants: detector coordinates of
    the photons both precisions put on the same facet and CCD within
    0.5 um RMS of the FP64 launch."""
    from marxs_amd.batch import BatchedTrace
    cis, inputs, unis, hosts = c5
    b64 = BatchedTrace(cis, inputs, unis, precision=64)
    b64.run()
    b32 = BatchedTrace(cis, inputs, unis, precision=32)
    b32.run()
    errs, n_same = [], 0
    for v in range(len(cis)):
        a, b = _columns(b64.outputs[v]), _columns(b32.outputs[v])
        ok = (a['CCD'] >= 0) & (a['CCD'] == b['CCD']) & (a['facet'] == b['facet'])
        n_same += int(ok.sum())
        errs.append(np.hypot(a['det_x'] - b['det_x'], a['det_y'] - b['det_y'])[ok])
    e = np.concatenate(errs) * 1e3
    print('C5 FP32 vs FP64: rms', float(np.sqrt(np.mean(e ** 2))), 'um over', n_same)
    assert n_same > 10_000 and float(np.sqrt(np.mean(e ** 2))) <= 0.5
