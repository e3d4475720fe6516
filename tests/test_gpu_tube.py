"""GPU parity of tube stages (CircularDetector on a Cylinder) against the
oracle (needs an MI355X).  FP64: every float column within 1e-9 relative,
ids exact; FP32 path: the tube test and its outputs run in float64 too."""
import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import configs
from marxs_amd.geometry import Cylinder
from marxs_amd.mathutils import axangle2mat
from tests.helpers import gpu_trace, oracle_trace, compare, make_inputs

pytestmark = pytest.mark.gpu

RTOL = 1e-9


def _ci(inst, **kw):
    from marxs_amd.tracer import CompiledInstrument
    return CompiledInstrument(inst, **kw)


def test_rowland_circle_detector_parity(gpu):
    """The resolving-power set-up of analysis/gratings.py:89-102: a tube on
    the Rowland circle (half-pipe facing the gratings) behind 200 CAT
    facets."""
    inst0, parts = configs.cat_rowland_instrument(200)
    det = mx.CircularDetector(geometry=Cylinder.from_rowland(
        parts['rowland'], width=1e5, rotation=np.pi, kwargs={'phi_lim': [-np.pi / 2, np.pi / 2]}),
        pixsize=0.024)
    inst = mx.Sequence(elements=[parts['aperture'], parts['lens'], parts['gratings'], det])
    for prec in (64, 32):
        ci = _ci(inst)
        ph, u = make_inputs(40000, 41 + prec, ci.n_uniform_slots, 'lines')
        ref = oracle_trace(inst, ph, u)
        got, st = gpu_trace(ci, ph, u, precision=prec)
        assert np.isfinite(ref['detpix_x']).sum() > 20000
        if prec == 64:
            compare(ref, got, rtol=RTOL)
        else:
            ok = (got['facet'] == ref['facet']) & np.isfinite(ref['det_phi'])
            err = np.abs(got['det_phi'] - ref['det_phi'])[ok] * 6000. * 1e3   # um along the circle
            assert ok.sum() > 20000 and np.sqrt(np.mean(err ** 2)) <= 0.5


def test_tubes_random_rays_two_elements(gpu):
    """Random rays against a rotated, shifted half-pipe and a second, wider
    tube in the same stage (hit after the first: last writer wins)."""
    rng = np.random.default_rng(5)
    t1 = mx.CircularDetector(pixsize=0.01, zoom=[3., 3., 2.], position=[0.5, -0.2, 0.1],
                             orientation=axangle2mat([0.2, 1, 0.3], 0.4),
                             geometry=Cylinder, phi_lim=[-1., 2.], id_col='tube', id_num=1)
    t2 = mx.CircularDetector(pixsize=0.02, zoom=[5., 5., 4.], id_col='tube', id_num=2)
    inst = mx.Sequence(elements=[t1, t2])
    n = 50000
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1)[:, None]
    pol = np.cross(d, rng.normal(size=(n, 3)))
    pol /= np.linalg.norm(pol, axis=1)[:, None]
    ph = {'pos': rng.uniform(-1, 1, (n, 3)), 'dir': d, 'pol': pol,
          'energy': np.ones(n), 'prob': np.ones(n)}
    ci = _ci(inst)
    u = np.zeros((ci.n_uniform_slots, n))
    ref = oracle_trace(inst, ph, u)
    got, st = gpu_trace(ci, ph, u, precision=64)
    compare(ref, got, rtol=RTOL)
    # photons that met both tubes, from the oracle element by element
    first = oracle_trace(mx.Sequence(elements=[t1]), ph, u)
    h1 = np.isfinite(first['det_phi'])
    second = oracle_trace(mx.Sequence(elements=[t2]), {**ph, 'pos': first['pos'], 'dir': first['dir']}, u)
    both = int((h1 & np.isfinite(second['det_phi'])).sum())
    assert (got['tube'] == 2).sum() > 10000 and both > 100 and st['multi_hit'] == both
