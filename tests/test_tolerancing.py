"""Tolerancing helpers (CPU): the movers, wiggle, period and attribute
changes and the 6-DOF parameter lists, following
marxs/design/tests/test_tolerancing.py."""
import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import tolerancing as T


def gsa():
    pos = [[0, y, z] for y in (-30., 0., 30.) for z in (-30., 0., 30.)]
    return mx.Parallel(elem_class=mx.FlatGrating,
                       elem_args={'d': 2e-4, 'order_selector': mx.OrderSelector([1]),
                                  'zoom': [1, 10, 10]},
                       elem_pos={'position': pos})


def poses(g):
    return np.stack([e.pos4d for e in g.elements])


@pytest.mark.parametrize('function', [T.moveglobal, T.moveindividual])
def test_change_parallel_elements(function):
    """test_tolerancing.py:75-89: zero changes keep the poses, any one
    parameter changes them."""
    g = gsa()
    ref = poses(g)
    function(g, 0., 0., 0.)
    assert np.all(poses(g) == ref)
    for key in ['dx', 'dy', 'dz', 'rx', 'ry', 'rz']:
        function(g, **{key: 1.23})
        assert not np.all(poses(g) == ref)


def test_wiggle_changes_and_is_random():
    g = gsa()
    ref = poses(g)
    T.wiggle(g, dx=10, dy=.1, rng=np.random.default_rng(0))
    diff = ref - poses(g)
    assert np.std(diff[:, 0, 3]) > np.std(diff[:, 1, 3])
    T.wiggle(g, rng=np.random.default_rng(0))
    assert np.allclose(poses(g), ref)


def test_moveelements_translate_and_rotate():
    """test_tolerancing.py:92-123: a global and an individual translation
    agree; rotations do not (different centres)."""
    g0, g1, g2 = gsa(), gsa(), gsa()
    T.moveglobal(g1, dy=-20)
    T.moveindividual(g2, dy=-20)
    assert np.allclose(poses(g1), poses(g2))
    assert not np.allclose(poses(g0), poses(g2))
    g1, g2 = gsa(), gsa()
    T.moveglobal(g1, rz=-1, ry=.2)
    T.moveindividual(g2, rz=-1, ry=.2)
    assert not np.allclose(poses(g1), poses(g2))


def test_moveelem():
    """test_tolerancing.py:126-132"""
    det = mx.FlatDetector(zoom=[1, 100, 100])
    assert det.geometry['center'][2] == 0
    T.moveelem(det, dz=5)
    assert det.geometry['center'][2] == 5
    assert np.all(det.geometry.pos4d[:3, :3] == np.eye(3) * [1, 100, 100])
    T.moveelem(det, dz=1)          # relative to the stored original pose
    assert det.geometry['center'][2] == 1


def test_lists_of_elements_and_attribute_errors():
    g1, g2 = gsa(), gsa()
    T.moveglobal([g1, g2], dx=3)
    assert np.allclose(poses(g1), poses(g2))
    gr = mx.FlatGrating(d=1e-3, order_selector=mx.OrderSelector([1]))
    T.varyperiod(gr, 2e-4, 1e-6, rng=np.random.default_rng(1))
    assert abs(gr._d - 2e-4) < 1e-5
    with pytest.raises(ValueError, match='grating period'):
        T.varyperiod(mx.FlatDetector(), 1., 1.)
    with pytest.raises(ValueError, match='does not have'):
        T.varyattribute(gr, not_an_attribute=3)
    T.varyattribute(gr, transmission=False)
    assert gr.transmission is False
    T.varyorderselector(gr, mx.OrderSelector, [2])
    assert list(gr.order_selector.orderlist) == [2]


def test_generate_6d_wigglelist():
    """tolerancing.py:454-516: one degree of freedom at a time, both signs for
    the movers, one side for wiggle, the all-zero row once."""
    glob, ind = T.generate_6d_wigglelist([0., .1, .2], [0., .01])
    assert len(glob) == 1 + 3 * 2 * 2 + 3 * 1 * 2
    assert len(ind) == 1 + 3 * 2 + 3 * 1
    for row in glob:
        assert sum(v != 0 for v in row.values()) <= 1
    assert T.reset_6d in glob
    one = T.select_1dof_changed(glob, 'dy')
    assert len(one) == 5 and all(r['dx'] == 0 for r in one)
