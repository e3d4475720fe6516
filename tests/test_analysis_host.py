"""Host-side analysis helpers (no GPU): sub-aperture sectors, the
probability-weighted average over orders and moving photons along their rays."""
import numpy as np
import pytest

import marxs_amd as mx
from marxs_amd import analysis as A


def test_identify_photon_in_subaperture():
    """marxs/analysis/tests/test_grating.py:99-104"""
    angle = [0, 59, 61, 180, 289]
    inaper = A.identify_photon_in_subaperture(np.deg2rad(angle), np.deg2rad(30), ang_0=np.pi / 2)
    assert np.all(inaper == [False, False, True, False, True])
    # angles are taken modulo 2 pi
    assert np.all(A.identify_photon_in_subaperture(np.deg2rad([61 + 360, -71]), np.deg2rad(30)))


def test_weighted_per_order():
    eff = mx.EfficiencyFile.from_arrays([2., 1., 3.], [[.2, .2, 0.], [.1, .1, 0.], [.3, .3, 0.]],
                                        [-1, 0, 1])
    data = np.array([[10., 20., 30.], [30., 40., 50.], [1e6, 1e6, 1e6]])
    # equal weights of orders -1 and 0, none for order 1: the plain mean of
    # the first two rows at every energy (also between grid energies)
    out = A.weighted_per_order(data, [-1, 0, 1], [1., 1.5, 3.], eff)
    assert np.allclose(out, [20., 30., 40.])
    with pytest.raises(ValueError, match='First dimension'):
        A.weighted_per_order(data[:2], [-1, 0, 1], [1., 1.5, 3.], eff)
    with pytest.raises(ValueError, match='Second dimension'):
        A.weighted_per_order(data, [-1, 0, 1], [1., 1.5], eff)
    with pytest.raises(KeyError):
        A.weighted_per_order(data, [-1, 0, 2], [1., 1.5, 3.], eff)


def test_propagate():
    p = mx.generate_test_photons(3)
    p['dir'] = np.tile([-1., 0.5, 0., 0.], (3, 1))
    mx.propagate(p, -10.)
    assert np.allclose(p['pos'], [[11., -5., 0., 1.]] * 3)
