"""The parity comparison itself (CPU): planted differences just outside
its bounds must fail and differences inside them must pass, so that the GPU
parity tests that use it mean what they state (1e-9 relative per value,
exact integers, a counted and capped pixel-edge exemption)."""
import numpy as np
import pytest

from helpers import compare, near_pixel_boundary


def _case(n=1000, seed=0):
    rng = np.random.default_rng(seed)
    ref = {'pos': rng.normal(0, 5000, (n, 3)), 'blaze': rng.uniform(1e-4, 0.05, n),
           'prob': rng.uniform(1e-6, 1, n), 'facet': rng.integers(-1, 600, n),
           'detpix_x': rng.uniform(0, 1023, n), 'pix_x': np.zeros(n, dtype=np.int64)}
    ref['pix_x'] = np.floor(ref['detpix_x'] + 0.5).astype(np.int64)
    return ref, {k: v.copy() for k, v in ref.items()}


def test_identical_passes():
    ref, got = _case()
    compare(ref, got, verbose=False)


def test_relative_bound_is_relative_for_small_values():
    """A difference of 2e-9 of a small value (1e-2) fails although it is far
    below 1e-9 in absolute terms (the bound scales with the value, not
    max(1, |a|)); only values below 1e-2 reach the 1e-11 floor."""
    ref, got = _case()
    k = int(np.argmin(np.abs(ref['blaze'] - 3e-2)))
    got['blaze'][k] = ref['blaze'][k] * (1 + 2e-9)
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
    got['blaze'][k] = ref['blaze'][k] * (1 + 0.5e-9)
    compare(ref, got, verbose=False)
    ref['prob'][3] = 0.05
    got['prob'][3] = 0.05 * (1 - 3e-9)
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
    # a value zero up to rounding: 5e-14 against -5e-14 passes on the floor
    ref['prob'][3] = got['prob'][3] = 0.5
    ref['blaze'][k], got['blaze'][k] = 5e-14, -5e-14
    rep = compare(ref, got, verbose=False)
    assert rep['blaze'][4] == 1


def test_vectors_against_their_length_and_zero_floor():
    ref, got = _case()
    got['pos'][5, 1] += 2e-9 * np.linalg.norm(ref['pos'][5])
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
    # a vector that is zero up to rounding: only the 1e-12 floor
    ref, got = _case()
    ref['pos'][7] = [3e-14, 0., -2e-14]
    got['pos'][7] = [1e-13, -5e-14, 0.]
    rep = compare(ref, got, verbose=False)
    assert rep['pos'][4] == 3
    got['pos'][7, 0] = 3e-11
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)


def test_nan_pattern_must_match():
    ref, got = _case()
    ref['blaze'][9] = np.nan
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
    got['blaze'][9] = np.nan
    compare(ref, got, verbose=False)


def test_integers_exact_and_edge_exemption_capped():
    ref, got = _case()
    got['facet'][11] += 1
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
    # pixel coordinates on an edge: the index may differ there, counted
    ref, got = _case(n=10000)
    edge = np.arange(0, 30)
    for t in (ref, got):
        t['detpix_x'][edge] = 100.5
        t['pix_x'][edge] = 101
    got['pix_x'][edge[:5]] = 100
    near = near_pixel_boundary(ref)
    rep = compare(ref, got, near=near, verbose=False)
    assert rep['pix_x'] == ('int', 0, 5)
    # more than max_near_frac of the photons: fails
    got['pix_x'][edge] = 100
    with pytest.raises(AssertionError):
        compare(ref, got, near=near, max_near_frac=1e-3, verbose=False)
    compare(ref, got, near=near, max_near_frac=5e-3, verbose=False)
    # without the exemption every index must match
    with pytest.raises(AssertionError):
        compare(ref, got, verbose=False)
