"""Committed golden vectors (tests/golden, written by make_golden.py): the
oracle reproduces them bit for bit from the stored description, and the
workload builders still produce that description (CPU)."""
import numpy as np
import pytest

from helpers import load_golden, encode_description
from oracle import trace as OT
from oracle.describe import describe

CASES = ['flat_grid_1000', 'cat_rowland_200_2000', 'cat_stack_200_2000']


@pytest.mark.parametrize('name', CASES)
def test_oracle_reproduces_golden(name):
    meta, els, ph, u, out = load_golden(name)
    got = OT.trace(els, ph, u)
    assert set(got) == set(out)
    for k, v in out.items():
        assert np.array_equal(np.asarray(got[k]), v, equal_nan=v.dtype.kind == 'f'), k


@pytest.mark.parametrize('name', CASES)
def test_builder_still_gives_golden_description(name):
    from marxs_amd import configs
    meta, els, ph, u, out = load_golden(name)
    inst = getattr(configs, meta['builder'])(**meta['kwargs'])
    inst = inst[0] if isinstance(inst, tuple) else inst
    now = describe(inst)
    assert len(now) == len(els)

    def same(a, b, key=''):
        # poses to 1e-9 mm (placement solves on the torus), all else exact;
        # stack layers nest as lists of dicts
        if isinstance(a, dict):
            assert isinstance(b, dict) and set(a) == set(b), key
            for k in a:
                same(a[k], b[k], k)
        elif isinstance(a, (list, tuple)) and not (a and isinstance(a[0], (int, float, str))):
            assert len(a) == len(b), key
            for x, y in zip(a, b):
                same(x, y, key)
        elif key == 'pos4d':
            assert np.allclose(a, b, rtol=0, atol=1e-9), key
        else:
            assert np.array_equal(np.asarray(a), np.asarray(b)), key
    for a, b in zip(now, els):
        same(a, b)


def test_golden_cases_exercise_the_path():
    for name in CASES:
        meta, els, ph, u, out = load_golden(name)
        det = [k for k in ('CCD', 'det_x') if k in out]
        assert np.isfinite(out['det_x']).sum() > 100, name
        assert (out[meta['description']['elements'][2]['id_col']] >= 0).sum() > 100
