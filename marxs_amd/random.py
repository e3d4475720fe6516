"""Injectable random stream for the trace.

The reference draws from numpy's global legacy stream inside each element
(``marxs/optics/grating.py:55-57``, ``aperture.py:93-94,139-142``).  Here the
draws are made up front, one set of uniforms per photon, so that a trace is a
pure function of (photons, uniforms).  By default they come from a per-device
torch generator that :func:`seed` resets.
"""
import torch

__all__ = ['seed', 'generator']

_generators = {}
_seed = None


def seed(s):
    """Seed the default uniform stream on every device."""
    global _seed
    _seed = int(s)
    for g in _generators.values():
        g.manual_seed(_seed)


def generator(device):
    device = torch.device(device)
    key = str(device)
    if key not in _generators:
        g = torch.Generator(device=device)
        if _seed is not None:
            g.manual_seed(_seed)
        else:
            g.seed()
        _generators[key] = g
    return _generators[key]
