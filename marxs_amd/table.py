"""A minimal column table for photon lists.

The reference keeps photons in an ``astropy.table.Table`` with one row per
photon; ``pos``, ``dir`` and ``polarization`` are (N, 4) homogeneous float64
columns, ``energy`` (keV) and ``probability`` are (N,) float64
(``marxs/utils.py:16-51``, ``marxs/base/base.py:197-226``).  astropy is not a
dependency here, so :class:`PhotonTable` offers the subset of the Table API
that the trace uses (``colnames``, item access, ``add_column``, ``meta``,
``copy``).  Any object with that interface (an astropy ``Table`` included) can
be passed to the trace instead.
"""
import copy as _copy

import numpy as np

__all__ = ['PhotonTable', 'generate_test_photons']


class PhotonTable:
    """Ordered mapping of column name -> numpy array, all of equal length."""

    def __init__(self, columns=None, meta=None):
        self._cols = {}
        self.meta = dict(meta) if meta else {}
        if columns:
            for k, v in columns.items():
                self[k] = v

    @property
    def colnames(self):
        return list(self._cols.keys())

    def __len__(self):
        if not self._cols:
            return 0
        return len(next(iter(self._cols.values())))

    def __contains__(self, name):
        return name in self._cols

    def __getitem__(self, key):
        if isinstance(key, str):
            return self._cols[key]
        # row selection (slice, mask or index array) -> new table
        return PhotonTable({k: v[key] for k, v in self._cols.items()},
                           meta=self.meta)

    def __setitem__(self, name, value):
        value = np.asarray(value)
        if value.ndim == 0:
            value = np.full(len(self), value)
        if self._cols and len(value) != len(self):
            raise ValueError(f'Column {name!r} has length {len(value)}, '
                             f'table has {len(self)} rows.')
        self._cols[name] = value

    def add_column(self, values, name=None):
        if name is None:
            name = values.name
        if name in self._cols:
            raise ValueError(f'Column {name} already exists')
        self[name] = values

    def remove_column(self, name):
        del self._cols[name]

    def keys(self):
        return self.colnames

    def copy(self):
        return PhotonTable({k: v.copy() for k, v in self._cols.items()},
                           meta=_copy.deepcopy(self.meta))

    def __repr__(self):
        return f'<PhotonTable rows={len(self)} cols={self.colnames}>'


def generate_test_photons(n=1):
    """``n`` identical photons at x=1 moving along -x, 1 keV, polarised along +y
    (the test-photon convention of ``marxs/utils.py:16-51``)."""
    return PhotonTable({
        'pos': np.tile(np.array([1., 0., 0., 1.]), (n, 1)),
        'dir': np.tile(np.array([-1., 0., 0., 0.]), (n, 1)),
        'energy': np.ones(n),
        'polarization': np.tile(np.array([0., 1., 0., 0.]), (n, 1)),
        'probability': np.ones(n),
    })
