"""Element protocol shared by optical elements and containers.

Mirrors ``marxs/base/base.py``: ``SimulationSequenceElement`` (``194-289``),
its output-column defaults (``249-286``: new float columns start as NaN, id
columns as int -1) and ``_parse_position_keywords`` (``292-336``: a pose is
either ``pos4d`` or ``compose(position, orientation, zoom)``).
"""
from collections import OrderedDict
import numpy as np

from .mathutils import compose

__all__ = ['GeometryError', 'SimulationSetupError', 'SimulationSetupWarning',
           'MarxsElement', 'SimulationSequenceElement',
           'parse_position_keywords']


class GeometryError(Exception):
    pass


class SimulationSetupError(Exception):
    pass


class SimulationSetupWarning(Warning):
    pass


class MarxsElement:
    """Base class: every element has a ``name``."""

    def __init__(self, **kwargs):
        self.name = kwargs.pop('name', self.__class__.__name__)
        if kwargs:
            raise ValueError('Initialization arguments {0} not understood'
                             .format(', '.join(kwargs.keys())))

    def describe(self):
        """Description for the photon list's metadata (``base/base.py:89-90``)."""
        return OrderedDict(element=self.name)


class SimulationSequenceElement(MarxsElement):
    """An element that processes a photon table.

    ``id_col``/``id_num``: when ``id_col`` is a string, photons interacting with
    this element get ``id_num`` written into that int column (default -1).
    """
    output_columns = []
    id_col = None

    def __init__(self, **kwargs):
        self.id_num = kwargs.pop('id_num', -9)
        if 'id_col' in kwargs:
            self.id_col = kwargs.pop('id_col')
        super().__init__(**kwargs)

    def __call__(self, photons, *args, **kwargs):
        return self.process_photons(photons, *args, **kwargs)


def parse_position_keywords(kwargs):
    """Pop ``pos4d`` or (``position``, ``orientation``, ``zoom``) from kwargs
    and return the 4x4 pose."""
    pos4d = kwargs.pop('pos4d', None)
    if pos4d is None:
        position = kwargs.pop('position', np.zeros(3))
        orientation = kwargs.pop('orientation', np.eye(3))
        zoom = kwargs.pop('zoom', 1.)
        if np.isscalar(zoom):
            zoom = np.ones(3) * zoom
        if len(zoom) != 3:
            raise ValueError('zoom must have three elements for x,y,z or be a '
                             'scalar (global zoom).')
        if np.any(np.asarray(zoom) <= 0):
            raise ValueError('All values in zoom must be positive-definite to '
                             'keep pos4d matrix valid.')
        pos4d = compose(position, orientation, zoom)
    else:
        if ('position' in kwargs) or ('orientation' in kwargs) or ('zoom' in kwargs):
            raise ValueError('If pos4d is specificed, the following keywords '
                             'cannot be given at the same time: position, '
                             'orientation, zoom.')
        pos4d = np.array(pos4d, dtype=float)
    if np.linalg.det(pos4d) == 0:
        raise ValueError('pos4d matrix is invalid (determinant is 0).')
    return pos4d
