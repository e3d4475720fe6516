"""Statistical scatter elements (``marxs/optics/scatter.py``).

``RandomGaussianScatter`` (``scatter.py:88-145``) turns the direction of
every photon that meets its plane by a Gaussian angle around an axis
perpendicular to the direction, then by a uniform angle around the original
direction, and parallel-transports the polarization; the position does not
change.  On the GPU it is a stack layer (see ``FlatStack``): the normal
deviate comes from the photon's uniform draw by the inverse normal CDF.
"""
from .optics import FlatOpticalElement

__all__ = ['RandomGaussianScatter']


class RandomGaussianScatter(FlatOpticalElement):
    """Gaussian scatter with sigma ``scatter`` (rad).  A sigma of 0 leaves
    the photons unchanged and adds no column (``scatter.py:124-129``)."""
    scattername = 'scatter'

    def __init__(self, **kwargs):
        if 'scatter' in kwargs:
            self.scatter = kwargs.pop('scatter')
        elif not hasattr(self, 'scatter'):
            raise ValueError('Keyword "scatter" missing.')
        if callable(self.scatter):
            raise NotImplementedError('a callable scatter has no GPU implementation')
        super().__init__(**kwargs)
