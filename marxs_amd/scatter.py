"""Statistical scatter elements (``marxs/optics/scatter.py``).

``RadialMirrorScatter`` (``scatter.py:24-78``) blurs within and across the
plane of reflection of a radial mirror.  ``RandomGaussianScatter``
(``scatter.py:88-145``) turns the direction of every photon that meets its
plane by a Gaussian angle around an axis perpendicular to the direction,
then by a uniform angle around the original direction.  Both
parallel-transport the polarization and leave the position unchanged.  On
the GPU they are stack layers (see ``FlatStack``): the normal deviates come
from the photon's uniform draws by the inverse normal CDF.
"""
from .optics import FlatOpticalElement

__all__ = ['RadialMirrorScatter', 'RandomGaussianScatter']


class RadialMirrorScatter(FlatOpticalElement):
    """Scatter of a radial mirror (``scatter.py:24-78``): in the plane of
    reflection, spanned by the direction and the radial vector from the
    element's centre to the photon's last interaction point, by a Gaussian
    angle of sigma ``inplanescatter`` (rad) about dir x radial; then
    perpendicular to that plane by a Gaussian angle of sigma
    ``perpplanescatter`` (rad, default 0) about the radial vector.  A sigma of
    0 leaves that rotation out and records an angle of 0.  On the GPU it is a
    stack layer; the normal deviates come from the photon's uniform draws by
    the inverse normal CDF."""
    inplanescattercol = 'inplanescatter'
    perpplanescattercol = 'perpplanescatter'

    def __init__(self, **kwargs):
        self.inplanescatter = float(kwargs.pop('inplanescatter'))
        self.perpplanescatter = float(kwargs.pop('perpplanescatter', 0.))
        super().__init__(**kwargs)


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
