"""marxs_amd -- MI355X-native photon tracer for the hot path of MARXS.

A Sequence/Parallel tree of flat optical elements (apertures, PerfectLens /
ThinLens, FlatGrating / CATGrating facet arrays, FlatDetector CCD arrays) is
compiled into structure-of-arrays tables and traced by hand-written HIP
kernels for gfx950 (libmarxs_hip.so, called through ctypes).  Set-up classes
(elements, containers, Rowland-torus placement) mirror the reference's
constructors; importing them does not need a GPU.  Tracing does.
"""
energy2wave = 1.2398419292004202e-06   # keV -> mm (marxs/__init__.py:14)

from .table import PhotonTable, generate_test_photons  # noqa: E402
from .base import (GeometryError, SimulationSetupError,  # noqa: E402
                   SimulationSetupWarning)
from .optics import (OrderSelector, EfficiencyFile, InterpolateEfficiencyTable,  # noqa: E402
                     FlatGrating, CATGrating,
                     NonParallelCATGrating, CircleAperture, RectangleAperture,
                     PerfectLens, ThinLens, FlatDetector, FlatOpticalElement,
                     OpticalElement, FlatStack, CircularDetector)
from .scatter import RandomGaussianScatter  # noqa: E402
from .catgrating import CATL1L2Stack  # noqa: E402
from .simulator import Sequence, Parallel, ParallelCalculated, propagate  # noqa: E402
from .rowland import (RowlandTorus, GratingArrayStructure, RectangularGrid,  # noqa: E402
                      CircularMeshGrid, ElementsOnTorus, design_tilted_torus)

__version__ = '0.1.0'


def CompiledInstrument(*args, **kwargs):
    """See :class:`marxs_amd.tracer.CompiledInstrument` (imported lazily so
    that set-up code works without the GPU library)."""
    from .tracer import CompiledInstrument as _CI
    return _CI(*args, **kwargs)
