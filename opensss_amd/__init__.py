"""MI355X-native TOF single scatter simulation (drop-in for openSSS's SSS path).

Host code is Python (this package); the compute runs in hand-written HIP
kernels of ``_lib/libsss_hip.so`` reached through ctypes (``_native``).
Module map (reference file -> module):

* ``SingleScatterSimulationTOF.py`` -> :mod:`opensss_amd.sss`
* ``RayTracing3DTOF.py``            -> :mod:`opensss_amd.raytracing`
* ``Sinogram.py``                   -> :mod:`opensss_amd.sinogram`
* ``DataReduction.py``              -> :mod:`opensss_amd.data_reduction`
* ``Geometry.py``                   -> :mod:`opensss_amd.geometry`
* ``Interpolation.py``              -> :mod:`opensss_amd.interpolation`
"""
from .data_reduction import ComputeLORCoordinates, SinogramLUT, SumLORsCounts
from .geometry import (ImportCASToRGeometry, PrepareGeometryAndSinograms, expandRings,
                       mashDetectors)
from .sinogram import SinogramCoordinates

__all__ = [
    'ComputeLORCoordinates', 'SinogramLUT', 'SumLORsCounts', 'ImportCASToRGeometry',
    'PrepareGeometryAndSinograms', 'expandRings', 'mashDetectors', 'SinogramCoordinates',
    'SingleScatterSimulationTOF', 'ScatterSubject', 'RayTracing3DTOF', 'RayTracing3DTOFBatch',
    'LineIntegrals', 'InterpolateScatters',
]


def __getattr__(name):
    # GPU entry points load libsss_hip.so on first use (raises if it is missing)
    if name in ('SingleScatterSimulationTOF', 'ScatterSubject'):
        from . import sss
        return getattr(sss, name)
    if name in ('RayTracing3DTOF', 'RayTracing3DTOFBatch', 'LineIntegrals'):
        from . import raytracing
        return getattr(raytracing, name)
    if name in ('InterpolateScatters', 'Interpolator'):
        from . import interpolation
        return getattr(interpolation, name)
    raise AttributeError(name)
