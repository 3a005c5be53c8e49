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
* ``TailsMask.py``                  -> :mod:`opensss_amd.tails_mask`
* ``BackProjection.py``             -> :mod:`opensss_amd.backprojection`
* ``Scaling.py``                    -> :mod:`opensss_amd.scaling`
"""
from .data_reduction import ComputeLORCoordinates, SinogramLUT, SumLORsCounts
from .geometry import (ImportCASToRGeometry, PrepareGeometryAndSinograms, expandRings,
                       mashDetectors)
from .sinogram import SinogramCoordinates

__all__ = [
    'ComputeLORCoordinates', 'SinogramLUT', 'SumLORsCounts', 'ImportCASToRGeometry',
    'PrepareGeometryAndSinograms', 'expandRings', 'mashDetectors', 'SinogramCoordinates',
    'SingleScatterSimulationTOF', 'ScatterSubject', 'RayTracing3DTOF', 'RayTracing3DTOFBatch',
    'LineIntegrals', 'InterpolateScatters', 'GenerateTailsMask', 'MaskGenerator',
    'MaskGeneratorMashing', 'BackProjectionScatterRandFit', 'BackProjectionScatterRandFitMashing',
    'ScaleScattersToPrompts',
]

_LAZY = {
    'SingleScatterSimulationTOF': 'sss', 'ScatterSubject': 'sss',
    'RayTracing3DTOF': 'raytracing', 'RayTracing3DTOFBatch': 'raytracing', 'LineIntegrals': 'raytracing',
    'InterpolateScatters': 'interpolation', 'Interpolator': 'interpolation',
    'GenerateTailsMask': 'tails_mask', 'MaskGenerator': 'tails_mask', 'MaskGeneratorMashing': 'tails_mask',
    'BackProjectionScatterRandFit': 'backprojection',
    'BackProjectionScatterRandFitMashing': 'backprojection',
    'ScaleScattersToPrompts': 'scaling',
}


def __getattr__(name):
    # GPU entry points load libsss_hip.so on first use (raises if it is missing)
    if name in _LAZY:
        import importlib
        return getattr(importlib.import_module(f'.{_LAZY[name]}', __name__), name)
    raise AttributeError(name)
