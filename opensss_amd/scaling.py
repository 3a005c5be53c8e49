"""Scale of the scatter estimate to the prompts, per TOF bin.

Drop-in for ``Python/openSSS/Scaling.py:27-121`` (``ScaleScattersToPrompts``):
the same files are read from and written to ``SavePath`` --
``PromptsSinogram.npz`` (or ``PromptsSinogramMashed.npz`` with span),
optional ``NormSinogram.npz`` and ``RandomsSinogram.npz``,
``AttenuationMask.npy`` and ``SSS_mashed_bin{b}.npz`` -- and the float32
scale factors go to ``ScaleFactors.npy``.  Each bin's fit runs in
``sss_tail_fit`` (see :mod:`opensss_amd.backprojection`).
"""
from __future__ import annotations

import os
import time

import numpy as np

from .backprojection import BackProjectionScatterRandFit, BackProjectionScatterRandFitMashing


def _load_npz(path):
    with np.load(path) as z:
        return z['arr_0']


def ScaleScattersToPrompts(SinogramIndex, SinogramCoordinates, SavePath, DesiredDimensions, FittingSize,
                           Geometry, ExtendedGeometry, AccelerationFactor, SpanFlag, LORCounts, *,
                           context=None):
    """float32 (TOF bins,) scale factors; also saved as ``ScaleFactors.npy``."""
    prompts = _load_npz(os.path.join(SavePath, 'PromptsSinogramMashed.npz' if SpanFlag
                                     else 'PromptsSinogram.npz'))
    n_bins = prompts.shape[-1]
    sino_shape = prompts.shape[:-1]
    norm_path = os.path.join(SavePath, 'NormSinogram.npz')
    norm = _load_npz(norm_path) if os.path.exists(norm_path) else np.ones(sino_shape, np.single)
    rand_path = os.path.join(SavePath, 'RandomsSinogram.npz')
    randoms = (_load_npz(rand_path) * norm) if os.path.exists(rand_path) else np.zeros(sino_shape, np.single)
    mask = np.float32(np.load(os.path.join(SavePath, 'AttenuationMask.npy')) >= 1)

    factors = np.zeros(n_bins, dtype=np.float32)
    print("Scaling for all bins")
    begin = time.time()
    for b in range(n_bins):
        start = time.time()
        scatter = _load_npz(os.path.join(SavePath, f'SSS_mashed_bin{b}.npz'))
        bin_prompts = prompts[:, :, :, b] / LORCounts
        if not SpanFlag:
            NR = np.asarray(SinogramIndex).shape[0]
            order = np.asarray(SinogramIndex)[0:NR, 0:NR].T
            f = BackProjectionScatterRandFit(bin_prompts, scatter, randoms, mask, DesiredDimensions,
                                             FittingSize, Geometry, SinogramCoordinates, order,
                                             AccelerationFactor, context=context)
        else:
            f = BackProjectionScatterRandFitMashing(bin_prompts, scatter, randoms, mask, DesiredDimensions,
                                                    FittingSize, ExtendedGeometry, SinogramCoordinates,
                                                    AccelerationFactor, SinogramIndex, context=context)
        print("Bin {} took {:.2f} seconds".format(b, time.time() - start))
        factors[b] = f[0]
    print("Time for Scaling: {:.2f} seconds".format(time.time() - begin))
    np.save(os.path.join(SavePath, 'ScaleFactors.npy'), factors)
    return factors
