"""Image preparation before the scatter estimate, on the GPU.

Drop-ins for ``Python/openSSS/Utils.py``: ``CropAndDownscale`` (:212-267),
``PrepareImages`` (:29-76), ``LoadImage`` (:104-140) and
``ComputePhisicalDimensions`` (:78-102).  The crop window (or the symmetric
zero pad when the crop is larger than the image on both sides) is worked out
here exactly as the reference slices its arrays; the smoothing, the spline
resize and the clip run in ``sss_crop_and_downscale``.

The resize is scikit-image's ``resize(order=3, mode='edge',
anti_aliasing=True, preserve_range=True, clip=True)`` as that library states
it (scikit-image is not installed here): a Gaussian of sigma (f - 1) / 2 per
axis downscaled by f (taps as scipy.ndimage's, truncated at 4 sigma, edge
values repeated), then the cubic B-spline interpolation of scipy.ndimage's
zoom in grid mode (coefficients of the image padded by 12 repeated edge
voxels), then the clip to the cropped image's range.  Volumes are float32
(what ``LoadImage`` returns; other dtypes are converted).
"""
from __future__ import annotations

import os
import re

import numpy as np

from . import _native


class _ResizeDesc(_native.ctypes.Structure):
    _fields_ = [('in_shape', _native.ctypes.c_int64 * 3), ('crop_start', _native.ctypes.c_int64 * 3),
                ('crop_shape', _native.ctypes.c_int64 * 3), ('out_shape', _native.ctypes.c_int64 * 3),
                ('radius', _native.ctypes.c_int32 * 3), ('weights', _native.ctypes.c_void_p * 3),
                ('clip', _native.ctypes.c_int32), ('zero_below', _native.ctypes.c_float)]


def crop_window(shape, voxel_size, crop_size):
    """Per axis (first input index, length) of the reference's crop
    (Utils.py:243-262): half width ceil(crop / voxel / 2) around shape // 2;
    a symmetric zero pad when the window passes both ends, else the slice
    [centre - half : centre + half] with Python's slice rules."""
    half = np.ceil(np.array(crop_size) / np.array(voxel_size) / 2).astype(int)
    centre = np.array(shape) // 2
    starts, lengths = [], []
    for d in range(3):
        s, e = int(centre[d] - half[d]), int(centre[d] + half[d])
        if s < 0 and e > shape[d]:
            pad = int(half[d] - centre[d])
            starts.append(-pad)
            lengths.append(int(shape[d]) + 2 * pad)
        else:
            a, b, _ = slice(s, e).indices(int(shape[d]))
            starts.append(a)
            lengths.append(max(0, b - a))
    return starts, lengths


def gaussian_taps(sigma, truncate=4.0):
    """scipy.ndimage's Gaussian taps (radius int(truncate sigma + 1/2),
    exp(-x^2 / (2 sigma^2)) normalised), or None when sigma <= 1e-15 (the
    axis is not filtered)."""
    if sigma <= 1e-15:
        return None
    r = int(truncate * float(sigma) + 0.5)
    x = np.arange(-r, r + 1)
    phi = np.exp(-0.5 / (sigma * sigma) * x ** 2)
    return np.ascontiguousarray((phi / phi.sum())[::-1], dtype=np.float64)


def CropAndDownscale(image, voxel_size, crop_size, downscaled_dimensions, clip=True, mode='edge',
                     preserve_range=True, anti_aliasing=True, order=3, *, zero_below=None, context=None):
    """Cropped and resized float32 (X, Y, Z) image (Utils.py:212-267).
    zero_below: values below it become 0 afterwards (PrepareImages' floor)."""
    if mode != 'edge' or order != 3 or not preserve_range:
        raise NotImplementedError("only the reference's defaults (mode='edge', order=3, preserve_range=True)")
    img = np.asfortranarray(np.asarray(image, dtype=np.float32))
    if img.ndim != 3:
        raise ValueError('image must be 3-D')
    starts, lengths = crop_window(img.shape, voxel_size, crop_size)
    if min(lengths) < 1:
        raise ValueError(f'the crop leaves an empty image (lengths {lengths})')
    out_shape = [int(v) for v in downscaled_dimensions]
    if min(out_shape) < 1:
        raise ValueError('downscaled dimensions must be positive')
    factors = np.divide(lengths, out_shape)
    taps = [gaussian_taps(max(0.0, (f - 1) / 2)) if anti_aliasing else None for f in factors]
    d = _ResizeDesc()
    for k in range(3):
        d.in_shape[k] = img.shape[k]
        d.crop_start[k] = starts[k]
        d.crop_shape[k] = lengths[k]
        d.out_shape[k] = out_shape[k]
        d.radius[k] = -1 if taps[k] is None else (taps[k].size - 1) // 2
        d.weights[k] = None if taps[k] is None else taps[k].ctypes.data
    d.clip = int(bool(clip))
    d.zero_below = np.float32(-np.inf if zero_below is None else zero_below)
    out = np.empty(out_shape, dtype=np.float32, order='F')
    ctx = context or _native.default_context()
    lib = _native.load_library()
    flat = img.ravel(order='K')   # x fastest
    _native.check(lib.sss_crop_and_downscale(ctx.handle, _native.ptr(flat), _native.ctypes.byref(d),
                                             out.ctypes.data_as(_native.ctypes.c_void_p)), 'sss_crop_and_downscale')
    return out


def LoadImage(Filename: str):
    """(image float32 (X, Y, Z), matrix size, voxel size mm) of an interfile
    header and its data file (Utils.py:104-140)."""
    with open(Filename, 'r') as fh:
        header = fh.read()
    scale = np.zeros(3, dtype=float)
    size = np.zeros(3, dtype=int)
    for i in range(3):
        size[i] = int(re.findall(rf'.*matrix size \[{i + 1}\] := (\d*).*', header)[0])
        scale[i] = float(re.findall(rf'.*scaling factor \(mm/pixel\) \[{i + 1}\] := (\d*\.*\d+).*', header)[0])
    datafile = re.findall(r'.*name of data file := (\S*).*', header)[0]
    image = np.fromfile(os.path.dirname(Filename) + f'/{datafile}', dtype=np.float32)
    image = image.reshape(size[2], size[1], size[0]).transpose((2, 1, 0))
    return image, size, scale


def PrepareImages(ActivityFilename, AttenuationFilename, SizeDesired, ScaleDesired, *, context=None):
    """(activity, attenuation (mu < 0.001 set to 0), size mm, dimensions) of
    the reference's image preparation (Utils.py:29-76)."""
    activity, dims_act, size_act = LoadImage(ActivityFilename)
    attenuation, _, size_atn = LoadImage(AttenuationFilename)
    SizeDesired = list(SizeDesired)
    for i, dim in enumerate(SizeDesired):
        if dim is None:
            SizeDesired[i] = dims_act[i] * size_act[i]
    SizeDesired = np.array(SizeDesired, dtype=float)
    ScaleDesired = np.asarray(ScaleDesired)
    DimsDesired = (SizeDesired // ScaleDesired).astype(int)
    SizeDesired = DimsDesired * ScaleDesired
    act = CropAndDownscale(activity, size_act, SizeDesired, DimsDesired, context=context)
    atn = CropAndDownscale(attenuation, size_atn, SizeDesired, DimsDesired, zero_below=0.001, context=context)
    return act, atn, SizeDesired, DimsDesired


def ComputePhisicalDimensions(DesiredSize, DeviceSize):
    """(ImageSize, FittingSize) in cm from sizes in mm (Utils.py:78-102)."""
    image = np.array([x / 10 for x in [-DesiredSize[0] / 2, -DesiredSize[1] / 2, -DesiredSize[2] / 2,
                                       DesiredSize[0] / 2, DesiredSize[1] / 2, DesiredSize[2] / 2]])
    fitting = np.array([x / 10 for x in [-DesiredSize[0] / 2, -DesiredSize[1] / 2, -DeviceSize[2] / 2,
                                         DesiredSize[0] / 2, DesiredSize[1] / 2, DeviceSize[2] / 2]])
    return image, fitting
