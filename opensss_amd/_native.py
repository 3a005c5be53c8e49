"""ctypes binding of libsss_hip.so (declarations in include/sss_hip.h).

The library is built in-tree by ``python -m opensss_amd.build``.  There is no
CPU fallback: every entry point of the package raises if the library is
missing or no HIP device is available.
"""
from __future__ import annotations

import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, '_lib', 'libsss_hip.so')

SSS_OK = 0
SSS_PRECISION_FP32 = 0
SSS_PRECISION_FP64 = 1
SSS_VOLUME_X_FASTEST = 0
SSS_VOLUME_Z_FASTEST = 1

_c_i32p = ctypes.POINTER(ctypes.c_int32)
_c_f64p = ctypes.POINTER(ctypes.c_double)


class SssError(RuntimeError):
    """A libsss_hip.so call returned a non-zero status."""

    def __init__(self, code: int, where: str, message: str):
        super().__init__(f'{where} failed ({code}): {message}')
        self.code = code


class SubjectDesc(ctypes.Structure):
    _fields_ = [
        ('grid', ctypes.c_int32 * 3), ('bounds', ctypes.c_double * 6),
        ('mu', ctypes.c_void_p),
        ('n_rings', ctypes.c_int32), ('n_det', ctypes.c_int32),
        ('geom', ctypes.c_void_p), ('normals', ctypes.c_void_p),
        ('det_size', ctypes.c_double * 2),
        ('ratio', ctypes.c_void_p), ('eff', ctypes.c_void_p),
        ('n_energy', ctypes.c_int32), ('eff_ref_row', ctypes.c_int32),
        ('nru', ctypes.c_int32), ('ndu', ctypes.c_int32),
        ('rings', ctypes.c_void_p), ('dets', ctypes.c_void_p),
        ('n_sp', ctypes.c_int32), ('sp_xyz', ctypes.c_void_p),
        ('n_slices', ctypes.c_int32), ('slice_ptr', ctypes.c_void_p),
        ('n_bins', ctypes.c_int32),
        ('tof_bin_width', ctypes.c_double), ('tof_kappa', ctypes.c_double),
        ('tof_beta', ctypes.c_double), ('tof_half_window', ctypes.c_int32),
        ('tof_t_lo', ctypes.c_int32), ('tof_t_hi', ctypes.c_int32),
        ('tof_trivial', ctypes.c_int32), ('precision', ctypes.c_int32),
        ('n_chunks', ctypes.c_int32), ('tile_rings', ctypes.c_int32), ('volume_order', ctypes.c_int32),
    ]


class SubjectInfo(ctypes.Structure):
    _fields_ = [
        ('n_halfpaths', ctypes.c_int64), ('n_samples', ctypes.c_int64),
        ('n_cells', ctypes.c_int64), ('n_pairs_enumerated', ctypes.c_int64),
        ('evals', ctypes.c_int64), ('device_bytes', ctypes.c_int64),
        ('n_tiles', ctypes.c_int32), ('n_chunks', ctypes.c_int32),
        ('n_truncated_walks', ctypes.c_int32), ('tof_fold', ctypes.c_int32),
    ]


# (name, restype, argtypes) of every exported symbol declared in include/sss_hip.h
SIGNATURES = [
    ('sss_last_error', ctypes.c_char_p, []),
    ('sss_version', ctypes.c_int, []),
    ('sss_ctx_create', ctypes.c_int, [ctypes.c_int, ctypes.POINTER(ctypes.c_void_p)]),
    ('sss_ctx_destroy', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_ctx_synchronize', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_ctx_stream', ctypes.c_void_p, [ctypes.c_void_p]),
    ('sss_ctx_release_cache', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_ctx_cached_bytes', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)]),
    ('sss_ctx_memory', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ('sss_host_alloc', ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    ('sss_host_free', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_trace_segments', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_line_integrals', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_subject_create', ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(SubjectDesc), ctypes.POINTER(ctypes.c_void_p)]),
    ('sss_subject_destroy', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_trace', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_trace_begin', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_trace_end', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_set_activity', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_subject_set_activity_device', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_subject_scatter', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_run', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_info_get', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(SubjectInfo)]),
    ('sss_subject_work_stats', ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    ('sss_subject_scatter_kernel_ms', ctypes.c_int,
     [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ('sss_subject_output_device', ctypes.c_void_p, [ctypes.c_void_p]),
    ('sss_subject_output_compact', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_subject_output_compact_async', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_subject_wait', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_set_tiling', ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]),
    ('sss_batch_create', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p), ctypes.c_int,
                                        ctypes.POINTER(ctypes.c_void_p)]),
    ('sss_batch_set_activity', ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    ('sss_batch_submit', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_batch_graph_launch', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_batch_wait', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_batch_output', ctypes.c_int, [ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    ('sss_batch_elapsed_ms', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_float)]),
    ('sss_batch_info', ctypes.c_int, [ctypes.c_void_p, _c_i32p, _c_i32p]),
    ('sss_batch_destroy', ctypes.c_int, [ctypes.c_void_p]),
    ('sss_subject_query', ctypes.c_int, [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int32)]),
    ('sss_subject_output_dense', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_buffer_alloc', ctypes.c_int, [ctypes.c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_void_p)]),
    ('sss_buffer_free', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_buffer_upload', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ('sss_buffer_download', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]),
    ('sss_interpolate', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_void_p,
      ctypes.c_int64, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_tail_mask', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p]),
    ('sss_tail_fit', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ('sss_forward_project', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p]),
    ('sss_poisson', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_double, ctypes.c_uint64, ctypes.c_void_p]),
    ('sss_tail_fit_noalloc', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p,
      ctypes.c_void_p]),
    ('sss_scale_to_prompts', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p]),
    ('sss_crop_and_downscale', ctypes.c_int, [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_listmode_histogram', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
      ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_int, ctypes.c_void_p,
      ctypes.c_void_p, ctypes.c_void_p]),
    ('sss_listmode_inject', ctypes.c_int,
     [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64, ctypes.c_void_p, ctypes.c_int64, ctypes.c_int64,
      ctypes.c_int, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
      ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int,
      ctypes.c_void_p, ctypes.c_void_p]),
]

SSS_LM_RANDOMS = 1
SSS_LM_NORM = 2
SSS_LM_DEVICE = 4
SSS_PROMPTS_F32 = 0
SSS_PROMPTS_U16 = 1
SSS_FIT_DEVICE = 1


class DeviceBuffer:
    """A device allocation owned by a context (``sss_buffer_*``)."""

    def __init__(self, ctx: 'Context', nbytes: int):
        self._lib = load_library()
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        check(self._lib.sss_buffer_alloc(ctx.handle, self.nbytes, ctypes.byref(p)), 'sss_buffer_alloc')
        self.ptr = p

    @classmethod
    def from_array(cls, ctx, arr: np.ndarray) -> 'DeviceBuffer':
        a = np.ascontiguousarray(arr)
        buf = cls(ctx, a.nbytes)
        buf.upload(a)
        return buf

    def upload(self, arr: np.ndarray) -> None:
        a = np.ascontiguousarray(arr)
        if a.nbytes > self.nbytes:
            raise ValueError('array larger than the device buffer')
        check(self._lib.sss_buffer_upload(self.ctx.handle, self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                          a.nbytes), 'sss_buffer_upload')

    def download(self, out: np.ndarray) -> np.ndarray:
        if not out.flags['C_CONTIGUOUS'] or out.nbytes > self.nbytes:
            raise ValueError('bad host buffer')
        check(self._lib.sss_buffer_download(self.ctx.handle, out.ctypes.data_as(ctypes.c_void_p),
                                            self.ptr, out.nbytes), 'sss_buffer_download')
        return out

    def free(self) -> None:
        if getattr(self, 'ptr', None):
            self._lib.sss_buffer_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass

_lib = None
_lock = threading.Lock()


def load_library() -> ctypes.CDLL:
    """Load libsss_hip.so (raises if it has not been built)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise RuntimeError(f'{LIB_PATH} is missing: build it with `python -m opensss_amd.build`; '
                                   'there is no CPU fallback')
            lib = ctypes.CDLL(LIB_PATH)
            for name, res, args in SIGNATURES:
                fn = getattr(lib, name)
                fn.restype = res
                fn.argtypes = args
            _lib = lib
    return _lib


def check(rc: int, where: str) -> None:
    if rc != SSS_OK:
        msg = load_library().sss_last_error()
        raise SssError(rc, where, msg.decode() if msg else '')


def ptr(a: np.ndarray) -> ctypes.c_void_p:
    return ctypes.c_void_p(a.ctypes.data)


class Context:
    """One HIP device and one stream (``sss_ctx``)."""

    def __init__(self, device: int = 0):
        self._lib = load_library()
        h = ctypes.c_void_p()
        check(self._lib.sss_ctx_create(int(device), ctypes.byref(h)), 'sss_ctx_create')
        self.handle = h
        self.device = device

    @property
    def stream(self) -> int:
        return int(self._lib.sss_ctx_stream(self.handle) or 0)

    def synchronize(self) -> None:
        check(self._lib.sss_ctx_synchronize(self.handle), 'sss_ctx_synchronize')

    def memory(self):
        """(free, total) device bytes of the context's GPU."""
        f, t = ctypes.c_int64(), ctypes.c_int64()
        check(self._lib.sss_ctx_memory(self.handle, ctypes.byref(f), ctypes.byref(t)), 'sss_ctx_memory')
        return f.value, t.value

    def release_cache(self) -> None:
        """Free the device blocks kept for later subjects."""
        check(self._lib.sss_ctx_release_cache(self.handle), 'sss_ctx_release_cache')

    def host_alloc(self, nbytes: int) -> int:
        """A pinned host buffer (address) from the context's cache."""
        p = ctypes.c_void_p()
        check(self._lib.sss_host_alloc(self.handle, int(nbytes), ctypes.byref(p)), 'sss_host_alloc')
        return int(p.value)

    def host_free(self, addr: int) -> None:
        check(self._lib.sss_host_free(self.handle, ctypes.c_void_p(addr)), 'sss_host_free')

    @property
    def cached_bytes(self) -> int:
        v = ctypes.c_int64()
        check(self._lib.sss_ctx_cached_bytes(self.handle, ctypes.byref(v)), 'sss_ctx_cached_bytes')
        return int(v.value)

    def close(self) -> None:
        if getattr(self, 'handle', None):
            self._lib.sss_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default_ctx = {}


def default_context(device: int = 0) -> Context:
    with _lock:
        ctx = _default_ctx.get(device)
    if ctx is None:
        ctx = Context(device)
        with _lock:
            _default_ctx[device] = ctx
    return ctx
