"""The C-ABI library loads and exports every function include/sss_hip.h declares
(no compute: runs without a GPU)."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = []
    for fn in sorted(os.listdir(os.path.join(ROOT, 'include'))):
        if not fn.endswith('.h'):
            continue
        text = open(os.path.join(ROOT, 'include', fn)).read()
        text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
        names += re.findall(r'^\s*(?:const\s+)?[\w\s\*]+?\b(sss_\w+)\s*\(', text, flags=re.M)
    return sorted(set(names))


def test_header_declares_the_abi():
    names = declared_functions()
    for must in ('sss_ctx_create', 'sss_subject_create', 'sss_subject_scatter',
                 'sss_trace_segments', 'sss_last_error'):
        assert must in names


def test_library_exports_every_declared_symbol():
    from opensss_amd import _native
    if not os.path.exists(_native.LIB_PATH):
        pytest.fail('libsss_hip.so is not built (run python -m opensss_amd.build)')
    lib = ctypes.CDLL(_native.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    bound = {name for name, _, _ in _native.SIGNATURES}
    assert set(declared_functions()) <= bound


def test_last_error_is_callable_without_device():
    from opensss_amd import _native
    lib = _native.load_library()
    assert isinstance(lib.sss_last_error(), bytes)
    assert lib.sss_version() >= 1
