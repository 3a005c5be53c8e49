"""Build libsss_hip.so (gfx950) in-tree with hipcc.

``python -m opensss_amd.build`` compiles every ``csrc/*.hip`` translation unit
and links ``opensss_amd/_lib/libsss_hip.so``.  Packed-FP32 instruction
selection is disabled for all device code.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, '_lib')
OBJ_DIR = os.path.join(OUT_DIR, 'obj')
LIB_NAME = 'libsss_hip.so'
ARCH = os.environ.get('SSS_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-munsafe-fp-atomics', '-Wno-unused-command-line-argument',
]


def hipcc() -> str:
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found')
    return path


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith('.hip'))


def _compile(src: str) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src)[:-4] + '.o')
    cmd = [hipcc(), *HIPCC_FLAGS, '-c', src, '-o', obj]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{res.stderr[-6000:]}')
    return obj


def build(verbose: bool = False) -> str:
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    with ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(OUT_DIR, LIB_NAME)
    tmp = out + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', f'--offload-arch={ARCH}', *objs, '-o', tmp]
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'link failed:\n{res.stderr[-6000:]}')
    os.replace(tmp, out)
    if verbose:
        print(f'built {out}')
    return out


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
