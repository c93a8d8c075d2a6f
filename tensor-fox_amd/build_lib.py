"""Builds libtfxhip.so (the HIP sources and host C++ sources under csrc/) for gfx950 with hipcc,
in-tree.

    python tensor-fox_amd/build_lib.py            # -> tensor-fox_amd/TensorFox/libtfxhip.so
"""

import concurrent.futures
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, 'csrc')
BUILD = os.path.join(HERE, '_build')
OUT = os.path.join(HERE, 'TensorFox', 'libtfxhip.so')

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('TFX_OFFLOAD_ARCH', 'gfx950')
FLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH,
         '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
         '-I' + os.path.join(ROOT, 'include'), '-I' + CSRC,
         '-Wno-unused-result']


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + '.o')
    cmd = [HIPCC] + FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('hipcc failed for %s:\n%s\n%s' % (src, ' '.join(cmd), r.stderr[-6000:]))
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))
    workers = min(len(srcs), max(1, min(8, os.cpu_count() or 1)))
    with concurrent.futures.ThreadPoolExecutor(workers) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = OUT + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n%s\n%s' % (' '.join(cmd), r.stderr[-4000:]))
    os.replace(tmp, OUT)
    if verbose:
        print('built', OUT, flush=True)
    return OUT


if __name__ == '__main__':
    build()
    sys.exit(0)
