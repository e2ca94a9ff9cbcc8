"""Build libfe_hip.so for gfx950 with hipcc (in-tree, next to the package).

Usage: python -m finite_elements_amd.csrc.build  [--verbose]
"""

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
INCLUDE = os.path.join(ROOT, 'include')
LIB_DIR = os.path.join(os.path.dirname(HERE), 'lib')
LIB = os.path.join(LIB_DIR, 'libfe_hip.so')
OBJ_DIR = os.path.join(ROOT, '_build')

SOURCES = ['fe_api.hip', 'fe_element.hip', 'fe_assembly.hip', 'fe_solve.hip', 'fe_pcg.hip', 'fe_coarse.hip', 'fe_gnn.hip', 'fe_lobpcg.hip', 'fe_post.hip',
           'fe_pattern.cpp']

HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('FE_OFFLOAD_ARCH', 'gfx950')

# Packed FP32 instructions are disabled for all device code (they have produced silent wrong results
# on these machines); hipcc -O3 would otherwise form them from ordinary float code.
DEVICE_FLAGS = ['-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
COMMON = ['-O3', '-std=c++17', '-fPIC', f'-I{INCLUDE}', f'-I{HERE}', '-Wall', '-Wno-unused-function',
          '-Wno-unused-variable', '-Wno-unused-but-set-variable']


def _compile(src, verbose=False):
    out = os.path.join(OBJ_DIR, os.path.splitext(src)[0] + '.o')
    path = os.path.join(HERE, src)
    if src.endswith('.hip'):
        cmd = [HIPCC, '-x', 'hip', f'--offload-arch={ARCH}', *DEVICE_FLAGS, *COMMON, '-c', path,
               '-o', out]
    else:
        cmd = ['g++', *COMMON, '-c', path, '-o', out]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'compiling {src} failed:\n{res.stdout}\n{res.stderr}')
    if verbose and res.stderr.strip():
        print(res.stderr, file=sys.stderr)
    return out


def build(verbose=False):
    """Compile every source and link libfe_hip.so; returns the library path."""
    os.makedirs(OBJ_DIR, exist_ok=True)
    os.makedirs(LIB_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *objs]
    if verbose:
        print(' '.join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f'linking libfe_hip.so failed:\n{res.stdout}\n{res.stderr}')
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose='--verbose' in sys.argv))
