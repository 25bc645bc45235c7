"""Build libemu_mi355x.so (HIP kernels + C-ABI) for gfx950 with hipcc.

Usage: python -m emustack_amd.build   (or emustack_amd.build.build()).
Objects go to emustack_amd/_build/, the library next to this file.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OBJ = os.path.join(HERE, '_build')
LIB = os.path.join(HERE, 'libemu_mi355x.so')
SOURCES = ['mesh.cpp', 'mesh1d.cpp', 'symbolic.cpp', 'capi.cpp', 'kernels_fem.hip', 'kernels_lu.hip',
           'kernels_krylov.hip', 'kernels_post.hip', 'kernels_1d.hip', 'cascade.hip', 'kernels_material.hip', 'kernels_fields.hip', 'noalloc.hip', 'engine.hip']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = os.environ.get('EMU_OFFLOAD_ARCH', 'gfx950')
# packed-fp32 instructions (v_pk_*_f32) are disabled for all device code:
# some of their forms have produced silently wrong results on this hardware
DEVFLAGS = ['-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
# host complex arithmetic: multiplication by the plain formula (no NaN
# recovery call path per product, which keeps the projected-matrix loops of the
# host Schur step from vectorising) and division by Smith's scaled algorithm
CXXFLAGS = ['-O3', '-fPIC', '-std=c++17', '-Wall', '-Wno-unused-function', '-Wno-unused-variable',
            '-Wno-unused-but-set-variable', '-Wno-unused-result', '-fcomplex-arithmetic=improved']


def _compile(src):
    path = os.path.join(CSRC, src)
    out = os.path.join(OBJ, src + '.o')
    if src.endswith('.hip'):
        cmd = [HIPCC, '-c', '-x', 'hip', path, '-o', out, '--offload-arch=' + ARCH] + DEVFLAGS
    else:
        cmd = [HIPCC, '-c', path, '-o', out, '-D__HIP_PLATFORM_AMD__']
    cmd += CXXFLAGS + ['-I' + CSRC]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    return out, r.stderr


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(_compile, SOURCES))
    for _, err in results:
        if verbose and err.strip():
            sys.stderr.write(err)
    objs = [o for o, _ in results]
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs + ['-lpthread']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    return LIB


if __name__ == '__main__':
    print(build(verbose=True))
