"""Build libnii2mesh_amd.so for gfx950 with hipcc (in-tree, no JIT cache).

    python -m nii2mesh_amd.build          # or nii2mesh_amd.build.build()

Every .hip / .cpp under csrc/ is compiled to an object in _build/ (parallel),
then linked into nii2mesh_amd/lib/libnii2mesh_amd.so; only the C ABI of
include/nii2mesh_amd.h is exported (csrc/exports.map).
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, 'csrc')
LIBDIR = os.path.join(PKG, 'lib')
LIB = os.path.join(LIBDIR, 'libnii2mesh_amd.so')
BUILD = os.path.join(ROOT, '_build')
ARCH = os.environ.get('N2M_OFFLOAD_ARCH', 'gfx950')

# -ffp-contract=off: no fused multiply-add, so float/double expressions round
# like the reference's separate operations; denormals kept (IEEE).
# packed-fp32-ops disabled: the compiler may not emit v_pk_*_f32.
HIPFLAGS = ['-O3', '-std=c++17', '-fPIC', '--offload-arch=' + ARCH, '-ffp-contract=off',
            '-fno-gpu-flush-denormals-to-zero', '-fno-fast-math',
            '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
            '-Wall', '-Wno-unused-function', '-Wno-unused-variable']


def hipcc():
    for c in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', shutil.which('hipcc')):
        if c and os.path.exists(c):
            return c
    raise RuntimeError('hipcc not found')


def _compile(src):
    obj = os.path.join(BUILD, os.path.basename(src) + '.o')
    cmd = [hipcc()] + HIPFLAGS + ['-I' + os.path.join(ROOT, 'include'), '-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))
    with cf.ThreadPoolExecutor(max_workers=min(8, len(srcs))) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs + [
        '-Wl,--version-script=' + os.path.join(CSRC, 'exports.map'), '-Wl,-Bsymbolic', '-lz']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
