"""Build libtopgpu.so (gfx950) in-tree: hipcc for the HIP sources, g++ for host-only C++.

    python -m top_gpu_on_matlab_amd.build            # or top_gpu_on_matlab_amd.build.build()

Objects go to top_gpu_on_matlab_amd/_build/, the library to top_gpu_on_matlab_amd/libtopgpu.so.
"""
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(PKG, 'csrc')
OUT = os.path.join(PKG, '_build')
LIB = os.path.join(PKG, 'libtopgpu.so')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = os.environ.get('TOPGPU_ARCH', 'gfx950')

HIP_FLAGS = ['-O3', '-std=c++17', '--offload-arch=' + ARCH, '-fPIC', '-Wall', '-Wno-unused-function',
             # packed FP32 instructions are disabled for all device code
             '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
             '-I' + os.path.join(PKG, '..', 'include')]
CXX_FLAGS = ['-O3', '-std=c++17', '-fPIC', '-Wall', '-fopenmp', '-D__HIP_PLATFORM_AMD__',
             '-I' + os.path.join(ROCM, 'include'), '-I' + os.path.join(PKG, '..', 'include')]


def _hipcc():
    p = shutil.which('hipcc') or os.path.join(ROCM, 'bin', 'hipcc')
    if not os.path.exists(p):
        raise RuntimeError('hipcc not found (ROCm required to build libtopgpu.so)')
    return p


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src) + '.o')
    if src.endswith('.hip'):
        cmd = [_hipcc()] + HIP_FLAGS + ['-c', src, '-o', obj]
    else:
        cmd = ['g++'] + CXX_FLAGS + ['-c', src, '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('compile failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    return obj


def build(verbose=True, jobs=None):
    """Compile every source in csrc/ and link libtopgpu.so.  Always rebuilds."""
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(('.hip', '.cpp')))
    jobs = jobs or min(8, os.cpu_count() or 4)
    with concurrent.futures.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, srcs))
    tmp = LIB + '.tmp'
    cmd = [_hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs + [
        '-L' + os.path.join(ROCM, 'lib'), '-fopenmp',
        '-Wl,-rpath,' + os.path.join(ROCM, 'lib')]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed: %s\n%s\n%s' % (' '.join(cmd), r.stdout, r.stderr))
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB, '(%d sources)' % len(srcs))
    return LIB


if __name__ == '__main__':
    build(verbose=True)
    sys.exit(0)
