"""Build libprojector.so for MI355X (gfx950) with hipcc.

    python -m ctprojector_amd.build [--jobs N] [--verbose]

Every source under ctprojector_amd/csrc is compiled on each call (hipcc
cross-compiles without a GPU); the library is linked in-tree at
ctprojector_amd/kernel/bin/libprojector.so, the path the reference's ctypes
loaders use (src/ct_projector/projector/numpy/ct_projector.py:15-17).
"""

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT_DIR = os.path.join(HERE, 'kernel', 'bin')
LIB = os.path.join(OUT_DIR, 'libprojector.so')
OBJ_DIR = os.path.join(HERE, 'kernel', '_obj')
ARCH = os.environ.get('CTP_OFFLOAD_ARCH', 'gfx950')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')

COMMON = ['-O3', '-std=c++17', '-fPIC', '-fvisibility=hidden', '-Wall', '-Wno-unused-function',
          '-Wno-unused-result']
# packed FP32 VALU instructions are disabled for every device compilation
DEVICE = ['--offload-arch=' + ARCH, '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
          '-fgpu-flush-denormals-to-zero']


def hipcc() -> str:
    path = shutil.which('hipcc') or '/opt/rocm/bin/hipcc'
    if not os.path.exists(path):
        raise RuntimeError('hipcc not found: the ROCm toolchain is required to build libprojector.so')
    return path


def sources():
    out = []
    for f in sorted(os.listdir(CSRC)):
        if f.endswith('.hip') or f.endswith('.cpp'):
            out.append(os.path.join(CSRC, f))
    return out


def _compile(src: str, verbose: bool) -> str:
    obj = os.path.join(OBJ_DIR, os.path.basename(src) + '.o')
    cmd = [hipcc()] + COMMON + ['-I', CSRC]
    if src.endswith('.hip'):
        cmd += DEVICE + ['-x', 'hip']
    else:
        cmd += ['-x', 'c++', '-D__HIP_PLATFORM_AMD__', '-I', os.path.join(ROCM, 'include')]
    cmd += ['-c', src, '-o', obj]
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError('compilation failed: %s\n%s' % (src, r.stdout))
    if verbose and r.stdout.strip():
        print(r.stdout)
    return obj


def build(jobs: int = 0, verbose: bool = False) -> str:
    os.makedirs(OUT_DIR, exist_ok=True)
    os.makedirs(OBJ_DIR, exist_ok=True)
    srcs = sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, verbose), srcs))
    tmp = LIB + '.tmp'
    cmd = [hipcc(), '-shared', '-fPIC', '--offload-arch=' + ARCH, '-o', tmp] + objs
    if verbose:
        print(' '.join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError('link failed:\n' + r.stdout)
    os.replace(tmp, LIB)
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument('--jobs', type=int, default=0)
    ap.add_argument('--verbose', action='store_true')
    a = ap.parse_args(argv)
    print(build(a.jobs, a.verbose))


if __name__ == '__main__':
    sys.exit(main())
