"""Ahead-of-time build of the HIP parts for gfx950 (the role OCCA's JIT plays in the
reference, benchmark/src/kinetix.cpp:409-475, done once at build time instead).

  * kinetix_amd/lib/libkinetix_mi.so  -- host library (C ABI of include/kinetix_mi.h) plus the
                                         mechanism-independent kernels (surrogate MLP)
  * kinetix_amd/mechs/<name>/kernels.hsaco -- one code object per generated mechanism
"""
from __future__ import annotations

import glob
import os
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
ARCH = 'gfx950'
# -packed-fp32-ops: no packed FP32 (v_pk_*_f32) code generation.
COMMON = ['-O3', '-std=c++17', f'--offload-arch={ARCH}', '-ffp-contract=fast',
          '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops']
HOST_SOURCES = ['kx_host.cpp', 'kx_selftest.hip', 'kx_mlp.hip']


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError('build failed: ' + ' '.join(cmd))
    return r


def build_mechanism(mech_dir, extra=()):
    src = os.path.join(mech_dir, 'kernels.hip')
    out = os.path.join(mech_dir, 'kernels.hsaco')
    inc = '-I' + os.path.join(PKG, 'csrc')          # kx_mlp_wave.h (surrogate inference)
    rinc = '-I' + os.path.join(mech_dir, 'routines')   # the generated routine files
    _run([HIPCC, '--genco', *COMMON, inc, rinc, *extra, src, '-o', out])
    return out


def build_host(extra=()):
    libdir = os.path.join(PKG, 'lib')
    os.makedirs(libdir, exist_ok=True)
    out = os.path.join(libdir, 'libkinetix_mi.so')
    srcs = [os.path.join(PKG, 'csrc', s) for s in HOST_SOURCES]
    _run([HIPCC, *COMMON, '-fPIC', '-shared', *extra, *srcs, '-o', out])
    return out


def mechanism_dirs():
    return sorted(os.path.dirname(p) for p in glob.glob(os.path.join(PKG, 'mechs', '*', 'kernels.hip')))


def build_all(jobs=None):
    """Host library and every mechanism's code object, compiled concurrently (one hipcc
    process each; the large straight-line mechanisms take minutes apiece)."""
    from concurrent.futures import ThreadPoolExecutor
    jobs = jobs or min(8, os.cpu_count() or 1)
    with ThreadPoolExecutor(jobs) as pool:
        futs = [pool.submit(build_host)] + [pool.submit(build_mechanism, d) for d in mechanism_dirs()]
        for f in futs:
            f.result()


if __name__ == '__main__':
    build_all()
