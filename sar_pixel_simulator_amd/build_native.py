"""Compile libfdtd_hip.so for gfx950 in-tree with hipcc.

    python -m sar_pixel_simulator_amd.build_native

Flags that matter for results:
  -ffp-contract=off     one rounding per operation in the field updates, so
                        results are bit-identical to the CPU path's loops
  -packed-fp32-ops off  no packed FP32 instructions are generated
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB_NAME = 'libfdtd_hip.so'
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ['fdtd_api.hip', 'fdtd_kernels.hip', 'fdtd_beam.hip', 'fdtd_materials.hip']
LIBS = ['-L/opt/rocm/lib', '-lhipfft', '-Wl,-rpath,/opt/rocm/lib']
ARCH = os.environ.get('FDTD_OFFLOAD_ARCH', 'gfx950')

HIPCC_FLAGS = [
    f'--offload-arch={ARCH}', '-O3', '-std=c++17', '-fPIC', '-shared',
    '-ffp-contract=off',
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    '-Wall', '-Wno-unused-function',
]


def hipcc():
    for cand in (os.environ.get('HIPCC'), '/opt/rocm/bin/hipcc', 'hipcc'):
        if cand and (os.path.isabs(cand) and os.path.exists(cand) or not os.path.isabs(cand)):
            return cand
    raise RuntimeError('hipcc not found')


def build(verbose=False):
    """Compile the HIP library; raises CalledProcessError on failure."""
    srcs = [os.path.join(CSRC, s) for s in SOURCES]
    tmp = LIB_PATH + '.tmp'
    cmd = [hipcc()] + HIPCC_FLAGS + ['-o', tmp] + srcs + LIBS
    if verbose:
        print(' '.join(cmd), flush=True)
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == '__main__':
    build(verbose='-q' not in sys.argv)
