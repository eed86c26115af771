"""Build libuw12g.so (HIP, gfx950) in-tree with hipcc.

Usage: python -m libuw12_amd.build  (or libuw12_amd.build.build()).
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
OUT = os.path.join(HERE, 'libuw12g.so')
BUILD_DIR = os.path.join(os.path.dirname(HERE), '_build', 'libuw12g')
ROCM = os.environ.get('ROCM_PATH', '/opt/rocm')
ARCH = os.environ.get('UW12G_ARCH', 'gfx950')

SOURCES = ['gemm_f64.hip', 'gemm_planes_wide.hip', 'kernels.hip', 'uw12g.cpp', 'uw12g_stream.cpp', 'testing.cpp']

COMMON = [
    '-O3', '-std=c++17', '-fPIC', f'--offload-arch={ARCH}',
    # keep hipcc from emitting packed FP32 instructions anywhere in this library
    '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
    # MFMA accumulators in ordinary VGPRs: no per-iteration AGPR copies, 2 waves/SIMD at 128x128 tiles
    '-mllvm', '-amdgpu-mfma-vgpr-form',
    '-Wall', '-Wno-unused-result', '-Wno-unused-variable',
]


def _hipcc():
    p = os.path.join(ROCM, 'bin', 'hipcc')
    return p if os.path.exists(p) else shutil.which('hipcc')


def _compile(src):
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + '.o')
    cmd = [_hipcc(), *COMMON, '-x', 'hip', '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(BUILD_DIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(len(SOURCES), 8)) as ex:
        results = list(ex.map(_compile, SOURCES))
    if verbose:
        for obj, err in results:
            if err.strip():
                print(err, file=sys.stderr)
    tmp = OUT + '.tmp'
    cmd = [_hipcc(), f'--offload-arch={ARCH}', '-shared', '-fPIC', '-o', tmp, *[o for o, _ in results],
           f'-L{ROCM}/lib', '-lrocsolver', '-lrocblas', f'-Wl,-rpath,{ROCM}/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, OUT)
    return OUT


SYNTH_DEV_OUT = os.path.join(HERE, 'libuw12_synth_dev.so')


def build_synth_dev():
    """Device generator of the synthetic s-Gaussian integrals (csrc/synth_dev.hip): the
    three-index blocks of the benchmark inputs, written straight into the library's
    device buffers through the uw12g_three_index_gen_fn callback."""
    os.makedirs(BUILD_DIR, exist_ok=True)
    tmp = SYNTH_DEV_OUT + '.tmp'
    cmd = [_hipcc(), *COMMON, '-shared', '-x', 'hip', os.path.join(CSRC, 'synth_dev.hip'), '-o', tmp,
           f'-Wl,-rpath,{ROCM}/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for synth_dev.hip:\n{r.stderr}')
    os.replace(tmp, SYNTH_DEV_OUT)
    return SYNTH_DEV_OUT


def build_cpp_driver():
    """C++ caller of include/uw12g.hpp used by the C++ API test (tests/cpp)."""
    root = os.path.dirname(HERE)
    out = os.path.join(root, 'tests', 'cpp', 'uw12_cpp_driver')
    cmd = ['g++', '-O2', '-std=c++17', '-I', os.path.join(root, 'include'),
           os.path.join(root, 'tests', 'cpp', 'uw12_cpp_driver.cpp'), '-o', out, f'-L{HERE}', '-luw12g',
           f'-Wl,-rpath,{HERE}', f'-Wl,-rpath,{ROCM}/lib']
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'C++ driver build failed:\n{r.stderr}')
    return out


if __name__ == '__main__':
    print(build(verbose=True))
    print(build_synth_dev())
    print(build_cpp_driver())
