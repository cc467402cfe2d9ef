"""Build of the HIP library ``libabf.so`` (gfx950) from ``csrc/``, in-tree."""

from __future__ import annotations

import os
import subprocess
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, 'csrc')
LIB = os.path.join(HERE, 'libabf.so')
# (source, extra flags, object name); the streamed kernels' files build once per block count
STREAM = [(f'abf_bf_mfma_{kind}.hip', [f'-DABF_ST_NK={nk}'], f'abf_bf_mfma_{kind}_{nk}.o') for nk in range(16, 1, -1)
          for kind in ('stream', 'wide') if kind == 'wide' or nk <= 8]
SOURCES = ['abf_bf_mfma_bf16.hip', 'abf_bf_mfma_f16.hip', 'abf_bf_mfma_f32.hip', 'abf_bf.hip', 'abf_steer.hip', 'abf_csm.hip',
           'abf_eig.hip', 'abf_deconv.hip', 'abf_sources.hip', 'abf_clean.hip', 'abf_resample.hip',
           'abf_welch_fft.hip', 'abf_api.hip']
# the batched Hermitian eigensolver of the eigenvalue path; the FFTs of the signal resampling
LIBS = ['-L/opt/rocm/lib', '-Wl,-rpath,/opt/rocm/lib', '-lrocsolver', '-lrocblas', '-lrocfft']
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')
# packed fp32 instructions (v_pk_*_f32) are disabled on purpose: some of their forms have given
# silently wrong results on MI355X boxes
FLAGS = ['--offload-arch=gfx950', '-O3', '-fPIC', '-std=c++17', '-Xclang', '-target-feature', '-Xclang',
         '-packed-fp32-ops', '-I', os.path.join(HERE, '..', 'include')]


def _compile(unit, build_dir: str) -> str:
    src, extra, name = unit if isinstance(unit, tuple) else (unit, [], unit.replace('.hip', '.o'))
    obj = os.path.join(build_dir, name)
    cmd = [HIPCC, *FLAGS, *extra, '-c', os.path.join(CSRC, src), '-o', obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'hipcc failed for {src}:\n{r.stderr}')
    return obj


def build(verbose: bool = False) -> str:
    build_dir = os.path.join(HERE, '_build')
    os.makedirs(build_dir, exist_ok=True)
    # the slowest units first
    with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, build_dir), STREAM + SOURCES))
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=gfx950', '-shared', '-fPIC', '-o', tmp, *objs, *LIBS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f'link failed:\n{r.stderr}')
    os.replace(tmp, LIB)
    if verbose:
        print('built', LIB)
    return LIB


if __name__ == '__main__':
    build(verbose=True)
