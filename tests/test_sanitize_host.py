"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only, no GPU).

``tests/sanitize/Makefile`` builds libabf with sanitized host code (its gfx950 code is built as
usual), the oracle's C restatement with the same sanitizers, and ``abi_host_check``, a C program
that drives the ABI's host paths: workspace sizes, launch plans, batch layouts and the argument
checks of every entry point.  The host tests of this directory then run in a child process
against those builds, with the sanitizer runtime preloaded.  Any sanitizer report fails the
test (UndefinedBehaviorSanitizer is built without recovery, AddressSanitizer aborts on its first
report).
"""

import glob
import os
import shutil
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, 'tests', 'sanitize')
BUILD = os.path.join(SAN, '_build')
HOST_TESTS = ['tests/test_host.py', 'tests/test_deconv_host.py', 'tests/test_autotune.py', 'tests/test_serving.py',
              'tests/test_oracle.py::test_c_restatement_matches_numpy_oracle', 'tests/test_serving_host.py']
REPORTS = ('ERROR: AddressSanitizer', 'runtime error:', 'ERROR: LeakSanitizer')


def _asan_runtime():
    hits = sorted(glob.glob('/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so'))
    return hits[-1] if hits else None


@pytest.fixture(scope='module')
def sanitized_build():
    if not os.path.exists('/opt/rocm/bin/hipcc') or shutil.which('make') is None or _asan_runtime() is None:
        pytest.skip('hipcc, make or the clang sanitizer runtime is missing')
    jobs = str(min(8, os.cpu_count() or 1))
    r = subprocess.run(['make', '-C', SAN, '-j', jobs], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    return BUILD


def _clean(out: str) -> None:
    for pat in REPORTS:
        assert pat not in out, out[-6000:]


def test_abi_host_paths_under_sanitizers(sanitized_build):
    env = dict(os.environ, ASAN_OPTIONS='detect_leaks=0:abort_on_error=1', UBSAN_OPTIONS='print_stacktrace=1')
    r = subprocess.run([os.path.join(sanitized_build, 'abi_host_check')], capture_output=True, text=True, env=env,
                       timeout=300)
    out = r.stdout + r.stderr
    _clean(out)
    assert r.returncode == 0, out[-4000:]
    assert 'abi host checks passed' in out


def test_host_tests_under_sanitizers(sanitized_build):
    """The host tests with libabf and the C oracle replaced by their sanitizer builds."""
    env = dict(os.environ,
               LD_PRELOAD=_asan_runtime(),
               ASAN_OPTIONS='detect_leaks=0:verify_asan_link_order=0',
               UBSAN_OPTIONS='print_stacktrace=1',
               ACOULAR_AMD_LIB=os.path.join(sanitized_build, 'libabf_asan.so'),
               ORACLE_CPU_REF_LIB=os.path.join(sanitized_build, 'libcpu_ref_asan.so'))
    probe = ("import sys; sys.path.insert(0, '.'); from acoular_amd import _lib; from oracle import cref; "
             "_lib.lib(); cref.lib(); maps = open('/proc/self/maps').read(); "
             "print('libabf_asan.so' in maps, 'libcpu_ref_asan.so' in maps, 'asan' in maps)")
    r = subprocess.run([sys.executable, '-c', probe], capture_output=True, text=True, env=env, cwd=ROOT, timeout=300)
    assert r.stdout.split() == ['True', 'True', 'True'], r.stdout + r.stderr[-3000:]
    tests = [t for t in HOST_TESTS if os.path.exists(os.path.join(ROOT, t.split('::')[0]))]
    cmd = [sys.executable, '-m', 'pytest', '-x', '-q', '-m', 'not gpu', '-p', 'no:cacheprovider', *tests]
    r = subprocess.run(cmd, capture_output=True, text=True, env=env, cwd=ROOT, timeout=900)
    out = r.stdout + r.stderr
    _clean(out)
    assert r.returncode == 0, out[-6000:]
    # the child really loaded the sanitizer builds
    assert 'passed' in out
