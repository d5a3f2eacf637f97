"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer.

`make -C oracle san` builds oracle/_san/liboracle_san.so from the oracle's sources with
-fsanitize=address,undefined and no recovery; a child pytest process loads it in place of
liboracle.so (SWR_ORACLE_LIB) with the ASan runtime preloaded and runs the oracle's own tests
(golden vectors, the reference comparisons, texture sampling, visibility, clipping, the FP64
differentiation reference).  Any sanitizer report aborts the child, which fails this test.
The slowest cases (the 1M-triangle frame, the scene digests, the Benchmark scene and the
scissor sweep) stay with the uninstrumented suite."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(ROOT, "oracle")
SAN_SO = os.path.join(ORACLE, "_san", "liboracle_san.so")
TESTS = ["tests/test_golden.py", "tests/test_oracle_texture.py", "tests/test_oracle_visibility.py",
         "tests/test_clip_early_out.py", "tests/test_diff_reference.py",
         "tests/test_oracle_vs_ref.py"]
SLOW = "not frame1m and not scene_digests and not benchmark_scene and not scissor"


def _runtime(name):
    try:
        p = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True,
                           check=True).stdout.strip()
    except (OSError, subprocess.CalledProcessError):
        return None
    return p if os.path.isabs(p) and os.path.exists(p) else None


def test_oracle_under_asan_ubsan():
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc's ASan runtime is not installed")
    b = subprocess.run(["make", "-C", ORACLE, "san"], capture_output=True, text=True)
    assert b.returncode == 0, b.stdout + b.stderr
    env = dict(os.environ, SWR_ORACLE_LIB=SAN_SO, LD_PRELOAD=asan,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        "-m", "not gpu", "-k", SLOW, *TESTS],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    tail = (r.stdout + r.stderr)[-4000:]
    assert r.returncode == 0, tail
    assert " passed" in r.stdout, tail
