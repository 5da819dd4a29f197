"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only): every function of the
FP64 oracle (oracle/san_check.c, built by `make -C oracle san`), and the kept host headers' helpers
(the npy reader and writer, the size policy, shape_tr, init_pot_line: tests/cpp/host_helpers.cpp).
GPU code is not instrumented: the sanitizers apply to the host side of each build only."""
import os
import subprocess

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


def _clean(r):
    text = r.stdout + r.stderr
    assert r.returncode == 0, text[-2000:]
    assert "AddressSanitizer" not in text and "runtime error" not in text and "LeakSanitizer" not in text, text[-2000:]


def test_oracle_under_asan_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "san"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", OMP_NUM_THREADS="2")
    r = subprocess.run([os.path.join(REPO, "oracle", "_san", "oracle_san")], capture_output=True, text=True, env=env,
                       timeout=600)
    _clean(r)
    assert "oracle sanitizer check done" in r.stdout


def test_host_helpers_under_asan_ubsan(tmp_path):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not available")
    exe = str(tmp_path / "host_helpers_san")
    # host-only translation unit; each sanitizer flag is scoped to the host compilation
    subprocess.run([HIPCC, "-O1", "-g", "-std=c++17", "-fno-omit-frame-pointer",
                    "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
                    "-I", os.path.join(REPO, "include"), os.path.join(REPO, "tests", "cpp", "host_helpers.cpp"),
                    "-o", exe], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1")
    r = subprocess.run([exe, str(tmp_path / "a.npy")], capture_output=True, text=True, env=env, timeout=300)
    _clean(r)
    assert "npy 24 2 3 4 -3 8.5" in r.stdout and "npy dtype mismatch rejected" in r.stdout
