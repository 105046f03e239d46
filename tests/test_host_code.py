"""CPU tests of host-side library code built on its own: the serving result
pool (first fit, merging of released blocks) against a byte-map model, and the
oracle and host C++ sources under AddressSanitizer and UndefinedBehaviorSanitizer
(host code only; no GPU here)."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


@pytest.mark.parametrize("san", [[], ["-fsanitize=address,undefined", "-fno-sanitize-recover=all"]])
def test_result_pool_model(tmp_path, san):
    exe = tmp_path / "pool_test"
    _run(["g++", "-std=c++17", "-O1", "-g", *san, os.path.join(ROOT, "tests", "cpu", "pool_test.cpp"),
          "-o", str(exe)])
    assert "pool ok" in _run([str(exe)])


def test_oracle_under_sanitizers(tmp_path):
    exe = tmp_path / "oracle_san"
    _run(["gcc", "-std=c11", "-O1", "-g", "-ffp-contract=off", "-fsanitize=address,undefined",
          "-fno-sanitize-recover=all", "-fno-sanitize=float-divide-by-zero",
          os.path.join(ROOT, "tests", "cpu", "oracle_san.c"), "-o", str(exe), "-lm"])
    assert "oracle ok" in _run([str(exe)])


def test_library_host_code_under_sanitizers(tmp_path):
    # built by build() (tests/cpu/Makefile): the library's sources with
    # ASan/UBSan on the host side, driven through the paths that need no GPU
    exe = os.path.join(ROOT, "tests", "cpu", "host_san")
    assert os.path.exists(exe), "build first: make -C tests/cpu"
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0")
    assert "host ok" in _run([exe, str(tmp_path)], env=env, timeout=120)
