"""The C++ host interface (include/exocubed.hpp) through its test program
tests/cxx/test_host.cpp, built by __graft_entry__.build(): the host-only checks
on the CPU, all checks on the GPU."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
PROG = os.path.join(HERE, "cxx", "test_host")


def _run(mode):
    if not os.path.exists(PROG):
        pytest.fail("tests/cxx/test_host is not built: run __graft_entry__.build()")
    r = subprocess.run([PROG, mode], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def test_host_checks():
    out = _run("host")
    assert "FAILED" not in out and "3 tests, 0 failed checks" in out


@pytest.mark.gpu
def test_all_checks_on_device():
    out = _run("all")
    assert "FAILED" not in out and "8 tests, 0 failed checks" in out
