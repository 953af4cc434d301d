"""The oracle's sources under AddressSanitizer and UndefinedBehaviorSanitizer
(tests/cpp/oracle_asan.cpp): t_eval rows, event points, point outputs with
short slots, the ragged spring-chain layouts, SDE paths and CSV documents,
each written into buffers of exactly the documented sizes, so any access past
a layout's end (or undefined arithmetic) aborts the driver.  Host code only.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "_build", "oracle_asan")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_layouts_are_clean_under_the_sanitizers():
    subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp"), "_build/oracle_asan"])
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=600, env=env)
    assert out.returncode == 0, out.stderr[-4000:]
    assert out.stdout.strip().endswith("ok")
