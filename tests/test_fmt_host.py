"""The device CSV formatter (csrc/de_fmt.h: Ryu shortest digits in Rust's
`{:?}` layout) built for the host with AddressSanitizer and
UndefinedBehaviorSanitizer, on the inputs of the GPU CSV tests (40 000 x 5
random bit patterns from seed 101 and the special values) and more random
doubles: every text equals the oracle's (whose search-based formatter is
pinned against Python's repr), no loop cap is reached, no sanitizer report.
The host build runs the device's integer arithmetic (the 64 x 64-bit high
product is formed from 32-bit products on both)."""
import os
import subprocess

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(HERE, "cpp", "_build", "fmt_check")


@pytest.fixture(scope="module")
def fmt_check():
    if not os.path.exists(EXE):
        subprocess.check_call(["make", "-s", "-C", os.path.join(HERE, "cpp"), "_build/fmt_check"])
    def run(vals, count_only=False):
        out = subprocess.run([EXE] + (["-c"] if count_only else []),
                             input=np.ascontiguousarray(vals, dtype=np.float64).tobytes(),
                             capture_output=True, timeout=600)
        assert out.returncode == 0, out.stderr.decode()[-2000:]
        assert b"runtime error" not in out.stderr and b"Sanitizer" not in out.stderr
        return out.stdout.decode()
    return run


def gpu_csv_test_inputs():
    n = 40000
    rng = np.random.default_rng(101)
    bits = rng.integers(0, 2 ** 64, 4 * n, dtype=np.uint64)
    vals = bits.view(np.float64).copy()
    special = np.array([0.0, -0.0, np.nan, -np.nan, np.inf, -np.inf, 5e-324, 1e-4, 1e16,
                        9999999999999998.0, 1.7976931348623157e308, 2.2250738585072014e-308,
                        0.1, 1.0, 123456789.0, 1e-5])
    vals[:len(special)] = special
    vals[n:n + len(special)] = special[::-1]
    return vals


def test_formatter_equals_oracle_on_the_gpu_test_inputs(orc, fmt_check):
    vals = gpu_csv_test_inputs()
    got = fmt_check(vals).split("\n")[:-1]
    assert len(got) == len(vals)
    for x, g in zip(vals, got):
        assert g == orc.fmt_debug(x), (x, g)


def test_formatter_reaches_no_cap_on_random_and_edge_doubles(fmt_check):
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2 ** 64, 400000, dtype=np.uint64)
    p2 = np.ldexp(1.0, np.arange(-1074, 1024))
    p10 = 10.0 ** np.arange(-323, 309)
    edges = np.concatenate([p2, np.nextafter(p2, 0), np.nextafter(p2, np.inf), p10,
                            np.nextafter(p10, 0), np.nextafter(p10, np.inf)])
    vals = np.concatenate([bits.view(np.float64), edges, -edges,
                           np.arange(1, 1 << 16, dtype=np.float64)])
    assert fmt_check(vals, count_only=True).strip() == "0"
