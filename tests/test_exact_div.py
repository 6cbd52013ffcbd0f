"""The kernels' shared-divisor quotient (cgli::div_r, cgli_internal.hpp):
x * (1/d) with one fused multiply-add correction must give the IEEE
division's bits, in binary64 and binary32.  A host build of the helper
against x / d on 4.2e7 operand pairs (binary64: random exponents over +-250,
one binade, divisors a few ulps from powers of two, exact quotients;
binary32: random exponents over +-30, one binade); the uncorrected product
fails the same check, so the check can fail."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HIPCC = "/opt/rocm/bin/hipcc"


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    if not os.path.exists(HIPCC):
        pytest.skip("hipcc not installed")
    exe = str(tmp_path_factory.mktemp("exact_div") / "check")
    subprocess.run([HIPCC, "--offload-arch=gfx950", "-O2", "-ffp-contract=off", "-I", os.path.join(ROOT, "include"),
                    "-I", os.path.join(ROOT, "cgli_fu_flow_amd", "csrc"), "-o", exe,
                    os.path.join(ROOT, "tests", "exact_div_check.hip")], check=True, capture_output=True)
    return exe


def run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True)
    last = r.stdout.strip().splitlines()[-1].split()
    return r.returncode, int(last[1]), int(last[3])


def test_corrected_quotient_is_the_division(checker):
    rc, total, bad = run(checker, "10000000")
    assert total > 4e7 and bad == 0 and rc == 0


def test_uncorrected_product_is_not(checker):
    rc, total, bad = run(checker, "100000", "plain")
    assert bad > 1000 and rc != 0
