"""The fused powf of the colourspace chain kernels (zmath.h powf_fused: fused
multiply-adds in the FP64 log2/exp2 polynomials, plus a table of the few
inputs where the fused form would round differently) equals the reference's
unfused powf for every float input: tests/powf_fma_check.cpp compares the
two over all 2^32 bit patterns of x for each exponent powf_fused accepts
(about ten seconds per exponent on 8 cores). The check plants nothing; a
wrong table entry, a missing entry or an exponent outside the verified list
(a compile error) fails it."""
import os
import re
import shutil
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _fma_flag():
    try:
        flags = open("/proc/cpuinfo").read()
    except OSError:
        return []
    return ["-mfma"] if re.search(r"\bfma\b", flags) else []


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_fused_powf_equals_reference_for_every_input():
    with tempfile.TemporaryDirectory() as d:
        exe = os.path.join(d, "check")
        subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", *_fma_flag(), "-I", os.path.join(ROOT, "zimg_amd", "csrc"),
                        os.path.join(ROOT, "tests", "powf_fma_check.cpp"), "-o", exe, "-lpthread"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=1500).stdout
    lines = [l for l in out.splitlines() if l.startswith("y=")]
    assert len(lines) == 10, out
    for l in lines:
        assert l.endswith("mismatches 0"), out
