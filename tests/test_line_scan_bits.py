"""Host unit test of the kernels' word-level line scan.

accrqa_amd/csrc/rqa_bits.hpp holds the integer logic every line kernel runs
per 32-step word: packed threshold ranks -> bit planes -> per-threshold
"not recurrent" words through a comparator tree, and run extraction with the
segment prefix/suffix hand-off that k_stitch joins.  The same header is
compiled here for the host and checked against a direct scan of random and
structured 0/1 sequences (line_scan_bits_test.cpp), so the GPU parity tests
are not the first place a bit-trick mistake can show up.
"""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(os.path.dirname(HERE), "accrqa_amd", "csrc")


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_line_scan_bits(tmp_path):
    exe = tmp_path / "line_scan_bits_test"
    subprocess.run(["g++", "-O2", "-std=c++17", "-w", "-I", CSRC, os.path.join(HERE, "line_scan_bits_test.cpp"),
                    "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.strip().endswith("ok")
