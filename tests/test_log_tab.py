"""The table-driven logarithm of the near-field, P2M and L2P kernels, checked on the CPU.

tests/log_tab_check.cpp builds pykifmm2d_amd/csrc/kf_logt.h with g++ (the same FMA-based
arithmetic the device executes) and measures it against the 80-bit logl: log-uniform samples
over every positive normal double and over the squared distances of the near field, the
neighbourhood of 1, and the end points of every table interval.  The table is checked
against -log(c_j) and the interval holding 1 against c = 1, L = 0; zero, subnormal, inf, nan
and negative inputs take the library log.  A dropped polynomial term, a wrong table entry or
a wrong exponent offset shows up as errors of many ulps.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def result(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("logt") / "log_tab_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "log_tab_check.cpp"),
                    "-o", exe], check=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split()
    return {out[i]: float(out[i + 1]) for i in range(0, len(out), 2)}


def test_log_table_accuracy(result):
    print(result)
    assert result["max_ulp"] < 2.0        # measured 1.81
    assert result["near1_ulp"] <= 1.0     # relative accuracy kept around x = 1
    assert result["bound_ulp"] < 2.0


def test_log_table_entries_and_specials(result):
    assert result["table_ok"] == 1
    assert result["special_ok"] == 1
