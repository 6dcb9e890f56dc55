"""The plan kernels' K0 (kf_k0.h: k0_r2), checked on the CPU against scipy.special.k0.

k0_r2 evaluates K0(k sqrt(r2)) from the squared distance: on the power-series branch it runs
the series in x^2 / 4 = k^2 r2 / 4 and takes log(x / 2) = log(r2) / 2 + log(k / 2) with the
table-driven log; beyond it, the Chebyshev branch of the library form.  The same 2e-15 relative
bound as the device test of the library form (test_gpu_parity.py), over both branches, the
split point and three wavenumbers; K0(0) = inf.  A wrong log offset, a dropped series term or a
wrong branch bound shows up as errors far above the bound."""
import os
import subprocess

import numpy as np
import pytest
import scipy.special

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def k0_r2(tmp_path_factory):
    exe = str(tmp_path_factory.mktemp("k0") / "k0_check")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", os.path.join(ROOT, "tests", "k0_check.cpp"),
                    "-o", exe], check=True)

    def run(k, r2):
        inp = "".join("%.17g %.17g\n" % (a, b) for a, b in zip(k, r2))
        out = subprocess.run([exe], input=inp, capture_output=True, text=True, check=True).stdout.split()
        return np.array([float(v) for v in out])
    return run


@pytest.mark.parametrize("k", [0.3, 1.0, 10.0])
def test_k0_r2_matches_scipy(k0_r2, k):
    x = np.concatenate([np.logspace(-9, -0.5, 800), np.linspace(0.31, 60.0, 1500),
                        np.nextafter(1.0, [0.0, 2.0]), [1.0]])
    r2 = (x / k) ** 2
    got = k0_r2(np.full_like(r2, k), r2)
    ref = scipy.special.k0(k * np.sqrt(r2))
    err = np.abs(got - ref) / ref
    print(k, float(err.max()))
    assert float(err.max()) < 2e-15


def test_k0_r2_at_zero_is_inf(k0_r2):
    assert np.isinf(k0_r2(np.array([10.0]), np.array([0.0]))[0])
