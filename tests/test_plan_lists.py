"""CPU replay of the planned evaluation's indexing (no GPU).

tests/plan_index_check.cpp is compiled with g++ against the library's host sources (the box
tree and the work lists) and walks every pass of one evaluation the way its kernel indexes
memory: every slab row in bounds of what the library allocates for the layout, every row
read written earlier in the evaluation, and every result the caller reads complete.  Trees
of points on curves and in annuli, where most boxes hold no sources, are the cases that
matter: rows of boxes without sources exist only for the downward pass."""
import os
import struct
import subprocess

import numpy as np
import pytest

from pykifmm2d_amd import workloads
from tests import treekeys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "pykifmm2d_amd", "csrc")


@pytest.fixture(scope="module")
def checker(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("plancheck") / "plan_index_check")
    cmd = ["g++", "-O2", "-std=c++17", "-I" + CSRC, "-I" + os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "plan_index_check.cpp"), os.path.join(CSRC, "tree_host.cpp"),
           os.path.join(CSRC, "plan_lists.cpp"), "-o", out]
    subprocess.run(cmd, check=True)
    return out


def _keys_file(path, x, y, cutoff):
    keys, _, mmin, W = treekeys.sorted_keys(x, y)
    with open(path, "wb") as f:
        f.write(struct.pack("<q", x.size) + struct.pack("<i", cutoff) + struct.pack("<d", mmin)
                + struct.pack("<d", W) + struct.pack("<i", 30))
        f.write(keys.astype("<u8").tobytes())


def _grid():
    g = np.arange(0, 33) / 32.0
    return [a.ravel() for a in np.meshgrid(g, g)]


CASES = {
    "uniform": (lambda: workloads.uniform(10000)[:2], 50),
    "clustered": (lambda: workloads.clustered(20000)[:2], 64),
    "annulus": (lambda: workloads.annulus(12000)[:2], 64),
    "star": (lambda: workloads.star_curve(42432, 1.0), 50),
    "grid": (_grid, 4),
}


@pytest.mark.parametrize("case", sorted(CASES))
def test_every_pass_stays_in_bounds_and_reads_written_rows(checker, tmp_path, case):
    gen, cutoff = CASES[case]
    x, y = gen()
    f = str(tmp_path / "keys.bin")
    _keys_file(f, np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), cutoff)
    for lg in (0, 3, 6):
        for mode in (0, 1):  # source targets (kf_plan_eval), all targets (kf_plan_expand)
            r = subprocess.run([checker, str(lg), str(mode), f], capture_output=True, text=True)
            assert r.returncode == 0, "%s R=%d mode %d:\n%s" % (case, 1 << lg, mode, r.stdout)
            assert r.stdout.startswith("OK")


def test_forest_passes_stay_in_bounds(checker, tmp_path):
    """A forest of uneven trees with three root-width classes ([0,1], [0,2] and the [-2,2] root
    of star curves) and trees of different depth sharing every launch."""
    rng = np.random.default_rng(4)
    files = []
    for i in range(12):
        n = int(round(np.exp(rng.uniform(np.log(1e3), np.log(2e4)))))
        if i % 3 == 0:
            x, y = workloads.star_curve(n, rng.uniform(0, 2 * np.pi))
        else:
            x, y = rng.random(n) * (1 + i % 3 - 1), rng.random(n) * (1 + i % 3 - 1)
            if i % 3 == 2:
                x, y = 2 * rng.random(n), 2 * rng.random(n)
        f = str(tmp_path / ("t%d.bin" % i))
        _keys_file(f, np.asarray(x, dtype=np.float64), np.asarray(y, dtype=np.float64), 50)
        files.append(f)
    for lg in (0, 3, 6):
        for mode in (0, 1):
            r = subprocess.run([checker, str(lg), str(mode)] + files, capture_output=True, text=True)
            assert r.returncode == 0, "forest R=%d mode %d:\n%s" % (1 << lg, mode, r.stdout)
            assert r.stdout.startswith("OK trees 12 classes 3")


def test_xcd_workgroup_map_is_a_contiguous_bijection(checker):
    """The workgroup -> item map of the XCD-aware launches covers every item exactly once and
    gives each of the 8 XCDs one contiguous range."""
    r = subprocess.run([checker, "xcd"], capture_output=True, text=True)
    assert r.returncode == 0 and r.stdout.startswith("OK xcd"), r.stdout
