"""Point generators of the reference's misc/utils.py (misc/utils.py:3-6), which its examples use
(examples/laplace.py:24): uniform random coordinates in [rmin, rmax)."""
import numpy as np


def random(N, rmin=0, rmax=1):
    return np.random.rand(N) * (rmax - rmin) + rmin


def random2(N, rmin=0, rmax=1):
    return random(N, rmin, rmax), random(N, rmin, rmax)
