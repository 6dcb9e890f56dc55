"""oracle -- TEST INFRASTRUCTURE ONLY.

CPU restatement of the reference 2-D KIFMM (pykifmm2d: tree.py, class_fmm.py, fmm.py)
used as the parity checker for the HIP product path, and timed as the CPU baseline.
Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py``
may import it.  The product package (``pykifmm2d_amd``) never does.

Pinning (see tests/test_oracle.py): FP64 direct sums with compensated summation, the
Laplace closed form, scipy's K0, the closed-form eigenvalues of the circulant C2E matrix,
tree invariants and brute force on tiny inputs, hand-built trees.  Functions without such
a check are marked "parity unpinned" in their docstring.
"""
