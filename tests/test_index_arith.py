"""Index arithmetic of the staged (de)contraction kernels
(cp2k_amd/csrc/grid_kernels.hip, fdiv): e / n for 0 <= e < 4096 and
1 <= n <= 256 as (int)((e + 0.5) * rcp(n)) in single precision, where the
hardware reciprocal is within one ulp of 1/n.  Checked for every e and n of
that range and for the reciprocal rounded to nearest and one ulp either side
(the host routes larger matrices to the other kernels)."""
import numpy as np


def test_fdiv_exact_in_range():
    e = np.arange(4096, dtype=np.int64)
    for n in range(1, 257):
        r = np.float32(1.0) / np.float32(n)
        for rr in (r, np.nextafter(r, np.float32(0)), np.nextafter(r, np.float32(1))):
            q = ((e.astype(np.float32) + np.float32(0.5)) * rr).astype(np.int64)
            np.testing.assert_array_equal(q, e // n, err_msg=f"n={n}")
