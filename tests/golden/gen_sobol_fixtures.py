"""Write sobol_columns.npz: golden Sobol generator columns and sample points from SciPy.

SciPy's `scipy.stats.qmc.Sobol` is built on the same Joe-Kuo new-joe-kuo-6.21201 direction numbers
that the reference sampler names (Sampler/ScrambledSobolSampler/ScrambledSobolSampler.cpp:26-27),
but through SciPy's own implementation of the direction-number recurrence.

- `points`: the first 64 points of all 1024 dimensions, drawn through SciPy's public API
  (unscrambled, 32-bit) and put in natural order. SciPy emits Gray-code order: its k-th point
  is x_n with n = k ^ (k >> 1). Before writing, all 2^13 drawn points are checked to equal the
  XOR of `columns` over the set bits of n, which pins columns 0..12 through the public API.
- `columns`: generator columns 0..51 (u32) of all 1024 dimensions taken from SciPy's 64-bit
  generator's direction numbers (`_sv`, high 32 bits). Columns >= 32 only matter for sample
  indices >= 2^32; a 32-bit column is taken as the top half of the 64-bit one.

Run from the repo root: python tests/golden/gen_sobol_fixtures.py  (takes a few seconds)
"""
import os

import numpy as np
import scipy.stats as st

NDIM = 1024
NCOL = 52
NPTS = 1 << 13


def main() -> None:
    x = st.qmc.Sobol(NDIM, scramble=False, bits=32).random(NPTS)
    gray = np.round(x * 2.0**32).astype(np.uint64).astype(np.uint32)
    k = np.arange(NPTS)
    points = np.empty_like(gray)
    points[k ^ (k >> 1)] = gray
    sv = st.qmc.Sobol(NDIM, scramble=False, bits=64)._sv
    cols = (sv[:, :NCOL] >> np.uint64(32)).astype(np.uint32)
    xor = np.zeros_like(points)
    for j in range(13):
        xor[(k >> j) & 1 == 1] ^= cols[:, j]
    assert np.array_equal(xor, points), "public points differ from the XOR of the columns"
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sobol_columns.npz")
    np.savez_compressed(out, columns=cols, points=points[:64])


if __name__ == "__main__":
    main()
