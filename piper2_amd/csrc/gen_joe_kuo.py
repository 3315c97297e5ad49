"""Write joe_kuo_1024.inc: the Joe-Kuo "new-joe-kuo-6.21201" primitive polynomials and initial
direction numbers m_1..m_s for the first 1024 Sobol dimensions.

The reference sampler names this direction-number set (ScrambledSobolSampler.cpp:26-27) and uses
1024 dimensions of it (Shared.hpp:22). Its generated matrix file is absent from the reference tree,
so the table is rebuilt from the published Joe-Kuo data, read here from the copy SciPy ships with
its own Sobol generator (scipy/stats/_sobol_direction_numbers.npz, same data set).

Each line of the output is `{s, a, {m_1, ..., m_s}},` where s is the polynomial degree and `a`
holds the s-1 interior coefficient bits. Dimension 0 (van der Corput) is written as s=0.
Run: python piper2_amd/csrc/gen_joe_kuo.py  (rewrites the .inc beside this script).
"""
import os

import numpy as np
import scipy

NDIM = 1024


def main() -> None:
    path = os.path.join(os.path.dirname(scipy.__file__), "stats", "_sobol_direction_numbers.npz")
    data = np.load(path)
    poly, vinit = data["poly"], data["vinit"]
    lines = ["// Generated by gen_joe_kuo.py from the Joe-Kuo new-joe-kuo-6.21201 direction numbers.",
             "// {degree s, interior coefficient bits a, initial direction numbers m_1..m_s}"]
    lines.append("{0, 0, {0}},")
    for d in range(1, NDIM):
        p = int(poly[d])
        s = p.bit_length() - 1
        a = (p >> 1) & ((1 << (s - 1)) - 1) if s > 1 else 0
        m = [int(v) for v in vinit[d, :s]]
        assert all(v & 1 and v < (1 << (i + 1)) for i, v in enumerate(m)), (d, m)
        lines.append("{%d, %d, {%s}}," % (s, a, ", ".join(str(v) for v in m)))
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "joe_kuo_1024.inc")
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


if __name__ == "__main__":
    main()
