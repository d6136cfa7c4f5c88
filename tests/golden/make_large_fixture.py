#!/usr/bin/env python3
"""Write large_c350.npz: the oracle's single point of a C350H702 alkane
(2102 AOs, 1052 atoms; workloads.alkane(350) with seed 350).

The oracle needs about a minute for this system, too long for a GPU test to
recompute, so its result is committed as a fixture: inputs (Z, xyz) and
expected outputs (energy, charges, gradient, SCC iterations).  Rerun after any
change to the oracle's single point.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import oracle as O  # noqa: E402
from xtbloom_amd import workloads  # noqa: E402

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "large_c350.npz")


def main():
    s = workloads.alkane(350, np.random.default_rng(350))
    ref = O.singlepoint(s["Z"], s["xyz"])
    assert ref.status == 0, ref.status
    np.savez_compressed(OUT, Z=np.asarray(s["Z"], np.int32), xyz=np.asarray(s["xyz"], np.float64),
                        energy=np.float64(ref.energy), charges=np.asarray(ref.charges, np.float64),
                        gradient=np.asarray(ref.gradient, np.float64),
                        iterations=np.int32(ref.iterations))
    print(f"wrote {OUT} ({os.path.getsize(OUT)} bytes), E = {ref.energy:.12f}")


if __name__ == "__main__":
    main()
