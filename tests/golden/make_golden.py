"""Writes the golden vectors of tests/golden/*.npz from the FP64 oracle.

The reference ships no golden vectors for this path, so these fixtures are
the oracle's outputs on small seeded workloads (inputs are regenerated from
the recorded builders and seeds); they pin the oracle against regressions
and give the GPU tests fixed expected values.  Run from the repository root:

    python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from muspinsim_amd import workloads as W  # noqa: E402
from oracle import pipeline  # noqa: E402

HERE = os.path.dirname(os.path.abspath(__file__))

CASES = {
    # name: (builder, kwargs)
    "zero_field_n16": (W.c1_zero_field_powder, dict(seed=1, nori=12, ntimes=200)),
    "alc_n192": (W.c2_alc_sweep, dict(seed=2, nfields=4, nori=3)),
    "tf_n64": (W.c3_tf_stress, dict(seed=3, nori=4, ntimes=300, nnuc=5)),
    "fit_n64": (W.c5_fit, dict(seed=5, nori=4, ntimes=250)),
    "dissipative_n8": (W.c4_dissipative, dict(seed=44, ninst=6, ntimes=150)),
}


def main():
    for name, (builder, kw) in CASES.items():
        systems, batch = builder(**kw)
        res = pipeline.run(systems, batch)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez(path, results=res, builder=builder.__name__,
                 kwargs=np.array(repr(sorted(kw.items()))))
        print(f"{path}: results {res.shape}")


if __name__ == "__main__":
    main()
