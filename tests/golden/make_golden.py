"""Write tests/golden/ginterp_vectors.npz from the compiled reference interpolator.

Runs the reference's own GridInterp.cpp (built into oracle/_ref by
oracle/Makefile from the reference sources) on seeded random inputs and saves
inputs and outputs.  The GPU box has no reference tree; these vectors carry the
reference's answers there.  Usage: python tests/golden/make_golden.py
"""

import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from oracle import native  # noqa: E402


def main():
    ref = native.reference_ginterp()
    if ref is None:
        raise SystemExit("oracle/_ref is not built (make -C oracle)")
    rng = np.random.default_rng(2024)
    v2 = rng.standard_normal((31, 17))
    p2 = rng.uniform(-1.5, 32.5, (3000, 2))
    p2[:8] = [[30, 0], [0, 16], [29.999999999, 15.5], [0, 0], [30.0, 16.0], [-1e-300, 3],
              [np.nan, 2], [5e15, 1]]
    o2 = np.empty(len(p2))
    ref.interp2d_vec_inplace(v2, p2, o2)
    v3 = rng.standard_normal((7, 9, 5))
    p3 = rng.uniform(-1.5, 9.5, (3000, 3))
    p3[:3] = [[6, 8, 4], [0, 0, 0], [5.5, 7.5, 3.999999]]
    o3 = np.empty(len(p3))
    ref.interp3d_vec_inplace(v3, p3, o3)
    out = os.path.join(os.path.dirname(os.path.abspath(__file__)), "ginterp_vectors.npz")
    np.savez_compressed(out, v2=v2, p2=p2, o2=o2, v3=v3, p3=p3, o3=o3)
    print("wrote", out)


if __name__ == "__main__":
    main()
