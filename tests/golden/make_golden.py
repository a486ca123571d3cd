"""Write tests/golden/golden_{64,32}.npz: input/output vectors of the
reference minfft (compiled from its own sources by oracle/Makefile into
oracle/_ref) for every transform on small 1D/2D/3D shapes.

Run in the build container (where the reference sources exist):
    make -C oracle && python tests/golden/make_golden.py
Inputs come from numpy's PCG64 with fixed seeds; for the inverse real DFT the
input is the reference's own real DFT of a random real array (a consistent
half-spectrum).
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)
from oracle.oracle import KINDS, Reference, in_shape, is_complex_in  # noqa: E402

SHAPES = [(1,), (2,), (4,), (8,), (16,), (32,), (64,), (128,), (1024,),
          (2, 2), (4, 8), (8, 4), (16, 2), (1, 8), (8, 1), (2, 4, 8), (4, 1, 2), (8, 8, 8)]


def shape_key(s):
    return "x".join(str(n) for n in s)


def main():
    for prec in (64, 32):
        ref = Reference(prec)
        rdt = np.float64 if prec == 64 else np.float32
        out = {}
        for si, shape in enumerate(SHAPES):
            for ki, kind in enumerate(KINDS):
                rng = np.random.default_rng(1000 * si + ki + (7 if prec == 32 else 0))
                if kind == "invrealdft":
                    x = ref.run("realdft", rng.uniform(-1, 1, shape).astype(rdt), shape)
                elif is_complex_in(kind):
                    ish = in_shape(kind, shape)
                    x = (rng.uniform(-1, 1, ish) + 1j * rng.uniform(-1, 1, ish)).astype(ref.cdt)
                else:
                    x = rng.uniform(-1, 1, shape).astype(rdt)
                y = ref.run(kind, x, shape)
                key = f"{kind}__{shape_key(shape)}"
                out[key + "__in"] = x
                out[key + "__out"] = y
        path = os.path.join(os.path.dirname(os.path.abspath(__file__)), f"golden_{prec}.npz")
        np.savez_compressed(path, **out)
        print(path, len(out) // 2, "vectors")


if __name__ == "__main__":
    main()
