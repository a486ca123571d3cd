"""Shared helpers of the test-suite: seeded inputs, error measures and the
tolerances the transforms are held to."""
import math

import numpy as np

from oracle.oracle import in_shape, is_complex_in


def rel_l2(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    dt = np.complex128 if (np.iscomplexobj(a) or np.iscomplexobj(b)) else np.float64
    a = a.astype(dt)
    b = b.astype(dt)
    nb = np.linalg.norm(b.ravel())
    if nb == 0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / nb)


def tol64(shape):
    """FP64 plain path: rel-L2 <= 1e-14 * log2 N (N = total points)."""
    n = int(np.prod(shape))
    return 1e-14 * max(1.0, math.log2(max(n, 2)))


TOL32 = 1e-6  # FP32 forms vs the FP64 oracle


def make_input(kind, shape, prec, seed, oracle=None):
    """Seeded input for `kind` on logical shape `shape` (numpy, host)."""
    rng = np.random.default_rng(seed)
    rdt = np.float64 if prec == 64 else np.float32
    cdt = np.complex128 if prec == 64 else np.complex64
    if kind == "invrealdft":
        assert oracle is not None
        x = rng.uniform(-1, 1, shape).astype(rdt)
        z = oracle.run("realdft", x, shape)
        return np.ascontiguousarray(z.astype(cdt))
    ish = in_shape(kind, shape)
    if is_complex_in(kind):
        return (rng.uniform(-1, 1, ish) + 1j * rng.uniform(-1, 1, ish)).astype(cdt)
    return rng.uniform(-1, 1, ish).astype(rdt)
