"""Pins the CPU oracle (oracle/minfft_oracle.c) before it is trusted:
against the reference's own outputs (golden fixtures written from the
compiled reference, and the compiled reference itself where it is built),
against independent FP64 libraries (numpy / scipy), closed forms and
long-double brute force.  CPU only."""
import math
import os

import numpy as np
import pytest
import scipy.fft as sfft

from oracle.oracle import KINDS, Oracle, Reference, ref_available, in_shape
from tests.helpers import make_input, rel_l2

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _golden(prec):
    return np.load(os.path.join(GOLDEN, f"golden_{prec}.npz"))


@pytest.mark.parametrize("prec", [64, 32])
def test_oracle_matches_reference_golden_vectors(prec):
    o = Oracle(prec)
    g = _golden(prec)
    keys = sorted({k.rsplit("__", 1)[0] for k in g.files})
    assert len(keys) == 180
    tol = 1e-15 if prec == 64 else 1e-7
    worst = 0.0
    for key in keys:
        kind, sk = key.split("__")
        shape = tuple(int(s) for s in sk.split("x"))
        y = o.run(kind, g[key + "__in"], shape)
        e = rel_l2(y, g[key + "__out"])
        worst = max(worst, e)
        assert e <= tol, (key, e)
    # the restatement keeps the reference's operation order: on this image's
    # libm the agreement is (close to) exact
    assert worst <= tol


@pytest.mark.skipif(not ref_available(64), reason="oracle/_ref not built")
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", [(2048,), (64, 32), (4, 8, 16), (2, 4, 8, 2), (2, 1, 4, 2, 8)])
def test_oracle_matches_compiled_reference(prec, shape):
    o, r = Oracle(prec), Reference(prec)
    for i, kind in enumerate(KINDS):
        x = make_input(kind, shape, prec, 17 + i, o)
        assert rel_l2(o.run(kind, x, shape), r.run(kind, x, shape)) <= (1e-15 if prec == 64 else 1e-7)


SHAPES = [(1,), (2,), (8,), (16,), (512,), (4096,), (8, 16), (32, 1), (4, 2, 8), (2, 4, 8, 2), (2, 2, 1, 4, 4),
          (1, 2, 1, 4, 2, 2)]


@pytest.mark.parametrize("shape", SHAPES)
def test_oracle_vs_numpy_scipy(oracle64, shape):
    o = oracle64
    n = int(np.prod(shape))
    tol = 1e-14 * max(1, math.log2(max(n, 2)))
    x = make_input("dft", shape, 64, 1)
    assert rel_l2(o.run("dft", x, shape), np.fft.fftn(x)) <= tol
    assert rel_l2(o.run("invdft", x, shape), np.fft.ifftn(x) * n) <= tol
    xr = make_input("dct2", shape, 64, 2)
    assert rel_l2(o.run("realdft", xr, shape), np.fft.rfftn(xr)) <= tol
    z = o.run("realdft", xr, shape)
    assert rel_l2(o.run("invrealdft", z, shape), np.fft.irfftn(z, s=shape, axes=list(range(len(shape)))) * n) <= tol
    axes = list(range(len(shape)))
    for t in (2, 3, 4):
        assert rel_l2(o.run(f"dct{t}", xr, shape), sfft.dctn(xr, type=t, axes=axes)) <= tol
        assert rel_l2(o.run(f"dst{t}", xr, shape), sfft.dstn(xr, type=t, axes=axes)) <= tol


@pytest.mark.parametrize("N", [2, 4, 8, 16, 32, 64])
def test_oracle_vs_long_double_brute_force(oracle64, N):
    x = make_input("dft", (N,), 64, N)
    for inv, kind in ((False, "dft"), (True, "invdft")):
        assert rel_l2(oracle64.run(kind, x, (N,)), oracle64.brute_dft(x, inverse=inv)) <= 1e-15


@pytest.mark.parametrize("N", [1, 2, 4, 16, 256, 8192])
def test_oracle_closed_forms(oracle64, N):
    o = oracle64
    d = np.zeros(N, complex)
    d[0] = 1
    assert np.allclose(o.run("dft", d, (N,)), np.ones(N), atol=1e-15, rtol=0)
    c = np.ones(N, complex)
    y = o.run("dft", c, (N,))
    e = np.zeros(N)
    e[0] = N
    assert np.allclose(y, e, atol=1e-12 * N, rtol=0)
    for k0 in sorted({0, min(1, N - 1), N // 3, N - 1}):
        tone = np.exp(2j * np.pi * k0 * np.arange(N) / N)
        t = np.zeros(N, complex)
        t[k0] = N
        assert np.allclose(o.run("dft", tone, (N,)), t, atol=1e-11 * N, rtol=0)


def test_oracle_length_one_specials(oracle64):
    o = oracle64
    x = np.array([0.75])
    assert o.run("dct2", x, (1,))[0] == 1.5 and o.run("dst2", x, (1,))[0] == 1.5
    assert o.run("dct3", x, (1,))[0] == 0.75 and o.run("dst3", x, (1,))[0] == 0.75
    assert o.run("dct4", x, (1,))[0] == pytest.approx(0.75 * math.sqrt(2), abs=0)
    z = o.run("realdft", x, (1,))
    assert z[0] == 0.75 + 0j


@pytest.mark.parametrize("shape", [(16,), (8, 4), (2, 4, 8)])
def test_oracle_inverse_identities(oracle64, shape):
    o = oracle64
    n = int(np.prod(shape))
    d = len(shape)
    x = make_input("dft", shape, 64, 3)
    assert rel_l2(o.run("invdft", o.run("dft", x, shape), shape), n * x) <= 1e-14
    xr = make_input("dct2", shape, 64, 4)
    assert rel_l2(o.run("invrealdft", o.run("realdft", xr, shape), shape), n * xr) <= 1e-14
    assert rel_l2(o.run("dct3", o.run("dct2", xr, shape), shape), (2 ** d) * n * xr) <= 1e-14
    assert rel_l2(o.run("dst3", o.run("dst2", xr, shape), shape), (2 ** d) * n * xr) <= 1e-14
    assert rel_l2(o.run("dct4", o.run("dct4", xr, shape), shape), (2 ** d) * n * xr) <= 1e-14
    assert rel_l2(o.run("dst4", o.run("dst4", xr, shape), shape), (2 ** d) * n * xr) <= 1e-14


def test_checks_discriminate_planted_mistakes(oracle64):
    """The checks above fail on plausible mistakes: wrong sign convention,
    transposed axes, DCT-2 taken for DCT-3, dropped factor 2."""
    o = oracle64
    x = make_input("dft", (64,), 64, 5)
    assert rel_l2(o.run("invdft", x, (64,)), np.fft.fftn(x)) > 0.5  # sign
    x2 = make_input("dft", (8, 16), 64, 6)
    # dims passed in the wrong order (Fortran order instead of C order)
    assert rel_l2(o.run("dft", x2, (16, 8)).reshape(8, 16), np.fft.fftn(x2)) > 0.5
    xr = make_input("dct2", (32,), 64, 7)
    assert rel_l2(o.run("dct2", xr, (32,)), sfft.dct(xr, type=3)) > 0.1
    assert rel_l2(o.run("dct4", xr, (32,)) / 2, sfft.dct(xr, type=4)) > 0.4


@pytest.mark.parametrize("prec", [64, 32])
def test_oracle_rejects_invalid_lengths(prec):
    o = Oracle(prec)
    with pytest.raises(ValueError):
        o.run("dft", np.zeros(3, complex), (3,))
    with pytest.raises(ValueError):
        o.run("dct2", np.zeros(6), (6,))
