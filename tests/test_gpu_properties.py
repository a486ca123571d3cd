"""GPU checks that do not go through the oracle: closed forms, algebraic
properties (hypothesis-generated inputs and shapes) and round trips at the
full benchmark sizes, where a CPU reference would take too long.

Conventions (docs/*.png of the reference): unnormalised transforms, forward
sign -1.  Identities used: invdft(dft(x)) = N x, invrealdft(realdft(x)) = N x,
dct3(dct2(x)) = dst3(dst2(x)) = 2N x and dct4(dct4(x)) = dst4(dst4(x)) = 2N x
per dimension of length N."""
import math

import numpy as np
import pytest
from hypothesis import HealthCheck, given, settings, strategies as st

import minfft_amd as mf
from oracle.oracle import out_shape
from tests.helpers import TOL32, rel_l2, tol64

pytestmark = pytest.mark.gpu

INVERSE = {"dft": "invdft", "realdft": "invrealdft", "dct2": "dct3", "dst2": "dst3",
           "dct4": "dct4", "dst4": "dst4"}


@pytest.fixture(scope="module")
def libs():
    return {64: mf.Minfft(64), 32: mf.Minfft(32)}


def run(lib, kind, x, shape):
    aux = mf.Aux(lib, mf.FAMILY[kind], shape)
    try:
        rdt = np.float64 if lib.prec == 64 else np.float32
        cdt = np.complex128 if lib.prec == 64 else np.complex64
        odt = cdt if kind in ("dft", "invdft", "realdft") else rdt
        y = np.empty(out_shape(kind, shape), dtype=odt)
        aux(kind, np.ascontiguousarray(x), y)
        return y
    finally:
        aux.close()


def pair_scale(kind, shape):
    """y = inverse(forward(x)) / scale for the kind's inverse pair."""
    n = int(np.prod(shape))
    if kind in ("dft", "realdft"):
        return float(n)
    return float(np.prod([2 * s for s in shape]))


# ------------------------------------------------------------ closed forms
@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("n", [1, 2, 8, 64, 1024, 1 << 14, 1 << 15, 1 << 17])
def test_dft_closed_forms(libs, prec, n):
    lib = libs[prec]
    cdt = np.complex128 if prec == 64 else np.complex64
    tol = 1e-12 if prec == 64 else 2e-5
    # delta -> all ones
    x = np.zeros(n, cdt)
    x[0] = 1
    y = run(lib, "dft", x, (n,))
    assert np.max(np.abs(y - 1)) < tol
    # constant -> N delta_0
    x = np.ones(n, cdt)
    y = run(lib, "dft", x, (n,))
    want = np.zeros(n, np.complex128)
    want[0] = n
    assert np.max(np.abs(y - want)) / n < tol
    # tone exp(+2 pi i k0 n / N) -> N delta_k0 (forward sign is negative)
    k0 = (3 * n) // 7
    ramp = np.exp(2j * np.pi * k0 * np.arange(n) / n)
    y = run(lib, "dft", ramp.astype(cdt), (n,))
    want = np.zeros(n, np.complex128)
    want[k0] = n
    assert np.max(np.abs(y - want)) / n < tol * 10
    # and the inverse tone -> N delta_k0 under invdft
    y = run(lib, "invdft", np.conj(ramp).astype(cdt), (n,))
    assert np.max(np.abs(y - want)) / n < tol * 10


@pytest.mark.parametrize("prec", [64, 32])
def test_length_one_specials(libs, prec):
    """N = 1: DCT-2/DST-2 give 2x, DCT-3/DST-3 give x, DCT-4/DST-4 give
    sqrt(2) x (minfft.c:700-704, 749-753, 798-802, 848-852, 897-901, 944-948)."""
    lib = libs[prec]
    rdt = np.float64 if prec == 64 else np.float32
    x = np.array([0.75], rdt)
    for kind, f in (("dct2", 2), ("dst2", 2), ("dct3", 1), ("dst3", 1), ("dct4", math.sqrt(2)),
                    ("dst4", math.sqrt(2))):
        y = run(lib, kind, x, (1,))
        assert abs(float(y[0]) - f * 0.75) < 1e-6, kind


# ------------------------------------------------------------ properties
SHAPES = st.sampled_from([(2,), (16,), (256,), (4096,), (1 << 15,), (8, 4), (64, 128), (4, 8, 16), (2, 256, 4)])
KINDS = st.sampled_from(["dft", "realdft", "dct2", "dst2", "dct4", "dst4"])


def draw_input(rng, kind, shape, prec):
    rdt = np.float64 if prec == 64 else np.float32
    cdt = np.complex128 if prec == 64 else np.complex64
    if kind == "dft":
        return (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(cdt)
    return rng.uniform(-1, 1, shape).astype(rdt)


@settings(max_examples=25, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(kind=KINDS, shape=SHAPES, prec=st.sampled_from([64, 32]), seed=st.integers(0, 2 ** 31 - 1),
       a=st.floats(-4, 4), b=st.floats(-4, 4))
def test_linearity_and_inverse_pairs(libs, kind, shape, prec, seed, a, b):
    lib = libs[prec]
    rng = np.random.default_rng(seed)
    x = draw_input(rng, kind, shape, prec)
    z = draw_input(rng, kind, shape, prec)
    dt = x.dtype
    fx, fz = run(lib, kind, x, shape), run(lib, kind, z, shape)
    fxz = run(lib, kind, (a * x.astype(np.complex128 if kind == "dft" else np.float64) +
                          b * z.astype(np.complex128 if kind == "dft" else np.float64)).astype(dt), shape)
    lin = a * fx.astype(np.complex128 if np.iscomplexobj(fx) else np.float64) + \
        b * fz.astype(np.complex128 if np.iscomplexobj(fz) else np.float64)
    bar = 20 * (tol64(shape) if prec == 64 else TOL32)
    if np.linalg.norm(lin) > 1e-3 * (np.linalg.norm(fx) + np.linalg.norm(fz)):
        assert rel_l2(fxz, lin) <= bar
    # inverse pair
    back = run(lib, INVERSE[kind], fx, shape) / pair_scale(kind, shape)
    assert rel_l2(back, x) <= 4 * (tol64(shape) if prec == 64 else TOL32), (kind, shape, prec)


@settings(max_examples=20, deadline=None, suppress_health_check=[HealthCheck.function_scoped_fixture])
@given(shape=SHAPES, prec=st.sampled_from([64, 32]), seed=st.integers(0, 2 ** 31 - 1))
def test_parseval_and_hermitian(libs, shape, prec, seed):
    """sum |X|^2 = N sum |x|^2; the real DFT output is the non-redundant half
    of a Hermitian spectrum: its DC (and Nyquist) rows are real in 1D."""
    lib = libs[prec]
    rng = np.random.default_rng(seed)
    n = int(np.prod(shape))
    x = draw_input(rng, "dft", shape, prec)
    X = run(lib, "dft", x, shape).astype(np.complex128)
    e_x = float(np.sum(np.abs(x.astype(np.complex128)) ** 2))
    e_X = float(np.sum(np.abs(X) ** 2))
    tol = 1e-12 if prec == 64 else 3e-6
    assert abs(e_X - n * e_x) <= tol * n * e_x
    r = draw_input(rng, "realdft", shape, prec)
    R = run(lib, "realdft", r, shape).astype(np.complex128)
    full = np.fft.fftn(r.astype(np.float64))
    # the half spectrum of the complex transform of the same real input
    Xr = run(lib, "dft", r.astype(x.dtype), shape).astype(np.complex128)
    h = shape[-1] // 2 + 1
    assert rel_l2(R, Xr[..., :h]) <= (tol64(shape) if prec == 64 else TOL32) * 4
    assert rel_l2(R, full[..., :h]) <= (tol64(shape) if prec == 64 else TOL32) * 4
    if len(shape) == 1:
        scale = max(1.0, float(np.max(np.abs(R))))
        assert abs(R[0].imag) <= tol * scale and abs(R[-1].imag) <= tol * scale


# ------------------------------------------------ full benchmark sizes
def _torch():
    import torch
    return torch


def test_round_trip_512_cubed_fp32(libs):
    """512^3 complex64 (the 3D configuration): invdft(dft(x)) = N x, on the
    device, FP32 and split-FP16 forms."""
    torch = _torch()
    shape = (512, 512, 512)
    n = 512 ** 3
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    x = torch.complex(torch.rand(shape, generator=g, device="cuda") * 2 - 1,
                      torch.rand(shape, generator=g, device="cuda") * 2 - 1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    print("[case] 512^3 FP32 round trip", flush=True)
    aux = mf.Aux(libs[32], "dft", shape)
    aux.enqueue("dft", x, y)
    aux.enqueue("invdft", y, z)
    torch.cuda.synchronize()
    aux.close()
    err = float(torch.linalg.vector_norm(z / n - x) / torch.linalg.vector_norm(x))
    assert err <= 2 * TOL32, err
    # Parseval on the forward result
    ex = float(torch.sum(torch.abs(x.to(torch.complex128)) ** 2))
    ey = float(torch.sum(torch.abs(y.to(torch.complex128)) ** 2))
    assert abs(ey - n * ex) <= 1e-5 * n * ex
    print("[case] 512^3 split-FP16 round trip", flush=True)
    for kind, src, dst in (("dft", x, y), ("invdft", y, z)):
        b = mf.Batch([(kind, shape, 0, 0)], mf.PREC_SPLIT_FP16)
        b.run(src.data_ptr(), dst.data_ptr(), None)
        torch.cuda.synchronize()
        b.close()
    err = float(torch.linalg.vector_norm(z / n - x) / torch.linalg.vector_norm(x))
    assert err <= 2 * TOL32, err


def test_round_trip_default_batch_fp32():
    """The 8192-entry uneven batch (lengths 2^6..2^16, seed 3): forward batch
    then inverse batch = N x per entry, on the device."""
    torch = _torch()
    rng = np.random.default_rng(3)
    ks = 6 + np.floor(11 * rng.random(8192)).astype(np.int64)
    sizes = (1 << ks).astype(np.int64)
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sizes.sum())
    g = torch.Generator(device="cuda")
    g.manual_seed(3)
    x = torch.complex(torch.rand(total, generator=g, device="cuda") * 2 - 1,
                      torch.rand(total, generator=g, device="cuda") * 2 - 1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    print("[case] default batch round trip", flush=True)
    fwd = mf.Batch([("dft", (int(n),), int(o) * 8, int(o) * 8) for n, o in zip(sizes, offs)], mf.PREC_FP32)
    inv = mf.Batch([("invdft", (int(n),), int(o) * 8, int(o) * 8) for n, o in zip(sizes, offs)], mf.PREC_FP32)
    fwd.run(x.data_ptr(), y.data_ptr(), None)
    inv.run(y.data_ptr(), z.data_ptr(), None)
    torch.cuda.synchronize()
    fwd.close()
    inv.close()
    scale = torch.from_numpy(np.repeat(sizes, sizes).astype(np.float32)).cuda()
    err = torch.abs(z / scale - x)
    # per-entry relative error, worst entry
    seg = torch.from_numpy(np.repeat(np.arange(len(sizes)), sizes)).cuda()
    num = torch.zeros(len(sizes), device="cuda", dtype=torch.float64).index_add_(0, seg, (err.double() ** 2))
    den = torch.zeros(len(sizes), device="cuda", dtype=torch.float64).index_add_(0, seg, (torch.abs(x).double() ** 2))
    worst = float(torch.sqrt(num / den).max())
    assert worst <= 2 * TOL32, worst


def test_round_trip_dct_images_fp32():
    """1024^2 images, mixed kinds in one batch (the DCT configuration's shape):
    dct3(dct2) = dst3(dst2) = dct4(dct4) = (2*1024)^2 x, on the device."""
    torch = _torch()
    n = 1024
    count = 24
    kinds = ["dct2", "dst2", "dct4"]
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    x = torch.rand(count * n * n, generator=g, device="cuda")
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    nb = n * n * 4
    print("[case] 1024^2 DCT batch round trip", flush=True)
    fwd = mf.Batch([(kinds[i % 3], (n, n), i * nb, i * nb) for i in range(count)], mf.PREC_FP32)
    inv = mf.Batch([(INVERSE[kinds[i % 3]], (n, n), i * nb, i * nb) for i in range(count)], mf.PREC_FP32)
    fwd.run(x.data_ptr(), y.data_ptr(), None)
    inv.run(y.data_ptr(), z.data_ptr(), None)
    torch.cuda.synchronize()
    fwd.close()
    inv.close()
    err = float(torch.linalg.vector_norm(z / (2.0 * n) ** 2 - x) / torch.linalg.vector_norm(x))
    assert err <= 2 * TOL32, err


@pytest.mark.parametrize("prec,logn", [(64, 26), (32, 28)])
def test_largest_1d_lengths(libs, prec, logn):
    """The longest 1D transforms a plan accepts (the two-step limit: FP64
    2^26, FP32 2^28 points): two output bins against direct FP64 sums on
    the host (exact integer phases), the round trip invdft(dft(x)) = N x and
    Parseval on the device; twice the length is refused by the maker."""
    torch = _torch()
    n = 1 << logn
    rdt = torch.float64 if prec == 64 else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(logn)
    x = torch.complex(torch.rand(n, generator=g, device="cuda", dtype=rdt) * 2 - 1,
                      torch.rand(n, generator=g, device="cuda", dtype=rdt) * 2 - 1)
    y = torch.empty_like(x)
    z = torch.empty_like(x)
    print(f"[case] 1D length 2^{logn} prec={prec}", flush=True)
    aux = mf.Aux(libs[prec], "dft", (n,))
    aux.enqueue("dft", x, y)
    aux.enqueue("invdft", y, z)
    torch.cuda.synchronize()
    aux.close()
    err = float(torch.linalg.vector_norm(z / n - x) / torch.linalg.vector_norm(x))
    assert err <= (tol64((n,)) if prec == 64 else 2 * TOL32), err
    ex = float(torch.sum(torch.abs(x.to(torch.complex128)) ** 2))
    ey = float(torch.sum(torch.abs(y.to(torch.complex128)) ** 2))
    assert abs(ey - n * ex) <= (1e-12 if prec == 64 else 1e-5) * n * ex
    rms = math.sqrt(ey / n)
    xh = x.cpu().numpy().astype(np.complex128)
    yh = y.cpu().numpy()
    del x, y, z
    chunk = 1 << 24
    for k in (1, n // 2 + 3):
        acc = 0j
        for s in range(0, n, chunk):
            m = (np.arange(s, s + chunk, dtype=np.int64) * k) % n  # exact phase index
            acc += np.sum(xh[s:s + chunk] * np.exp(-2j * np.pi * (m / n)))
        e = abs(complex(yh[k]) - acc) / rms
        assert e <= (1e-12 if prec == 64 else 1e-5), (k, e)
    assert not libs[prec].mkaux_raw("minfft_mkaux_dft_1d", 2 * n)
    assert "limit" in mf.last_error(), mf.last_error()
