"""Every benchmark configuration checked at its own size on the GPU.

- 2D DCT-2 / DCT-3 (and DST-2 / DST-3) on 2048x2048 FP64, N(0,1) input:
  against the FP64 oracle (which agrees with the reference's own FP64 output
  to the last bit, tests/test_oracle.py) at rel-L2 <= 1e-14 * log2(2048^2),
  and the inverse pair dct3(dct2(x)) = 4 * 2048^2 * x.  Reference routines:
  minfft.c:735-738 (DCT-2), 834-837 (DCT-3).
- The 256-image 1024x1024 FP32 batch of mixed DCT-2 / DST-2 / DCT-4
  (image i gets type i mod 3) on 8x8-block-structured input, as one batch
  plan: the first and last images and one of each type against the FP64
  oracle (rel-L2 <= 1e-6), with gaps between the images that must stay
  untouched.  Reference routines: minfft.c:735-738, 784-787, 931-934.
- 512^3 complex64, radix-16 stages as split-FP16 MFMA products: on a field of
  64 pure plane waves (built on the device from separable FP64 exponentials)
  the exact spectrum is 512^3 e^{i phi_j} at the 64 wavevectors and zero
  elsewhere -- rel-L2 over the whole array <= 1e-6; with the configuration's
  N(0, 1e-3) noise added, against the FP64 oracle composed from 1D
  transforms along each axis (the reference's 3D maker overflows int at this
  size in FP64, minfft.c:1039) at rel-L2 <= 1e-6, and the 64 spikes within
  1e-6 of the closed form.  Reference routine: minfft.c:287-290 per axis.
"""
import math

import numpy as np
import pytest

import minfft_amd as mf
from tests.helpers import TOL32, rel_l2

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


@pytest.fixture(scope="module")
def lib64():
    return mf.Minfft(64)


@pytest.fixture(scope="module")
def lib32():
    return mf.Minfft(32)


# ------------------------------------------------------------- 2048^2 FP64
@pytest.mark.parametrize("pair", [("dct2", "dct3"), ("dst2", "dst3")])
def test_t2t3_2048_squared_fp64(lib64, oracle64, pair):
    torch = _torch()
    fwd, inv = pair
    n = 2048
    shape = (n, n)
    print(f"[case] {fwd}/{inv} {n}x{n} FP64", flush=True)
    rng = np.random.default_rng(22)
    x = rng.standard_normal(shape)
    xd = torch.from_numpy(x).cuda()
    yd = torch.empty_like(xd)
    zd = torch.empty_like(xd)
    aux = mf.Aux(lib64, "t2t3", shape)
    aux.enqueue(fwd, xd, yd)
    aux.enqueue(inv, yd, zd)
    torch.cuda.synchronize()
    aux.close()
    y = yd.cpu().numpy()
    tol = 1e-14 * math.log2(n * n)
    yo = oracle64.run(fwd, x, shape)
    e = rel_l2(y, yo)
    print(f"  {fwd} rel-L2 vs oracle {e:.3e} (bar {tol:.1e})", flush=True)
    assert e <= tol, e
    # the inverse kind on the same input, against the oracle
    zo = oracle64.run(inv, y, shape)
    zi = torch.empty_like(xd)
    aux = mf.Aux(lib64, "t2t3", shape)
    aux.enqueue(inv, yd, zi)
    torch.cuda.synchronize()
    aux.close()
    e3 = rel_l2(zi.cpu().numpy(), zo)
    print(f"  {inv} rel-L2 vs oracle {e3:.3e}", flush=True)
    assert e3 <= tol, e3
    # inverse pair: 2N per dimension
    ep = rel_l2(zd.cpu().numpy() / (4.0 * n * n), x)
    assert ep <= tol, ep


def test_realdft_pair_2048_squared_fp64(lib64, oracle64):
    """The same grid's real DFT (2048x1025 complex output) and its inverse,
    both against the oracle, N(0,1) input."""
    torch = _torch()
    n = 2048
    shape = (n, n)
    print("[case] realdft/invrealdft 2048x2048 FP64", flush=True)
    rng = np.random.default_rng(21)
    x = rng.standard_normal(shape)
    xd = torch.from_numpy(x).cuda()
    zd = torch.empty((n, n // 2 + 1), dtype=torch.complex128, device="cuda")
    bd = torch.empty_like(xd)
    aux = mf.Aux(lib64, "realdft", shape)
    aux.enqueue("realdft", xd, zd)
    aux.enqueue("invrealdft", zd, bd)
    torch.cuda.synchronize()
    aux.close()
    tol = 1e-14 * 22
    z = zd.cpu().numpy()
    e = rel_l2(z, oracle64.run("realdft", x, shape))
    assert e <= tol, e
    e2 = rel_l2(bd.cpu().numpy(), oracle64.run("invrealdft", z, shape))
    assert e2 <= tol, e2


# ------------------------------------------- 256 x 1024^2 FP32 DCT batch
def test_dct_image_batch_full_size(oracle64):
    torch = _torch()
    n, count = 1024, 256
    kinds = ("dct2", "dst2", "dct4")
    gap = 1024  # floats between images, on both sides
    stride = n * n + gap
    print(f"[case] {count} x {n}^2 FP32 mixed DCT-2/DST-2/DCT-4 batch", flush=True)
    g = torch.Generator(device="cuda")
    g.manual_seed(5)
    m = torch.rand((count, n // 8, n // 8), generator=g, device="cuda", dtype=torch.float32)
    img = m.repeat_interleave(8, 1).repeat_interleave(8, 2)
    img += (torch.rand((count, n, n), generator=g, device="cuda", dtype=torch.float32) - 0.5) * 0.1
    sentinel = -12345.5
    xin = torch.full((count * stride,), sentinel, device="cuda", dtype=torch.float32)
    xout = torch.full((count * stride,), sentinel, device="cuda", dtype=torch.float32)
    xin.view(count, stride)[:, : n * n] = img.reshape(count, n * n)
    del img, m
    xin_copy = xin.clone()
    plan = mf.Batch([(kinds[i % 3], (n, n), i * stride * 4, i * stride * 4) for i in range(count)],
                    mf.PREC_FP32)
    plan.run(xin.data_ptr(), xout.data_ptr(), None)
    torch.cuda.synchronize()
    plan.close()
    assert torch.equal(xin, xin_copy), "input changed"
    gaps = xout.view(count, stride)[:, n * n:]
    assert bool(torch.all(gaps == sentinel)), "gap between images written"
    worst = 0.0
    for i in sorted({0, 1, 2, 127, 128, 129, count - 1}):
        k = kinds[i % 3]
        xi = xin.view(count, stride)[i, : n * n].reshape(n, n).cpu().numpy().astype(np.float64)
        yi = xout.view(count, stride)[i, : n * n].reshape(n, n).cpu().numpy()
        r = oracle64.run(k, xi, (n, n))
        e = rel_l2(yi, r)
        mx = float(np.max(np.abs(yi - r)))
        print(f"  image {i} {k}: rel-L2 {e:.3e} max-abs {mx:.3e}", flush=True)
        worst = max(worst, e)
        assert e <= TOL32, (i, k, e)
    print(f"  worst rel-L2 {worst:.3e}", flush=True)


# ------------------------------------------------- 512^3 split-FP16 MFMA
def _plane_waves(torch, n, seed=4, count=64):
    """Sum of `count` unit plane waves with distinct integer wavevectors and
    random phases, in FP64 on the device (separable exponentials with exact
    integer phase reduction); returns (field complex128, wavevectors, phases)."""
    rng = np.random.default_rng(seed)
    ks = set()
    while len(ks) < count:
        ks.add(tuple(int(v) for v in rng.integers(0, n, 3)))
    ks = sorted(ks)
    phis = rng.uniform(0, 2 * np.pi, count)
    r = torch.arange(n, device="cuda", dtype=torch.int64)
    field = torch.zeros((n, n, n), dtype=torch.complex128, device="cuda")

    def ex(k):
        ang = ((k * r) % n).to(torch.float64) * (2 * math.pi / n)
        return torch.complex(torch.cos(ang), torch.sin(ang))

    for (kx, ky, kz), p in zip(ks, phis):
        a = ex(kx) * complex(math.cos(p), math.sin(p))
        field += a[:, None, None] * (ex(ky)[:, None] * ex(kz)[None, :])[None, :, :]
    return field, ks, phis


def _split16(torch, x, y, spec=None):
    plan = mf.Batch([("dft", tuple(x.shape), 0, 0)], mf.PREC_SPLIT_FP16)
    if spec:
        plan.schedule(0, spec)
    plan.run(x.data_ptr(), y.data_ptr(), None)
    torch.cuda.synchronize()
    desc = plan.describe()
    plan.close()
    return desc


def test_512_cubed_split_fp16_plane_waves(lib32):
    torch = _torch()
    n = 512
    print("[case] 512^3 split-FP16 on 64 pure plane waves vs the exact spectrum", flush=True)
    field, ks, phis = _plane_waves(torch, n)
    x = field.to(torch.complex64)
    del field
    torch.cuda.empty_cache()
    y = torch.empty_like(x)
    _split16(torch, x, y)
    idx = torch.tensor(ks, device="cuda")
    want = torch.tensor((n ** 3) * np.exp(1j * phis), device="cuda", dtype=torch.complex128)

    def err_vs_closed_form(out):
        spikes = out[idx[:, 0], idx[:, 1], idx[:, 2]].to(torch.complex128)
        tot = float(torch.sum(torch.abs(out.to(torch.complex128)) ** 2))
        # ||out - exact||^2 = ||out||^2 - ||out at spikes||^2 + ||spikes - want||^2
        off = tot - float(torch.sum(torch.abs(spikes) ** 2))
        on = float(torch.sum(torch.abs(spikes - want) ** 2))
        return math.sqrt(max(off, 0.0) + on) / (n ** 3 * math.sqrt(len(ks)))

    e16 = err_vs_closed_form(y)
    print(f"  split-FP16 rel-L2 vs closed form {e16:.3e}", flush=True)
    assert e16 <= TOL32, e16
    # both radix-16 stages on the matrix cores (schedule key mstages=2)
    _split16(torch, x, y, "mstages=2")
    e16b = err_vs_closed_form(y)
    print(f"  split-FP16, both radix-16 stages, rel-L2 vs closed form {e16b:.3e}", flush=True)
    assert e16b <= TOL32, e16b
    # the FP32 form of the plain API on the same input
    aux = mf.Aux(lib32, "dft", (n, n, n))
    aux.enqueue("dft", x, y)
    torch.cuda.synchronize()
    aux.close()
    e32 = err_vs_closed_form(y)
    print(f"  FP32 rel-L2 vs closed form {e32:.3e}", flush=True)
    assert e32 <= TOL32, e32


def test_512_cubed_split_fp16_noisy_vs_oracle(oracle64):
    torch = _torch()
    n = 512
    print("[case] 512^3 split-FP16 on plane waves + N(0,1e-3) noise vs the FP64 oracle", flush=True)
    field, ks, phis = _plane_waves(torch, n)
    g = torch.Generator(device="cuda")
    g.manual_seed(4)
    field += torch.complex(torch.randn(field.shape, generator=g, device="cuda", dtype=torch.float64),
                           torch.randn(field.shape, generator=g, device="cuda", dtype=torch.float64)) * 1e-3
    x = field.to(torch.complex64)
    del field
    torch.cuda.empty_cache()
    y = torch.empty_like(x)
    desc = _split16(torch, x, y)
    print("  plan: " + desc.strip(), flush=True)
    yh = y.cpu().numpy()
    spikes = np.array([yh[k] for k in ks])
    srel = float(np.max(np.abs(spikes - (n ** 3) * np.exp(1j * phis))) / n ** 3)
    print(f"  spikes vs closed form: max rel {srel:.3e}", flush=True)
    assert srel <= TOL32, srel
    xh = x.cpu().numpy()
    del x, y
    print("  FP64 oracle along each axis ...", flush=True)
    ref = oracle64.run("dft", xh.astype(np.complex128), (n, n, n))
    e = rel_l2(yh, ref)
    mx = float(np.max(np.abs(yh - ref)))
    print(f"  split-FP16 rel-L2 vs FP64 oracle {e:.3e}, max-abs {mx:.3e}", flush=True)
    assert e <= TOL32, e


# ------------------------------------- long 1D real kinds, FP32 (two-step)
@pytest.mark.parametrize("kind", ["realdft", "invrealdft", "dct2", "dst2", "dct3", "dct4", "dst4"])
def test_long_real_kinds_fp32(lib32, oracle64, kind):
    """FP32 real/DCT/DST transforms of 2^20 points (their half-length FFT
    needs the two-step path; the final post-twiddle is computed in FP64)
    against the FP64 oracle on the same float input: rel-L2 <= 1e-6, with
    rel-L2 and max error printed."""
    from tests.helpers import make_input
    n = 1 << 20
    shape = (n,)
    if kind == "invrealdft":
        xr = make_input("dct2", shape, 32, 31)
        x = oracle64.run("realdft", xr.astype(np.float64), shape).astype(np.complex64)
    else:
        x = make_input(kind, shape, 32, 30)
    aux = mf.Aux(lib32, mf.FAMILY[kind], shape)
    odt = np.complex64 if kind == "realdft" else np.float32
    y = np.empty(n // 2 + 1 if kind == "realdft" else n, dtype=odt)
    aux(kind, x, y)
    aux.close()
    ref = oracle64.run(kind, x.astype(np.complex128 if kind == "invrealdft" else np.float64), shape)
    e = rel_l2(y, ref)
    mx = float(np.max(np.abs(y.astype(ref.dtype) - ref)) / np.max(np.abs(ref)))
    print(f"[case] {kind} 2^20 FP32: rel-L2 {e:.3e}, max error / max |y| {mx:.3e}", flush=True)
    assert e <= TOL32, e
