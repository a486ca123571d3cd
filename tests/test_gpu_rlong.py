"""Real, DCT and DST transforms whose half-length FFT is longer than one tile,
in two HBM passes (rlong.hpp: pre-processing fused into step 1 of the
two-step FFT, post-processing into step 2), against the FP64 oracle, and
against the previous four-pass form (pre-kernel, two-step FFT, post-kernel;
schedule key rlong=0) on the same input.

Bars: FP64 rel-L2 <= 1e-14*log2(N); FP32 rel-L2 <= 1e-6 (both vs the FP64
oracle on the same input)."""
import numpy as np
import pytest

import minfft_amd as mf
from oracle.oracle import KINDS, out_shape, is_complex_in, is_complex_out
from tests.helpers import TOL32, make_input, rel_l2, tol64

pytestmark = pytest.mark.gpu

REAL_KINDS = [k for k in KINDS if k not in ("dft", "invdft")]


def _input(kind, shape, prec, seed, oracle64):
    if kind == "invrealdft":
        # a Hermitian-consistent half spectrum: the real DFT of real data
        xr = make_input("dct2", shape, prec, seed)
        return oracle64.run("realdft", xr.astype(np.float64), shape).astype(
            np.complex128 if prec == 64 else np.complex64)
    return make_input(kind, shape, prec, seed)


def _run(entries, xs, prec, spec=None):
    """One batch of (kind, shape) entries on packed device buffers; returns
    the outputs (numpy) and the host copy of the output buffer."""
    import torch
    rs = 8 if prec == 64 else 4
    ioffs, ooffs = [], []
    ip = op = 0
    for (kind, shape), x in zip(entries, xs):
        ioffs.append(ip)
        ip += (x.nbytes + 255) // 256 * 256 + 256
        ooffs.append(op)
        n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
        op += (n_out * rs + 255) // 256 * 256 + 256
    hin = np.zeros(ip, dtype=np.uint8)
    for x, o in zip(xs, ioffs):
        hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    dout = torch.full((op,), 0x7f, dtype=torch.uint8, device="cuda")
    b = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)],
                 mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
    if spec is not None:
        for g in range(len(b.describe().strip().split("\n"))):
            b.schedule(g, spec)
    st = torch.cuda.Stream()
    b.run(din.data_ptr(), dout.data_ptr(), st.cuda_stream)
    torch.cuda.synchronize()
    hout = dout.cpu().numpy()
    b.close()
    assert np.array_equal(din.cpu().numpy(), hin), "input modified"
    rdt = np.float64 if prec == 64 else np.float32
    cdt = np.complex128 if prec == 64 else np.complex64
    ys = []
    for (kind, shape), o in zip(entries, ooffs):
        odt = cdt if is_complex_out(kind) else rdt
        n_out = int(np.prod(out_shape(kind, shape)))
        ys.append(hout[o:o + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(kind, shape)).copy())
    return ys, hout, ooffs


def _check(kind, shape, prec, y, x, oracle64, what):
    ref = oracle64.run(kind, x.astype(np.complex128 if is_complex_in(kind) else np.float64), shape)
    e = rel_l2(y, ref)
    bar = tol64(shape) if prec == 64 else TOL32
    print(f"  {what} {kind} {shape} prec={prec}: rel-L2 {e:.3e}", flush=True)
    assert e <= bar, f"{what} {kind} {shape} prec={prec}: rel-L2 {e:.3e} > {bar:.1e}"


# (precision, N): half-length FFTs from just past one tile (FP32 2^15, FP64
# 2^14) to 2^20 points, every split log2 N1 = log2 L / 2 from 7 to 10
CASES = [(32, 1 << 16), (32, 1 << 19), (32, 1 << 21), (64, 1 << 15), (64, 1 << 17), (64, 1 << 20)]


@pytest.mark.parametrize("prec,N", CASES, ids=lambda v: str(v))
def test_two_pass_long_real_kinds(oracle64, prec, N):
    """Every real kind at N points, two entries each (an entry table), in
    two passes against the oracle; the four-pass form agrees with it to the
    same bar; the bytes between and after the entries stay untouched."""
    print(f"[case] two-pass long real kinds prec={prec} N={N}", flush=True)
    entries, xs = [], []
    for i, kind in enumerate(REAL_KINDS):
        for j in range(2):
            entries.append((kind, (N,)))
            xs.append(_input(kind, (N,), prec, 500 + 10 * i + j, oracle64))
    ys, hout, ooffs = _run(entries, xs, prec)
    for (kind, shape), x, y in zip(entries, xs, ys):
        _check(kind, shape, prec, y, x, oracle64, "two-pass")
    # untouched gaps (filled with 0x7f before the run)
    rs = 8 if prec == 64 else 4
    for (kind, shape), o in zip(entries, ooffs):
        n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
        end = o + n_out * rs
        nxt = (end + 255) // 256 * 256 + 256
        assert np.all(hout[end:min(nxt, len(hout))] == 0x7f), (kind, "gap written")
    ys4, _, _ = _run(entries, xs, prec, "rlong=0")
    for (kind, shape), x, y4 in zip(entries, xs, ys4):
        _check(kind, shape, prec, y4, x, oracle64, "four-pass")


@pytest.mark.parametrize("prec", [32, 64])
def test_two_pass_long_real_axes_of_multidim(oracle64, prec):
    """Long real axes inside multi-dimensional transforms: the real DFT and
    its inverse with a long last axis (4 rows), DCT-2 and DST-4 along a
    long first axis (items 2 apart), against the oracle."""
    n = 1 << (16 if prec == 32 else 15)
    print(f"[case] two-pass multi-dimensional prec={prec} n={n}", flush=True)
    entries = [("realdft", (4, n)), ("invrealdft", (4, n)), ("dct2", (n, 2)), ("dst4", (n, 2))]
    xs = [_input(k, s, prec, 700 + i, oracle64) for i, (k, s) in enumerate(entries)]
    ys, _, _ = _run(entries, xs, prec)
    for (kind, shape), x, y in zip(entries, xs, ys):
        _check(kind, shape, prec, y, x, oracle64, "two-pass")


def test_two_pass_plain_api_in_place(oracle64):
    """The 27-symbol API on host arrays (plan schedule rlong=1 by default),
    out of place and in place, FP32 real DFT and DCT-3 of 2^18 points."""
    lib = mf.Minfft(32)
    n = 1 << 18
    print(f"[case] two-pass plain API n={n}", flush=True)
    for kind in ("realdft", "dct3"):
        x = make_input(kind, (n,), 32, 77)
        aux = mf.Aux(lib, mf.FAMILY[kind], (n,))
        odt = np.complex64 if kind == "realdft" else np.float32
        y = np.empty(out_shape(kind, (n,)), dtype=odt)
        aux(kind, x, y)
        _check(kind, (n,), 32, y, x, oracle64, "out of place")
        buf = np.zeros(max(x.nbytes, y.nbytes), dtype=np.uint8)
        buf[:x.nbytes] = np.frombuffer(x.tobytes(), dtype=np.uint8)
        aux(kind, buf, buf)
        yi = buf[:y.nbytes].view(odt).reshape(y.shape)
        assert np.array_equal(yi, y), kind
        aux.close()
