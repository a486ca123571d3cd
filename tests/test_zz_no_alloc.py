"""minfft_enqueue_noalloc (include/minfft_ext.h) against minfft_enqueue, the
entry point it is a variant of: the same kernels on the same aux, with the
scratch in the library's fixed device array instead of the aux's buffers,
so the outputs must be identical bit for bit.  Each case makes two calls in
a row, the second with a smaller input (on its own aux), on transforms that
use each kind of scratch: the two-step intermediate (long complex axes), the
complex staging of long real/DCT/DST axes, and the full-array temp of the
2D inverse real DFT.  A transform needing more scratch than the array holds
is refused with a reason and its output left untouched."""
import numpy as np
import pytest

import minfft_amd as mf

pytestmark = pytest.mark.gpu

SCRATCH = 64 << 20

# (prec, family, kind, large shape, smaller shape)
CASES = [
    (32, "dft", "dft", (1 << 20,), (1 << 16,)),
    (64, "dft", "invdft", (1 << 19,), (1 << 15,)),
    (32, "realdft", "realdft", (1 << 21,), (1 << 17,)),
    (64, "t2t3", "dct2", (1 << 19,), (1 << 16,)),
    (32, "t4", "dst4", (1 << 20,), (1 << 17,)),
    (64, "realdft", "invrealdft", (512, 1024), (128, 256)),
    (32, "dft", "dft", (4, 1 << 18), (2, 1 << 17)),
]


def _torch():
    import torch
    return torch


def _io(torch, prec, kind, shape, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    rdt = torch.float64 if prec == 64 else torch.float32
    cdt = torch.complex128 if prec == 64 else torch.complex64
    last = shape[-1]
    ish = shape[:-1] + (last // 2 + 1,) if kind == "invrealdft" else shape
    osh = shape[:-1] + (last // 2 + 1,) if kind == "realdft" else shape
    cin = kind in ("dft", "invdft", "invrealdft")
    cout = kind in ("dft", "invdft", "realdft")
    x = torch.rand(ish + ((2,) if cin else ()), generator=g, device="cuda", dtype=rdt) * 2 - 1
    if cin:
        x = torch.view_as_complex(x.contiguous()).to(cdt)
    y = torch.full(osh, float("nan"), device="cuda", dtype=cdt if cout else rdt)
    return x, y


@pytest.mark.parametrize("prec,family,kind,big,small", CASES,
                         ids=[f"{c[2]}-{c[0]}-{'x'.join(map(str, c[3]))}" for c in CASES])
def test_noalloc_matches_enqueue(prec, family, kind, big, small):
    torch = _torch()
    lib = mf.Minfft(prec)
    for call, shape in enumerate((big, small)):
        print(f"[case] noalloc {kind} prec={prec} shape={shape} call {call + 1}", flush=True)
        aux = mf.Aux(lib, family, shape)
        x, y_ref = _io(torch, prec, kind, shape, 100 + call)
        x_keep = x.clone()
        y = torch.full_like(y_ref, float("nan"))
        aux.enqueue(kind, x, y_ref)
        aux.enqueue_noalloc(kind, x, y)
        torch.cuda.synchronize()
        aux.close()
        assert torch.equal(x, x_keep), "input changed"
        assert not bool(torch.isnan(y).any()), "output not fully written"
        assert torch.equal(y, y_ref), float((y - y_ref).abs().max())


def test_noalloc_refuses_oversized_scratch():
    """An FP32 complex 1D transform of 2^23 points needs 2 x 64 MiB of
    two-step scratch: refused, with the sizes in the reason, nothing
    launched; the same aux then still runs through minfft_enqueue."""
    torch = _torch()
    lib = mf.Minfft(32)
    n = 1 << 23
    aux = mf.Aux(lib, "dft", (n,))
    x, y = _io(torch, 32, "dft", (n,), 7)
    with pytest.raises(mf.MinfftError, match="bytes of scratch"):
        aux.enqueue_noalloc("dft", x, y)
    torch.cuda.synchronize()
    assert bool(torch.isnan(y).all())
    aux.enqueue("dft", x, y)
    torch.cuda.synchronize()
    aux.close()
    assert not bool(torch.isnan(y).any())


def test_noalloc_refuses_host_pointers():
    lib = mf.Minfft(64)
    aux = mf.Aux(lib, "dft", (1 << 15,))
    x = np.ones(1 << 15, dtype=np.complex128)
    y = np.zeros_like(x)
    with pytest.raises(mf.MinfftError, match="device"):
        aux.enqueue_noalloc("dft", x, y)
    aux.close()
    assert not np.any(y)
