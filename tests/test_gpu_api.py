"""GPU parity of the 27-symbol API (libminfft.so / libminfft_single.so)
against the CPU oracle and the reference's golden vectors.

Bars: FP64 rel-L2 <= 1e-14*log2(N) vs the FP64 oracle (which agrees with the
reference to the last bit); FP32 rel-L2 <= 1e-6 vs the FP64 oracle on the same
(float-rounded) input."""
import os

import numpy as np
import pytest

import minfft_amd as mf
from oracle.oracle import KINDS, in_shape, out_shape, is_complex_in, is_complex_out
from tests.helpers import TOL32, make_input, rel_l2, tol64

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


@pytest.fixture(scope="module")
def libs():
    return {64: mf.Minfft(64), 32: mf.Minfft(32)}


def gpu_run(lib, kind, x, shape, inplace=False):
    """Run through the C API on host (numpy) buffers."""
    aux = mf.Aux(lib, mf.FAMILY[kind], shape)
    rdt = np.float64 if lib.prec == 64 else np.float32
    cdt = np.complex128 if lib.prec == 64 else np.complex64
    odt = cdt if is_complex_out(kind) else rdt
    try:
        if not inplace:
            y = np.empty(out_shape(kind, shape), dtype=odt)
            aux(kind, np.ascontiguousarray(x), y)
            return y
        nbytes = max(x.nbytes, int(np.prod(out_shape(kind, shape))) * np.dtype(odt).itemsize)
        buf = np.zeros(nbytes, dtype=np.uint8)
        buf[: x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
        aux(kind, buf, buf)
        n = int(np.prod(out_shape(kind, shape)))
        return buf.view(odt)[:n].reshape(out_shape(kind, shape)).copy()
    finally:
        aux.close()


def check(kind, shape, prec, y, x, oracle64):
    """y (GPU) vs the FP64 oracle on the same input."""
    yo = oracle64.run(kind, x.astype(np.complex128 if is_complex_in(kind) else np.float64), shape)
    e = rel_l2(y, yo)
    bar = tol64(shape) if prec == 64 else TOL32
    assert e <= bar, f"{kind} {shape} prec={prec}: rel-L2 {e:.3e} > {bar:.1e}"
    return e


SHAPES = [(1,), (2,), (4,), (8,), (16,), (32,), (64,), (128,), (256,), (512,), (1024,), (2048,),
          (4096,), (8192,), (16384,),
          (2, 2), (4, 8), (8, 1), (1, 8), (64, 32), (32, 64), (256, 256), (2, 4, 8), (8, 8, 8),
          (16, 1, 32), (32, 16, 8), (512, 32), (1024, 16), (512, 2, 8),
          # the generic makers take any number of dimensions (reference minfft.h:46-61)
          (2, 4, 8, 2), (8, 16, 4, 32), (2, 2, 1, 4, 4), (1, 2, 1, 4, 2, 2)]


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", SHAPES, ids=lambda s: "x".join(map(str, s)))
def test_all_kinds_vs_oracle(libs, oracle64, prec, shape):
    lib = libs[prec]
    print(f"[case] all kinds prec={prec} shape={shape}", flush=True)
    for i, kind in enumerate(KINDS):
        x = make_input(kind, shape, prec, 100 + i, oracle64 if prec == 64 else None) \
            if kind != "invrealdft" else None
        if kind == "invrealdft":
            xr = make_input("dct2", shape, prec, 100 + i)
            z = oracle64.run("realdft", xr.astype(np.float64), shape)
            x = z.astype(np.complex128 if prec == 64 else np.complex64)
        y = gpu_run(lib, kind, x, shape)
        check(kind, shape, prec, y, x, oracle64)


@pytest.mark.parametrize("prec", [64, 32])
def test_golden_vectors(libs, oracle64, prec):
    """Every golden vector of the reference (tests/golden) through the GPU."""
    g = np.load(os.path.join(GOLDEN, f"golden_{prec}.npz"))
    keys = sorted({k.rsplit("__", 1)[0] for k in g.files})
    print(f"[case] golden prec={prec} ({len(keys)} vectors)", flush=True)
    for key in keys:
        kind, sk = key.split("__")
        shape = tuple(int(s) for s in sk.split("x"))
        y = gpu_run(libs[prec], kind, g[key + "__in"], shape)
        e = rel_l2(y, g[key + "__out"])
        bar = tol64(shape) if prec == 64 else 4 * TOL32
        assert e <= bar, (key, e)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", [(1,), (2,), (16,), (1024,), (8, 16), (4, 2, 8), (64, 64)],
                         ids=lambda s: "x".join(map(str, s)))
def test_in_place(libs, oracle64, prec, shape):
    print(f"[case] in-place prec={prec} shape={shape}", flush=True)
    for i, kind in enumerate(KINDS):
        if kind == "invrealdft":
            xr = make_input("dct2", shape, prec, 200 + i)
            x = oracle64.run("realdft", xr.astype(np.float64), shape).astype(
                np.complex128 if prec == 64 else np.complex64)
        else:
            x = make_input(kind, shape, prec, 200 + i)
        y = gpu_run(libs[prec], kind, x, shape, inplace=True)
        check(kind, shape, prec, y, x, oracle64)


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", [(2, 2), (1, 16), (4096, 4), (16384, 2), (8, 4, 2), (2, 1, 8), (2048, 2048)],
                         ids=lambda s: "x".join(map(str, s)))
def test_realdft_packed_dc_nyquist(libs, oracle64, prec, shape):
    """The first complex pass of a multi-dimensional real DFT transforms the
    real DC and Nyquist columns as one complex column and separates them
    after: both columns and the rest against the oracle (column lengths 1 to
    2^14; FP64 2^14 takes the two-step path, which does not pack)."""
    print(f"[case] realdft packed pair prec={prec} shape={shape}", flush=True)
    x = make_input("realdft", shape, prec, 700 + len(shape))
    y = gpu_run(libs[prec], "realdft", x, shape)
    yo = oracle64.run("realdft", x.astype(np.float64), shape)
    bar = tol64(shape) if prec == 64 else TOL32
    for name, sl in (("all", np.s_[...]), ("DC", np.s_[..., 0]), ("Nyquist", np.s_[..., -1])):
        e = rel_l2(y[sl], yo[sl])
        assert e <= bar, f"{name} column(s) {shape} prec={prec}: rel-L2 {e:.3e} > {bar:.1e}"
    # the DC and Nyquist columns of a real input's spectrum are Hermitian along the other axes
    for c in (0, -1):
        col = y[..., c]
        flip = np.conj(np.roll(np.flip(col, axis=tuple(range(col.ndim))), 1, axis=tuple(range(col.ndim))))
        assert rel_l2(col, flip) <= 10 * bar


@pytest.mark.parametrize("prec", [64, 32])
@pytest.mark.parametrize("shape", [(2, 2), (1, 16), (4096, 4), (16384, 2), (8, 4, 2), (2, 1, 8), (4, 8, 16),
                                   (2048, 2048)], ids=lambda s: "x".join(map(str, s)))
def test_invrealdft_packed_dc_nyquist(libs, oracle64, prec, shape):
    """The last complex pass of a multi-dimensional inverse real DFT
    transforms the Hermitian parts of columns h = 0 and h = H-1 as one
    complex column (the real pass after it reads only their real parts).
    Inputs are arbitrary complex arrays, not spectra of real arrays, so the
    imaginary parts the reference ignores are not zero; against the oracle,
    with the input preserved."""
    print(f"[case] invrealdft packed pair prec={prec} shape={shape}", flush=True)
    rng = np.random.default_rng(710 + len(shape) + shape[-1])
    ish = in_shape("invrealdft", shape)
    cdt = np.complex128 if prec == 64 else np.complex64
    x = (rng.uniform(-1, 1, ish) + 1j * rng.uniform(-1, 1, ish)).astype(cdt)
    x0 = x.copy()
    y = gpu_run(libs[prec], "invrealdft", x, shape)
    assert np.array_equal(x, x0)
    yo = oracle64.run("invrealdft", x.astype(np.complex128), shape)
    bar = tol64(shape) if prec == 64 else TOL32
    e = rel_l2(y, yo)
    assert e <= bar, f"{shape} prec={prec}: rel-L2 {e:.3e} > {bar:.1e}"


@pytest.mark.parametrize("prec,N", [(64, 16384), (64, 32768), (64, 131072), (32, 32768), (32, 65536),
                                    (32, 262144)])
def test_long_1d_two_step(libs, oracle64, prec, N):
    """Lengths beyond one work-group's LDS take the two-step path."""
    print(f"[case] two-step prec={prec} N={N}", flush=True)
    for i, kind in enumerate(KINDS):
        if kind == "invrealdft":
            xr = make_input("dct2", (N,), prec, 300 + i)
            x = oracle64.run("realdft", xr.astype(np.float64), (N,)).astype(
                np.complex128 if prec == 64 else np.complex64)
        else:
            x = make_input(kind, (N,), prec, 300 + i)
        y = gpu_run(libs[prec], kind, x, (N,))
        check(kind, (N,), prec, y, x, oracle64)


def test_device_pointers_zero_copy(libs, oracle64):
    import torch
    print("[case] device pointers", flush=True)
    dev = torch.device("cuda:0")
    for prec in (64, 32):
        cdt = torch.complex128 if prec == 64 else torch.complex64
        rdt = torch.float64 if prec == 64 else torch.float32
        shape = (32, 64)
        x = make_input("dft", shape, prec, 7)
        xt = torch.from_numpy(x).to(dev)
        yt = torch.empty(shape, dtype=cdt, device=dev)
        aux = mf.Aux(libs[prec], "dft", shape)
        aux("dft", xt, yt)
        check("dft", shape, prec, yt.cpu().numpy(), x, oracle64)
        aux.close()
        xr = make_input("dct2", shape, prec, 8)
        xrt = torch.from_numpy(xr).to(dev)
        zt = torch.empty((32, 33), dtype=cdt, device=dev)
        aux = mf.Aux(libs[prec], "realdft", shape)
        aux("realdft", xrt, zt)
        check("realdft", shape, prec, zt.cpu().numpy(), xr, oracle64)
        back = torch.empty(shape, dtype=rdt, device=dev)
        aux("invrealdft", zt, back)
        n = 32 * 64
        assert rel_l2(back.cpu().numpy() / n, xr) <= (1e-14 if prec == 64 else 1e-6)
        aux.close()


def test_round_trips_and_input_preserved(libs):
    print("[case] round trips", flush=True)
    lib = libs[64]
    shape = (16, 32)
    n = 16 * 32
    x = make_input("dft", shape, 64, 9)
    x0 = x.copy()
    aux = mf.Aux(lib, "dft", shape)
    y = np.empty_like(x)
    aux("dft", x, y)
    assert np.array_equal(x, x0)
    back = np.empty_like(x)
    aux("invdft", y, back)
    assert rel_l2(back / n, x) <= 1e-14
    aux.close()
    xr = make_input("dct2", shape, 64, 10)
    for a, b, fac in (("dct2", "dct3", 4 * n), ("dst2", "dst3", 4 * n), ("dct4", "dct4", 4 * n),
                      ("dst4", "dst4", 4 * n)):
        aux = mf.Aux(lib, mf.FAMILY[a], shape)
        y1 = np.empty_like(xr)
        aux(a, xr, y1)
        y2 = np.empty_like(xr)
        aux(b, y1, y2)
        assert rel_l2(y2 / fac, xr) <= 1e-14, (a, b)
        aux.close()
    # multi-dimensional inverse real DFT preserves its input here
    z = np.fft.rfftn(xr)
    z0 = z.copy()
    aux = mf.Aux(lib, "realdft", shape)
    out = np.empty(shape)
    aux("invrealdft", z, out)
    assert np.array_equal(z, z0)
    assert rel_l2(out / n, xr) <= 1e-14
    aux.close()


@pytest.mark.parametrize("prec", [64, 32])
def test_makers_reject_invalid_lengths(libs, prec):
    print(f"[case] maker errors prec={prec}", flush=True)
    lib = libs[prec]
    for fam in ("dft", "realdft", "t2t3", "t4"):
        for N in (0, -4, 3, 12, 1000):
            assert not lib.mkaux_raw(f"minfft_mkaux_{fam}_1d", N)
            assert "power" in mf.last_error() or "invalid" in mf.last_error()
        assert not lib.mkaux_raw(f"minfft_mkaux_{fam}_2d", 8, 6)
        assert not lib.mkaux_raw(f"minfft_mkaux_{fam}_3d", 8, 8, 0)
        import ctypes
        Ns = (ctypes.c_int * 1)(8)
        assert not lib.mkaux_raw(f"minfft_mkaux_{fam}", 0, Ns)
    lib.free_aux(None)


def test_mismatched_aux_is_reported(libs):
    lib = libs[64]
    aux = mf.Aux(lib, "dft", (16,))
    x = np.zeros(16)
    with pytest.raises(mf.MinfftError):
        lib.run_ptr("dct2", x.ctypes.data, x.ctypes.data, aux.h)
    aux.close()
    # an aux of the FP64 library is refused by the FP32 library
    aux = mf.Aux(lib, "dft", (16,))
    with pytest.raises(mf.MinfftError):
        libs[32].run_ptr("dft", x.ctypes.data, x.ctypes.data, aux.h)
    aux.close()


@pytest.mark.parametrize("single", [False, True])
def test_c_application_runs_on_the_dropin(tmp_path, single):
    """The C89 program of tests/c_app (written against minfft.h, host
    arrays) runs on the GPU through libminfft.so / libminfft_single.so and
    meets its checks: a direct O(N^2) DFT and the inverse-pair identities
    of every kind."""
    import subprocess
    from tests.test_abi import _build_c_app
    exe = _build_c_app(tmp_path, single)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout, r.stderr)


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("which", ["c99", "cpp"])
def test_c99_and_cpp_clients_run_on_the_dropin(tmp_path, which, single):
    """A C99 program with C99 complex numbers (2D DFT against a direct sum,
    real-DFT round trip) and a C++ program with std::complex arrays (3D DFT
    against a direct sum, DCT-2/DCT-3 round trip) run on the GPU."""
    import subprocess
    from tests.test_abi import _build_client
    exe = _build_client(tmp_path, which, single)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0 and r.stdout.strip().endswith("ok"), (r.stdout, r.stderr)


def test_mixed_host_and_device_pointers(libs, oracle64):
    """One side a host array, the other a device array (staged through the
    aux's pinned bounce buffer on the host side); in place on a host array;
    and a host array larger than the bounce buffer's 64 MiB limit."""
    import torch
    dev = torch.device("cuda:0")
    print("[case] mixed host/device pointers", flush=True)
    for prec in (64, 32):
        cdt = np.complex128 if prec == 64 else np.complex64
        shape = (64, 128)
        x = make_input("dft", shape, prec, 17)
        aux = mf.Aux(libs[prec], "dft", shape)
        # host in, device out
        yt = torch.empty(shape, dtype=torch.complex128 if prec == 64 else torch.complex64, device=dev)
        aux("dft", x, yt)
        check("dft", shape, prec, yt.cpu().numpy(), x, oracle64)
        # device in, host out
        y = np.empty(shape, dtype=cdt)
        aux("dft", torch.from_numpy(x).to(dev), y)
        check("dft", shape, prec, y, x, oracle64)
        # in place on a host array
        z = x.copy()
        aux("dft", z, z)
        check("dft", shape, prec, z, x, oracle64)
        aux.close()
    # 2^23 complex128 points = 128 MiB: pageable copies
    shape = (1 << 23,)
    x = make_input("dft", shape, 64, 19)
    y = np.empty_like(x)
    aux = mf.Aux(libs[64], "dft", shape)
    aux("dft", x, y)
    back = np.empty_like(x)
    aux("invdft", y, back)
    aux.close()
    assert rel_l2(back / shape[0], x) <= tol64(shape)


def test_threads_with_distinct_aux(libs, oracle64):
    """The reference is re-entrant across distinct aux objects: eight host
    threads make their own aux (one large enough to be tuned when made),
    run 10 transforms each through host and device pointers at once, and
    every result matches the oracle."""
    import threading
    import torch
    print("[case] threads with distinct aux", flush=True)
    cases = [("dft", (1024,), 64), ("invdft", (64, 64), 32), ("realdft", (256, 128), 64),
             ("dct2", (32, 32), 32), ("dst4", (4096,), 64), ("dft", (8, 16, 32), 32),
             ("dct3", (128, 64), 64), ("realdft", (1024, 1024), 64)]
    want, inputs = [], []
    for i, (kind, shape, prec) in enumerate(cases):
        x = make_input(kind, shape, prec, 800 + i)
        inputs.append(x)
        want.append(oracle64.run(kind, x.astype(np.complex128 if is_complex_in(kind) else np.float64), shape))
    errors = [None] * len(cases)
    start = threading.Barrier(len(cases))

    def work(i):
        kind, shape, prec = cases[i]
        try:
            aux = mf.Aux(libs[prec], mf.FAMILY[kind], shape)
            start.wait(timeout=60)
            x = inputs[i]
            odt = (np.complex128 if prec == 64 else np.complex64) if is_complex_out(kind) else \
                (np.float64 if prec == 64 else np.float32)
            xd = torch.from_numpy(np.ascontiguousarray(x)).cuda()
            for r in range(10):
                if r % 2:
                    yd = torch.empty(out_shape(kind, shape), dtype=torch.from_numpy(np.empty(0, odt)).dtype,
                                     device="cuda")
                    aux(kind, xd, yd)
                    y = yd.cpu().numpy()
                else:
                    y = np.empty(out_shape(kind, shape), dtype=odt)
                    aux(kind, x, y)
                bar = tol64(shape) if prec == 64 else TOL32
                e = rel_l2(y, want[i])
                if e > bar:
                    raise AssertionError(f"{kind} {shape} prec={prec} repeat {r}: rel-L2 {e:.3e}")
            aux.close()
        except BaseException as exc:  # reported by the main thread
            errors[i] = exc
            start.abort()  # no thread keeps waiting for this one

    threads = [threading.Thread(target=work, args=(i,)) for i in range(len(cases))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=100)
        assert not t.is_alive(), "a worker thread did not finish"
    for i, e in enumerate(errors):
        assert e is None, f"case {cases[i]}: {e!r}"


@pytest.mark.parametrize("shape", [(256,), (16, 16), (4, 8, 16), (32768,)], ids=lambda s: "x".join(map(str, s)))
def test_back_to_back_synchronous_calls(libs, shape):
    """Every synchronous call returns with its own result in place: 60
    consecutive calls on fresh host arrays (short ones: kernels on the
    pinned buffer, then the polled completion word; 32768 points: DMA
    staging) and on device arrays, each checked against numpy at once."""
    import torch
    print(f"[case] back-to-back calls {shape}", flush=True)
    lib = libs[64]
    aux = mf.Aux(lib, "dft", shape)
    rng = np.random.default_rng(31 + len(shape))
    n = int(np.prod(shape))
    for i in range(60):
        x = (rng.uniform(-1, 1, shape) + 1j * rng.uniform(-1, 1, shape)).astype(np.complex128)
        ref = np.fft.fftn(x)
        if i % 3 == 2:
            xd = torch.from_numpy(x).cuda()
            yd = torch.empty_like(xd)
            aux("dft", xd, yd)
            y = yd.cpu().numpy()
        else:
            y = np.empty_like(x)
            aux("dft", x, y)
        assert rel_l2(y, ref) <= tol64(shape), (i, rel_l2(y, ref))
        back = np.empty_like(x)
        aux("invdft", y, back)
        assert rel_l2(back / n, x) <= tol64(shape), i
    aux.close()


def test_aux_stream_ordered_after_default_stream(libs, oracle64):
    """The aux's own stream (minfft_enqueue with a NULL stream, and the
    synchronous calls on device arrays) is ordered after work queued on the
    legacy default stream: an input written there by a long chain of torch
    kernels just before the call is what the transform reads."""
    import torch
    n = 1 << 20
    for prec in (32, 64):
        cdt = torch.complex64 if prec == 32 else torch.complex128
        aux = mf.Aux(libs[prec], "dft", (n,))  # made first: no maker time between the writes and the call
        print(f"[case] default-stream ordering prec={prec}", flush=True)
        x = torch.zeros(n, dtype=cdt, device="cuda")
        for k in range(64):  # ~64 kernels still queued when the transform is enqueued
            x.add_(1.0 / 64)
        y = torch.empty_like(x)
        aux.enqueue("dft", x, y)  # NULL stream: the aux's own
        z = torch.empty_like(x)
        aux("dft", x, z)          # synchronous call on device arrays
        torch.cuda.synchronize()
        aux.close()
        # x = 1 everywhere: the spectrum is N at bin 0 and 0 elsewhere
        for out in (y, z):
            assert abs(complex(out[0].item()) - n) <= 1e-3 * n
            assert float(out[1:].abs().max()) <= 1e-2


def test_set_stream_runs_calls_on_caller_stream(libs):
    """minfft_set_stream: after it, the aux's synchronous calls and NULL-stream
    enqueues are ordered after work queued on the caller's (non-blocking)
    stream -- an input written there by a long chain of kernels just before
    the call is what the transform reads -- and the synchronous call has
    written its result on return.  NULL restores the aux's own stream; a bad
    aux is refused."""
    import torch
    n = 1 << 20
    s = torch.cuda.Stream()
    for prec in (32, 64):
        cdt = torch.complex64 if prec == 32 else torch.complex128
        aux = mf.Aux(libs[prec], "dft", (n,))
        print(f"[case] set_stream prec={prec}", flush=True)
        aux.set_stream(s.cuda_stream)
        x = torch.empty(n, dtype=cdt, device="cuda")
        y = torch.full_like(x, float("nan"))
        z = torch.full_like(x, float("nan"))
        torch.cuda.synchronize()
        with torch.cuda.stream(s):
            x.zero_()
            for k in range(64):  # ~64 kernels still queued on s when the calls are made
                x.add_(1.0 / 64)
        aux.enqueue("dft", x, y)  # NULL stream: now s
        aux("dft", x, z)          # synchronous, on s
        zc = z.cpu()              # no synchronization of s before this read
        torch.cuda.synchronize()
        for out in (y.cpu(), zc):
            assert abs(complex(out[0].item()) - n) <= 1e-3 * n
            assert float(out[1:].abs().max()) <= 1e-2
        aux.set_stream(None)
        w = torch.empty_like(x)
        aux("dft", x, w)
        assert torch.equal(w, z)
        aux.close()
    rc = mf.engine().minfft_set_stream(None, s.cuda_stream)
    assert rc == -1 and "invalid aux" in mf.last_error()


def test_host_arrays_pipelined_bounce(libs, oracle64, monkeypatch):
    """Host arrays of 4-64 MiB go through the aux's pinned bounce buffer in
    4 MiB pieces (the host copy of a piece overlapping the DMA of the one
    before): a 2048x2048 FP64 real DFT (32 MiB in, 2048x1025 complex out --
    not a whole number of pieces) and a 2^21-point FP64 complex DFT, with the
    pipeline on and off (MINFFT_HOST_PIPE=0), give identical results, which
    match the oracle."""
    rng = np.random.default_rng(5)
    n = 2048
    x = rng.standard_normal((n, n))
    c = rng.uniform(-1, 1, 1 << 21) + 1j * rng.uniform(-1, 1, 1 << 21)
    outs = {}
    for pipe in ("1", "0"):
        monkeypatch.setenv("MINFFT_HOST_PIPE", pipe)
        print(f"[case] host arrays, pipelined bounce={pipe}", flush=True)
        z = np.full((n, n // 2 + 1), np.nan + 0j)
        aux = mf.Aux(libs[64], "realdft", (n, n))
        aux("realdft", x, z)
        back = np.empty_like(x)
        aux("invrealdft", z, back)
        aux.close()
        y = np.full_like(c, np.nan)
        aux = mf.Aux(libs[64], "dft", (1 << 21,))
        aux("dft", c, y)
        aux.close()
        outs[pipe] = (z, back, y)
    for a, b in zip(outs["1"], outs["0"]):
        assert np.array_equal(a, b)
    z, back, y = outs["1"]
    assert rel_l2(z, oracle64.run("realdft", x, (n, n))) <= tol64((n, n))
    assert rel_l2(back / (n * n), x) <= tol64((n, n))
    assert rel_l2(y, oracle64.run("dft", c, (1 << 21,))) <= tol64((1 << 21,))
