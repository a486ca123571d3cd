"""GPU parity of the batch extension (minfft_mkaux_batch / minfft_batch_run):
every entry must equal the single-transform routine applied to it, i.e.
the CPU oracle, to the same bars as the plain API."""
import numpy as np
import pytest

import minfft_amd as mf
from oracle.oracle import in_shape, out_shape, is_complex_in, is_complex_out
from tests.helpers import TOL32, make_input, rel_l2, tol64

pytestmark = pytest.mark.gpu


def _torch():
    import torch
    return torch


def run_batch(entries, prec, xs, oracle64, check_idx=None, seed=0, mode=None):
    """entries: list of (kind, shape); xs: list of numpy inputs.  Packs the
    inputs into one device buffer with gaps, runs the batch, checks entries."""
    torch = _torch()
    rdt = np.float64 if prec == 64 else np.float32
    cdt = np.complex128 if prec == 64 else np.complex64
    rs = np.dtype(rdt).itemsize
    ioffs, ooffs = [], []
    ipos = opos = 0
    for (kind, shape), x in zip(entries, xs):
        ipos = (ipos + 15) & ~15  # entries start on 16-byte boundaries (any item size)
        opos = (opos + 15) & ~15
        ioffs.append(ipos)
        ipos += x.nbytes + 64 * rs  # gap between entries
        n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
        ooffs.append(opos)
        opos += n_out * rs + 32 * rs
    hin = np.zeros(ipos, dtype=np.uint8)
    for x, o in zip(xs, ioffs):
        hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    dout = torch.zeros(opos, dtype=torch.uint8, device="cuda")
    if mode is None:
        mode = mf.PREC_FP64 if prec == 64 else mf.PREC_FP32
    b = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)], mode)
    st = torch.cuda.Stream()
    b.run(din.data_ptr(), dout.data_ptr(), st.cuda_stream)
    torch.cuda.synchronize()
    b.run(din.data_ptr(), dout.data_ptr(), st.cuda_stream)  # graph replay
    torch.cuda.synchronize()
    hout = dout.cpu().numpy()
    b.close()
    idx = range(len(entries)) if check_idx is None else check_idx
    worst = 0.0
    for i in idx:
        kind, shape = entries[i]
        odt = cdt if is_complex_out(kind) else rdt
        n_out = int(np.prod(out_shape(kind, shape)))
        y = hout[ooffs[i]:ooffs[i] + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(kind, shape))
        x = xs[i]
        ref = oracle64.run(kind, x.astype(np.complex128 if is_complex_in(kind) else np.float64), shape)
        e = rel_l2(y, ref)
        bar = tol64(shape) if prec == 64 else TOL32
        assert e <= bar, (i, kind, shape, prec, e)
        worst = max(worst, e)
    # the gaps between outputs stay untouched
    mask = np.ones(opos, dtype=bool)
    for (kind, shape), oo in zip(entries, ooffs):
        n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
        mask[oo:oo + n_out * rs] = False
    assert not hout[mask].any()
    return worst


@pytest.mark.parametrize("prec", [32, 64])
def test_uneven_1d_c2c_batch(oracle64, prec):
    """Mixed lengths 2^0..2^17, forward and inverse, small counts (tile and
    chunked two-step paths)."""
    rng = np.random.default_rng(5)
    entries, xs = [], []
    for k in list(range(0, 18)) * 2:
        kind = "dft" if rng.random() < 0.6 else "invdft"
        entries.append((kind, (1 << k,)))
        xs.append(make_input(kind, (1 << k,), prec, int(rng.integers(1 << 30))))
    print(f"[case] uneven 1D batch prec={prec} entries={len(entries)}", flush=True)
    run_batch(entries, prec, xs, oracle64)


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("logn", [14, 15, 16])
def test_fused_long_batch(oracle64, prec, logn):
    """>= 256 entries of one long length: the fused two-step kernel (one
    work-group per transform, private scratch slot)."""
    n = 1 << logn
    cnt = 300
    rng = np.random.default_rng(logn)
    entries, xs = [], []
    for i in range(cnt):
        kind = "dft" if i % 3 else "invdft"
        entries.append((kind, (n,)))
        xs.append(make_input(kind, (n,), prec, int(rng.integers(1 << 30))))
    print(f"[case] fused batch prec={prec} N=2^{logn} count={cnt}", flush=True)
    # every entry computed; check a spread including first/last of each
    # work-group round (static striding over 2*256 work-groups)
    idx = sorted(set([0, 1, 2, 255, 256, 257, cnt - 2, cnt - 1] + list(range(0, cnt, 37))))
    run_batch(entries, prec, xs, oracle64, check_idx=idx)


@pytest.mark.parametrize("prec", [32, 64])
def test_mixed_real_kinds_2d_batch(oracle64, prec):
    rng = np.random.default_rng(11)
    kinds = ["dct2", "dst2", "dct4", "dst4", "dct3", "dst3", "realdft", "invrealdft"]
    shapes = [(64, 64), (16, 32), (8, 8, 4), (128, 2), (1, 16), (32,)]
    entries, xs = [], []
    for i in range(40):
        kind = kinds[i % len(kinds)]
        shape = shapes[(i // len(kinds) + i) % len(shapes)]
        if kind == "invrealdft":
            xr = make_input("dct2", shape, prec, 1000 + i)
            x = oracle64.run("realdft", xr.astype(np.float64), shape).astype(
                np.complex128 if prec == 64 else np.complex64)
        else:
            x = make_input(kind, shape, prec, 1000 + i)
        entries.append((kind, shape))
        xs.append(x)
    print(f"[case] mixed real kinds prec={prec}", flush=True)
    run_batch(entries, prec, xs, oracle64)


@pytest.mark.parametrize("prec", [32, 64])
def test_real_kinds_odd_offsets(oracle64, prec):
    """Real rows that start at an odd real index (not 16-byte aligned): the
    staged loads and stores take the scalar path, and the real DFT's input
    pairs straddle complex boundaries."""
    torch = _torch()
    rdt = np.float32 if prec == 32 else np.float64
    cdt = np.complex64 if prec == 32 else np.complex128
    rs = np.dtype(rdt).itemsize
    kinds = ["realdft", "dct2", "dst3", "dct4", "dst2", "realdft"]
    shapes = [(64,), (8, 32), (256,), (4, 64), (1024,), (2, 16)]
    entries, xs, ioffs, ooffs = [], [], [], []
    ipos, opos = 3 * rs, 0
    for i, (kind, shape) in enumerate(zip(kinds, shapes)):
        x = make_input(kind, shape, prec, 500 + i)
        entries.append((kind, shape))
        xs.append(x)
        ioffs.append(ipos)
        ipos += x.nbytes + 5 * rs  # odd real count between entries
        n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
        if is_complex_out(kind):
            opos = -(-opos // (2 * rs)) * (2 * rs)  # complex items: complex-aligned
        elif (opos // rs) % 2 == 0:
            opos += rs  # real rows: start at an odd real index
        ooffs.append(opos)
        opos += n_out * rs + rs
    hin = np.zeros(ipos, dtype=np.uint8)
    for x, o in zip(xs, ioffs):
        hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    dout = torch.zeros(opos, dtype=torch.uint8, device="cuda")
    b = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)],
                 mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
    print(f"[case] odd offsets prec={prec}", flush=True)
    b.run(din.data_ptr(), dout.data_ptr(), None)
    torch.cuda.synchronize()
    hout = dout.cpu().numpy()
    b.close()
    for (kind, shape), x, oo in zip(entries, xs, ooffs):
        odt = cdt if is_complex_out(kind) else rdt
        n_out = int(np.prod(out_shape(kind, shape)))
        y = hout[oo:oo + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(kind, shape))
        ref = oracle64.run(kind, x.astype(np.float64), shape)
        bar = tol64(shape) if prec == 64 else TOL32
        assert rel_l2(y, ref) <= bar, (kind, shape, prec)


def test_batch_host_pointers(oracle64):
    """minfft_batch: host buffers, synchronous."""
    sizes = [64, 1024, 4096, 256]
    esz = 8
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    x = make_input("dft", (int(sum(sizes)),), 32, 3)
    y = np.zeros_like(x)
    b = mf.Batch([("dft", (n,), int(o) * esz, int(o) * esz) for n, o in zip(sizes, offs)], mf.PREC_FP32)
    b.run_sync(x, y)
    # again, into a larger output whose tail no entry covers (kept as is):
    # the plan's staging buffers grow and are reused
    y2 = np.full(len(y) + 100, 7 + 7j, dtype=y.dtype)
    b.run_sync(x, y2)
    b.run_sync(x, y2)
    b.close()
    for n, o in zip(sizes, offs):
        ref = oracle64.run("dft", x[o:o + n].astype(np.complex128), (n,))
        assert rel_l2(y[o:o + n], ref) <= TOL32
        assert rel_l2(y2[o:o + n], ref) <= TOL32
    assert np.all(y2[len(y):] == 7 + 7j)


def test_batch_rejects_bad_entries():
    with pytest.raises(mf.MinfftError):
        mf.Batch([("dft", (12,), 0, 0)], mf.PREC_FP32)
    with pytest.raises(mf.MinfftError):
        mf.Batch([("dft", (16,), 4, 0)], mf.PREC_FP32)  # misaligned offset
    with pytest.raises(mf.MinfftError):
        mf.Batch([], mf.PREC_FP32)


def test_split_fp16_mfma_batch(oracle64):
    """MINFFT_SPLIT_FP16: 512-point axes take the split-FP16 MFMA radix-16
    tile (contiguous and strided axes, forward and inverse); other lengths
    the FP32 path.  Bar: rel-L2 <= 1e-6 vs the FP64 oracle."""
    shapes = [(512,), (4, 512), (512, 4), (2, 512, 8), (512, 8, 2), (16, 512), (1024,), (512, 64)]
    rng = np.random.default_rng(21)
    entries, xs = [], []
    for i in range(24):
        shape = shapes[i % len(shapes)]
        kind = "dft" if i % 2 == 0 else "invdft"
        x = make_input(kind, shape, 32, int(rng.integers(1 << 30)))
        if i % 3 == 1:
            x = (x * np.float32(3e4)).astype(np.complex64)  # large values: per-column scaling
        entries.append((kind, shape))
        xs.append(x)
    print("[case] split-FP16 batch", flush=True)
    worst = run_batch(entries, 32, xs, oracle64, mode=mf.PREC_SPLIT_FP16)
    print(f"split-FP16 worst rel-L2 {worst:.2e}")


def test_split_fp16_split_column_tile(oracle64):
    """The matrix-core form of the radix-2-split column tile (32 transforms
    per work-group) on 512-point columns, pinned on (csplit=1), off and by
    default: 2D and 3D entries, a partial last tile (3 x 16 columns), large
    values (per-column scaling), forward and inverse, every entry against
    the FP64 oracle at rel-L2 <= 1e-6; the three forms agree to 1e-6."""
    torch = _torch()
    shapes = [(512, 64), (512, 16), (512, 8, 16), (512, 32)]
    rng = np.random.default_rng(33)
    entries, xs = [], []
    for i in range(10):
        shape = shapes[i % len(shapes)]
        kind = "dft" if i % 2 == 0 else "invdft"
        x = make_input(kind, shape, 32, int(rng.integers(1 << 30)))
        if i % 3 == 1:
            x = (x * np.float32(3e4)).astype(np.complex64)
        entries.append((kind, shape))
        xs.append(x)
    ioffs, ooffs, pos = [], [], 0
    for x in xs:
        ioffs.append(pos)
        ooffs.append(pos)
        pos += x.nbytes + 256
    hin = np.zeros(pos, dtype=np.uint8)
    for x, o in zip(xs, ioffs):
        hin[o:o + x.nbytes] = np.frombuffer(x.tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    b = mf.Batch([(k, sh, io, oo) for (k, sh), io, oo in zip(entries, ioffs, ooffs)], mf.PREC_SPLIT_FP16)
    ngroups = b.describe().count("\n")
    refs = [oracle64.run(k, x.astype(np.complex128), sh) for (k, sh), x in zip(entries, xs)]
    outs = {}
    for spec in ("csplit=1", "csplit=0", ""):
        for gi in range(ngroups):
            b.schedule(gi, spec)
        dout = torch.zeros(pos, dtype=torch.uint8, device="cuda")
        print(f"[case] split-FP16 column tile schedule '{spec}'", flush=True)
        b.run(din.data_ptr(), dout.data_ptr(), None)
        torch.cuda.synchronize()
        h = dout.cpu().numpy()
        got = [h[o:o + x.nbytes].view(np.complex64).reshape(sh) for (k, sh), x, o in zip(entries, xs, ooffs)]
        for i, (y, r) in enumerate(zip(got, refs)):
            assert rel_l2(y, r) <= TOL32, (spec, i, entries[i], rel_l2(y, r))
        outs[spec] = got
    b.close()
    for i in range(len(entries)):
        assert rel_l2(outs["csplit=1"][i], outs["csplit=0"][i]) <= 1e-6, i


@pytest.mark.parametrize("mode", ["fp32", "split-fp16", "fp64"])
def test_split_column_tiles_in_place(oracle64, mode):
    """The radix-2-split column tiles (pinned on: csplit=1) run in place give
    the same bits as out of place -- every work-group reads all of its
    columns before it writes -- and match the oracle."""
    torch = _torch()
    prec = 64 if mode == "fp64" else 32
    pm = {"fp32": mf.PREC_FP32, "split-fp16": mf.PREC_SPLIT_FP16, "fp64": mf.PREC_FP64}[mode]
    isz = 16 if prec == 64 else 8
    shapes = [("dft", (512, 64)), ("invdft", (512, 16)), ("dft", (1024, 32)), ("invdft", (512, 8, 16))] * 2
    entries, off = [], 0
    for kind, sh in shapes:
        entries.append((kind, sh, off, off))
        off += int(np.prod(sh)) * isz
    rdt = torch.float64 if prec == 64 else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(9)
    x = torch.complex(torch.rand(off // isz, generator=g, device="cuda", dtype=rdt) * 2 - 1,
                      torch.rand(off // isz, generator=g, device="cuda", dtype=rdt) * 2 - 1)
    b = mf.Batch(entries, pm)
    for gi in range(b.describe().count("\n")):
        b.schedule(gi, "csplit=1")
    print(f"[case] split column tiles in place, {mode}", flush=True)
    y = torch.zeros_like(x)
    b.run(x.data_ptr(), y.data_ptr(), None)
    z = x.clone()
    b.run(z.data_ptr(), z.data_ptr(), None)
    torch.cuda.synchronize()
    b.close()
    assert torch.equal(z, y)
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for i in (0, 1, 2, 3):
        kind, sh, o, _ = entries[i]
        n = int(np.prod(sh))
        ref = oracle64.run(kind, xh[o // isz:o // isz + n].astype(np.complex128).reshape(sh), sh)
        e = rel_l2(yh[o // isz:o // isz + n].reshape(sh), ref)
        assert e <= (tol64(sh) if prec == 64 else TOL32), (mode, i, kind, sh, e)


@pytest.mark.parametrize("prec", [32, 64])
def test_dct_columns_of_1024_points(oracle64, prec):
    """FP32 and FP64 DCT/DST columns of 1024 points (the height of the DCT image
    batch), every kind, match the oracle: partial last tiles (16 columns;
    48 through an offset table), a 3D array, two-column arrays, offset
    tables with even and with odd item offsets; in-place runs give the bits
    of out-of-place ones."""
    torch = _torch()
    cases = [("dct2", (1024, 64)), ("dst2", (1024, 16)), ("dct2", (1024, 4, 8)), ("dst2", (1024, 2)),
             ("dct3", (1024, 32)), ("dst3", (1024, 2)), ("dct4", (1024, 16)), ("dst4", (1024, 4, 8))]
    entries, off = [], 0
    isz = prec // 8
    for kind, sh in cases:
        entries.append((kind, sh, off * isz, off * isz))
        off += int(np.prod(sh))
    # a group of three entries at even item offsets, one of two at odd ones
    for kind, sh, o0, cnt in (("dct2", (1024, 16), 0, 3), ("dst2", (1024, 32), 1, 2)):
        n = int(np.prod(sh))
        for _ in range(cnt):
            start = off + o0
            entries.append((kind, sh, start * isz, start * isz))
            off = start + n + 2
    g = torch.Generator(device="cuda")
    g.manual_seed(21)
    x = torch.rand(off, generator=g, device="cuda", dtype=torch.float64 if prec == 64 else torch.float32) * 2 - 1
    b = mf.Batch(entries, mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
    print(f"[case] dct/dst columns of 1024 points, FP{prec}", flush=True)
    y = torch.zeros_like(x)
    b.run(x.data_ptr(), y.data_ptr(), None)
    z = x.clone()
    b.run(z.data_ptr(), z.data_ptr(), None)
    torch.cuda.synchronize()
    b.close()
    xh, yh, zh = x.cpu().numpy(), y.cpu().numpy(), z.cpu().numpy()
    for i, (kind, sh, o, _) in enumerate(entries):
        o //= isz
        n = int(np.prod(sh))
        assert np.array_equal(zh[o:o + n], yh[o:o + n]), (i, kind, sh)
        ref = oracle64.run(kind, xh[o:o + n].astype(np.float64).reshape(sh), sh)
        e = rel_l2(yh[o:o + n].reshape(sh), ref)
        assert e <= (tol64(sh) if prec == 64 else TOL32), (prec, i, kind, sh, e)


@pytest.mark.parametrize("prec", [32, 64])
def test_persistent_short_lengths(oracle64, prec, monkeypatch):
    """Short 1D complex groups (2^4 .. 2^12, both directions) run in one
    persistent launch per direction (multi.hpp): entries at offsets one
    complex item off the 16-byte grid, counts that leave the last item of a
    group partial; the results of the persistent launch and of one launch
    per group both agree with the oracle (they need not agree bit for bit:
    the compiler may contract multiply-adds differently in the two
    kernels)."""
    torch = _torch()
    rng = np.random.default_rng(40 + prec)
    cdt = np.complex128 if prec == 64 else np.complex64
    cs = np.dtype(cdt).itemsize
    entries, xs = [], []
    for lg in range(4, 13):
        for kind in ("dft", "invdft"):
            for _ in range(int(rng.integers(1, 90))):
                n = 1 << lg
                entries.append((kind, (n,)))
                xs.append((rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(cdt))
    order = rng.permutation(len(entries))
    entries = [entries[i] for i in order]
    xs = [xs[i] for i in order]
    offs, pos = [], cs
    for x in xs:
        offs.append(pos)
        pos += x.nbytes + cs
    hin = np.zeros(pos + cs, dtype=np.uint8)
    for x, o in zip(xs, offs):
        hin[o:o + x.nbytes] = np.frombuffer(x.tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    mode = mf.PREC_FP64 if prec == 64 else mf.PREC_FP32
    outs = {}
    for flag in ("1", "2", "0"):
        monkeypatch.setenv("MINFFT_MULTI", flag)
        b = mf.Batch([(k, s, o, o) for (k, s), o in zip(entries, offs)], mode)
        desc = b.describe()
        assert ("persistent" in desc) == (flag != "0"), desc
        dout = torch.zeros(pos + cs, dtype=torch.uint8, device="cuda")
        print(f"[case] persistent short lengths MINFFT_MULTI={flag} prec={prec}", flush=True)
        for _ in range(2):  # capture, then replay
            b.run(din.data_ptr(), dout.data_ptr(), None)
            torch.cuda.synchronize()
        outs[flag] = dout.cpu().numpy()
        b.close()
    for i in range(0, len(entries), 3):
        kind, shape = entries[i]
        ref = oracle64.run(kind, xs[i].astype(np.complex128), shape)
        bar = tol64(shape) if prec == 64 else TOL32
        for f in ("1", "2", "0"):
            y = outs[f][offs[i]:offs[i] + xs[i].nbytes].view(cdt)
            assert rel_l2(y, ref) <= bar, (f, i, kind, shape)


def test_split_fp16_keeps_512_on_matrix_cores():
    """In a split-FP16 plan the 512-point 1D groups keep their launch (the
    matrix-core tile); other short lengths join the persistent launch."""
    b = mf.Batch([("dft", (512,), 0, 0), ("dft", (256,), 8192, 8192)], mf.PREC_SPLIT_FP16)
    lines = b.describe().strip().split("\n")
    b.close()
    l512 = [ln for ln in lines if "dft 512 " in ln]
    l256 = [ln for ln in lines if "dft 256 " in ln]
    assert l512 and "persistent" not in l512[0], lines
    assert l256 and "persistent" in l256[0], lines


def _warm_device():
    """Load the libraries' kernels and the runtime's own buffers once, so
    that device-memory readings before and after a test compare plans only."""
    torch = _torch()
    x = torch.zeros(1 << 16, dtype=torch.float64, device="cuda")
    y = torch.empty_like(x)
    for prec in (mf.PREC_FP64, mf.PREC_FP32):
        b = mf.Batch([("invrealdft", (64, 64), 0, 0), ("dft", (1 << 15,), 0, 0)], prec)
        b.run(x.data_ptr(), y.data_ptr(), None)
        torch.cuda.synchronize()
        b.close()
    torch.cuda.synchronize()


def test_batch_allocation_failure_returns_null_without_leak():
    """A plan whose scratch exceeds the device memory (3000 inverse real DFTs
    of 4096^2 FP64 need ~400 GB of complex staging) is refused: the maker
    returns NULL with a reason, and frees what it had allocated."""
    import torch
    _warm_device()
    free0, total = torch.cuda.mem_get_info()
    need = 3000 * 4096 * 2049 * 16
    assert need > total
    with pytest.raises(mf.MinfftError, match="hipMalloc"):
        mf.Batch([("invrealdft", (4096, 4096), 0, 0)] * 3000, mf.PREC_FP64)
    torch.cuda.synchronize()
    free1, _ = torch.cuda.mem_get_info()
    assert free1 >= free0 - (64 << 20), (free0, free1)


@pytest.mark.parametrize("prec,logn", [(mf.PREC_FP64, 27), (mf.PREC_FP32, 29)])
def test_batch_refuses_lengths_beyond_the_two_step_limit(prec, logn, monkeypatch):
    """Batch entries longer than the plain makers accept (1D FP64 2^26, FP32
    2^28; real kinds twice that) are refused when the plan is made, with the
    plain makers' reason, before anything is allocated or tuned."""
    _warm_device()
    for kind, n in [("dft", 1 << logn), ("invdft", 1 << logn), ("realdft", 1 << (logn + 1)),
                    ("dct2", 1 << (logn + 1))]:
        print(f"[case] batch refuses {kind} 2^{n.bit_length() - 1} prec={prec}", flush=True)
        with pytest.raises(mf.MinfftError, match="two-step limit"):
            mf.Batch([("dft", (64,), 0, 0), (kind, (n,), 0, 0)], prec)
    # the longest accepted complex length still makes a plan (untuned: no
    # data is touched; the real kinds' longest plans would need twiddle
    # tables of 8N entries, ~200 GB for FP32, so they are not made here)
    monkeypatch.setenv("MINFFT_AUTOTUNE", "0")
    b = mf.Batch([("dft", (1 << (logn - 1),), 0, 0)], prec)
    b.close()


def test_batch_inside_a_callers_graph_capture(oracle64):
    """A plan with several units of work run on a stream that the caller is
    capturing (torch.cuda.graph) puts its launches, branches included, into
    the caller's graph; every replay, with the plan's own eager runs in
    between, equals the eager result bit for bit, and sampled entries match
    the oracle."""
    torch = _torch()
    spec = [("dft", 65536)] * 3 + [("dft", 16384)] * 4 + [("dft", 1024)] * 20 + [("dft", 64)] * 30 + \
        [("invdft", 256)] * 10
    entries, off = [], 0
    for kind, n in spec:
        entries.append((kind, (n,), off, off))
        off += n * 8
    npts = off // 8
    g0 = torch.Generator(device="cuda")
    g0.manual_seed(11)
    x = torch.complex(torch.rand(npts, generator=g0, device="cuda") * 2 - 1,
                      torch.rand(npts, generator=g0, device="cuda") * 2 - 1)
    b = mf.Batch(entries, mf.PREC_FP32)
    desc = b.describe()
    own = [ln for ln in desc.splitlines() if "persistent" not in ln]
    assert "persistent" in desc and len(own) == 2, desc  # four units: a graph of its own
    s = torch.cuda.Stream()
    want = torch.zeros_like(x)
    b.run(x.data_ptr(), want.data_ptr(), s.cuda_stream)
    y = torch.zeros_like(x)
    b.run(x.data_ptr(), y.data_ptr(), s.cuda_stream)  # the pointers checked, the branch streams made
    torch.cuda.synchronize()
    print("[case] batch captured into a caller's graph", flush=True)
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph, stream=s):
        b.run(x.data_ptr(), y.data_ptr(), s.cuda_stream)
    torch.cuda.synchronize()
    for r in range(4):
        y.zero_()
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        assert torch.equal(y, want), r
        z = torch.zeros_like(x)
        b.run(x.data_ptr(), z.data_ptr(), None)  # the plan's own graph, eagerly
        torch.cuda.synchronize()
        assert torch.equal(z, want), r
    b.close()
    xh, yh = x.cpu().numpy(), want.cpu().numpy()
    for i in (0, 3, 7, 27, 57, len(spec) - 1):
        kind, (n,), o, _ = entries[i]
        ref = oracle64.run(kind, xh[o // 8:o // 8 + n].astype(np.complex128), (n,))
        assert rel_l2(yh[o // 8:o // 8 + n], ref) <= TOL32, (i, kind, n)


@pytest.mark.parametrize("prec", [32, 64])
def test_uneven_batch_in_place_equals_out_of_place(oracle64, prec):
    """An uneven 1D C2C batch like the benchmark's (2^6..2^16, forward and
    inverse, every kernel of the batch: the short-length launch, tiles,
    radix-2-split and fused two-step kernels) run in place gives the same
    bits as out of place, and sampled entries match the oracle."""
    torch = _torch()
    rng = np.random.default_rng(77)
    lens = [1 << int(6 + 11 * rng.random()) for _ in range(420)] + [1 << 16] * 270
    kinds = ["dft" if rng.random() < 0.7 else "invdft" for _ in lens]
    entries, off = [], 0
    for k, n in zip(kinds, lens):
        entries.append((k, (n,), off, off))
        off += n * (16 if prec == 64 else 8)
    npts = sum(lens)
    rdt = torch.float64 if prec == 64 else torch.float32
    g = torch.Generator(device="cuda")
    g.manual_seed(prec)
    x = torch.complex(torch.rand(npts, generator=g, device="cuda", dtype=rdt) * 2 - 1,
                      torch.rand(npts, generator=g, device="cuda", dtype=rdt) * 2 - 1)
    b = mf.Batch(entries, mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
    print(f"[case] in-place uneven batch prec={prec}: {len(entries)} entries\n{b.describe()}", flush=True)
    y = torch.zeros_like(x)
    b.run(x.data_ptr(), y.data_ptr(), None)
    z = x.clone()
    b.run(z.data_ptr(), z.data_ptr(), None)
    torch.cuda.synchronize()
    b.close()
    assert torch.equal(z, y)
    isz = 16 if prec == 64 else 8
    xh, yh = x.cpu().numpy(), y.cpu().numpy()
    for i in (0, 1, 2, 100, 419, 420, len(entries) - 1):
        kind, (n,), o, _ = entries[i]
        ref = oracle64.run(kind, xh[o // isz:o // isz + n].astype(np.complex128), (n,))
        e = rel_l2(yh[o // isz:o // isz + n], ref)
        assert e <= (tol64((n,)) if prec == 64 else TOL32), (i, kind, n, e)


@pytest.mark.parametrize("single_unit", [False, True], ids=["several-units", "one-unit"])
def test_graph_replay_and_direct_launch_agree(oracle64, monkeypatch, single_unit):
    """The same plan run as a replayed graph (MINFFT_GRAPH=1), as direct
    launches (0) and by the default rule gives identical outputs, replay
    after replay with other device work in between, and they match the
    oracle."""
    torch = _torch()
    prec = 32
    if single_unit:
        entries = [("dft", (1024,))] * 24
    else:
        entries = [("dft", (n,)) for n in (64, 512, 4096, 16384, 65536)] * 3 + \
            [("realdft", (64, 32)), ("dct2", (32, 64)), ("invdft", (256,))]
    xs = [make_input(k, s, prec, 900 + i) for i, (k, s) in enumerate(entries)]
    ioffs, ooffs, ipos, opos = [], [], 0, 0
    for (k, s), x in zip(entries, xs):
        ioffs.append(ipos)
        ipos += x.nbytes
        ooffs.append(opos)
        opos += int(np.prod(out_shape(k, s))) * (8 if is_complex_out(k) else 4)
    hin = np.zeros(ipos, dtype=np.uint8)
    for x, o in zip(xs, ioffs):
        hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    b = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)], mf.PREC_FP32)
    st = torch.cuda.Stream()
    outs = {}
    for mode in ("1", "0", None):
        if mode is None:
            monkeypatch.delenv("MINFFT_GRAPH", raising=False)
        else:
            monkeypatch.setenv("MINFFT_GRAPH", mode)
        dout = torch.zeros(opos, dtype=torch.uint8, device="cuda")
        got = []
        for rep in range(3):
            dout.zero_()
            b.run(din.data_ptr(), dout.data_ptr(), st.cuda_stream)
            torch.cuda.synchronize()
            got.append(dout.cpu().numpy().copy())
            torch.randn(1 << 20, device="cuda").sum().item()  # other device work between replays
        for g in got[1:]:
            assert np.array_equal(g, got[0]), f"mode {mode}: replays differ"
        outs[mode] = got[0]
    b.close()
    assert np.array_equal(outs["1"], outs["0"]) and np.array_equal(outs["1"], outs[None])
    hout = outs["1"]
    for i, ((k, s), x) in enumerate(zip(entries, xs)):
        odt = np.complex64 if is_complex_out(k) else np.float32
        n_out = int(np.prod(out_shape(k, s)))
        y = hout[ooffs[i]:ooffs[i] + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(k, s))
        ref = oracle64.run(k, x.astype(np.complex128 if is_complex_in(k) else np.float64), s)
        assert rel_l2(y, ref) <= TOL32, (i, k, s)


ALL_KINDS = ["dft", "invdft", "realdft", "invrealdft", "dct2", "dst2", "dct3", "dst3", "dct4", "dst4"]


@pytest.mark.parametrize("prec", [32, 64])
@pytest.mark.parametrize("seed", [1, 2, 3])
def test_random_mixed_batches(oracle64, prec, seed):
    """Seeded random batches: every kind, 1 to 3 dimensions, 2^0..2^11
    points per axis (at most 2^16 per entry), shapes repeated 1 to 12 times
    (groups of one or several entries: persistent launch, small-pass tiles,
    packed DC/Nyquist columns, graph or direct launch), every entry against
    the oracle and the gaps between outputs untouched."""
    rng = np.random.default_rng(1000 + seed)
    entries, xs = [], []
    while len(entries) < 60:
        kind = ALL_KINDS[int(rng.integers(len(ALL_KINDS)))]
        d = int(rng.integers(1, 4))
        while True:
            shape = tuple(1 << int(rng.integers(0, 12)) for _ in range(d))
            if int(np.prod(shape)) <= 1 << 16:
                break
        for _ in range(int(rng.choice([1, 1, 2, 5, 12]))):
            entries.append((kind, shape))
            xs.append(make_input(kind, shape, prec, int(rng.integers(1 << 30)), oracle64))
    print(f"[case] random batch seed={seed} prec={prec}: {len(entries)} entries", flush=True)
    run_batch(entries, prec, xs, oracle64)


def test_allocation_failure_hook(oracle64, monkeypatch):
    """With MINFFT_FAIL_ALLOC=k the k-th device allocation and every later
    one fail as out of memory.  For each k, the plain maker (twiddles,
    two-step scratch, real-DFT temporary, tuning buffers) and the batch
    maker (offset tables, scratch, temporaries, counter, tuning buffers)
    either return NULL with the reason or succeed; a failing transform
    reports its error; nothing is left on the device, and plans made
    afterwards compute correctly."""
    torch = _torch()
    lib = mf.Minfft(64)
    _warm_device()
    lib.free_aux(lib.mkaux_raw("minfft_mkaux_realdft_2d", 64, 64))
    free0 = torch.cuda.mem_get_info()[0]
    n16 = 1 << 16
    bent = [("dft", (n16,), i * n16 * 8, i * n16 * 8) for i in range(260)]
    bent += [("invrealdft", (64, 128), 0, 0), ("dft", (256,), 0, 0), ("dft", (4096,), 0, 0)]

    def make_plain():
        return lib.mkaux_raw("minfft_mkaux_realdft_2d", 1 << 15, 64)

    def make_batch():
        try:
            return mf.Batch(bent, mf.PREC_FP32)
        except mf.MinfftError:
            return None

    for name, make, free in (("plain", make_plain, lib.free_aux), ("batch", make_batch, lambda b: b.close())):
        outcomes = []
        for k in range(1, 41):
            monkeypatch.setenv("MINFFT_FAIL_ALLOC", str(k))
            h = make()
            if h:
                free(h)
            else:
                err = mf.last_error()
                assert "hipMalloc" in err or "memory" in err.lower(), (name, k, err)
            outcomes.append(bool(h))
        monkeypatch.delenv("MINFFT_FAIL_ALLOC")
        print(f"[case] {name}: succeeds from k = {outcomes.index(True) + 1 if True in outcomes else None}",
              flush=True)
        assert not outcomes[0] and outcomes[-1], (name, outcomes)
        first = outcomes.index(True)
        assert all(outcomes[first:]), (name, outcomes)

    # a transform whose host-array staging cannot be allocated reports it
    shape = (64, 64)
    x = make_input("dft", shape, 64, 17)
    aux = mf.Aux(lib, "dft", shape)
    y = np.empty_like(x)
    monkeypatch.setenv("MINFFT_FAIL_ALLOC", "1")
    with pytest.raises(mf.MinfftError):
        aux("dft", x, y)
    monkeypatch.delenv("MINFFT_FAIL_ALLOC")
    aux("dft", x, y)
    assert rel_l2(y, oracle64.run("dft", x, shape)) <= tol64(shape)
    aux.close()

    torch.cuda.synchronize()
    free1 = torch.cuda.mem_get_info()[0]
    assert free1 >= free0 - (64 << 20), (free0, free1)


def test_distinct_plans_run_concurrently(oracle64):
    """Two batch plans (each with its own scratch, counters and graph) run
    at the same time on two streams, five times each without waiting in
    between, give the same outputs as one run of each alone, and those match
    the oracle."""
    torch = _torch()
    rng = np.random.default_rng(44)
    plans = []
    for p in range(2):
        entries = [("dft", (1 << 15,))] * 8 + [("dft", (1 << 9,))] * 64 + [("dct2", (64, 64))] * 4 + \
            [("invdft", (1 << (12 + p),))] * 16
        xs = [make_input(k, s, 32, int(rng.integers(1 << 30))) for k, s in entries]
        ioffs, ooffs, ipos, opos = [], [], 0, 0
        for (k, s), x in zip(entries, xs):
            ioffs.append(ipos)
            ipos += x.nbytes
            ooffs.append(opos)
            opos += int(np.prod(out_shape(k, s))) * (8 if is_complex_out(k) else 4)
        hin = np.zeros(ipos, dtype=np.uint8)
        for x, o in zip(xs, ioffs):
            hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
        b = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)], mf.PREC_FP32)
        plans.append(dict(entries=entries, xs=xs, ooffs=ooffs, b=b, din=torch.from_numpy(hin).cuda(),
                          dout=torch.zeros(opos, dtype=torch.uint8, device="cuda"), st=torch.cuda.Stream()))
    alone = []
    for p in plans:
        p["b"].run(p["din"].data_ptr(), p["dout"].data_ptr(), p["st"].cuda_stream)
        torch.cuda.synchronize()
        alone.append(p["dout"].cpu().numpy().copy())
        p["dout"].zero_()
    torch.cuda.synchronize()
    for _ in range(5):
        for p in plans:
            p["b"].run(p["din"].data_ptr(), p["dout"].data_ptr(), p["st"].cuda_stream)
    torch.cuda.synchronize()
    for p, a in zip(plans, alone):
        assert np.array_equal(p["dout"].cpu().numpy(), a)
        for i, ((k, s), x) in enumerate(zip(p["entries"], p["xs"])):
            if i % 7:
                continue
            odt = np.complex64 if is_complex_out(k) else np.float32
            n_out = int(np.prod(out_shape(k, s)))
            y = a[p["ooffs"][i]:p["ooffs"][i] + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(k, s))
            ref = oracle64.run(k, x.astype(np.complex128 if is_complex_in(k) else np.float64), s)
            assert rel_l2(y, ref) <= TOL32, (i, k, s)
        p["b"].close()


def test_batch_refuses_short_buffers():
    """minfft_batch checks the buffers against the extents its entries
    address (largest offset plus the entry's items) and refuses a short one
    before any copy or launch: the output stays as it was."""
    sizes = [64, 1024, 256]
    offs = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    total = int(sum(sizes))
    b = mf.Batch([("dft", (n,), int(o) * 8, int(o) * 8) for n, o in zip(sizes, offs)], mf.PREC_FP32)
    x = make_input("dft", (total,), 32, 9)
    y = np.full(total, 3 + 4j, dtype=np.complex64)
    with pytest.raises(mf.MinfftError, match="smaller than the entries"):
        b.run_sync(x[: total - 1], y)
    with pytest.raises(mf.MinfftError, match="smaller than the entries"):
        b.run_sync(x, y[: total - 1])
    assert np.all(y == 3 + 4j)
    # device buffers: the same check
    torch = _torch()
    xd = torch.from_numpy(x).cuda()
    yd = torch.full((total,), 3 + 4j, dtype=torch.complex64, device="cuda")
    with pytest.raises(mf.MinfftError, match="smaller than the entries"):
        b.run_sync(xd, yd[: total - 16])
    torch.cuda.synchronize()
    assert bool(torch.all(yd == 3 + 4j))
    b.run_sync(x, y)  # the exact sizes are accepted
    b.close()
    assert not np.any(y == 3 + 4j)


def test_async_entry_points_refuse_host_pointers():
    """minfft_enqueue and minfft_batch_run take device pointers only: host
    (pageable) arrays are refused with a reason before any launch, and the
    output array stays as it was."""
    lib = mf.Minfft(32)
    aux = mf.Aux(lib, "dft", (256,))
    x = make_input("dft", (256,), 32, 5)
    y = np.full(256, 1 + 1j, dtype=np.complex64)
    with pytest.raises(mf.MinfftError, match="device"):
        aux.enqueue("dft", x, y)
    aux.close()
    assert np.all(y == 1 + 1j)
    b = mf.Batch([("dft", (256,), 0, 0)], mf.PREC_FP32)
    with pytest.raises(mf.MinfftError, match="device"):
        b.run(x.ctypes.data, y.ctypes.data)
    assert np.all(y == 1 + 1j)
    # one device and one host pointer: refused as well
    torch = _torch()
    xd = torch.from_numpy(x).cuda()
    with pytest.raises(mf.MinfftError, match="device"):
        b.run(xd.data_ptr(), y.ctypes.data)
    torch.cuda.synchronize()
    assert np.all(y == 1 + 1j)
    b.close()


@pytest.mark.parametrize("prec,logn", [(32, 16), (32, 15), (64, 16), (64, 15), (64, 14)])
def test_cooperative_kernel_matches_fused(oracle64, prec, logn):
    """The XCD-cooperative two-step kernel (coop=1; the default for FP64) against the
    fused one-work-group-per-transform kernel (coop=0): the same arithmetic,
    so every entry's output is identical bit for bit, with the lookahead at
    1 (most step-2 chunks deferred to the work-group finishing step 1), 5 and
    64 (none deferred), out of place and in place; a spread of entries also
    against the oracle.  FP32 2^15 and FP64 2^14 run with the radix-2-split
    kernel off so that the two-step kernels take the group."""
    torch = _torch()
    n = 1 << logn
    cnt = 300
    cdt = torch.complex64 if prec == 32 else torch.complex128
    g = torch.Generator(device="cuda")
    g.manual_seed(logn)
    x = torch.complex(torch.rand(cnt * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1,
                      torch.rand(cnt * n, generator=g, device="cuda", dtype=torch.float64) * 2 - 1).to(cdt)
    esz = 8 if prec == 32 else 16
    kinds = ["dft" if i % 3 else "invdft" for i in range(cnt)]
    b = mf.Batch([(kinds[i], (n,), i * n * esz, i * n * esz) for i in range(cnt)],
                 mf.PREC_FP32 if prec == 32 else mf.PREC_FP64)
    extra = ",split2=0" if (prec, logn) in ((32, 15), (64, 14)) else ""
    groups = b.describe().count("\n")
    outs = {}
    for spec in ("coop=0", "coop=1,coopla=1", "coop=1,coopla=8", "coop=1,coopla=64"):
        for gi in range(groups):
            b.schedule(gi, spec + extra)
        print(f"[case] coop prec={prec} N=2^{logn} {spec}", flush=True)
        y = torch.full_like(x, float("nan"))
        b.run(x.data_ptr(), y.data_ptr())
        b.run(x.data_ptr(), y.data_ptr())  # replayed
        z = x.clone()
        b.run(z.data_ptr(), z.data_ptr())  # in place
        torch.cuda.synchronize()
        assert torch.equal(y, z), spec
        outs[spec] = y
    for spec, y in outs.items():
        assert torch.equal(y, outs["coop=0"]), spec
    b.close()
    yh = outs["coop=1,coopla=8"].cpu().numpy().reshape(cnt, n)
    xh = x.cpu().numpy().reshape(cnt, n)
    bar = TOL32 if prec == 32 else tol64((n,))
    for i in sorted({0, 1, 7, 8, 9, 150, cnt - 2, cnt - 1}):
        ref = oracle64.run(kinds[i], xh[i].astype(np.complex128), (n,))
        assert rel_l2(yh[i], ref) <= bar, i
