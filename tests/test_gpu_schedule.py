"""GPU tests of the batch autotuner and of pinned schedules: every schedule
the tuner can choose must give the same results (to the parity bars) as the
default one, i.e. agree with the CPU oracle."""
import numpy as np
import pytest

import minfft_amd as mf
from oracle.oracle import out_shape, is_complex_in, is_complex_out
from tests.helpers import TOL32, make_input, rel_l2, tol64

pytestmark = pytest.mark.gpu

# every knob value the tuner tries (autotune.cpp), alone and in the
# combinations that change the long-axis path
SPECS = ["", "tpw=128", "tpw=512", "cfb=64", "cfb=256", "xcd=0", "wave=0",
         "cfcore=0", "cfcore=1", "cfcore=0,cfb=256", "cfcore=0,wave=0", "e32=1", "e32=1,split2=1",
         "e32=1,split2=0,fused=0", "e32=1,cfcore=0",
         "fused=0", "falt=1", "wpc=1", "wpc=4", "split=-1", "split=1",
         "split2=0", "split2=0,split=-1", "split2=0,split=1", "split2=1", "split2=2", "split2=3",
         "fused=0,split=-1", "fused=0,split=1",
         "rsplit=11", "rsplit=12", "rsplit=21", "rsplit=22", "rsplit=31", "rsplit=32", "rsplit=41", "e4=0",
         "coop=0", "coop=1", "coop=1,coopla=1", "coop=1,coopla=4", "coop=1,coopla=16", "coop=1,coopla=64",
         "mall=0", "mall=1", "mall=2", "mall=32", "csplit=0", "csplit=1", "csplit=1,cfb=64", "csplit=1,xcd=0"]


class Packed:
    """Entries packed into one device input and one device output buffer."""

    def __init__(self, entries, xs, prec):
        import torch
        self.torch = torch
        self.prec = prec
        rs = 8 if prec == 64 else 4
        self.rs = rs
        ioffs, ooffs = [], []
        ipos = opos = 0
        for (kind, shape), x in zip(entries, xs):
            ioffs.append(ipos)
            ipos += x.nbytes + 16 * rs
            n_out = int(np.prod(out_shape(kind, shape))) * (2 if is_complex_out(kind) else 1)
            ooffs.append(opos)
            opos += n_out * rs + 16 * rs
        hin = np.zeros(ipos, dtype=np.uint8)
        for x, o in zip(xs, ioffs):
            hin[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
        self.din = torch.from_numpy(hin).cuda()
        self.dout = torch.zeros(opos, dtype=torch.uint8, device="cuda")
        self.entries, self.xs, self.ooffs = entries, xs, ooffs
        self.batch = mf.Batch([(k, s, io, oo) for (k, s), io, oo in zip(entries, ioffs, ooffs)],
                              mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)

    def run(self):
        self.dout.zero_()
        self.batch.run(self.din.data_ptr(), self.dout.data_ptr(), None)
        self.torch.cuda.synchronize()
        return self.dout.cpu().numpy()

    def output(self, hout, i):
        kind, shape = self.entries[i]
        rdt = np.float64 if self.prec == 64 else np.float32
        cdt = np.complex128 if self.prec == 64 else np.complex64
        odt = cdt if is_complex_out(kind) else rdt
        n_out = int(np.prod(out_shape(kind, shape)))
        o = self.ooffs[i]
        return hout[o:o + n_out * np.dtype(odt).itemsize].view(odt).reshape(out_shape(kind, shape))


# the longest single-tile length + 1 (FP32 2^15, FP64 2^14): single-pass
# radix-2-split kernel or two-step; + 2: fused two-step kernel (>= 256
# entries) or two-step
def _long(prec):
    return (1 << 16, 1 << 15) if prec == 32 else (1 << 15, 1 << 14)


def _cases(prec):
    rng = np.random.default_rng(77 + prec)
    fused_n, split_n = _long(prec)
    entries = []
    entries += [("dft", (64, 256))] * 24 + [("invdft", (128, 32, 16))] * 4
    entries += [("dft", (8, 2048))] * 4 + [("invdft", (1024, 8))] * 4
    entries += [("dft", (fused_n,))] * 260 + [("invdft", (split_n,))] * 40
    entries += [("dct2", (128, 128))] * 8 + [("dst4", (64, 32))] * 8
    # real kinds whose half-length FFTs reach the radix-32 tiles (rows and columns)
    entries += [("dct2", (4, 1024)), ("dst3", (1024, 2)), ("dct4", (2, 2048)), ("realdft", (2, 1024)),
                ("dst2", (1024,)), ("dct3", (2048,)), ("dst4", (1024, 4))] * 2
    entries += [("dft", (split_n // 2,))] * 6 + [("invdft", (split_n // 4,))] * 6
    entries += [("realdft", (1 << 17,))] * 3 + [("realdft", (2 * split_n,))] * 2
    # complex column passes of 512 and 1024 points (the radix-2-split column
    # tile; 3 x 16 columns leave a partial last tile of 32), and along axis 0
    # of a 3D entry
    entries += [("dft", (512, 64))] * 3 + [("invdft", (512, 16))] * 3 + [("dft", (1024, 32))] * 2
    entries += [("invdft", (512, 8, 16))] * 2 + [("realdft", (512, 4, 32))] * 2
    # one-entry multi-dimensional groups (run as single transforms: slabs of axis 0)
    entries += [("dft", (64, 32, 16)), ("dct2", (128, 64)), ("invdft", (16, 2048)), ("dst4", (4, 8, 32))]
    xs = [make_input(k, s, prec, int(rng.integers(1 << 30))) for k, s in entries]
    return entries, xs


@pytest.mark.parametrize("prec", [32, 64])
def test_every_schedule_matches_oracle(oracle64, prec):
    entries, xs = _cases(prec)
    pk = Packed(entries, xs, prec)
    ngroups = pk.batch.describe().count("\n")
    # entries to check: all but the fused-length ones, of which a spread
    long_idx = [i for i, (k, s) in enumerate(entries) if s == (_long(prec)[0],)]
    check = [i for i in range(len(entries)) if i not in long_idx] + long_idx[::41] + [long_idx[-1]]
    refs = {}
    for i in check:
        kind, shape = entries[i]
        x = xs[i].astype(np.complex128 if is_complex_in(kind) else np.float64)
        refs[i] = oracle64.run(kind, x, shape)
    for spec in SPECS:
        for g in range(ngroups):
            pk.batch.schedule(g, spec)
        print(f"[case] schedule '{spec}' prec={prec}", flush=True)
        hout = pk.run()
        for i in check:
            kind, shape = entries[i]
            bar = tol64(shape) if prec == 64 else TOL32
            e = rel_l2(pk.output(hout, i), refs[i])
            assert e <= bar, (spec, i, kind, shape, e)
    pk.batch.close()


def test_autotuner_tunes_large_groups(oracle64):
    """A group with enough work is timed at plan time; its schedule is
    reported and the results agree with the oracle."""
    n = 1 << 13
    cnt = 1024
    entries = [("dft", (n,))] * cnt + [("dct2", (16,))] * 3
    rng = np.random.default_rng(5)
    xs = [make_input(k, s, 32, int(rng.integers(1 << 30))) for k, s in entries]
    pk = Packed(entries, xs, 32)
    desc = pk.batch.describe()
    print(desc)
    lines = desc.strip().split("\n")
    assert len(lines) == 2
    big = [ln for ln in lines if f"dft {n} count={cnt}" in ln]
    assert big and " tuned:" in big[0], desc
    small = [ln for ln in lines if "dct2 16 count=3" in ln]
    assert small and "untuned" in small[0], desc
    hout = pk.run()
    for i in [0, 1, 777, cnt - 1, cnt, cnt + 2]:
        kind, shape = entries[i]
        ref = oracle64.run(kind, xs[i].astype(np.complex128 if is_complex_in(kind) else np.float64), shape)
        assert rel_l2(pk.output(hout, i), ref) <= TOL32
    with pytest.raises(mf.MinfftError):
        pk.batch.schedule(0, "nonsense=3")
    with pytest.raises(mf.MinfftError):
        pk.batch.schedule(99, "")
    pk.batch.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_residue_split_every_code(oracle64, prec):
    """Every residue-split kernel (rsplit.hpp) on forward and inverse entries of
    each length it serves, with the entries one complex item off the 16-byte
    grid (so the FP32 paired store takes its unaligned branch on odd entries)
    and counts that are not multiples of the 8-transform grid groups."""
    import torch
    rs = 8 if prec == 64 else 4
    cs = 2 * rs
    cdt = np.complex128 if prec == 64 else np.complex64
    rng = np.random.default_rng(11 + prec)
    for logn in (14, 15, 16):
        n = 1 << logn
        kinds = ["dft"] * 11 + ["invdft"] * 5
        xs = [(rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(cdt) for _ in kinds]
        offs = [(i * (n + 1) + 1) * cs for i in range(len(kinds))]
        total = offs[-1] + n * cs + cs
        hin = np.zeros(total, dtype=np.uint8)
        for x, o in zip(xs, offs):
            hin[o:o + n * cs] = np.frombuffer(x.tobytes(), dtype=np.uint8)
        din = torch.from_numpy(hin).cuda()
        dout = torch.zeros(total, dtype=torch.uint8, device="cuda")
        batch = mf.Batch([(k, (n,), o, o) for k, o in zip(kinds, offs)],
                         mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
        refs = [oracle64.run(k, x.astype(np.complex128), (n,)) for k, x in zip(kinds, xs)]
        for code in (11, 12, 21, 22, 31, 32, 41, 42):
            for g in range(2):
                batch.schedule(g, f"rsplit={code}")
            print(f"[case] rsplit={code} n=2^{logn} prec={prec}", flush=True)
            dout.zero_()
            batch.run(din.data_ptr(), dout.data_ptr(), None)
            torch.cuda.synchronize()
            hout = dout.cpu().numpy()
            for i, (k, o) in enumerate(zip(kinds, offs)):
                y = hout[o:o + n * cs].view(cdt)
                bar = tol64((n,)) if prec == 64 else TOL32
                e = rel_l2(y, refs[i])
                assert e <= bar, (code, logn, i, k, e)
        batch.close()


def test_autotuner_single_group_plan(oracle64):
    """A plan with one group is tuned with the residue-split kernels among
    the candidates (its cache key differs from a multi-group plan's); the
    chosen schedule is reported and the results agree with the oracle."""
    n = 1 << 14
    cnt = 320
    entries = [("dft", (n,))] * cnt
    rng = np.random.default_rng(9)
    xs = [make_input(k, s, 32, int(rng.integers(1 << 30))) for k, s in entries]
    pk = Packed(entries, xs, 32)
    desc = pk.batch.describe()
    print(desc)
    lines = desc.strip().split("\n")
    assert len(lines) == 1 and " tuned:" in lines[0] and "rsplit=" in lines[0], desc
    hout = pk.run()
    for i in (0, 1, cnt // 2, cnt - 1):
        kind, shape = entries[i]
        ref = oracle64.run(kind, xs[i].astype(np.complex128), shape)
        assert rel_l2(pk.output(hout, i), ref) <= TOL32
    pk.batch.close()


def test_split_fp16_schedules(oracle64):
    """The split-FP16 MFMA tiles under the schedules that change their launch
    shape: column passes on the column variant (c-fast core) or on the row
    variant with LDS-staged column input and output (cfcore=0); and with
    both radix-16 stages on the matrix cores (mstages=2)."""
    import torch
    shapes = [(512, 64), (64, 512), (512, 8, 2), (8, 512, 4)]
    kinds = ["dft", "invdft", "dft", "invdft"]
    rng = np.random.default_rng(33)
    xs = [make_input(k, sh, 32, int(rng.integers(1 << 30))) for k, sh in zip(kinds, shapes)]
    offs, pos = [], 0
    for x in xs:
        offs.append(pos)
        pos += x.nbytes + 64
    hin = np.zeros(pos, dtype=np.uint8)
    for x, o in zip(xs, offs):
        hin[o:o + x.nbytes] = np.frombuffer(x.tobytes(), dtype=np.uint8)
    din = torch.from_numpy(hin).cuda()
    dout = torch.zeros(pos, dtype=torch.uint8, device="cuda")
    batch = mf.Batch([(k, sh, o, o) for k, sh, o in zip(kinds, shapes, offs)], mf.PREC_SPLIT_FP16)
    refs = [oracle64.run(k, x.astype(np.complex128), sh) for k, x, sh in zip(kinds, xs, shapes)]
    for spec in ("", "cfcore=0", "cfcore=1", "cfcore=0,cfb=256", "xcd=0", "mstages=2", "mstages=2,cfcore=0",
                 "mstages=2,cfcore=1"):
        for g in range(len(shapes)):
            batch.schedule(g, spec)
        print(f"[case] split-FP16 schedule '{spec}'", flush=True)
        dout.zero_()
        batch.run(din.data_ptr(), dout.data_ptr(), None)
        torch.cuda.synchronize()
        hout = dout.cpu().numpy()
        for i, (x, o) in enumerate(zip(xs, offs)):
            y = hout[o:o + x.nbytes].view(np.complex64).reshape(shapes[i])
            e = rel_l2(y, refs[i])
            assert e <= TOL32, (spec, i, kinds[i], shapes[i], e)
    batch.close()


@pytest.mark.parametrize("prec", [32, 64])
def test_residue_split_never_in_place(oracle64, prec):
    """The residue-split kernel's work-groups all read a transform's whole
    input while writing their outputs, so an in-place run (input and output
    the same buffer) with the residue-split schedule pinned takes another
    single-pass or two-step kernel: the results still match the oracle, and
    the groups beside it (sharing the two-step scratch) are not disturbed."""
    import torch
    n = 1 << 15
    entries = [("dft", (n,))] * 6 + [("realdft", (1 << 17,))] * 2 + [("invdft", (1 << 14,))] * 4
    xs = [make_input(k, s, prec, 950 + i) for i, (k, s) in enumerate(entries)]
    cs = 16 if prec == 64 else 8
    offs, pos = [], 0
    for (k, s), x in zip(entries, xs):
        offs.append(pos)
        pos += max(x.nbytes, int(np.prod(out_shape(k, s))) * (cs if is_complex_out(k) else cs // 2))
        pos = (pos + 255) & ~255
    buf = np.zeros(pos, dtype=np.uint8)
    for x, o in zip(xs, offs):
        buf[o:o + x.nbytes] = np.frombuffer(np.ascontiguousarray(x).tobytes(), dtype=np.uint8)
    d = torch.from_numpy(buf).cuda()
    b = mf.Batch([(k, s, o, o) for (k, s), o in zip(entries, offs)], mf.PREC_FP64 if prec == 64 else mf.PREC_FP32)
    ngroups = b.describe().count("\n")
    for g in range(ngroups):
        b.schedule(g, "rsplit=22" if prec == 32 else "rsplit=31")
    print(f"[case] in-place residue-split schedule prec={prec}", flush=True)
    b.run(d.data_ptr(), d.data_ptr(), None)
    torch.cuda.synchronize()
    out = d.cpu().numpy()
    b.close()
    bar = lambda s: tol64(s) if prec == 64 else TOL32
    for i, ((k, s), x) in enumerate(zip(entries, xs)):
        odt = (np.complex128 if prec == 64 else np.complex64) if is_complex_out(k) else \
            (np.float64 if prec == 64 else np.float32)
        nout = int(np.prod(out_shape(k, s)))
        y = out[offs[i]:offs[i] + nout * np.dtype(odt).itemsize].view(odt).reshape(out_shape(k, s))
        ref = oracle64.run(k, x.astype(np.complex128 if is_complex_in(k) else np.float64), s)
        assert rel_l2(y, ref) <= bar(s), (i, k, s, rel_l2(y, ref))


def test_plain_in_place_with_residue_split_schedule(tmp_path):
    """A plain plan whose process-wide schedule asks for the residue-split
    kernel (MINFFT_TUNE, read once per process: a fresh process) runs out
    of place on device arrays through the residue-split kernel, and in
    place (device arrays, or host arrays, which are transformed in place in
    the device staging buffer) through the other single-pass kernel; all
    match numpy."""
    import os
    import subprocess
    import sys
    script = tmp_path / "inplace.py"
    script.write_text(
        "import sys, numpy as np\n"
        f"sys.path.insert(0, {os.path.dirname(os.path.dirname(os.path.abspath(__file__)))!r})\n"
        "import minfft_amd as mf\n"
        "n = 1 << 15\n"
        "rng = np.random.default_rng(5)\n"
        "x = (rng.uniform(-1, 1, n) + 1j * rng.uniform(-1, 1, n)).astype(np.complex64)\n"
        "ref = np.fft.fft(x.astype(np.complex128))\n"
        "import torch\n"
        "aux = mf.Aux(mf.Minfft(32), 'dft', (n,))\n"
        "xd = torch.from_numpy(x).cuda(); yd = torch.empty_like(xd); zd = xd.clone()\n"
        "aux('dft', xd, yd)\n"
        "aux('dft', zd, zd)\n"
        "h = x.copy()\n"
        "aux('dft', h, h)\n"
        "for name, r in (('out of place', yd.cpu().numpy()), ('in place', zd.cpu().numpy()), ('host', h)):\n"
        "    e = np.linalg.norm(r - ref) / np.linalg.norm(ref)\n"
        "    print(name, e)\n"
        "    assert e < 1e-6, (name, e)\n"
        "print('ok')\n")
    env = dict(os.environ, MINFFT_TUNE="rsplit=22")
    r = subprocess.run([sys.executable, str(script)], capture_output=True, text=True, timeout=180, env=env)
    print(r.stdout, r.stderr[-2000:])
    assert r.returncode == 0 and r.stdout.strip().endswith("ok")
