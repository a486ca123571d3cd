"""The autotuner's schedule file (MINFFT_TUNE_CACHE) across processes: a
first process tunes a batch group and writes the file (a header naming the
device and library build, then one line per group); a second process with
the same file takes the schedule from it without measuring (the plan
describes the group as "tuned (cache file)" with the first process's
timings) and its outputs match the FP64 oracle; a file whose header names
another build is ignored and replaced."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from tests.helpers import TOL32, rel_l2

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

CHILD = r"""
import json, sys
import numpy as np
import torch
sys.path.insert(0, sys.argv[1])
import minfft_amd as mf
n, count = 1 << 14, 64
rng = np.random.default_rng(11)
x = (rng.uniform(-1, 1, (count, n)) + 1j * rng.uniform(-1, 1, (count, n))).astype(np.complex64)
xd = torch.from_numpy(x).cuda()
yd = torch.empty_like(xd)
b = mf.Batch([("dft", (n,), i * n * 8, i * n * 8) for i in range(count)], mf.PREC_FP32)
b.run(xd.data_ptr(), yd.data_ptr())
torch.cuda.synchronize()
desc = b.describe()
b.close()
np.save(sys.argv[2], yd[:2].cpu().numpy())
print(json.dumps({"describe": desc.strip()}))
"""


def _child(cache, out_npy):
    env = dict(os.environ, MINFFT_TUNE_CACHE=cache)
    env.pop("MINFFT_TUNE", None)
    env.pop("MINFFT_AUTOTUNE", None)
    r = subprocess.run([sys.executable, "-c", CHILD, ROOT, out_npy], env=env, capture_output=True, text=True,
                       timeout=240)
    assert r.returncode == 0, r.stderr[-2000:]
    return json.loads(r.stdout.strip().splitlines()[-1])["describe"]


def _timings(desc):
    kv = dict(t.split("=", 1) for t in desc.split() if "=" in t)
    return kv["default_us"], kv["tuned_us"]


def test_tune_cache_reused_by_a_second_process(tmp_path, oracle64):
    cache = str(tmp_path / "tune.txt")
    print("[case] tune cache: first process tunes", flush=True)
    d1 = _child(cache, str(tmp_path / "y1.npy"))
    assert "tuned:" in d1 and "cache file" not in d1, d1
    lines = open(cache).read().splitlines()
    assert lines[0].startswith("#minfft-tune v5 arch=gfx950"), lines[0]
    assert any(ln.startswith("p1.k0.d1.16384x1x1.c6") for ln in lines[1:]), lines
    print("[case] tune cache: second process reuses it", flush=True)
    d2 = _child(cache, str(tmp_path / "y2.npy"))
    assert "tuned (cache file)" in d2, d2
    assert _timings(d2) == _timings(d1), (d1, d2)
    # the schedule line itself is the same
    assert d1.split(": ", 2)[2] == d2.split(": ", 2)[2], (d1, d2)
    y2 = np.load(str(tmp_path / "y2.npy"))
    rng = np.random.default_rng(11)
    n = 1 << 14
    x = (rng.uniform(-1, 1, (64, n)) + 1j * rng.uniform(-1, 1, (64, n))).astype(np.complex64)
    for i in range(2):
        e = rel_l2(y2[i], oracle64.run("dft", x[i].astype(np.complex128), (n,)))
        assert e <= TOL32, e
    print("[case] tune cache: a file of another build is ignored and replaced", flush=True)
    with open(cache, "w") as f:
        f.write("#minfft-tune v5 arch=gfx950 cus=256 build=another\n")
        f.write("p1.k0.d1.16384x1x1.c6 1 1 1 1 1 1 1 1 1 1 1 1 0 1 5 64 1.0 1.0\n")
    d3 = _child(cache, str(tmp_path / "y3.npy"))
    assert "tuned:" in d3 and "cache file" not in d3, d3
    assert open(cache).read().splitlines()[0] == lines[0]
    assert not [p for p in os.listdir(tmp_path) if ".tmp." in p]
