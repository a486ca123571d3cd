"""The bench harness contract on the CPU: one GPU only (more exits with a
message), and the default workload is the documented one."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def _run(args, env_extra=None):
    env = dict(os.environ)
    env.update(env_extra or {})
    return subprocess.run([sys.executable, os.path.join(ROOT, "bench.py")] + args, capture_output=True,
                          text=True, env=env, timeout=120)


def test_more_than_one_gpu_exits_with_a_message():
    r = _run(["--gpus", "2", "--steps", "1", "--warmup", "0"])
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "one GPU" in r.stderr
    assert r.stdout.strip() == ""  # no bench line


def test_world_size_above_one_exits_with_a_message():
    r = _run(["--steps", "1"], {"WORLD_SIZE": "2"})
    assert r.returncode == 2, (r.returncode, r.stdout, r.stderr)
    assert "one GPU" in r.stderr


def test_default_workload():
    import bench

    sizes = bench.workload()
    assert len(sizes) == 8192
    assert sizes.min() == 64 and sizes.max() == 65536
    assert np.all((sizes & (sizes - 1)) == 0)
    # the figures the bench line and the documentation quote
    assert int(sizes.sum()) == 97355840
    assert abs(float(sizes.sum()) * 16 / 1e9 - 1.558) < 1e-3
    vals, counts = np.unique(sizes, return_counts=True)
    assert dict(zip(vals.tolist(), counts.tolist()))[65536] == 747
    # the same seed gives the same workload
    assert np.array_equal(sizes, bench.workload())


def test_cpu_harness_times_the_reference():
    """The C timing harness behind the CPU baselines runs the reference's
    functions on threads and reports -1 on lengths it cannot run."""
    import pytest
    import bench
    from oracle.oracle import Reference, ref_available

    if not ref_available(32) or not os.path.exists(os.path.join(ROOT, "oracle", "libcpubatch.so")):
        pytest.skip("oracle/_ref or oracle/libcpubatch.so not built")
    run = bench._cpu_harness(Reference(32))
    assert run([64, 1024, 256, 65536, 8], 1, 2) > 0
    assert run([64, 1024, 256, 65536, 8] * 3, 4, 1) > 0
    assert run([64, 1024], 8, 1) > 0  # more threads than entries
    assert run([96], 1, 1) == -1.0  # not a power of two
    assert run([64], 0, 1) == -1.0
