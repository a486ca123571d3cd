"""The library's host code (plans, batch plans, tuner and its cache file,
staging of host arrays, streams, error and allocation-failure paths) under
AddressSanitizer + UndefinedBehaviorSanitizer (+ LeakSanitizer).

`python -m minfft_amd.build --sanitize` (also run by the build entry point)
builds lib/san/: the three libraries with instrumented host code and the
driver programs of tests/c_app/host_checks.cpp, one per precision.  Without a
GPU the drivers check the makers' failure paths; with one they run every host
path (see the driver's sections) and check their results."""
import os
import subprocess

import pytest

import minfft_amd as mf

SAN = os.path.join(os.path.dirname(mf.lib_path("libminfft.so")), "san")
SAN_ENV = {
    # the sanitizer runtime is linked into the driver; another preloaded
    # library may come first in the load order
    "ASAN_OPTIONS": "detect_leaks=1:verify_asan_link_order=0:abort_on_error=0:exitcode=23",
    "UBSAN_OPTIONS": "halt_on_error=1:print_stacktrace=1:exitcode=24",
    "LSAN_OPTIONS": "exitcode=25",
}


# allocations that the HIP and HSA runtimes keep until exit (their own
# caches and device objects) are not the library's
LSAN_SUPPRESSIONS = "leak:libhsa-runtime64.so\nleak:libamdhip64.so\n"


def _run(prec, tmp_path, args=(), timeout=600):
    exe = os.path.join(SAN, f"host_checks{prec}")
    if not os.path.exists(exe):
        pytest.skip("sanitizer variant not built (python -m minfft_amd.build --sanitize)")
    supp = tmp_path / "lsan.supp"
    supp.write_text(LSAN_SUPPRESSIONS)
    env = dict(os.environ)
    env.update(SAN_ENV)
    env["LSAN_OPTIONS"] += f":suppressions={supp}:print_suppressions=0"
    print(f"[case] host_checks{prec} under ASan/UBSan/LSan", flush=True)
    r = subprocess.run([exe] + list(args), capture_output=True, text=True, timeout=timeout, env=env)
    out = r.stdout + r.stderr
    print(r.stdout[-2000:])
    print(r.stderr[-4000:])
    for marker in ("AddressSanitizer", "LeakSanitizer", "runtime error:"):
        assert marker not in out, out[-4000:]
    assert r.returncode == 0, out[-4000:]
    return r.stdout


@pytest.mark.parametrize("prec", [64, 32])
def test_host_code_without_device_under_sanitizers(prec, tmp_path):
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: the device run covers this")
    assert _run(prec, tmp_path).strip().endswith("no device: ok")


@pytest.mark.gpu
@pytest.mark.parametrize("prec", [64, 32])
def test_host_code_under_sanitizers(prec, tmp_path):
    out = _run(prec, tmp_path, [str(tmp_path / "tune.cache")])
    assert out.strip().endswith("ok") and "FAIL" not in out
