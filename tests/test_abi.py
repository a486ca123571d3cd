"""The C-ABI boundary, checked without a GPU: the libraries are built, load,
and export every symbol include/minfft.h and include/minfft_ext.h declare,
with the reference's 27 names (aimukhin/minfft minfft.h:35-63)."""
import ctypes
import os
import re
import subprocess

import pytest

import minfft_amd as mf

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")


def _declared(header):
    txt = open(os.path.join(INC, header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return set(re.findall(r"\b(minfft_\w+)\s*\(", txt))


def _exported(lib):
    out = subprocess.run(["nm", "-D", "--defined-only", mf.lib_path(lib)],
                         capture_output=True, text=True, check=True).stdout
    return {line.split()[-1] for line in out.splitlines() if " T " in line}


def test_header_declares_the_reference_api():
    names = _declared("minfft.h")
    assert names == set(mf.SYMBOLS)
    assert len(names) == 27


@pytest.mark.parametrize("lib", ["libminfft.so", "libminfft_single.so"])
def test_api_libraries_export_all_27_symbols(lib):
    exp = {s for s in _exported(lib) if s.startswith("minfft_")}
    assert exp == set(mf.SYMBOLS)


def test_engine_exports_extension_symbols():
    exp = _exported("libminfft_engine.so")
    assert _declared("minfft_ext.h") <= exp
    assert set(mf.EXT_SYMBOLS) <= exp
    # the engine itself must not export the precision-specific API
    assert not ({s for s in exp if s.startswith("minfft_")} & set(mf.SYMBOLS))


@pytest.mark.parametrize("prec", [64, 32])
def test_libraries_load_through_ctypes(prec):
    lib = mf.Minfft(prec)
    for s in mf.SYMBOLS:
        assert getattr(lib.lib, s)
    assert mf.engine().minfft_last_error() is not None


def test_kernels_built_for_gfx950():
    """The engine's embedded code object targets gfx950 (this build's own
    output, read as bytes)."""
    data = open(mf.lib_path("libminfft_engine.so"), "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def _build_c_app(tmp_path, single):
    """A C89 program written against minfft.h (tests/c_app/app.c), compiled
    unchanged and linked against the drop-in library of its precision."""
    lib_dir = os.path.dirname(mf.lib_path("libminfft.so"))
    exe = os.path.join(str(tmp_path), f"app{int(single)}")
    cmd = ["gcc", "-std=c89", "-pedantic", "-Wall", "-Wextra", "-Werror", "-O2",
           f"-DMINFFT_SINGLE={int(single)}", "-I", INC, os.path.join(ROOT, "tests", "c_app", "app.c"),
           "-L", lib_dir, "-lminfft_single" if single else "-lminfft", f"-Wl,-rpath,{lib_dir}", "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


# the other client programs: C99 complex numbers (<complex.h> before
# minfft.h) and C++ std::complex arrays
CLIENTS = {"c99": ("gcc", ["-std=c99"], "app_c99.c", ["-lm"]),
           "cpp": ("g++", ["-std=c++11"], "app_cpp.cpp", [])}


def _build_client(tmp_path, which, single):
    cc, std, src, libs = CLIENTS[which]
    lib_dir = os.path.dirname(mf.lib_path("libminfft.so"))
    exe = os.path.join(str(tmp_path), f"{which}{int(single)}")
    cmd = [cc] + std + ["-pedantic", "-Wall", "-Wextra", "-Werror", "-O2", f"-DMINFFT_SINGLE={int(single)}",
                        "-I", INC, os.path.join(ROOT, "tests", "c_app", src), "-L", lib_dir,
                        "-lminfft_single" if single else "-lminfft", f"-Wl,-rpath,{lib_dir}"] + libs + ["-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.parametrize("single", [False, True])
def test_c_application_builds_against_the_dropin(tmp_path, single):
    assert os.path.exists(_build_c_app(tmp_path, single))


@pytest.mark.parametrize("single", [False, True])
@pytest.mark.parametrize("which", sorted(CLIENTS))
def test_c99_and_cpp_clients_build_against_the_dropin(tmp_path, which, single):
    """minfft.h with C99 complex numbers and from C++ (extern "C"), both
    precisions, warnings as errors."""
    assert os.path.exists(_build_client(tmp_path, which, single))


def test_makers_fail_loudly_without_a_device():
    """With no GPU visible the makers return NULL and give the reason: there
    is no CPU fallback behind the API."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is visible: the makers succeed")
    for prec in (64, 32):
        lib = mf.Minfft(prec)
        assert not lib.mkaux_raw("minfft_mkaux_dft_1d", 8)
        assert "device" in mf.last_error()
        assert not lib.mkaux_raw("minfft_mkaux_t2t3_2d", 8, 8)
    with pytest.raises(mf.MinfftError, match="device"):
        mf.Batch([("dft", (8,), 0, 0)], mf.PREC_FP32)


def test_size_checks_come_before_any_device_call():
    """Sizes that no device can hold are refused by the makers themselves,
    before any HIP call, so the checks hold with no GPU: a product of axis
    lengths beyond 2^40 points (including one that would wrap a 64-bit
    product: 2^20 x 2^20 x 2^30), batch entries the same, entry offsets
    beyond any buffer, and batches whose entries hold more than 2^44 points
    in all."""
    for prec in (64, 32):
        lib = mf.Minfft(prec)
        assert not lib.mkaux_raw("minfft_mkaux_dft_3d", 1 << 20, 1 << 20, 1 << 30)
        assert "too large" in mf.last_error(), mf.last_error()
        assert not lib.mkaux_raw("minfft_mkaux_realdft_3d", 1 << 14, 1 << 14, 1 << 14)
        assert "too large" in mf.last_error(), mf.last_error()
    for dims in [(1 << 20, 1 << 20, 1 << 30), (1 << 14, 1 << 14, 1 << 14)]:
        with pytest.raises(mf.MinfftError, match="entry too large"):
            mf.Batch([("dft", (8,), 0, 0), ("dft", dims, 0, 0)], mf.PREC_FP32)
    with pytest.raises(mf.MinfftError, match="offset beyond"):
        mf.Batch([("dft", (8,), 8 << 50, 0)], mf.PREC_FP32)
    with pytest.raises(mf.MinfftError, match="batch too large"):
        mf.Batch([("dft", (1 << 20, 1 << 20), 0, 0)] * 17, mf.PREC_FP64)
