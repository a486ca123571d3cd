"""Build the MI355X minfft libraries in-tree with hipcc (gfx950).

Produces, under minfft_amd/lib/:
  libminfft_engine.so  kernels (tile/pre/post/twiddle), plans, batch extension
  libminfft.so         the 27 minfft.h entry points, double precision
  libminfft_single.so  the same entry points, single precision (MINFFT_SINGLE)

and, with --sanitize, under minfft_amd/lib/san/ the same three libraries with
their host code (plans, batch, tuner, staging, API) instrumented by
AddressSanitizer and UndefinedBehaviorSanitizer (-Xarch_host: the kernels
are the ordinary objects), plus the driver programs tests/c_app/host_checks.cpp
linked against them (host_checks64, host_checks32).

Usage: python -m minfft_amd.build [--force] [-j N] [--sanitize]
"""
import argparse
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INC = os.path.join(ROOT, "include")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "lib")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MINFFT_ARCH", "gfx950")

COMMON = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # no packed FP32 math (v_pk_*_f32) in device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INC, "-I", CSRC,
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-but-set-variable",
]

# host-only instrumentation: each -fsanitize= applies to the host pass only
SAN_FLAGS = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
             "-Xarch_host", "-fno-omit-frame-pointer", "-g"]
SAN_OBJ = os.path.join(OBJ, "san")
SAN_LIB = os.path.join(LIB, "san")
CLANGXX = os.environ.get("MINFFT_CLANGXX", "/opt/rocm/llvm/bin/clang++")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HEADERS = ["fft_core.hpp", "kinds.hpp", "tile.hpp", "ktable.hpp", "engine.hpp", "big.hpp", "bigtable.hpp", "mfma.hpp",
           "autotune.hpp", "split2.hpp", "rsplit.hpp", "multi.hpp", "coop.hpp", "rlong.hpp", "csplit.hpp"]


def _newer(target, deps):
    if not os.path.exists(target):
        return True
    t = os.path.getmtime(target)
    return any(os.path.getmtime(d) > t for d in deps if os.path.exists(d))


def _units():
    """(object, source, extra flags) for every translation unit."""
    u = []
    for f64 in (0, 1):
        for kind in range(10):
            o = os.path.join(OBJ, f"inst_f{64 if f64 else 32}_k{kind}.o")
            u.append((o, "inst.hip", [f"-DINST_F64={f64}", f"-DINST_KIND={kind}"], "engine"))
    u.append((os.path.join(OBJ, "twiddle.o"), "twiddle.hip", [], "engine"))
    u.append((os.path.join(OBJ, "big_inst.o"), "big_inst.hip", [], "engine"))
    u.append((os.path.join(OBJ, "csplit_inst.o"), "csplit_inst.hip", [], "engine"))
    for part in range(3):
        u.append((os.path.join(OBJ, f"rsplit_inst{part}.o"), "rsplit_inst.hip", [f"-DRSPLIT_PART={part}"], "engine"))
    u.append((os.path.join(OBJ, "engine.o"), "engine.cpp", ["-x", "hip"], "engine"))
    u.append((os.path.join(OBJ, "batch.o"), "batch.cpp", ["-x", "hip"], "engine"))
    u.append((os.path.join(OBJ, "autotune.o"), "autotune.cpp", ["-x", "hip"], "engine"))
    u.append((os.path.join(OBJ, "noalloc.o"), "noalloc.cpp", ["-x", "hip"], "engine"))
    u.append((os.path.join(OBJ, "api64.o"), "api.cpp", ["-x", "hip"], "api64"))
    u.append((os.path.join(OBJ, "api32.o"), "api.cpp", ["-x", "hip", "-DMINFFT_SINGLE=1"], "api32"))
    return u


def _san_units(units):
    """The host-code units again, instrumented, into _build/san/."""
    return [(os.path.join(SAN_OBJ, os.path.basename(o)), src, extra + SAN_FLAGS, tag)
            for o, src, extra, tag in units if src.endswith(".cpp")]


def _compile(unit, force):
    obj, src, extra, _ = unit
    srcp = os.path.join(CSRC, src)
    deps = [srcp] + [os.path.join(CSRC, h) for h in HEADERS] + \
        [os.path.join(INC, "minfft.h"), os.path.join(INC, "minfft_ext.h")]
    if not force and not _newer(obj, deps):
        return obj, None
    cmd = [HIPCC] + COMMON + extra + ["-c", srcp, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        return obj, f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}"
    return obj, None


def _link(out, objs, extra=()):
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", out] + list(objs) + list(extra)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")


def build(force=False, jobs=None, verbose=True):
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIB, exist_ok=True)
    units = _units()
    jobs = jobs or min(8, os.cpu_count() or 4)
    errors = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda u: _compile(u, force), units):
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    eng = [u[0] for u in units if u[3] == "engine"]
    eng_so = os.path.join(LIB, "libminfft_engine.so")
    if force or _newer(eng_so, eng):
        _link(eng_so, eng)
    rpath = ["-Wl,-rpath,$ORIGIN", "-L" + LIB, "-lminfft_engine"]
    for name, tag in (("libminfft.so", "api64"), ("libminfft_single.so", "api32")):
        objs = [u[0] for u in units if u[3] == tag]
        so = os.path.join(LIB, name)
        if force or _newer(so, objs + [eng_so]):
            _link(so, objs, rpath)
    if verbose:
        print(f"built {eng_so}, libminfft.so, libminfft_single.so")
    return LIB


def build_sanitized(force=False, jobs=None, verbose=True):
    """The sanitizer variant of the libraries (after build(): it reuses the
    kernel objects) and the host_checks driver programs, in lib/san/.  The
    libraries leave the sanitizer runtimes' symbols to the program that loads
    them, which links the runtimes itself (clang++ -fsanitize=...)."""
    build(force=False, jobs=jobs, verbose=False)
    os.makedirs(SAN_OBJ, exist_ok=True)
    os.makedirs(SAN_LIB, exist_ok=True)
    units = _units()
    san = _san_units(units)
    jobs = jobs or min(8, os.cpu_count() or 4)
    errors = []
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        for obj, err in ex.map(lambda u: _compile(u, force), san):
            if err:
                errors.append(err)
    if errors:
        raise RuntimeError("hipcc failed:\n" + "\n".join(errors))
    by_src = {(u[1], u[3]): u[0] for u in san}
    eng = [by_src.get((u[1], u[3]), u[0]) for u in units if u[3] == "engine"]
    eng_so = os.path.join(SAN_LIB, "libminfft_engine.so")
    if force or _newer(eng_so, eng):
        _link(eng_so, eng)
    rpath = ["-Wl,-rpath,$ORIGIN", "-L" + SAN_LIB, "-lminfft_engine"]
    for name, tag in (("libminfft.so", "api64"), ("libminfft_single.so", "api32")):
        objs = [u[0] for u in san if u[3] == tag]
        so = os.path.join(SAN_LIB, name)
        if force or _newer(so, objs + [eng_so]):
            _link(so, objs, rpath)
    src = os.path.join(ROOT, "tests", "c_app", "host_checks.cpp")
    for single in (0, 1):
        exe = os.path.join(SAN_LIB, f"host_checks{32 if single else 64}")
        deps = [src, os.path.join(INC, "minfft.h"), os.path.join(INC, "minfft_ext.h"), eng_so]
        if not force and not _newer(exe, deps):
            continue
        cmd = [CLANGXX, "-std=c++17", "-O1", "-g", "-Wall", "-Wextra", "-Werror",
               "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
               "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__", f"-DMINFFT_SINGLE={single}",
               "-I", INC, "-I", os.path.join(ROCM, "include"), src, "-o", exe,
               "-L" + SAN_LIB, "-lminfft_single" if single else "-lminfft", "-lminfft_engine",
               "-L" + os.path.join(ROCM, "lib"), "-lamdhip64", "-lpthread",
               f"-Wl,-rpath,{SAN_LIB}", f"-Wl,-rpath,{os.path.join(ROCM, 'lib')}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"$ {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    if verbose:
        print(f"built the sanitizer variant in {SAN_LIB}")
    return SAN_LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("--sanitize", action="store_true")
    a = ap.parse_args(argv)
    build(force=a.force, jobs=a.j)
    if a.sanitize:
        build_sanitized(force=a.force, jobs=a.j)


if __name__ == "__main__":
    sys.exit(main())
