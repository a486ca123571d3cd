"""ctypes front end of the CPU oracle and of the compiled reference.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.

- ``Oracle(prec)``: the C restatement in oracle/minfft_oracle.c
  (functions o64_* / o32_*), operating on numpy arrays.
- ``Reference(prec, fast=False)``: the reference minfft.c compiled by
  oracle/Makefile into oracle/_ref (symbols renamed ref64_* / ref32_*),
  driven through its own API (minfft.h:35-63): make an aux, run, free.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))

KINDS = ("dft", "invdft", "realdft", "invrealdft",
         "dct2", "dst2", "dct3", "dst3", "dct4", "dst4")
R2R_CODE = {"dct2": 2, "dct3": 3, "dct4": 4, "dst2": -2, "dst3": -3, "dst4": -4}
AUX_FAMILY = {"dft": "dft", "invdft": "dft", "realdft": "realdft",
              "invrealdft": "realdft", "dct2": "t2t3", "dst2": "t2t3",
              "dct3": "t2t3", "dst3": "t2t3", "dct4": "t4", "dst4": "t4"}


def real_dtype(prec):
    return np.float64 if prec == 64 else np.float32


def out_shape(kind, shape):
    """Shape (in real units for real kinds, complex units for complex)."""
    shape = tuple(shape)
    if kind == "realdft":
        return shape[:-1] + (shape[-1] // 2 + 1,)
    return shape


def in_shape(kind, shape):
    shape = tuple(shape)
    if kind == "invrealdft":
        return shape[:-1] + (shape[-1] // 2 + 1,)
    return shape


def is_complex_in(kind):
    return kind in ("dft", "invdft", "invrealdft")


def is_complex_out(kind):
    return kind in ("dft", "invdft", "realdft")


def _lib(name):
    path = os.path.join(HERE, name)
    if not os.path.exists(path):
        raise RuntimeError(f"{path} is not built; run `make -C oracle`")
    return ctypes.CDLL(path, mode=ctypes.RTLD_LOCAL)


class Oracle:
    """The C restatement of minfft's algorithms (oracle/minfft_oracle.c)."""

    def __init__(self, prec=64):
        assert prec in (32, 64)
        self.prec = prec
        self.rdt = real_dtype(prec)
        self.cdt = np.complex128 if prec == 64 else np.complex64
        lib = _lib("liboracle.so")
        p = f"o{prec}_"
        self._dft = getattr(lib, p + "dft")
        self._realdft = getattr(lib, p + "realdft")
        self._invrealdft = getattr(lib, p + "invrealdft")
        self._r2r = getattr(lib, p + "r2r")
        self._brute = getattr(lib, p + "brute_dft")
        for f in (self._dft, self._realdft, self._invrealdft, self._r2r, self._brute):
            f.restype = ctypes.c_int

    @staticmethod
    def _dims(shape):
        d = len(shape)
        return d, (ctypes.c_int * d)(*shape)

    def run(self, kind, x, shape):
        """Apply transform `kind` to x of logical transform shape `shape`.

        Complex data is passed as numpy complex arrays; real data as real
        arrays.  Returns a new array.
        """
        shape = tuple(int(s) for s in shape)
        d, Ns = self._dims(shape)
        if kind in ("dft", "invdft"):
            xin = np.ascontiguousarray(x, dtype=self.cdt).reshape(shape)
            y = np.empty(shape, dtype=self.cdt)
            rc = self._dft(d, Ns, xin.ctypes.data_as(ctypes.c_void_p),
                           y.ctypes.data_as(ctypes.c_void_p), int(kind == "invdft"))
        elif kind == "realdft":
            xin = np.ascontiguousarray(x, dtype=self.rdt).reshape(shape)
            y = np.empty(out_shape(kind, shape), dtype=self.cdt)
            rc = self._realdft(d, Ns, xin.ctypes.data_as(ctypes.c_void_p),
                               y.ctypes.data_as(ctypes.c_void_p))
        elif kind == "invrealdft":
            xin = np.ascontiguousarray(x, dtype=self.cdt).reshape(in_shape(kind, shape))
            y = np.empty(shape, dtype=self.rdt)
            rc = self._invrealdft(d, Ns, xin.ctypes.data_as(ctypes.c_void_p),
                                  y.ctypes.data_as(ctypes.c_void_p))
        else:
            xin = np.ascontiguousarray(x, dtype=self.rdt).reshape(shape)
            y = np.empty(shape, dtype=self.rdt)
            rc = self._r2r(R2R_CODE[kind], d, Ns, xin.ctypes.data_as(ctypes.c_void_p),
                           y.ctypes.data_as(ctypes.c_void_p))
        if rc != 0:
            raise ValueError(f"oracle rejected {kind} {shape}")
        return y

    def brute_dft(self, x, inverse=False):
        x = np.ascontiguousarray(x, dtype=self.cdt)
        y = np.empty_like(x)
        self._brute(len(x), x.ctypes.data_as(ctypes.c_void_p),
                    y.ctypes.data_as(ctypes.c_void_p), int(inverse))
        return y


class Reference:
    """The reference minfft compiled from its own sources into oracle/_ref."""

    def __init__(self, prec=64, fast=False):
        assert prec in (32, 64)
        self.prec = prec
        self.rdt = real_dtype(prec)
        self.cdt = np.complex128 if prec == 64 else np.complex64
        name = f"_ref/libminfft_ref{prec}{'_fast' if fast else ''}.so"
        self.lib = _lib(name)
        self.pfx = f"ref{prec}_minfft_"
        for k in KINDS:
            f = self.fn(k)
            f.restype = None
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        for fam in ("dft", "realdft", "t2t3", "t4"):
            f = self.fn(f"mkaux_{fam}")
            f.restype = ctypes.c_void_p
            f.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
        self.fn("free_aux").argtypes = [ctypes.c_void_p]
        self.fn("free_aux").restype = None

    def fn(self, name):
        return getattr(self.lib, self.pfx + name)

    def mkaux(self, kind, shape):
        d = len(shape)
        Ns = (ctypes.c_int * d)(*shape)
        a = self.fn(f"mkaux_{AUX_FAMILY[kind]}")(d, Ns)
        if not a:
            raise ValueError(f"reference mkaux failed for {kind} {shape}")
        return a

    def free(self, a):
        self.fn("free_aux")(a)

    def run(self, kind, x, shape, aux=None):
        shape = tuple(int(s) for s in shape)
        own = aux is None
        a = self.mkaux(kind, shape) if own else aux
        try:
            idt = self.cdt if is_complex_in(kind) else self.rdt
            odt = self.cdt if is_complex_out(kind) else self.rdt
            xin = np.array(x, dtype=idt, copy=True).reshape(in_shape(kind, shape))
            y = np.empty(out_shape(kind, shape), dtype=odt)
            self.fn(kind)(xin.ctypes.data, y.ctypes.data, a)
            return y
        finally:
            if own:
                self.free(a)


def ref_available(prec=64, fast=False):
    return os.path.exists(os.path.join(
        HERE, f"_ref/libminfft_ref{prec}{'_fast' if fast else ''}.so"))


def rel_l2(a, b):
    a = np.asarray(a, dtype=np.complex128 if np.iscomplexobj(a) or np.iscomplexobj(b) else np.float64)
    b = np.asarray(b, dtype=a.dtype)
    nb = np.linalg.norm(b.ravel())
    if nb == 0:
        return float(np.linalg.norm(a.ravel()))
    return float(np.linalg.norm((a - b).ravel()) / nb)
