"""minfft_amd -- Python front end of the MI355X minfft drop-in.

Everything here goes through the C ABI of the in-tree shared libraries
(minfft_amd/lib/libminfft.so, libminfft_single.so, libminfft_engine.so); the
transforms themselves run in hand-written HIP kernels for gfx950.  There is
no CPU fallback: if the libraries are not built, importing the bindings
raises.

The names mirror the reference's C API (aimukhin/minfft minfft.h:35-63):
``mkaux_dft_1d``, ``dft``, ``invdft``, ``realdft`` ... ``free_aux``.  Arrays
may be numpy arrays (host memory: staged through the device by the library)
or torch tensors on the GPU (used in place).
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, "lib")

KINDS = ("dft", "invdft", "realdft", "invrealdft",
         "dct2", "dst2", "dct3", "dst3", "dct4", "dst4")
KIND_ID = {k: i for i, k in enumerate(KINDS)}
FAMILY = {"dft": "dft", "invdft": "dft", "realdft": "realdft", "invrealdft": "realdft",
          "dct2": "t2t3", "dst2": "t2t3", "dct3": "t2t3", "dst3": "t2t3",
          "dct4": "t4", "dst4": "t4"}
MAKERS = [f"minfft_mkaux_{fam}{suf}" for fam in ("dft", "realdft", "t2t3", "t4")
          for suf in ("_1d", "_2d", "_3d", "")]
TRANSFORMS = [f"minfft_{k}" for k in KINDS]
SYMBOLS = TRANSFORMS + MAKERS + ["minfft_free_aux"]
EXT_SYMBOLS = ["minfft_last_error", "minfft_mkaux_batch", "minfft_batch_run", "minfft_batch",
               "minfft_free_batch", "minfft_batch_schedule", "minfft_batch_describe", "minfft_enqueue",
               "minfft_enqueue_noalloc", "minfft_set_stream"]

PREC_FP64, PREC_FP32, PREC_SPLIT_FP16 = 0, 1, 2


def lib_path(name):
    return os.path.join(LIBDIR, name)


class MinfftError(RuntimeError):
    pass


class BatchEntry(ctypes.Structure):
    """struct minfft_batch_entry of include/minfft_ext.h"""
    _fields_ = [("kind", ctypes.c_int), ("d", ctypes.c_int), ("n", ctypes.c_int * 3),
                ("in_offset", ctypes.c_longlong), ("out_offset", ctypes.c_longlong)]


_ENGINE = None


def engine():
    """libminfft_engine.so (error channel, batch extension)."""
    global _ENGINE
    if _ENGINE is None:
        p = lib_path("libminfft_engine.so")
        if not os.path.exists(p):
            raise MinfftError(f"{p} is not built: run `python -m minfft_amd.build`")
        lib = ctypes.CDLL(p, mode=ctypes.RTLD_GLOBAL)
        lib.minfft_last_error.restype = ctypes.c_char_p
        lib.minfft_mkaux_batch.restype = ctypes.c_void_p
        lib.minfft_mkaux_batch.argtypes = [ctypes.c_int, ctypes.POINTER(BatchEntry), ctypes.c_int]
        lib.minfft_batch_run.restype = ctypes.c_int
        lib.minfft_batch_run.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        lib.minfft_batch.restype = ctypes.c_int
        lib.minfft_batch.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t,
                                     ctypes.c_void_p, ctypes.c_size_t]
        lib.minfft_free_batch.restype = None
        lib.minfft_free_batch.argtypes = [ctypes.c_void_p]
        lib.minfft_batch_schedule.restype = ctypes.c_int
        lib.minfft_batch_schedule.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p]
        lib.minfft_batch_describe.restype = ctypes.c_int
        lib.minfft_batch_describe.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_int]
        lib.minfft_enqueue.restype = ctypes.c_int
        lib.minfft_enqueue.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                       ctypes.c_void_p, ctypes.c_void_p]
        lib.minfft_enqueue_noalloc.restype = ctypes.c_int
        lib.minfft_enqueue_noalloc.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p,
                                               ctypes.c_void_p, ctypes.c_void_p]
        lib.minfft_set_stream.restype = ctypes.c_int
        lib.minfft_set_stream.argtypes = [ctypes.c_void_p, ctypes.c_void_p]
        _ENGINE = lib
    return _ENGINE


def last_error():
    return engine().minfft_last_error().decode()


class Minfft:
    """The 27-symbol minfft.h API of libminfft.so (prec=64) or
    libminfft_single.so (prec=32)."""

    def __init__(self, prec=64):
        assert prec in (32, 64)
        engine()
        name = "libminfft.so" if prec == 64 else "libminfft_single.so"
        p = lib_path(name)
        if not os.path.exists(p):
            raise MinfftError(f"{p} is not built: run `python -m minfft_amd.build`")
        self.prec = prec
        self.lib = ctypes.CDLL(p, mode=ctypes.RTLD_LOCAL)
        for k in KINDS:
            f = getattr(self.lib, f"minfft_{k}")
            f.restype = None
            f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
        for fam in ("dft", "realdft", "t2t3", "t4"):
            getattr(self.lib, f"minfft_mkaux_{fam}_1d").argtypes = [ctypes.c_int]
            getattr(self.lib, f"minfft_mkaux_{fam}_2d").argtypes = [ctypes.c_int] * 2
            getattr(self.lib, f"minfft_mkaux_{fam}_3d").argtypes = [ctypes.c_int] * 3
            getattr(self.lib, f"minfft_mkaux_{fam}").argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
            for suf in ("_1d", "_2d", "_3d", ""):
                getattr(self.lib, f"minfft_mkaux_{fam}{suf}").restype = ctypes.c_void_p
        self.lib.minfft_free_aux.restype = None
        self.lib.minfft_free_aux.argtypes = [ctypes.c_void_p]

    # -- makers (reference minfft.h:46-61) ---------------------------------
    def mkaux(self, family, shape):
        """Generic maker minfft_mkaux_<family>(d, Ns); returns a handle or
        raises MinfftError with the library's reason."""
        d = len(shape)
        Ns = (ctypes.c_int * max(d, 1))(*shape)
        a = getattr(self.lib, f"minfft_mkaux_{family}")(d, Ns)
        if not a:
            raise MinfftError(f"minfft_mkaux_{family}{tuple(shape)}: {last_error()}")
        return a

    def mkaux_raw(self, name, *args):
        """Call a maker by name; returns the raw pointer (None = NULL)."""
        return getattr(self.lib, name)(*args)

    def free_aux(self, a):
        self.lib.minfft_free_aux(a)

    # -- transforms (reference minfft.h:35-44) ------------------------------
    def run_ptr(self, kind, x, y, a):
        """Run transform `kind` on raw pointers (host or device)."""
        getattr(self.lib, f"minfft_{kind}")(x, y, a)
        err = last_error()
        if err:
            raise MinfftError(f"minfft_{kind}: {err}")


def _ptr(arr):
    """Data pointer of a numpy array or a torch tensor."""
    if hasattr(arr, "data_ptr"):
        return arr.data_ptr()
    return arr.ctypes.data


class Aux:
    """Owning wrapper of an aux made for `family` and `shape`."""

    def __init__(self, lib, family, shape):
        self.lib = lib
        self.family = family
        self.shape = tuple(shape)
        self.h = lib.mkaux(family, self.shape)

    def __call__(self, kind, x, y):
        assert FAMILY[kind] == self.family
        self.lib.run_ptr(kind, _ptr(x), _ptr(y), self.h)
        return y

    def enqueue(self, kind, x, y, stream=None):
        """Asynchronous run on device pointers (minfft_enqueue, minfft_ext.h)."""
        assert FAMILY[kind] == self.family
        rc = engine().minfft_enqueue(KIND_ID[kind], _ptr(x), _ptr(y), self.h, stream)
        if rc:
            raise MinfftError(f"minfft_enqueue({kind}): {last_error()}")
        return y

    def enqueue_noalloc(self, kind, x, y, stream=None):
        """minfft_enqueue_noalloc (minfft_ext.h): as enqueue, with the
        kernels' scratch in the library's fixed device array; allocates
        nothing."""
        assert FAMILY[kind] == self.family
        rc = engine().minfft_enqueue_noalloc(KIND_ID[kind], _ptr(x), _ptr(y), self.h, stream)
        if rc:
            raise MinfftError(f"minfft_enqueue_noalloc({kind}): {last_error()}")
        return y

    def set_stream(self, stream=None):
        """minfft_set_stream (minfft_ext.h): the synchronous calls and
        NULL-stream enqueues of this aux run on `stream` (a hipStream_t
        handle; None = the aux's own stream)."""
        rc = engine().minfft_set_stream(self.h, stream)
        if rc:
            raise MinfftError(f"minfft_set_stream: {last_error()}")

    def close(self):
        if self.h:
            self.lib.free_aux(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Batch:
    """A minfft_mkaux_batch plan (include/minfft_ext.h)."""

    def __init__(self, entries, prec=PREC_FP32):
        arr = (BatchEntry * len(entries))()
        for i, e in enumerate(entries):
            kind, dims, ioff, ooff = e
            arr[i].kind = KIND_ID[kind] if isinstance(kind, str) else int(kind)
            arr[i].d = len(dims)
            for k, n in enumerate(dims):
                arr[i].n[k] = n
            arr[i].in_offset = ioff
            arr[i].out_offset = ooff
        self._entries = arr
        self.h = engine().minfft_mkaux_batch(len(entries), arr, prec)
        if not self.h:
            raise MinfftError(f"minfft_mkaux_batch: {last_error()}")

    def run(self, in_ptr, out_ptr, stream=None):
        rc = engine().minfft_batch_run(self.h, in_ptr, out_ptr, stream)
        if rc:
            raise MinfftError(f"minfft_batch_run: {last_error()}")

    def run_sync(self, x, y):
        nb_in = x.nbytes if hasattr(x, "nbytes") else x.numel() * x.element_size()
        nb_out = y.nbytes if hasattr(y, "nbytes") else y.numel() * y.element_size()
        rc = engine().minfft_batch(self.h, _ptr(x), nb_in, _ptr(y), nb_out)
        if rc:
            raise MinfftError(f"minfft_batch: {last_error()}")
        return y

    def schedule(self, group, spec):
        """Pin group `group`'s launch schedule ("key=value,..." over the
        process-wide one; "" restores it)."""
        rc = engine().minfft_batch_schedule(self.h, group, spec.encode())
        if rc:
            raise MinfftError(f"minfft_batch_schedule: {last_error()}")

    def describe(self):
        """One line per group: kind, dims, count and the chosen schedule."""
        n = engine().minfft_batch_describe(self.h, None, 0)
        if n < 0:
            raise MinfftError(f"minfft_batch_describe: {last_error()}")
        buf = ctypes.create_string_buffer(n)
        engine().minfft_batch_describe(self.h, buf, n)
        return buf.value.decode()

    def close(self):
        if self.h:
            engine().minfft_free_batch(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
