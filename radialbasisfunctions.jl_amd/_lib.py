"""ctypes binding of librbfx.so (include/rbfx.h) plus device-memory plumbing.

torch is used only to own device buffers and to obtain HIP stream handles; every
kernel is launched by the C-ABI library. There is no CPU fallback: without the
library or without a gfx950 device the calls raise.
"""
import ctypes as C
import pathlib
import threading

from . import _abi as abi

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "librbfx.so"

_LIB = None
_LOCK = threading.Lock()
_INIT_DEVICES = set()


class RbfxError(RuntimeError):
    """A failure reported by the C ABI (code, message)."""

    def __init__(self, code, msg):
        super().__init__(f"rbfx error {code}: {msg}")
        self.code = code
        self.msg = msg


def lib():
    """Load librbfx.so (raises when it has not been built)."""
    global _LIB
    if _LIB is None:
        with _LOCK:
            if _LIB is None:
                if not LIB_PATH.exists():
                    raise RuntimeError(
                        f"{LIB_PATH} is missing: the HIP library has not been built "
                        "(run __graft_entry__.build())"
                    )
                L = C.CDLL(str(LIB_PATH))
                L.rbfx_init.restype = C.c_int
                L.rbfx_init.argtypes = [C.c_int32]
                L.rbfx_last_error.restype = C.c_char_p
                L.rbfx_abi_version.restype = C.c_int32
                L.rbfx_knn.restype = C.c_int
                L.rbfx_knn.argtypes = [C.c_void_p, C.c_int64, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                       C.c_void_p, C.c_void_p, C.c_void_p]
                L.rbfx_build_weights.restype = C.c_int
                L.rbfx_build_weights.argtypes = [C.POINTER(abi.BuildArgs)]
                L.rbfx_spmv.restype = C.c_int
                L.rbfx_spmv.argtypes = [C.POINTER(abi.SpmvArgs)]
                L.rbfx_last_build_counts.restype = C.c_int
                L.rbfx_last_build_counts.argtypes = [C.POINTER(C.c_int64)]
                _LIB = L
    return _LIB


def exported_symbols():
    return ["rbfx_init", "rbfx_last_error", "rbfx_abi_version", "rbfx_knn", "rbfx_build_weights",
            "rbfx_spmv", "rbfx_last_build_counts"]


def _check(rc):
    if rc != 0:
        msg = lib().rbfx_last_error().decode(errors="replace")
        raise RbfxError(rc, msg)


def torch():
    import torch as _t

    return _t


def device_for(device=None):
    """Resolve the torch device the engine runs on and initialise the library there."""
    t = torch()
    if not t.cuda.is_available():
        raise RuntimeError("the rbfx engine needs a ROCm GPU (gfx950); none is visible")
    dev = t.device(device) if device is not None else t.device("cuda", t.cuda.current_device())
    if dev.type != "cuda":
        raise ValueError(f"rbfx runs on a HIP device, got {dev}")
    idx = dev.index if dev.index is not None else t.cuda.current_device()
    if idx not in _INIT_DEVICES:
        with t.cuda.device(idx):
            _check(lib().rbfx_init(idx))
        _INIT_DEVICES.add(idx)
    return t.device("cuda", idx)


def stream_ptr(dev):
    t = torch()
    return C.c_void_p(t.cuda.current_stream(dev).cuda_stream)


def ptr(x):
    return None if x is None else C.c_void_p(x.data_ptr())


def knn(data, query, k, q_ptr=None):
    """data (N,d) f64 cuda, query (Q,d) f64 cuda -> (Q,k) int32 cuda (or flat with q_ptr)."""
    t = torch()
    dev = data.device
    n_q = query.shape[0]
    if q_ptr is None:
        out = t.empty((n_q, k), dtype=t.int32, device=dev)
    else:
        out = t.empty(int(q_ptr[-1].item()), dtype=t.int32, device=dev)
    with t.cuda.device(dev):
        _check(lib().rbfx_knn(ptr(data), data.shape[0], ptr(query), n_q, data.shape[1], int(k), ptr(q_ptr),
                              ptr(out), stream_ptr(dev)))
    return out


def build_weights(args, dev):
    t = torch()
    with t.cuda.device(dev):
        rc = lib().rbfx_build_weights(C.byref(args))
    counts = (C.c_int64 * 16)()
    lib().rbfx_last_build_counts(counts)
    return rc, list(counts)


def spmv(args, dev):
    t = torch()
    with t.cuda.device(dev):
        _check(lib().rbfx_spmv(C.byref(args)))


def last_error():
    return lib().rbfx_last_error().decode(errors="replace")
