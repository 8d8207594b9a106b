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
                L.rbfx_build_weights_async.restype = C.c_int
                L.rbfx_build_weights_async.argtypes = [C.POINTER(abi.BuildArgs)]
                L.rbfx_build_weights_noalloc.restype = C.c_int
                L.rbfx_build_weights_noalloc.argtypes = [C.POINTER(abi.BuildArgs)]
                L.rbfx_source_tiles.restype = C.c_int
                L.rbfx_source_tiles.argtypes = [C.c_int64, C.c_int32, C.c_void_p, C.c_int32, C.c_void_p, C.c_void_p,
                                                C.c_void_p, C.c_void_p, C.c_int64, C.POINTER(C.c_int64), C.c_void_p]
                L.rbfx_spmv_async.restype = C.c_int
                L.rbfx_spmv_async.argtypes = [C.POINTER(abi.SpmvArgs)]
                L.rbfx_last_build_counts.restype = C.c_int
                L.rbfx_last_build_counts.argtypes = [C.POINTER(C.c_int64)]
                L.rbfx_permute_rows.restype = C.c_int
                L.rbfx_permute_rows.argtypes = [C.c_int64, C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int64,
                                                C.c_int32, C.c_void_p, C.c_void_p, C.c_void_p]
                L.rbfx_spatial_order.restype = C.c_int
                L.rbfx_spatial_order.argtypes = [C.c_void_p, C.c_int64, C.c_int32, C.c_void_p, C.c_void_p]
                L.rbfx_last_build_stats.restype = C.c_int
                L.rbfx_last_build_stats.argtypes = [C.POINTER(C.c_int64)]
                L.rbfx_tune_plan.restype = C.c_int
                L.rbfx_tune_plan.argtypes = [C.c_int32]
                L.rbfx_tune_save.restype = C.c_int
                L.rbfx_tune_save.argtypes = [C.c_char_p]
                L.rbfx_tune_load.restype = C.c_int
                L.rbfx_tune_load.argtypes = [C.c_char_p]
                L.rbfx_tune_reset.restype = C.c_int
                L.rbfx_last_build_timing.restype = C.c_int
                L.rbfx_last_build_timing.argtypes = [C.POINTER(C.c_double)]
                L.rbfx_mlp_param_count.restype = C.c_int64
                L.rbfx_mlp_bf16_count.restype = C.c_int64
                L.rbfx_mlp_workspace_bytes.restype = C.c_int64
                L.rbfx_mlp_workspace_bytes.argtypes = [C.c_int32]
                L.rbfx_mlp_run.restype = C.c_int
                L.rbfx_mlp_run.argtypes = [C.POINTER(abi.MlpRunArgs)]
                L.rbfx_mlp_adam.restype = C.c_int
                L.rbfx_mlp_adam.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_float,
                                            C.c_float, C.c_float, C.c_float, C.c_int32, C.c_void_p]
                L.rbfx_mlp_sync_bf16.restype = C.c_int
                L.rbfx_mlp_sync_bf16.argtypes = [C.c_void_p, C.c_void_p, C.c_void_p]
                L.rbfx_stencil_features.restype = C.c_int
                L.rbfx_stencil_features.argtypes = [C.c_void_p, C.c_void_p, C.c_int64, C.c_int32, C.c_int32,
                                                    C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                                    C.c_void_p]
                _LIB = L
    return _LIB


def exported_symbols():
    return ["rbfx_init", "rbfx_last_error", "rbfx_abi_version", "rbfx_knn", "rbfx_build_weights",
            "rbfx_spmv", "rbfx_last_build_counts", "rbfx_last_build_timing", "rbfx_mlp_param_count", "rbfx_mlp_bf16_count",
            "rbfx_mlp_workspace_bytes", "rbfx_mlp_run", "rbfx_mlp_adam", "rbfx_mlp_sync_bf16",
            "rbfx_stencil_features", "rbfx_spatial_order", "rbfx_build_weights_async", "rbfx_spmv_async",
            "rbfx_build_weights_noalloc", "rbfx_source_tiles",
            "rbfx_permute_rows", "rbfx_last_build_buckets", "rbfx_tune_bucket", "rbfx_last_build_stats",
            "rbfx_tune_save", "rbfx_tune_load", "rbfx_tune_reset", "rbfx_tune_plan"]


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


def last_build_buckets():
    """(stencils per size bucket, kernel per bucket) of the last build on this thread:
    buckets n <= 32, 33..64, 65..; kernels 0 registers x2, 1 registers, 2 LDS, -1 none."""
    n = (C.c_int64 * 3)()
    k = (C.c_int32 * 3)()
    _check(lib().rbfx_last_build_buckets(n, k))
    return list(n), list(k)


def tune_bucket(bucket, kernel):
    """Pin the kernel of a size bucket (-1: automatic). Kernel ids as rbfx_tune_bucket
    documents them: 0 rook registers x2, 1 rook registers (pinned only), 2 rook LDS,
    3/4 those in FP64 within FP32 mode, 5 null-space Cholesky, 6 the same, 8 lanes of
    three rows per stencil, 7 rook LDS with the packed lower triangle."""
    _check(lib().rbfx_tune_bucket(int(bucket), int(kernel)))


def last_build_stats():
    """[stencils the null-space kernels handed to the rook kernels, autotuner sample
    launches, launch plan of an uneven build (1 per bucket, 2 one launch; 0 one size),
    0, ...] of the last build on this thread."""
    st = (C.c_int64 * 8)()
    _check(lib().rbfx_last_build_stats(st))
    return list(st)


def tune_plan(plan):
    """Pin the launch plan of uneven builds: 0 automatic, 1 one launch per size bucket,
    2 one launch of the packed LDS kernel over all stencils."""
    _check(lib().rbfx_tune_plan(int(plan)))


def tune_save(path):
    """Write the autotuner's choices (every device) to a text file."""
    _check(lib().rbfx_tune_save(str(path).encode()))


def tune_load(path):
    """Use the choices of a file written by tune_save instead of timing samples."""
    _check(lib().rbfx_tune_load(str(path).encode()))


def tune_reset():
    """Forget every autotuner choice made or loaded in this process."""
    _check(lib().rbfx_tune_reset())


def spmv(args, dev, asynchronous=False):
    t = torch()
    with t.cuda.device(dev):
        f = lib().rbfx_spmv_async if asynchronous else lib().rbfx_spmv
        _check(f(C.byref(args)))


def build_weights_async(args, dev):
    t = torch()
    with t.cuda.device(dev):
        _check(lib().rbfx_build_weights_async(C.byref(args)))


def last_error():
    return lib().rbfx_last_error().decode(errors="replace")


def mlp_run(x, params, wbf16, workspace, batch, target=None, grads=None, loss=None, y_out=None, n_out=0,
            row_weight=None):
    """One forward (and with target, backward) pass of the surrogate (device tensors)."""
    t = torch()
    dev = x.device
    a = abi.MlpRunArgs()
    a.batch = int(batch)
    a.n_out = int(n_out)
    a.x = x.data_ptr()
    a.target = None if target is None else target.data_ptr()
    a.params = params.data_ptr()
    a.wbf16 = wbf16.data_ptr()
    a.grads = None if grads is None else grads.data_ptr()
    a.loss = None if loss is None else loss.data_ptr()
    a.y_out = None if y_out is None else y_out.data_ptr()
    a.workspace = workspace.data_ptr()
    a.stream = stream_ptr(dev).value
    a.row_weight = None if row_weight is None else row_weight.data_ptr()
    with t.cuda.device(dev):
        _check(lib().rbfx_mlp_run(C.byref(a)))


def mlp_adam(params, grads, m, v, wbf16, lr, beta1, beta2, eps, step):
    t = torch()
    dev = params.device
    with t.cuda.device(dev):
        _check(lib().rbfx_mlp_adam(ptr(params), ptr(grads), ptr(m), ptr(v), ptr(wbf16), float(lr), float(beta1),
                                   float(beta2), float(eps), int(step), stream_ptr(dev)))


def mlp_sync_bf16(params, wbf16):
    t = torch()
    dev = params.device
    with t.cuda.device(dev):
        _check(lib().rbfx_mlp_sync_bf16(ptr(params), ptr(wbf16), stream_ptr(dev)))


def mlp_workspace_bytes(batch):
    return int(lib().rbfx_mlp_workspace_bytes(int(batch)))


def stencil_features(data, eval_points, adj, k, want_xi=True, want_x=True, want_h=True):
    """Canonical radius-scaled offsets of fixed-size 2-D stencils (device tensors):
    (xi (n,k,2) f64, network inputs (n,128) bf16, h (n,) f64, r1 (n,) f64)."""
    t = torch()
    dev = data.device
    n = eval_points.shape[0]
    xi = t.empty((n, k, 2), dtype=t.float64, device=dev) if want_xi else None
    x = t.empty((n, abi.MLP_IN_PAD), dtype=t.bfloat16, device=dev) if want_x else None
    h = t.empty(n, dtype=t.float64, device=dev) if want_h else None
    r1 = t.empty(n, dtype=t.float64, device=dev)
    with t.cuda.device(dev):
        _check(lib().rbfx_stencil_features(ptr(data), ptr(eval_points), n, data.shape[1], int(k), ptr(adj), ptr(xi),
                                           ptr(x), ptr(h), ptr(r1), stream_ptr(dev)))
    return xi, x, h, r1


def last_build_timing():
    """Device ms of the last build's stages: main kernel, FP32 list solve, FP64 fallback, total."""
    ms = (C.c_double * 4)()
    lib().rbfx_last_build_timing(ms)
    return list(ms)


def spatial_order(points):
    """Morton order of device points (n, dim) f64 -> int32 (n,) device permutation."""
    t = torch()
    dev = points.device
    out = t.empty(points.shape[0], dtype=t.int32, device=dev)
    with t.cuda.device(dev):
        _check(lib().rbfx_spatial_order(ptr(points), points.shape[0], points.shape[1], ptr(out), stream_ptr(dev)))
    return out


def source_tiles(n_rows, k, col, tile_rows):
    """Source-tile plan of fixed-length rows stored back to back (device col): per tile
    of tile_rows rows its distinct columns and each entry's position among them.
    Returns (tile_src_n int32, tile_src_off int64, tile_col int16 (read as uint16),
    tile_src int32)."""
    t = torch()
    dev = col.device
    n_tiles = (int(n_rows) + tile_rows - 1) // tile_rows
    tn = t.empty(n_tiles, dtype=t.int32, device=dev)
    toff = t.empty(n_tiles, dtype=t.int64, device=dev)
    tcol = t.empty(int(n_rows) * int(k), dtype=t.int16, device=dev)
    n_src = C.c_int64(0)
    with t.cuda.device(dev):
        _check(lib().rbfx_source_tiles(int(n_rows), int(k), ptr(col), int(tile_rows), ptr(tn), ptr(toff), ptr(tcol),
                                       None, 0, C.byref(n_src), stream_ptr(dev)))
        tsrc = t.empty(max(1, n_src.value), dtype=t.int32, device=dev)
        _check(lib().rbfx_source_tiles(int(n_rows), int(k), ptr(col), int(tile_rows), ptr(tn), ptr(toff), ptr(tcol),
                                       ptr(tsrc), n_src.value, C.byref(n_src), stream_ptr(dev)))
    return tn, toff, tcol, tsrc


def permute_rows(order, k, col, vals):
    """Apply plan: CSR rows of fixed length k copied into `order` (device tensors).
    vals is (n_valarrays, nnz). Returns (col_p, vals_p)."""
    t = torch()
    dev = col.device
    n = order.shape[0]
    col_p = t.empty_like(col)
    vals_p = t.empty_like(vals)
    with t.cuda.device(dev):
        _check(lib().rbfx_permute_rows(n, int(k), ptr(order), ptr(col), ptr(vals), vals.shape[1], vals.shape[0],
                                       ptr(col_p), ptr(vals_p), stream_ptr(dev)))
    return col_p, vals_p
