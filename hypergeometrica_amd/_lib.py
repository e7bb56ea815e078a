"""ctypes binding of libhgx.so (the C ABI of include/hgx.h).

There is no CPU path: if the library is missing or no device is usable,
every call raises HgxError.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import threading

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libhgx.so")

HGX_OK, HGX_EINVAL, HGX_ETOOBIG, HGX_ENOMEM, HGX_EDEVICE, HGX_ENOTINIT = 0, -1, -2, -3, -4, -5

# every symbol the header declares (checked by the CPU test suite)
EXPORTS = (
    "hgx_init", "hgx_shutdown", "hgx_last_error", "hgx_num_moduli", "hgx_modulus",
    "hgx_product_length", "hgx_max_transform_bits", "hgx_mul", "hgx_sqr", "hgx_mul_batch",
    "hgx_mul_device", "hgx_sqr_device", "hgx_mul_batch_device", "hgx_ntt_forward",
    "hgx_ntt_reverse", "hgx_submit", "hgx_poll", "hgx_wait", "hgx_fill_splitmix64",
    "hgx_timing_enable", "hgx_timing_read", "hgx_add_device", "hgx_sub_device", "hgx_graphs_enable",
    "hgx_plan_set", "hgx_plan_get", "hgx_autotune", "hgx_split_enable", "hgx_staging_limit",
    "hgx_server_stats", "hgx_mul_device_noalloc",
    "hgx_device_alloc", "hgx_device_free", "hgx_upload", "hgx_download", "hgx_device_copy", "hgx_device_fill",
    "hgx_device_significant", "hgx_device_compare", "hgx_shl_device", "hgx_host_alloc", "hgx_host_free",
    "hgx_synchronize", "hgx_stream_create", "hgx_stream_destroy", "hgx_modmul_probe",
    "hgx_kernel_set", "hgx_kernel_get", "hgx_series_split", "hgx_partial_get", "hgx_partial_get_device",
    "hgx_partial_free",
)


class HgxError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"hgx error {code}: {msg}")
        self.code = code


class MulDesc(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("na", ctypes.c_size_t), ("b", ctypes.c_void_p),
                ("nb", ctypes.c_size_t), ("out", ctypes.c_void_p)]


class ServerStats(ctypes.Structure):
    _fields_ = [("completed", ctypes.c_uint64), ("failed", ctypes.c_uint64), ("rounds", ctypes.c_uint64),
                ("staged", ctypes.c_uint64), ("queued", ctypes.c_uint64), ("in_flight", ctypes.c_uint64),
                ("uptime_s", ctypes.c_double), ("launch_s", ctypes.c_double), ("retire_s", ctypes.c_double),
                ("wait_s", ctypes.c_double), ("lanes", ctypes.c_int),
                ("split_downloads", ctypes.c_int), ("uploads", ctypes.c_int), ("downloads", ctypes.c_int)]


class SeriesDesc(ctypes.Structure):
    """hgx_series: the four term functions a, b, p, q as polynomials in k."""
    _fields_ = [("degree", ctypes.c_int * 4), ("coef", (ctypes.c_int64 * 8) * 4), ("at_zero", ctypes.c_int * 4),
                ("value0", ctypes.c_int64 * 4)]


class KernelTime(ctypes.Structure):
    _fields_ = [("name", ctypes.c_char * 32), ("launches", ctypes.c_uint64), ("total_ms", ctypes.c_double)]


_u64 = ctypes.c_uint64
_sz = ctypes.c_size_t
_vp = ctypes.c_void_p
_lib = None
_init_lock = threading.Lock()
_initialised = False


def load() -> ctypes.CDLL:
    """Load libhgx.so (no device call)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise HgxError(HGX_ENOTINIT, f"{LIB_PATH} is not built (run __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        L.hgx_init.argtypes = [ctypes.c_int, _sz]
        L.hgx_shutdown.argtypes = []
        L.hgx_shutdown.restype = None
        L.hgx_last_error.restype = ctypes.c_char_p
        L.hgx_num_moduli.restype = ctypes.c_int
        L.hgx_modulus.argtypes = [ctypes.c_int]
        L.hgx_modulus.restype = _u64
        L.hgx_product_length.argtypes = [_sz, _sz]
        L.hgx_product_length.restype = _sz
        L.hgx_max_transform_bits.restype = ctypes.c_int
        L.hgx_mul.argtypes = [_vp, _sz, _vp, _sz, _vp]
        L.hgx_sqr.argtypes = [_vp, _sz, _vp]
        L.hgx_mul_batch.argtypes = [ctypes.POINTER(MulDesc), _sz]
        L.hgx_mul_device.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp]
        L.hgx_sqr_device.argtypes = [_vp, _sz, _vp, _vp]
        L.hgx_mul_device_noalloc.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp]
        L.hgx_add_device.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp]
        L.hgx_plan_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
        _ip = ctypes.POINTER(ctypes.c_int)
        L.hgx_plan_get.argtypes = [ctypes.c_int, _ip, _ip, _ip]
        L.hgx_autotune.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.hgx_kernel_set.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
        L.hgx_kernel_get.argtypes = [ctypes.c_int, _ip, _ip]
        L.hgx_staging_limit.argtypes = [_sz]
        L.hgx_sub_device.argtypes = [_vp, _sz, _vp, _sz, _vp, _vp]
        L.hgx_mul_batch_device.argtypes = [ctypes.POINTER(MulDesc), _sz, _vp]
        L.hgx_ntt_forward.argtypes = [_vp, _sz, ctypes.c_int]
        L.hgx_ntt_reverse.argtypes = [_vp, _sz, ctypes.c_int]
        L.hgx_submit.argtypes = [ctypes.POINTER(MulDesc), ctypes.POINTER(_u64)]
        L.hgx_poll.argtypes = [_u64]
        L.hgx_wait.argtypes = [_u64]
        L.hgx_fill_splitmix64.argtypes = [_vp, _sz, _u64, ctypes.c_int, _vp]
        L.hgx_timing_enable.argtypes = [ctypes.c_int]
        L.hgx_timing_read.argtypes = [ctypes.POINTER(KernelTime), ctypes.c_int]
        L.hgx_server_stats.argtypes = [ctypes.POINTER(ServerStats)]
        L.hgx_device_alloc.argtypes = [_sz, ctypes.POINTER(_vp)]
        L.hgx_device_free.argtypes = [_vp, _vp]
        L.hgx_upload.argtypes = [_vp, _vp, _sz, _vp]
        L.hgx_download.argtypes = [_vp, _vp, _sz, _vp]
        L.hgx_device_copy.argtypes = [_vp, _vp, _sz, _vp]
        L.hgx_device_fill.argtypes = [_vp, _sz, _u64, _vp]
        L.hgx_device_significant.argtypes = [_vp, _sz, ctypes.POINTER(_sz), _vp]
        L.hgx_device_compare.argtypes = [_vp, _sz, _vp, _sz, ctypes.POINTER(ctypes.c_int), _vp]
        L.hgx_shl_device.argtypes = [_vp, _sz, ctypes.c_int, _vp, _vp]
        L.hgx_host_alloc.argtypes = [_sz, ctypes.POINTER(_vp)]
        L.hgx_host_free.argtypes = [_vp]
        L.hgx_synchronize.argtypes = [_vp]
        L.hgx_stream_create.argtypes = [ctypes.POINTER(_vp)]
        L.hgx_stream_destroy.argtypes = [_vp]
        L.hgx_modmul_probe.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
        L.hgx_series_split.argtypes = [ctypes.POINTER(SeriesDesc), _u64, _u64, ctypes.POINTER(_vp)]
        L.hgx_partial_get.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_sz), _vp]
        L.hgx_partial_get_device.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(_sz), _vp,
                                             _vp]
        L.hgx_partial_free.argtypes = [_vp]
        L.hgx_partial_free.restype = None
        _lib = L
    return _lib


def check(rc: int) -> int:
    if rc < 0:
        msg = load().hgx_last_error()
        raise HgxError(rc, msg.decode() if msg else "")
    return rc


_exit_hook = False


generation = 0
"""Count of library shutdowns in this process: device blocks handed out
before one (hgx_device_alloc) are gone after it, and their owners must not
release them again."""


def _shutdown_at_exit() -> None:
    """Release the library (request server, streams, device memory) while the
    HIP runtime is still fully up: Python's exit hooks run before the C
    runtime's and the loaded libraries' finalizers."""
    global _initialised, generation
    if _initialised and _lib is not None:
        _initialised = False
        generation += 1
        _lib.hgx_shutdown()


def shutdown() -> None:
    """hgx_shutdown now (the exit hook then has nothing left to do)."""
    _shutdown_at_exit()


def lib(device: int = 0, budget: int = 0) -> ctypes.CDLL:
    """The library, initialised on DEVICE (once per process)."""
    global _initialised, _exit_hook
    L = load()
    if not _initialised:
        with _init_lock:
            if not _initialised:
                check(L.hgx_init(device, budget))
                _initialised = True
                if not _exit_hook:
                    atexit.register(_shutdown_at_exit)
                    _exit_hook = True
    return L


def timing_enable(on: bool = True) -> None:
    check(lib().hgx_timing_enable(1 if on else 0))


def server_stats() -> dict:
    """The request server's counters (hgx_server_stats) as a dict."""
    st = ServerStats()
    check(load().hgx_server_stats(ctypes.byref(st)))
    return {name: getattr(st, name) for name, _ in ServerStats._fields_}


def timing_read() -> dict:
    """{kernel name: (launches, total ms)} of the calls since timing_enable."""
    buf = (KernelTime * 64)()
    n = check(lib().hgx_timing_read(buf, 64))
    return {buf[i].name.decode(): (int(buf[i].launches), float(buf[i].total_ms)) for i in range(n)}
