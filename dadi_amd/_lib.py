"""
ctypes binding of libdadi_hip.so (the C ABI declared in include/dadi_hip.h).

The library is the only compute backend of this package: if it is missing or
cannot be loaded, every integration and sampling call raises
``BackendUnavailable``; there is no CPU fallback.
"""
import ctypes
import os
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
# DADI_AMD_LIB names another build of the same library (e.g. one compiled
# with different DH_* tunables) to load instead of the in-tree one.
LIB_PATH = os.environ.get("DADI_AMD_LIB") or os.path.join(HERE, "libdadi_hip.so")

MAXD = 5
ABI_VERSION = 6
ORDER_TV, ORDER_PRE, ORDER_PRE1 = 0, 1, 2

#: numpy mirror of ``dh_segment``
SEG_DTYPE = np.dtype([
    ("nsteps", "<i8"), ("dt", "<f8"), ("inject", "<f8"),
    ("beta_p", "<f8"), ("beta_q", "<f8"),
    ("nu", "<f8", (MAXD,)), ("gamma", "<f8", (MAXD,)), ("h", "<f8", (MAXD,)),
    ("mig", "<f8", (MAXD, MAXD - 1)),
    ("beta_r", "<f8"), ("chrom", "<i4"), ("reserved", "<i4"),
])

#: numpy mirror of ``dh_instance``
INST_DTYPE = np.dtype([
    ("phi", "<u8"), ("xx", "<u8"), ("seg", "<u8"),
    ("nseg", "<i8"), ("cur_seg", "<i8"), ("cur_step", "<i8"),
    ("D", "<i4"), ("n", "<i4"), ("order", "<i4"), ("frozen", "<i4"),
    ("inject", "<i4"), ("status", "<i4"), ("ld", "<i4"), ("reserved", "<i4"),
])

#: numpy mirror of ``dh_sweep_item``
ITEM_DTYPE = np.dtype([
    ("inst", INST_DTYPE), ("seg_of_step", "<u8"), ("total_steps", "<i8"),
    ("first_axis", "<i4"), ("wg_begin", "<i4"), ("wgs", "<i4"), ("step_shift", "<i4"),
])


class BackendUnavailable(RuntimeError):
    """The HIP library or a GPU is not available."""


class DeviceError(RuntimeError):
    """A library call returned an error code."""


class DhInstance(ctypes.Structure):
    _fields_ = [("phi_dev", ctypes.c_void_p), ("xx_dev", ctypes.c_void_p),
                ("seg_dev", ctypes.c_void_p), ("nseg", ctypes.c_longlong),
                ("cur_seg", ctypes.c_longlong), ("cur_step", ctypes.c_longlong),
                ("D", ctypes.c_int), ("n", ctypes.c_int), ("order", ctypes.c_int),
                ("frozen", ctypes.c_int), ("inject", ctypes.c_int), ("status", ctypes.c_int),
                ("ld", ctypes.c_int), ("reserved", ctypes.c_int)]


class DhLaunch(ctypes.Structure):
    _fields_ = [("waves", ctypes.c_int), ("groups", ctypes.c_int),
                ("max_steps", ctypes.c_longlong), ("chunk", ctypes.c_int),
                ("recompute", ctypes.c_int), ("gamma_dev", ctypes.c_void_p),
                ("gamma_doubles", ctypes.c_longlong), ("delj", ctypes.c_int),
                ("max_us", ctypes.c_int)]


_vp = ctypes.c_void_p
_SIGS = {
    "dh_init": ([ctypes.c_int], ctypes.c_int),
    "dh_last_error": ([ctypes.c_char_p, ctypes.c_int], ctypes.c_int),
    "dh_abi_version": ([], ctypes.c_int),
    "dh_sizeof_segment": ([], ctypes.c_size_t),
    "dh_sizeof_instance": ([], ctypes.c_size_t),
    "dh_lds_bytes": ([ctypes.c_int] * 7, ctypes.c_size_t),
    "dh_nd_groups": ([ctypes.c_int] * 8, ctypes.c_int),
    "dh_max_lds": ([], ctypes.c_size_t),
    "dh_integrate_nd": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                         ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_integrate_packed": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                             ctypes.c_int, _vp, ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_packed_lds_bytes": ([ctypes.c_int] * 4, ctypes.c_size_t),
    "dh_integrate_1d": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, _vp,
                         ctypes.c_longlong, ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_integrate_1d_wave": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                              ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_1d_wave_lds_bytes": ([ctypes.c_int], ctypes.c_size_t),
    "dh_integrate_batch": ([_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                            ctypes.c_int, ctypes.c_int, _vp, ctypes.c_longlong,
                            ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int, _vp],
                           ctypes.c_int),
    "dh_sizeof_sweep_item": ([], ctypes.c_size_t),
    "dh_integrate_factored": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                               ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_factored_slice_doubles": ([ctypes.c_int, ctypes.c_int], ctypes.c_longlong),
    "dh_integrate_factored_static": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, ctypes.POINTER(DhLaunch), _vp],
                                     ctypes.c_int),
    "dh_static_scratch_doubles": ([], ctypes.c_longlong),
    "dh_integrate_resident": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                               ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_resident_rows": ([ctypes.c_int], ctypes.c_int),
    "dh_integrate_resident_latency": ([_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp,
                                       ctypes.POINTER(DhLaunch), _vp], ctypes.c_int),
    "dh_resident_latency_slice_doubles": ([ctypes.c_int], ctypes.c_longlong),
    "dh_resident_latency_rows": ([ctypes.c_int], ctypes.c_int),
    "dh_sweep_axis": ([ctypes.POINTER(DhInstance), ctypes.c_int, ctypes.c_int, _vp,
                       ctypes.c_longlong, _vp, ctypes.c_longlong, ctypes.c_int, _vp],
                      ctypes.c_int),
    "dh_advance_steps": ([_vp, ctypes.c_longlong, _vp], ctypes.c_int),
    "dh_integrate_graph": ([ctypes.POINTER(DhInstance), _vp, _vp, ctypes.c_longlong,
                            ctypes.c_int, ctypes.c_int, _vp], ctypes.c_int),
    "dh_integrate_graph_range": ([ctypes.POINTER(DhInstance), _vp, _vp, ctypes.c_longlong,
                                  ctypes.c_longlong, ctypes.c_longlong, ctypes.c_int,
                                  ctypes.c_int, _vp], ctypes.c_int),
    "dh_contract_axis": ([_vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp],
                         ctypes.c_int),
    "dh_sample_axis": ([_vp, ctypes.c_longlong, ctypes.c_int, _vp, ctypes.c_int, _vp, _vp,
                        _vp, _vp, _vp], ctypes.c_int),
}

EXPORTED = tuple(_SIGS)

_lock = threading.Lock()
_lib = None
_inited_device = None


def load(path=LIB_PATH):
    """Load the shared library and bind every entry point (no GPU needed)."""
    global _lib
    with _lock:
        if _lib is not None:
            return _lib
        if not os.path.exists(path):
            raise BackendUnavailable(
                "libdadi_hip.so is not built (run __graft_entry__.build() or make -C dadi_amd/csrc)")
        # torch ships its own HIP runtime (soname libamdhip64.so.7).  Load it
        # first so this library binds to that same runtime instead of mapping
        # a second HIP/HSA runtime into the process, which cannot share
        # device pointers with torch's allocations.
        import torch  # noqa: F401
        try:
            L = ctypes.CDLL(path)
        except OSError as e:
            raise BackendUnavailable("cannot load %s: %s" % (path, e))
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        if L.dh_abi_version() != ABI_VERSION:
            raise BackendUnavailable("libdadi_hip.so ABI version mismatch")
        if L.dh_sizeof_segment() != SEG_DTYPE.itemsize or \
                L.dh_sizeof_instance() != INST_DTYPE.itemsize or \
                L.dh_sizeof_sweep_item() != ITEM_DTYPE.itemsize:
            raise BackendUnavailable("libdadi_hip.so struct layout mismatch")
        _lib = L
        return L


def last_error():
    L = load()
    buf = ctypes.create_string_buffer(1024)
    L.dh_last_error(buf, len(buf))
    return buf.value.decode(errors="replace")


def check(rc, what):
    if rc != 0:
        raise DeviceError("%s failed (%d): %s" % (what, rc, last_error()))


def init(device=0):
    """Load the library and initialise it on ``device`` (needs a GPU)."""
    global _inited_device
    L = load()
    import torch
    if not torch.cuda.is_available():
        raise BackendUnavailable("no GPU visible to torch; the HIP backend needs one")
    if _inited_device != device:
        check(L.dh_init(device), "dh_init")
        _inited_device = device
    return L
