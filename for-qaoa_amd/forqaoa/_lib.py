"""ctypes binding of libforqaoa.so (the C ABI in include/forqaoa.h).

The library is loaded from the package's in-tree build directory
(for-qaoa_amd/lib).  There is no fallback: if the library is missing or was
built for another machine, every entry point raises.
"""
from __future__ import annotations

import ctypes
import os

_PKG_ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(_PKG_ROOT, "lib", "libforqaoa.so")

FQ_OK = 0
FQ_ERR_ARG = 1
FQ_ERR_OOM = 2
FQ_ERR_BUDGET = 3
FQ_ERR_DEVICE = 4
FQ_ERR_UNSUPPORTED = 5

FQ_C128 = 0
FQ_C64 = 1

FQ_SCHED_REFERENCE = 0
FQ_SCHED_FUSED = 1
FQ_SCHED_Z2 = 2

FQ_K_COST, FQ_K_RX, FQ_K_SWAP, FQ_K_FUSED, FQ_K_REDUCE, FQ_K_TOPK, FQ_K_ADJOINT = range(7)
KERNEL_CLASSES = ["cost", "rx", "swap", "fused", "reduce", "topk", "adjoint"]

_ERR_NAMES = {
    FQ_ERR_ARG: "bad argument",
    FQ_ERR_OOM: "out of memory",
    FQ_ERR_BUDGET: "over budget",
    FQ_ERR_DEVICE: "device error",
    FQ_ERR_UNSUPPORTED: "unsupported",
}


class FQError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{_ERR_NAMES.get(code, code)}] {msg}")
        self.code = code


_lib = None

# (name, restype, argtypes) of every exported entry point
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_up = ctypes.POINTER(ctypes.c_uint64)
_sz = ctypes.c_size_t
_i64p = ctypes.POINTER(ctypes.c_int64)
_szp = ctypes.POINTER(ctypes.c_size_t)

SIGNATURES = {
    "fq_last_error": (ctypes.c_char_p, []),
    "fq_version": (_i, []),
    "fq_state_create": (_i, [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "fq_state_destroy": (_i, [_vp]),
    "fq_state_info": (_i, [_vp, _ip, _ip, _ip, _szp]),
    "fq_state_device_ptr": (_i, [_vp, ctypes.POINTER(_vp)]),
    "fq_state_set_stream": (_i, [_vp, _vp]),
    "fq_state_init_zero": (_i, [_vp]),
    "fq_init_h": (_i, [_vp]),
    "fq_graph_create": (_i, [_i, ctypes.POINTER(_vp)]),
    "fq_graph_add_edge": (_i, [_vp, _i, _i, _d]),
    "fq_graph_destroy": (_i, [_vp]),
    "fq_graph_weights": (_i, [_vp, _dp]),
    "fq_graph_relabel": (_i, [_vp, _i, _i, _i]),
    "fq_cost_layer": (_i, [_vp, _vp, _d, _i]),
    "fq_multi_rx": (_i, [_vp, _ip, _i, _d]),
    "fq_swap_conti": (_i, [_vp, _i, _i, _i]),
    "fq_run": (_i, [_vp, _vp, _i, _dp, _dp, _i]),
    "fq_expect_cut": (_i, [_vp, _vp, _dp, _dp]),
    "fq_expect_cut_static": (_i, [_vp, _vp, _dp, _dp]),
    "fq_topk": (_i, [_vp, _i, _up, _dp]),
    "fq_qubit_map": (_i, [_vp, _ip]),
    "fq_state_copy_out": (_i, [_vp, _vp, _sz, _i]),
    "fq_state_copy_in": (_i, [_vp, _vp, _sz]),
    "fq_state_gather": (_i, [_vp, _up, _sz, _dp, _i]),
    "fq_sync": (_i, [_vp]),
    "fq_state_clone": (_i, [_vp, _vp]),
    "fq_backward": (_i, [_vp, _vp, _i, _dp, _dp, _dp, _dp, _dp]),
    "fq_set_memory_budget": (_i, [_sz]),
    "fq_memory_in_use": (_i, [_szp]),
    "fq_slab_create": (_i, [_i, _i, _i, ctypes.POINTER(_vp)]),
    "fq_slab_create_sched": (_i, [_i, _i, _i, _i, ctypes.POINTER(_vp)]),
    "fq_slab_destroy": (_i, [_vp]),
    "fq_slab_info": (_i, [_vp, _ip, _ip, _ip, _ip]),
    "fq_slab_assign": (_i, [_vp, _i, _vp, _i, _dp, _dp]),
    "fq_slab_step": (_i, [_vp, _ip]),
    "fq_slab_set_layers_per_step": (_i, [_vp, _i]),
    "fq_slab_poll": (_i, [_vp, _ip]),
    "fq_slab_collect": (_i, [_vp, _ip, _dp, _dp, _ip]),
    "fq_slab_copy_out": (_i, [_vp, _i, _vp, _sz]),
    "fq_timing_enable": (_i, [_vp, _i]),
    "fq_timing_get": (_i, [_vp, _i, _dp, _i64p, _dp]),
    "fq_timing_reset": (_i, [_vp]),
    "fq_stream_probe": (_i, [_sz, _i, _dp, _dp, _dp]),
    "fq_layer_passes": (_i, [_i, _i, _i, _i, ctypes.POINTER(_i), ctypes.POINTER(_i), ctypes.POINTER(_i),
                             ctypes.POINTER(ctypes.c_uint64)]),
}


def lib():
    """Load libforqaoa.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"libforqaoa.so not found at {LIB_PATH}: build it with `make -C for-qaoa_amd` "
                "(or __graft_entry__.build()); there is no CPU fallback"
            )
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def has_symbol(name):
    return getattr(lib(), name, None) is not None


def check(rc):
    if rc != FQ_OK:
        raise FQError(rc, lib().fq_last_error().decode(errors="replace"))
    return rc
