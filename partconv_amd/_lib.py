"""Loader for libpartconv_amd.so (the C ABI declared in include/partconv_amd.h).

The library is built in-tree by partconv_amd.build; nothing here falls back to
a CPU implementation: if the library is missing, loading raises.
"""
from __future__ import annotations

import atexit
import ctypes
import pathlib
import sys

PKG = pathlib.Path(__file__).resolve().parent
LIB_PATH = PKG / "libpartconv_amd.so"

PC_OK = 0
PC_BATCHED = 0
PC_PLAIN = 1

_lib = None


class PcMultiInfo(ctypes.Structure):
    _fields_ = [("FS", ctypes.c_int), ("buffer_size", ctypes.c_int), ("max_block_size", ctypes.c_int),
                ("verbosity", ctypes.c_int), ("numInputChannels", ctypes.c_int),
                ("numOutputChannels", ctypes.c_int), ("numPCs", ctypes.c_int),
                ("num_levels", ctypes.c_int), ("frameNum", ctypes.c_ulonglong)]


class PcStats(ctypes.Structure):
    _fields_ = [("ticks", ctypes.c_ulonglong), ("deadline_misses", ctypes.c_ulonglong),
                ("p50_us", ctypes.c_double), ("p99_us", ctypes.c_double), ("max_us", ctypes.c_double),
                ("mean_us", ctypes.c_double), ("period_us", ctypes.c_double),
                ("persistent_ticks", ctypes.c_ulonglong), ("stalls", ctypes.c_ulonglong)]


class PcInstance(ctypes.Structure):
    _fields_ = [("mute", ctypes.c_int), ("in_ch", ctypes.c_int), ("out_ch", ctypes.c_int),
                ("scale", ctypes.c_float), ("M", ctypes.c_int), ("N", ctypes.POINTER(ctypes.c_int)),
                ("P", ctypes.POINTER(ctypes.c_int)), ("ir", ctypes.POINTER(ctypes.c_float)),
                ("ir_len", ctypes.c_int)]


# name -> (restype, argtypes); every symbol declared in include/partconv_amd.h
_vp = ctypes.c_void_p
_ip = ctypes.POINTER(ctypes.c_int)
_fp = ctypes.POINTER(ctypes.c_float)
_dp = ctypes.POINTER(ctypes.c_double)
SIGNATURES = {
    "pc_last_error": (ctypes.c_char_p, []),
    "pc_shutdown": (None, []),
    "pc_read_ir": (ctypes.c_int, [ctypes.c_char_p, _fp, ctypes.c_longlong, ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "pc_device_name": (ctypes.c_int, [ctypes.c_char_p, ctypes.c_int]),
    "pc_greedy_partition": (ctypes.c_int, [_ip, ctypes.c_int, ctypes.c_int, _ip, _ip, _ip]),
    "pcm_create": (_vp, [ctypes.c_int, ctypes.c_int]),
    "pcm_set_mode": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_setup": (ctypes.c_int, [_vp, ctypes.c_char_p, ctypes.c_int, ctypes.c_int, ctypes.c_double]),
    "pcm_setup_mem": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(PcInstance)]),
    "pcm_run": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "pcm_run_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_longlong, ctypes.c_int]),
    "pcm_reset": (ctypes.c_int, [_vp]),
    "pcm_cleanup": (ctypes.c_int, [_vp]),
    "pcm_destroy": (None, [_vp]),
    "pcm_get_info": (ctypes.c_int, [_vp, ctypes.POINTER(PcMultiInfo)]),
    "pcm_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(PcStats)]),
    "pcm_clear_stats": (ctypes.c_int, [_vp]),
    "pcm_get_latencies": (ctypes.c_longlong, [_vp, _dp, ctypes.c_longlong]),
    "pcm_instance_partition": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, ctypes.c_int]),
    "pcm_probe": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_trace": (ctypes.c_int, [_vp, ctypes.c_longlong]),
    "pcm_set_persistent_workers": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_set_host_session": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_set_release_window": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_process": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_longlong]),
    "pcm_setup_max": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int,
                                     ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.c_int]),
    "pcm_run_channels": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_void_p), ctypes.POINTER(ctypes.c_void_p)]),
    "pcm_workers_info": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int)]),
    "pcm_release_info": (ctypes.c_int, [_vp]),
    "pcm_stall_level": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "pcm_set_partial_width": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_set_level_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcm_get_tuning": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pcm_trace_read": (ctypes.c_longlong, [_vp, ctypes.POINTER(ctypes.c_ulonglong), ctypes.c_longlong]),
    "pcm_probe_read": (ctypes.c_int, [_vp, _dp, ctypes.POINTER(ctypes.c_longlong), _dp]),
    "pcf_create": (_vp, [ctypes.c_int, ctypes.c_int]),
    "pcf_setup": (ctypes.c_int, [_vp, ctypes.c_int, _ip, _ip, _fp, ctypes.c_int, ctypes.c_int,
                                 ctypes.c_char_p]),
    "pcf_run_level": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcf_sync_levels": (ctypes.c_int, [_vp]),
    "pcf_sync_levels_relaxed": (ctypes.c_int, [_vp]),
    "pcf_start_sync_depth": (ctypes.c_int, [_vp]),
    "pcf_end_sync_depth": (ctypes.c_int, [_vp]),
    "pcf_input": (ctypes.c_int, [_vp, _fp]),
    "pcf_output": (ctypes.c_int, [_vp, _fp]),
    "pcf_output_relaxed": (ctypes.c_int, [_vp, _fp]),
    "pcf_reset": (ctypes.c_int, [_vp]),
    "pcf_bench": (ctypes.c_double, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_float, _dp]),
    "pcf_cleanup": (ctypes.c_int, [_vp]),
    "pcf_destroy": (None, [_vp]),
    "pcf_get_levels": (ctypes.c_int, [_vp, _ip, _ip, _ip, ctypes.c_int]),
    "pcf_get_fs": (ctypes.c_int, [_vp]),
    "pcf_set_fs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcc_create": (_vp, [ctypes.c_int, _ip, _ip, ctypes.c_int]),
    "pcc_levels": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int]),
    "pcc_set_filter": (ctypes.c_int, [_vp, _vp, _vp]),
    "pcc_forward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_backward": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_backward_scaled": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_mse": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_forward_noalloc": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_set_budget": (ctypes.c_int, [_vp, ctypes.c_longlong]),
    "pcc_recomputed_levels": (ctypes.c_longlong, [_vp]),
    "pcc_tune": (ctypes.c_int, [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp]),
    "pcc_level_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "pcc_set_level_tuning": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pcc_probe": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcc_probe_kernel": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_int]),
    "pcc_set_level_streams": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcc_set_call_graphs": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcc_call_graph_stats": (ctypes.c_int, [_vp] + [ctypes.POINTER(ctypes.c_longlong)] * 4),
    "pcc_probe_read": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong),
                                      ctypes.POINTER(ctypes.c_double)]),
    "pcc_destroy": (None, [_vp]),
    "pcx_create": (_vp, [ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
    "pcx_set_filters": (ctypes.c_int, [_vp, _vp, _vp]),
    "pcx_forward": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int, _vp]),
    "pcx_destroy": (None, [_vp]),
    "pcxr_create": (_vp, [ctypes.c_int, ctypes.c_int, ctypes.c_int, _ip, _ip, ctypes.c_int, ctypes.c_int,
                          ctypes.c_int]),
    "pcxr_set_filters": (ctypes.c_int, [_vp, _vp]),
    "pcxr_run_device": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_longlong]),
    "pcxr_run": (ctypes.c_int, [_vp, _vp, _vp, ctypes.c_int]),
    "pcxr_reset": (ctypes.c_int, [_vp]),
    "pcxr_get_stats": (ctypes.c_int, [_vp, ctypes.POINTER(PcStats)]),
    "pcxr_clear_stats": (ctypes.c_int, [_vp]),
    "pcxr_partition": (ctypes.c_int, [_vp, _ip, _ip, ctypes.c_int]),
    "pcxr_stall_level": (ctypes.c_int, [_vp, ctypes.c_int, ctypes.c_double]),
    "pcxr_set_release_window": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcxr_destroy": (None, [_vp]),
    "pcxr_probe": (ctypes.c_int, [_vp, ctypes.c_int]),
    "pcxr_probe_read": (ctypes.c_int, [_vp, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_longlong),
                                       ctypes.POINTER(ctypes.c_double)]),
}


def lib():
    """Load the HIP library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m partconv_amd.build`")
        L = ctypes.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
        # registered after torch (imported first by every caller that drives a
        # GPU), so it runs before torch's own exit hooks
        atexit.register(shutdown)
    return _lib


_shut = False


def shutdown():
    """Destroy every live handle (resident kernels stopped, launcher threads
    joined) and release the library's device resources: pc_shutdown().  Runs
    at interpreter exit; afterwards destroy() is a no-op."""
    global _shut
    if _shut:
        return
    _shut = True
    if _lib is not None:
        _lib.pc_shutdown()


def destroy(fn_name, handle):
    """Destroy a handle from a __del__: a no-op once the library has shut down
    or while the interpreter finalises (the library's exit hook owns teardown
    then), so no destructor ever runs against a torn-down runtime."""
    try:
        if _shut or not handle or _lib is None or sys.is_finalizing():
            return
        getattr(_lib, fn_name)(handle)
    except Exception:
        pass


def last_error() -> str:
    return lib().pc_last_error().decode(errors="replace")
