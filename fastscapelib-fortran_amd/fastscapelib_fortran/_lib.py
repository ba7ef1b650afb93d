"""ctypes binding of libfastscape_hip.so (the C ABI declared in include/fastscape_hip.h).

The library is the product path: there is no CPU fallback.  Importing this module on a machine
where the library is missing or cannot initialise HIP raises immediately.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libfastscape_hip.so")

# field ids (include/fastscape_hip.h)
FSX_H, FSX_B, FSX_U, FSX_PRECIP, FSX_ETOT, FSX_ERATE, FSX_AREA, FSX_LAKE_DEPTH, FSX_SEDFLUX, FSX_LENGTH, \
    FSX_HWATER, FSX_KF, FSX_KD, FSX_FMIX = range(14)
FSX_REC, FSX_NDON, FSX_DON, FSX_ORDER, FSX_MREC, FSX_MNREC, FSX_MSTACK = range(7)
FSX_SLOPE, FSX_CURVATURE, FSX_CHI, FSX_CATCHMENT = range(4)  # fsx_copy_derived
FSX_STRATI_REFLECTORS, FSX_STRATI_ERODED, FSX_STRATI_SLOPES, FSX_STRATI_SHORE = range(4)  # fsx_strati_read

# phase timers, in the library's order (include/fastscape_hip.h, fsx_get_timers)
PHASES = ["uplift", "receivers", "donors", "lm_labels", "lm_basins", "lm_walk", "lm_edges", "lm_sort", "lm_mst",
          "lm_orient", "lm_correct", "levels", "area_prep", "area_sweep", "spl_prep", "spl_fact", "spl_sweep",
          "spl_post", "diffusion", "mfd_route", "mfd_spl", "advect", "marine", "strati"]

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


FSX_OK, FSX_EINVAL, FSX_EHIP, FSX_ESTEP, FSX_ENOMEM = 0, -1, -2, -3, -4
FSX_INDEXX_NOALLOC_MAX = 524288  # keys per fsx_indexx_noalloc call (include/fastscape_hip.h)


class FastscapeError(RuntimeError):
    """A C-ABI call failed; the message is the library's last error."""


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with `make -C fastscapelib-fortran_amd/csrc` "
                          "(or __graft_entry__.build()); there is no CPU fallback")
    lib = C.CDLL(LIB_PATH)
    sig = {
        "fsx_create": [C.c_int, _ip, _ip, C.POINTER(_vp)],
        "fsx_destroy": [_vp],
        "fsx_set_xl_yl": [_vp, C.c_int, C.c_double, C.c_double],
        "fsx_set_dt": [_vp, C.c_int, C.c_double],
        "fsx_set_bc": [_vp, C.c_int, C.c_int],
        "fsx_set_erosional_parameters": [_vp, C.c_int, _vp, C.c_double, C.c_double, C.c_double, _vp, C.c_double,
                                         C.c_double, C.c_double, C.c_double, C.c_int],
        "fsx_set_newton_tol": [_vp, C.c_int, C.c_double],
        "fsx_set_marine_parameters": [_vp, C.c_int] + [C.c_double] * 9,
        "fsx_strati": [_vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double],
        "fsx_strati_set_active": [_vp, C.c_int, C.c_int],
        "fsx_strati_read": [_vp, C.c_int, C.c_int, _dp],
        "fsx_init_h": [_vp, C.c_int, _vp, C.c_int],
        "fsx_set_field": [_vp, C.c_int, C.c_int, _vp, C.c_int],
        "fsx_set_all_layers": [_vp, C.c_int, _vp, C.c_int],
        "fsx_reset_cumulative_erosion": [_vp, C.c_int],
        "fsx_set_active": [_vp, C.c_int, C.c_int],
        "fsx_execute_steps": [_vp, C.c_int, _vp, C.c_int],
        "fsx_synchronize": [_vp],
        "fsx_query": [_vp],
        "fsx_copy_field": [_vp, C.c_int, C.c_int, _vp, C.c_int],
        "fsx_copy_ifield": [_vp, C.c_int, C.c_int, _vp, C.c_int],
        "fsx_copy_derived": [_vp, C.c_int, C.c_int, _vp, C.c_int],
        "fsx_set_v": [_vp, C.c_int, _vp, _vp, C.c_int],
        "fsx_copy_derived_batch": [_vp, C.c_int, C.POINTER(C.c_void_p), C.c_int],
        "fsx_get_fluxes_batch": [_vp, _vp],
        "fsx_get_fluxes": [_vp, C.c_int, _dp],
        "fsx_get_step": [_vp, C.c_int, _ip],
        "fsx_get_sizes": [_vp, C.c_int, _ip, _ip],
        "fsx_route": [_vp, _vp, C.c_int],
        "fsx_indexx": [C.c_int, _dp, _ip],
        "fsx_indexx_noalloc": [C.c_int, _vp, _vp, _vp, C.c_int],
        "fsx_get_stats": [_vp, C.c_int, C.POINTER(C.c_longlong)],
        "fsx_memory": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
        "fsx_set_profiling": [_vp, C.c_int],
        "fsx_get_timers": [_vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong), C.c_int],
        "fsx_reset_timers": [_vp],
        "fsx_create_pool": [C.c_longlong, C.c_int, C.POINTER(_vp)],
        "fsx_add_instance": [_vp, C.c_int, C.c_int, _ip],
        "fsx_remove_instance": [_vp, C.c_int],
        "fsx_pool_info": [_vp, C.POINTER(C.c_longlong), C.POINTER(C.c_longlong), C.POINTER(C.c_int)],
        "fsx_device_memory": [C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)],
        "fsx_set_graphs": [_vp, C.c_int],
        "fsx_time_phase": [_vp, C.c_int],
        "fsx_autotune": [_vp, C.c_int, _ip],
        "fsx_get_tuning": [_vp, _ip],
        "fsx_set_tuning": [_vp, _ip],
    }
    for name, args in sig.items():
        f = getattr(lib, name)
        f.argtypes = args
        f.restype = C.c_int
    lib.fsx_last_error.restype = C.c_char_p
    lib.fsx_last_error.argtypes = []
    lib.fsx_version.restype = C.c_char_p
    return lib


lib = _load()


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib.fsx_last_error().decode(errors="replace")
        raise FastscapeError(f"{what}: {msg} (status {rc})" if what else f"{msg} (status {rc})")


def host_f64(x, n: int, name: str) -> np.ndarray:
    """A contiguous float64 host copy of x with at least n entries (broadcast from a scalar)."""
    a = np.asarray(x, dtype=np.float64)
    if a.ndim == 0:
        a = np.full(n, float(a))
    a = np.ascontiguousarray(a.reshape(-1))
    if a.size < n:
        raise ValueError(f"{name}: expected at least {n} values, got {a.size}")
    return a


def ptr(a: np.ndarray):
    return C.c_void_p(a.ctypes.data)


def dev_ptr(t) -> C.c_void_p:
    """Device pointer of a torch tensor (float64/int32, contiguous, on the GPU)."""
    if not t.is_contiguous():
        raise ValueError("device tensor must be contiguous")
    return C.c_void_p(t.data_ptr())


class DeviceBuffer:
    """Device memory from the HIP runtime the library itself is linked against (no torch: a PyTorch
    build may bundle a HIP runtime of its own), for callers that hand device arrays to the C ABI."""

    _hip = None

    def __init__(self, nbytes: int):
        if DeviceBuffer._hip is None:
            h = C.CDLL("libamdhip64.so.7")  # the runtime libfastscape_hip.so links (already loaded)
            h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
            h.hipFree.argtypes = [C.c_void_p]
            h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
            h.hipDeviceSynchronize.argtypes = []
            for f in (h.hipMalloc, h.hipFree, h.hipMemcpy, h.hipDeviceSynchronize):
                f.restype = C.c_int
            DeviceBuffer._hip = h
        self.nbytes = int(nbytes)
        self.ptr = C.c_void_p()
        rc = DeviceBuffer._hip.hipMalloc(C.byref(self.ptr), max(1, self.nbytes))
        if rc != 0:
            raise FastscapeError(f"hipMalloc({self.nbytes}) failed with status {rc}")

    def upload(self, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        if a.nbytes > self.nbytes:
            raise ValueError("array larger than the buffer")
        rc = DeviceBuffer._hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, 1)  # host to device
        if rc != 0:
            raise FastscapeError(f"hipMemcpy failed with status {rc}")
        return self

    def download(self, a: np.ndarray) -> np.ndarray:
        if not a.flags.c_contiguous or a.nbytes > self.nbytes:
            raise ValueError("target must be contiguous and fit the buffer")
        rc = DeviceBuffer._hip.hipMemcpy(a.ctypes.data, self.ptr, a.nbytes, 2)  # device to host
        if rc != 0:
            raise FastscapeError(f"hipMemcpy failed with status {rc}")
        return a

    def free(self):
        if self.ptr.value:
            DeviceBuffer._hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def device_synchronize() -> None:
    """hipDeviceSynchronize on the library's HIP runtime."""
    DeviceBuffer(0).free()  # loads the runtime handle
    rc = DeviceBuffer._hip.hipDeviceSynchronize()
    if rc != 0:
        raise FastscapeError(f"hipDeviceSynchronize failed with status {rc}")


def device_copy(dst: DeviceBuffer, src: DeviceBuffer, nbytes: int) -> None:
    rc = DeviceBuffer._hip.hipMemcpy(dst.ptr, src.ptr, int(nbytes), 3)  # device to device
    if rc != 0:
        raise FastscapeError(f"hipMemcpy failed with status {rc}")


def indexx(keys) -> np.ndarray:
    """loc_min_3_indexx on the device: the reference's 1-based sort permutation of keys."""
    k = np.ascontiguousarray(np.asarray(keys, dtype=np.float64).reshape(-1))
    out = np.empty(k.size, np.int32)
    check(lib.fsx_indexx(k.size, k.ctypes.data_as(_dp), out.ctypes.data_as(_ip)), "fsx_indexx")
    return out


def device_memory():
    """(free, total) device memory in bytes."""
    f, t = C.c_longlong(), C.c_longlong()
    check(lib.fsx_device_memory(C.byref(f), C.byref(t)), "device_memory")
    return f.value, t.value

