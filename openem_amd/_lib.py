"""ctypes binding of ``libopenem_hip.so`` (the C ABI declared in ``include/openem_hip.h``).

The library is built in-tree by ``__graft_entry__.build()`` (or ``openem_amd.build``). There is
no CPU fallback: a missing or unloadable library raises ``OpenEMLibraryError``.
"""

from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import numpy as np

LIB_NAME = "libopenem_hip.so"
LIB_PATH = Path(__file__).resolve().parent / LIB_NAME

OEM_OK, OEM_EINVAL, OEM_ENOMEM, OEM_EHIP, OEM_ESTATE, OEM_ENOTSUP = range(6)
KERNELS = ("update_h", "update_e", "dipole_e", "dipole_h", "disp_pre", "disp_post",
           "dft_snap", "dft_flush", "step_advance", "tfsf_h", "tfsf_e", "mode_h", "mode_e",
           "update_he", "step_tail")
K_COUNT = len(KERNELS)

#: every symbol include/openem_hip.h and include/openem_host.h declare (checked by the CPU
#: test suite)
EXPORTS = ("oem_version", "oem_last_error", "oem_device_count", "oem_ctx_create",
           "oem_ctx_destroy", "oem_set_axis", "oem_set_materials_lut",
           "oem_set_materials_dense", "oem_set_dispersion", "oem_add_plane_source",
           "oem_add_dipoles", "oem_add_tfsf_points", "oem_add_mode_points", "oem_add_flux_plane",
           "oem_add_field_box", "oem_add_time_monitor", "oem_read_time_monitor",
           "oem_add_flux_time_monitor", "oem_read_flux_time_monitor", "oem_prepare",
           "oem_run", "oem_profile_run", "oem_timed_run", "oem_step_index", "oem_read_monitor",
           "oem_read_field", "oem_write_field", "oem_set_chunk", "oem_set_step_form",
           "oem_launch_plan", "oem_set_energy_tables",
           "oem_energy", "oem_set_flux_weights", "oem_monitor_flux", "oem_monitor_flux_noalloc",
           "oem_set_overlap_weights",
           "oem_monitor_overlap", "oem_checkpoint_alloc", "oem_checkpoint_save",
           "oem_checkpoint_load", "oem_adjoint_setup", "oem_adjoint_objective",
           "oem_adjoint_record", "oem_adjoint_forward", "oem_adjoint_back", "oem_adjoint_gradient", "oem_pool_create",
           "oem_pool_destroy", "oem_pool_ctx_create", "oem_pool_attach", "oem_pool_launch",
           "oem_pool_ctx_done", "oem_pool_wait", "oem_pool_stats", "oem_host_factorize")


class OpenEMLibraryError(RuntimeError):
    """The HIP library is missing or failed to load."""


class OpenEMError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[oem status {code}] {msg}")
        self.code = code


class Dims(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32),
                ("n_steps", C.c_int32), ("pml_lo", C.c_int32 * 3), ("pml_hi", C.c_int32 * 3),
                ("complex_fields", C.c_int32), ("dft_batch", C.c_int32),
                ("ch", C.c_float), ("cw", C.c_float)]


_P = C.c_void_p
_SIGS = {
    "oem_version": ([], C.c_int),
    "oem_last_error": ([C.c_char_p, C.c_size_t], C.c_int),
    "oem_device_count": ([_P], C.c_int),
    "oem_ctx_create": ([_P, C.c_int32, _P], C.c_int),
    "oem_ctx_destroy": ([_P], C.c_int),
    "oem_set_axis": ([_P, C.c_int32] + [_P] * 15, C.c_int),
    "oem_set_materials_lut": ([_P, _P, C.c_int32, _P], C.c_int),
    "oem_set_materials_dense": ([_P, _P, _P], C.c_int),
    "oem_set_dispersion": ([_P, C.c_int32, _P, _P, _P, C.c_int32, _P, C.c_int32,
                            _P, _P, _P, _P, _P], C.c_int),
    "oem_add_plane_source": ([_P, C.c_int32, C.c_int32, C.c_int32, C.c_float, C.c_float,
                              _P, _P], C.c_int),
    "oem_add_dipoles": ([_P, C.c_int32, C.c_int32, _P, _P, _P, _P, C.c_int32, _P, C.c_int32],
                        C.c_int),
    "oem_add_tfsf_points": ([_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, C.c_int32, _P],
                            C.c_int),
    "oem_add_mode_points": ([_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P], C.c_int),
    "oem_add_flux_plane": ([_P, C.c_int32, C.c_int32, C.c_int32, _P, _P], C.c_int),
    "oem_add_field_box": ([_P, _P, _P, C.c_int32, _P, _P], C.c_int),
    "oem_add_time_monitor": ([_P, C.c_int32, _P, _P, _P, C.c_int32, C.c_int32, C.c_int32,
                              C.c_int32, _P], C.c_int),
    "oem_read_time_monitor": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_add_flux_time_monitor": ([_P, C.c_int32, C.c_int32, _P, _P, _P, _P, _P, C.c_int32, _P, _P,
                                   _P, _P, _P, C.c_int32, _P, _P, _P, _P, _P, C.c_double,
                                   C.c_int32, C.c_int32, C.c_int32, C.c_int32, _P], C.c_int),
    "oem_read_flux_time_monitor": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_prepare": ([_P], C.c_int),
    "oem_run": ([_P, C.c_int64], C.c_int),
    "oem_profile_run": ([_P, C.c_int64, _P, _P], C.c_int),
    "oem_timed_run": ([_P, C.c_int64, _P], C.c_int),
    "oem_step_index": ([_P, _P], C.c_int),
    "oem_read_monitor": ([_P, C.c_int32, _P, _P, C.c_size_t], C.c_int),
    "oem_read_field": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_write_field": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_set_chunk": ([_P, C.c_int32], C.c_int),
    "oem_set_step_form": ([_P, C.c_int32], C.c_int),
    "oem_launch_plan": ([_P, _P, _P], C.c_int),
    "oem_set_energy_tables": ([_P] + [_P] * 7 + [C.c_int32], C.c_int),
    "oem_energy": ([_P, _P], C.c_int),
    "oem_set_flux_weights": ([_P, C.c_int32, _P, _P], C.c_int),
    "oem_monitor_flux": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_monitor_flux_noalloc": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_set_overlap_weights": ([_P, C.c_int32, C.c_int32, _P], C.c_int),
    "oem_monitor_overlap": ([_P, C.c_int32, _P, C.c_size_t], C.c_int),
    "oem_checkpoint_alloc": ([_P, C.c_int32], C.c_int),
    "oem_checkpoint_save": ([_P, C.c_int32], C.c_int),
    "oem_checkpoint_load": ([_P, C.c_int32], C.c_int),
    "oem_adjoint_setup": ([_P, C.c_int32] + [_P] * 6 + [C.c_int32], C.c_int),
    "oem_adjoint_objective": ([_P, C.c_int32, C.c_int32, _P, _P, _P, _P], C.c_int),
    "oem_adjoint_record": ([_P, C.c_int64], C.c_int),
    "oem_adjoint_forward": ([_P, C.c_int64], C.c_int),
    "oem_adjoint_back": ([_P, C.c_int64], C.c_int),
    "oem_adjoint_gradient": ([_P, _P, C.c_size_t], C.c_int),
    "oem_pool_create": ([C.c_int32, C.c_uint64, C.c_int32, _P], C.c_int),
    "oem_pool_destroy": ([_P], C.c_int),
    "oem_pool_ctx_create": ([_P, _P, _P], C.c_int),
    "oem_pool_attach": ([_P, _P], C.c_int),
    "oem_pool_launch": ([_P, C.c_int32], C.c_int),
    "oem_pool_ctx_done": ([_P, _P, C.c_int32, _P], C.c_int),
    "oem_pool_wait": ([_P, C.c_int32], C.c_int),
    "oem_pool_stats": ([_P, _P, _P, _P, _P], C.c_int),
    "oem_host_factorize": ([C.c_int64, C.c_int32, _P, _P, _P, _P, C.c_int64, _P], C.c_int),
}

_lib = None


def factorize(cols, max_uniq: int):
    """Distinct rows over the columns ``cols`` (arrays of one length with 4- or 8-byte
    elements; an 8-byte column counts as two 32-bit words), numbered in order of first
    appearance: ``(inv int32 (n,), uniq uint32 (n_uniq, n_words))``, or ``None`` when there are
    more than ``max_uniq`` distinct rows."""
    lib = load()
    keep, ptrs, strides = [], [], []
    n = None
    for c in cols:
        a = np.ascontiguousarray(np.asarray(c).reshape(-1))
        if a.dtype.itemsize not in (4, 8):
            raise ValueError(f"factorize: {a.dtype} columns are not supported")
        if n is None:
            n = a.size
        elif a.size != n:
            raise ValueError("factorize: columns differ in length")
        keep.append(a)
        w = a.itemsize // 4
        for h in range(w):
            ptrs.append(a.ctypes.data + 4 * h)
            strides.append(w)
    pa = (C.c_void_p * len(ptrs))(*ptrs)
    sa = np.asarray(strides, np.int32)
    inv = np.empty(n, np.int32)
    cap = int(min(max_uniq, max(n, 1)))
    uniq = np.empty((cap, len(ptrs)), np.uint32)
    nu = C.c_int64(0)
    check(lib, lib.oem_host_factorize(n, len(ptrs), pa, ptr(sa), ptr(inv), ptr(uniq), cap,
                                      C.byref(nu)))
    if nu.value < 0:
        return None
    return inv, uniq[:nu.value].copy()


def load(path: str | os.PathLike | None = None):
    """Load (once) and return the library with argtypes set."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    if path is None and os.environ.get("OPENEM_HIP_LIB"):
        path = os.environ["OPENEM_HIP_LIB"]      # an alternative build of the same ABI
    p = Path(path) if path is not None else LIB_PATH
    if not p.exists():
        raise OpenEMLibraryError(
            f"{p} not found: build it with `python -c 'import __graft_entry__ as g; g.build()'`"
            " (hipcc --offload-arch=gfx950). There is no CPU fallback.")
    try:
        lib = C.CDLL(str(p), mode=C.RTLD_GLOBAL)
    except OSError as e:
        raise OpenEMLibraryError(f"cannot load {p}: {e}") from e
    for name, (args, res) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = res
    if lib.oem_version() != 1:
        raise OpenEMLibraryError(f"ABI version {lib.oem_version()} != 1")
    _lib = lib if (path is None or os.environ.get("OPENEM_HIP_LIB") == str(path)) else _lib
    return lib


def last_error(lib) -> str:
    buf = C.create_string_buffer(1024)
    lib.oem_last_error(buf, 1024)
    return buf.value.decode(errors="replace")


def check(lib, rc: int) -> None:
    if rc != OEM_OK:
        raise OpenEMError(rc, last_error(lib))


def ptr(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def device_count() -> int:
    lib = load()
    n = C.c_int32(0)
    check(lib, lib.oem_device_count(C.byref(n)))
    return int(n.value)
