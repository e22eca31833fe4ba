"""ctypes binding of libcrysplight_hip.so (include/crysplight.h).

The product path has no CPU fallback: if the shared object is missing or no HIP
device is present, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os
import re
from pathlib import Path

from . import params as P

HERE = Path(__file__).resolve().parent
# CRYSPLIGHT_HIP_LIB names another build of the library (the Julia binding reads the same)
LIB_PATH = Path(os.environ.get("CRYSPLIGHT_HIP_LIB", HERE / "libcrysplight_hip.so"))
HEADER = HERE.parent / "include" / "crysplight.h"

_lib = None

_vp = C.c_void_p
_i32, _i64, _u64, _f32 = C.c_int32, C.c_int64, C.c_uint64, C.c_float
_pbox, _pop, _pro = C.POINTER(P.Box), C.POINTER(P.OpticalParams), C.POINTER(P.Readout)
_ptb, _pem = C.POINTER(P.TimeBinning), C.POINTER(P.Emission)

SIGNATURES = {
    "crysp_create": (_vp, [C.c_int]),
    "crysp_destroy": (None, [_vp]),
    "crysp_last_error": (C.c_char_p, []),
    "crysp_abi_version": (C.c_int, []),
    "crysp_transport_records": (C.c_int, [_vp, _pbox, _pop, _pro, _vp, _vp, _vp, _vp, _i32,
                                          _i64, _u64, _u64, _i32, _pem, _vp, _vp, _vp, _vp]),
    "crysp_transport_accumulate": (C.c_int, [_vp, _pbox, _pop, _pro, _ptb, _vp, _vp, _vp, _vp,
                                             _i32, _i64, _u64, _u64, _i32, _pem, _vp, _vp, _vp]),
    "crysp_run_events": (C.c_int, [_vp, _pbox, _pop, _pro, _vp, _vp, _vp, _i64, _i64, _u64,
                                   _u64, _i32, _pem, _vp, _vp, _vp, _vp]),
    "crysp_set_schedule": (C.c_int, [_vp, _i32, _i32]),
    "crysp_get_schedule": (C.c_int, [_vp, _vp, _vp]),
    "crysp_set_queue": (C.c_int, [_vp, _i32, _i32]),
    "crysp_set_event_sessions": (C.c_int, [_vp, _i32, _i32]),
    "crysp_set_wave_timeline": (C.c_int, [_vp, _i32]),
    "crysp_get_wave_timeline": (C.c_int, [_vp, _vp, _i64, _vp, _vp]),
    "crysp_autotune_events": (C.c_int, [_vp, _pbox, _pop, _pro, _vp, _vp, _vp, _i64, _i64, _u64,
                                        _i32, _pem, _vp, _i32, _vp, _i32, _i32, _vp, _vp, _vp]),
    "crysp_autotune_events_queue": (C.c_int, [_vp, _pbox, _pop, _pro, _vp, _vp, _vp, _i64, _i64,
                                              _u64, _i32, _pem, _vp, _i32, _vp, _i32, _vp, _i32,
                                              _vp, _i32, _i32, _vp, _vp]),
    "crysp_autotune_accumulate": (C.c_int, [_vp, _pbox, _pop, _pro, _ptb, _vp, _vp, _vp, _vp,
                                            _i32, _i64, _u64, _i32, _pem, _vp, _i32, _vp, _i32,
                                            _vp, _i32, _vp, _i32, _i32, _vp, _vp]),
    "crysp_run_events_noalloc": (C.c_int, [_vp, _pbox, _pop, _pro, _vp, _vp, _vp, _i64, _i64,
                                           _u64, _u64, _i32, _pem, _vp, _vp, _vp, _vp]),
    "crysp_debug_draws": (C.c_int, [_vp, _u64, _u64, _i64, _i32, _vp]),
    "crysp_debug_math": (C.c_int, [_vp, _i32, _vp, _i64, _vp]),
    "crysp_set_timing": (C.c_int, [_vp, _i32]),
    "crysp_kernel_time": (C.c_int, [_vp, _vp, _vp, _i32]),
    "crysp_set_op_counting": (C.c_int, [_vp, _i32]),
    "crysp_get_op_counts": (C.c_int, [_vp, _vp, _i32]),
}


class CryspError(RuntimeError):
    pass


def header_symbols() -> list[str]:
    """Function names declared in include/crysplight.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(crysp_[a-z0-9_]+)\s*\(", text)))


def _share_hip_runtime() -> None:
    """One HIP runtime per process.  torch's ROCm wheel ships its own libamdhip64
    (SONAME libamdhip64.so.7) and loads it as "libamdhip64.so"; when torch is imported
    first, this library's DT_NEEDED libamdhip64.so.7 binds to that copy.  Loaded the
    other way round the process would hold two runtimes and torch could not see the GPU."""
    try:
        import torch  # noqa: F401
    except Exception:
        pass


def load() -> C.CDLL:
    """Load the shared object (raises if it was not built)."""
    global _lib
    if _lib is not None:
        return _lib
    _share_hip_runtime()
    if not LIB_PATH.exists():
        raise CryspError(f"{LIB_PATH} is missing: build it with `python -m crysplight_amd.build`"
                         " (there is no CPU fallback)")
    lib = C.CDLL(str(LIB_PATH))
    for name, (res, args) in SIGNATURES.items():
        if hasattr(lib, name):
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = load().crysp_last_error().decode(errors="replace")
        raise CryspError(f"{what} failed ({rc}): {msg}")


class Context:
    """One library context (HIP stream, device buffers, graphs, schedule) on a device.
    Replaces the ArrayT backend selector (src/kernel.jl:319)."""

    def __init__(self, device: int = 0):
        lib = load()
        h = lib.crysp_create(int(device))
        if not h:
            raise CryspError("crysp_create failed: " + lib.crysp_last_error().decode())
        self._h = h
        self.device = device

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None):
            load().crysp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # schedule / timing -------------------------------------------------------
    def set_schedule(self, wg: int, ppt: int):
        check(load().crysp_set_schedule(self._h, wg, ppt), "crysp_set_schedule")

    def set_queue(self, chunk: int, refill_min: int):
        check(load().crysp_set_queue(self._h, chunk, refill_min), "crysp_set_queue")

    def set_wave_timeline(self, enable: bool):
        check(load().crysp_set_wave_timeline(self._h, int(bool(enable))), "crysp_set_wave_timeline")

    def wave_timeline(self):
        """Seconds between each wave's end and the end of the last wave of the last launch
        (waves that never ran dropped), from the device wall clock."""
        import numpy as np
        n, hz = C.c_int64(), C.c_double()
        check(load().crysp_get_wave_timeline(self._h, None, 0, C.byref(n), C.byref(hz)),
              "crysp_get_wave_timeline")
        out = np.zeros(max(n.value, 0), np.uint64)
        check(load().crysp_get_wave_timeline(self._h, out.ctypes.data, n.value, C.byref(n),
                                             C.byref(hz)), "crysp_get_wave_timeline")
        out = out[out > 0].astype(np.float64)
        return (out.max() - out) / hz.value if out.size else out

    def set_event_sessions(self, enable: bool, piece: int = 1024):
        """Event mode: workgroups own whole events (one write per map row) when the
        deposits are ordered by event, or per-detection atomics (enable False)."""
        check(load().crysp_set_event_sessions(self._h, int(bool(enable)), int(piece)),
              "crysp_set_event_sessions")

    def get_schedule(self) -> tuple[int, int]:
        a, b = C.c_int32(), C.c_int32()
        check(load().crysp_get_schedule(self._h, C.byref(a), C.byref(b)), "crysp_get_schedule")
        return a.value, b.value

    def set_timing(self, on: bool):
        check(load().crysp_set_timing(self._h, int(on)), "crysp_set_timing")

    def kernel_time(self, reset: bool = False) -> tuple[float, int]:
        ms, n = C.c_double(), C.c_int64()
        check(load().crysp_kernel_time(self._h, C.byref(ms), C.byref(n), int(reset)),
              "crysp_kernel_time")
        return ms.value, n.value

    def set_op_counting(self, on: bool):
        check(load().crysp_set_op_counting(self._h, int(on)), "crysp_set_op_counting")

    def op_counts(self, reset: bool = False):
        import numpy as np
        out = np.zeros(10, np.int64)
        check(load().crysp_get_op_counts(self._h, out.ctypes.data, int(reset)),
              "crysp_get_op_counts")
        return out


_default: dict[int, Context] = {}


def default_context(device: int = 0) -> Context:
    ctx = _default.get(device)
    if ctx is None:
        ctx = _default[device] = Context(device)
    return ctx


def ptr(a) -> int | None:
    """Address of a numpy array or a torch tensor (device or host); None for None."""
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return a.data_ptr()
    return a.ctypes.data


def env_flag(name: str) -> bool:
    return os.environ.get(name, "") not in ("", "0")


def _is_torch(a) -> bool:
    return hasattr(a, "data_ptr") and hasattr(a, "is_contiguous")


def check_array(a, name: str, dtype, shape=None, *, numel=None, device: int | None = None):
    """Validate an array that crosses the C-ABI as a raw pointer: dtype, C-contiguity,
    shape (or element count) and, for a torch tensor on the GPU, its device.  The library
    trusts the sizes it is given, so a wrong dtype or a short buffer would otherwise turn
    into out-of-bounds device reads or writes.  Raises ValueError."""
    import numpy as np
    if a is None:
        raise ValueError(f"{name}: missing array")
    want = np.dtype(dtype)
    if _is_torch(a):
        import torch
        got = torch.empty(0, dtype=a.dtype).numpy().dtype if a.dtype != torch.bfloat16 else None
        if got != want:
            raise ValueError(f"{name}: dtype {a.dtype}, expected {want}")
        if not a.is_contiguous():
            raise ValueError(f"{name}: tensor is not contiguous")
        if a.device.type == "cuda" and device is not None and a.device.index != device:
            raise ValueError(f"{name}: tensor on cuda:{a.device.index}, the context is on "
                             f"device {device}")
        shp, n = tuple(a.shape), a.numel()
    else:
        if not isinstance(a, np.ndarray):
            raise ValueError(f"{name}: expected a numpy array or a torch tensor")
        if a.dtype != want:
            raise ValueError(f"{name}: dtype {a.dtype}, expected {want}")
        if not a.flags["C_CONTIGUOUS"]:
            raise ValueError(f"{name}: array is not C-contiguous")
        shp, n = a.shape, a.size
    if shape is not None and tuple(shp) != tuple(shape):
        raise ValueError(f"{name}: shape {tuple(shp)}, expected {tuple(shape)}")
    if numel is not None and n != numel:
        raise ValueError(f"{name}: {n} elements, expected {numel}")


def ready(*arrays) -> None:
    """The library works on its own HIP stream.  Device tensors handed to it must hold
    their final contents when the call starts, so work that torch still has queued for
    them (a fill, a copy, the kernel that produced them) is waited for first."""
    import sys
    if "torch" not in sys.modules:
        return
    import torch
    devs = {a.device for a in arrays if _is_torch(a) and a.device.type == "cuda"}
    for d in devs:
        torch.cuda.current_stream(d).synchronize()
