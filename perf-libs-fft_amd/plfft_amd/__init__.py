"""plfft_amd -- Python binding of the MI355X PLFFT C ABI (include/plfft.h).

The library ``libplfft_amd.so`` next to this file implements the reference's
23-function C interface (include/plfft.h:212-532 of arm/perf-libs-fft) plus the
GPU extensions of include/plfft_gpu.h. This module only wraps it with ctypes:
every transform runs in the HIP kernels of the library, and importing fails
loudly when the library has not been built.

Buffers passed to :meth:`Plan.execute` may be torch CUDA tensors, raw device
pointers (ints) or numpy arrays (host memory, staged by the library).
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libplfft_amd.so")

# --- enums, values of include/plfft.h -------------------------------------
PLFFT_OK = 0
PLFFT_NULL_ARGUMENT = 1
PLFFT_INVALID_SIZE = 2
PLFFT_UNSUPPORTED_BATCH = 3
PLFFT_INVALID_TRANSFORM_KIND = 4
PLFFT_INVALID_DATA_TYPE = 5
PLFFT_INVALID_IO_ALIAS = 6
PLFFT_PLAN_CREATION_FAILED = 7
PLFFT_INVALID_SME_MODE = 8
PLFFT_SME_NOT_ENABLED = 9
PLFFT_SME_UNSUPPORTED_R2R = 10
PLFFT_FP16_NOT_AVAILABLE = 11
PLFFT_SME_NOT_AVAILABLE = 12
PLFFT_SME_UNSUPPORTED_FIXED_POINT = 13
PLFFT_INVALID_PLANNER_RIGOR = 14

PLFFT_FORWARD = -1
PLFFT_BACKWARD = 1

R2R_KINDS = {
    "dct1": 0, "dct2": 1, "dct3": 2, "dct4": 3,
    "dst1": 4, "dst2": 5, "dst3": 6, "dst4": 7,
    "dht": 8, "r2hc": 9, "hc2r": 10,
}

DATA_TYPES = {"f32": 0, "f64": 1, "f16": 2, "q7": 3, "q15": 4, "bf16": 5}
REAL_BYTES = {"f32": 4, "f64": 8, "f16": 2, "bf16": 2, "q7": 1, "q15": 2}

PLFFT_IO_NO_ALIAS = 0
PLFFT_IO_MAY_ALIAS = 1
PLFFT_SME_DISABLED = 0
PLFFT_SME_ENABLED = 1
PLFFT_ESTIMATE, PLFFT_MEASURE, PLFFT_PATIENT, PLFFT_EXHAUSTIVE = 0, 1, 2, 3

# include/plfft_gpu.h: bytes of the fixed scratch behind plfft_plan_execute_noalloc
PLFFT_NOALLOC_SCRATCH_BYTES = 64 << 20

# Every symbol declared in include/plfft.h and include/plfft_gpu.h.
EXPORTS = [
    "plfft_status_to_string",
    "plfft_config_create", "plfft_config_create_r2c", "plfft_config_create_c2r",
    "plfft_config_create_r2r", "plfft_config_destroy",
    "plfft_config_set_data_type", "plfft_config_get_data_type",
    "plfft_config_set_io_alias", "plfft_config_get_io_alias",
    "plfft_config_set_sme_mode", "plfft_config_get_sme_mode",
    "plfft_config_set_rigor", "plfft_config_get_rigor",
    "plfft_config_set_batch", "plfft_config_get_howmany",
    "plfft_config_get_istride", "plfft_config_get_idist",
    "plfft_config_get_ostride", "plfft_config_get_odist",
    "plfft_plan_create_from_config", "plfft_plan_destroy", "plfft_plan_execute",
    # plfft_gpu.h
    "plfft_plan_get_workspace_size", "plfft_plan_execute_async",
    "plfft_plan_execute_noalloc",
    "plfft_plan_execute_adjoint", "plfft_plan_execute_adjoint_async",
    "plfft_plan_describe", "plfft_last_error",
    "plfft_mixed_plan_create", "plfft_mixed_plan_destroy",
    "plfft_mixed_plan_get_workspace_size", "plfft_mixed_plan_get_span",
    "plfft_mixed_plan_execute", "plfft_mixed_plan_execute_async",
    "plfft_mixed_plan_describe",
]

_lib = None


class PlfftError(RuntimeError):
    def __init__(self, status: int, what: str = ""):
        self.status = status
        msg = status_to_string(status) if _lib is not None else str(status)
        detail = last_error()
        super().__init__(f"{what}: {msg} ({status})" + (f": {detail}" if detail else ""))


def load() -> ctypes.CDLL:
    """Load libplfft_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"{LIB_PATH} is missing: build it with `make -C perf-libs-fft_amd` "
            "(or __graft_entry__.build())")
    lib = ctypes.CDLL(LIB_PATH)
    vp, i64, i32 = ctypes.c_void_p, ctypes.c_int64, ctypes.c_int
    pp = ctypes.POINTER(ctypes.c_void_p)
    sig = {
        "plfft_status_to_string": (ctypes.c_char_p, [i32]),
        "plfft_config_create": (i32, [pp, i64, i32]),
        "plfft_config_create_r2c": (i32, [pp, i64]),
        "plfft_config_create_c2r": (i32, [pp, i64]),
        "plfft_config_create_r2r": (i32, [pp, i64, i32]),
        "plfft_config_destroy": (None, [vp]),
        "plfft_config_set_data_type": (None, [vp, i32]),
        "plfft_config_get_data_type": (i32, [vp]),
        "plfft_config_set_io_alias": (None, [vp, i32]),
        "plfft_config_get_io_alias": (i32, [vp]),
        "plfft_config_set_sme_mode": (None, [vp, i32]),
        "plfft_config_get_sme_mode": (i32, [vp]),
        "plfft_config_set_rigor": (None, [vp, i32]),
        "plfft_config_get_rigor": (i32, [vp]),
        "plfft_config_set_batch": (None, [vp, i64, i64, i64, i64, i64]),
        "plfft_config_get_howmany": (i64, [vp]),
        "plfft_config_get_istride": (i64, [vp]),
        "plfft_config_get_idist": (i64, [vp]),
        "plfft_config_get_ostride": (i64, [vp]),
        "plfft_config_get_odist": (i64, [vp]),
        "plfft_plan_create_from_config": (i32, [pp, vp]),
        "plfft_plan_destroy": (None, [vp]),
        "plfft_plan_execute": (None, [vp, vp, vp]),
        "plfft_plan_get_workspace_size": (ctypes.c_size_t, [vp]),
        "plfft_plan_execute_async": (i32, [vp, vp, vp, vp, vp]),
        "plfft_plan_execute_noalloc": (i32, [vp, vp, vp, vp]),
        "plfft_plan_execute_adjoint": (None, [vp, vp, vp]),
        "plfft_plan_execute_adjoint_async": (i32, [vp, vp, vp, vp, vp]),
        "plfft_plan_describe": (ctypes.c_size_t, [vp, ctypes.c_char_p, ctypes.c_size_t]),
        "plfft_last_error": (ctypes.c_char_p, []),
        "plfft_mixed_plan_create": (i32, [pp, i64, vp, vp, vp, i32, i32, i32]),
        "plfft_mixed_plan_destroy": (None, [vp]),
        "plfft_mixed_plan_get_workspace_size": (ctypes.c_size_t, [vp]),
        "plfft_mixed_plan_get_span": (i64, [vp, i32]),
        "plfft_mixed_plan_execute": (i32, [vp, vp, vp]),
        "plfft_mixed_plan_execute_async": (i32, [vp, vp, vp, vp, vp]),
        "plfft_mixed_plan_describe": (ctypes.c_size_t, [vp, ctypes.c_char_p, ctypes.c_size_t]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    _lib = lib
    return lib


def status_to_string(status: int) -> str:
    return load().plfft_status_to_string(status).decode()


def last_error() -> str:
    if _lib is None:
        return ""
    return (_lib.plfft_last_error() or b"").decode()


def _ptr(buf) -> int:
    """Device or host address of a torch tensor, numpy array or int."""
    if buf is None:
        return 0
    if isinstance(buf, int):
        return buf
    if hasattr(buf, "data_ptr"):
        return buf.data_ptr()
    if hasattr(buf, "ctypes"):
        return buf.ctypes.data
    raise TypeError(f"cannot take the address of {type(buf)}")


class Config:
    """Owning wrapper of a plfft_config_t."""

    def __init__(self, kind: str, n: int, direction: int = PLFFT_FORWARD, r2r: int | str = 0):
        lib = load()
        h = ctypes.c_void_p()
        if kind == "c2c":
            st = lib.plfft_config_create(ctypes.byref(h), n, direction)
        elif kind == "r2c":
            st = lib.plfft_config_create_r2c(ctypes.byref(h), n)
        elif kind == "c2r":
            st = lib.plfft_config_create_c2r(ctypes.byref(h), n)
        elif kind == "r2r":
            k = R2R_KINDS[r2r] if isinstance(r2r, str) else int(r2r)
            st = lib.plfft_config_create_r2r(ctypes.byref(h), n, k)
        else:
            raise ValueError(kind)
        if st != PLFFT_OK:
            raise PlfftError(st, "config create")
        self.h = h
        self.kind = kind
        self.n = n

    def set(self, dtype=None, alias=None, rigor=None, sme=None, batch=None):
        lib = load()
        if dtype is not None:
            lib.plfft_config_set_data_type(self.h, DATA_TYPES[dtype] if isinstance(dtype, str) else int(dtype))
        if alias is not None:
            lib.plfft_config_set_io_alias(self.h, int(alias))
        if rigor is not None:
            lib.plfft_config_set_rigor(self.h, int(rigor))
        if sme is not None:
            lib.plfft_config_set_sme_mode(self.h, int(sme))
        if batch is not None:
            lib.plfft_config_set_batch(self.h, *[int(v) for v in batch])
        return self

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.plfft_config_destroy(self.h)
            self.h = None


@dataclass
class Layout:
    howmany: int = 1
    istride: int = 1
    idist: int = 0
    ostride: int = 1
    odist: int = 0


class Plan:
    """A GPU plan created through plfft_plan_create_from_config."""

    def __init__(self, kind: str, n: int, direction: int = PLFFT_FORWARD, dtype: str = "f32",
                 howmany: int = 1, istride: int = 1, idist: int | None = None,
                 ostride: int = 1, odist: int | None = None, may_alias: bool = False,
                 rigor: int = PLFFT_MEASURE, r2r: int | str = 0):
        lib = load()
        cfg = Config(kind, n, direction, r2r)
        in_count, out_count = _counts(kind, n)
        if idist is None:
            idist = in_count * istride
        if odist is None:
            odist = out_count * ostride
        cfg.set(dtype=dtype, alias=PLFFT_IO_MAY_ALIAS if may_alias else PLFFT_IO_NO_ALIAS,
                rigor=rigor, batch=(howmany, istride, idist, ostride, odist))
        h = ctypes.c_void_p()
        st = lib.plfft_plan_create_from_config(ctypes.byref(h), cfg.h)
        if st != PLFFT_OK:
            raise PlfftError(st, "plan create")
        self.h = h
        self.kind, self.n, self.direction, self.dtype = kind, n, direction, dtype
        self.layout = Layout(howmany, istride, idist, ostride, odist)
        self.in_count, self.out_count = in_count, out_count

    def execute(self, x, y):
        load().plfft_plan_execute(self.h, _ptr(x), _ptr(y))

    def execute_async(self, x, y, stream=None, workspace=None):
        st = load().plfft_plan_execute_async(self.h, _ptr(x), _ptr(y), _ptr(workspace),
                                             _stream_handle(stream))
        if st != PLFFT_OK:
            raise PlfftError(st, "execute_async")

    def execute_noalloc(self, x, y, stream=None):
        """plfft_plan_execute_noalloc: scratch from the library's fixed device
        array (PLFFT_NOALLOC_SCRATCH_BYTES), no allocation."""
        st = load().plfft_plan_execute_noalloc(self.h, _ptr(x), _ptr(y), _stream_handle(stream))
        if st != PLFFT_OK:
            raise PlfftError(st, "execute_noalloc")

    def adjoint(self, y, x):
        load().plfft_plan_execute_adjoint(self.h, _ptr(y), _ptr(x))

    def adjoint_async(self, y, x, stream=None, workspace=None):
        st = load().plfft_plan_execute_adjoint_async(self.h, _ptr(y), _ptr(x), _ptr(workspace),
                                                     _stream_handle(stream))
        if st != PLFFT_OK:
            raise PlfftError(st, "adjoint_async")

    @property
    def workspace_size(self) -> int:
        return int(load().plfft_plan_get_workspace_size(self.h))

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        load().plfft_plan_describe(self.h, buf, 4096)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.plfft_plan_destroy(self.h)
            self.h = None


class MixedPlan:
    """A batch of C2C transforms of different lengths run as one launch
    (plfft_mixed_plan_create of include/plfft_gpu.h).

    ``lengths[i]`` complex elements of member i are read at ``in_offsets[i]``
    and written at ``out_offsets[i]`` (complex elements; None packs the
    members back to back).
    """

    def __init__(self, lengths, direction: int = PLFFT_FORWARD, dtype: str = "f32",
                 in_offsets=None, out_offsets=None, rigor: int = PLFFT_MEASURE):
        lib = load()
        count = len(lengths)
        arr = (ctypes.c_int64 * max(count, 1))(*[int(v) for v in lengths])

        def offs(o):
            if o is None:
                return None
            return (ctypes.c_int64 * max(count, 1))(*[int(v) for v in o])

        self._keep = (arr, offs(in_offsets), offs(out_offsets))
        h = ctypes.c_void_p()
        st = lib.plfft_mixed_plan_create(
            ctypes.byref(h), count, ctypes.cast(arr, ctypes.c_void_p),
            ctypes.cast(self._keep[1], ctypes.c_void_p) if self._keep[1] is not None else None,
            ctypes.cast(self._keep[2], ctypes.c_void_p) if self._keep[2] is not None else None,
            int(direction), DATA_TYPES[dtype] if isinstance(dtype, str) else int(dtype),
            int(rigor))
        if st != PLFFT_OK:
            raise PlfftError(st, "mixed plan create")
        self.h = h
        self.lengths = [int(v) for v in lengths]
        self.direction, self.dtype = direction, dtype

    @property
    def in_span(self) -> int:
        return int(load().plfft_mixed_plan_get_span(self.h, 0))

    @property
    def out_span(self) -> int:
        return int(load().plfft_mixed_plan_get_span(self.h, 1))

    @property
    def workspace_size(self) -> int:
        return int(load().plfft_mixed_plan_get_workspace_size(self.h))

    def execute(self, x, y):
        st = load().plfft_mixed_plan_execute(self.h, _ptr(x), _ptr(y))
        if st != PLFFT_OK:
            raise PlfftError(st, "mixed execute")

    def execute_async(self, x, y, stream=None, workspace=None):
        st = load().plfft_mixed_plan_execute_async(self.h, _ptr(x), _ptr(y), _ptr(workspace),
                                                   _stream_handle(stream))
        if st != PLFFT_OK:
            raise PlfftError(st, "mixed execute_async")

    def describe(self) -> str:
        buf = ctypes.create_string_buffer(4096)
        load().plfft_mixed_plan_describe(self.h, buf, 4096)
        return buf.value.decode()

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.plfft_mixed_plan_destroy(self.h)
            self.h = None


def _counts(kind: str, n: int):
    if kind == "r2c":
        return n, n // 2 + 1
    if kind == "c2r":
        return n // 2 + 1, n
    return n, n


def _stream_handle(stream) -> int:
    if stream is None:
        return 0
    if isinstance(stream, int):
        return stream
    if hasattr(stream, "cuda_stream"):
        return stream.cuda_stream
    raise TypeError(type(stream))
