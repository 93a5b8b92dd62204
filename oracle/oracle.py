"""ctypes wrapper of liboracle.so -- the CPU restatement of the reference path.

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
bench.py's cpu_baseline leg, always as the checker (or the timed CPU baseline),
never by the product library.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int64)
_i64 = ctypes.c_int64


def build() -> None:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = ctypes.CDLL(LIB)
        L.orc_splitmix64.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.orc_splitmix64.restype = ctypes.c_uint64
        L.orc_fill_uniform_f64.argtypes = [ctypes.c_uint64, _i64, _dp]
        L.orc_fill_uniform_f32.argtypes = [ctypes.c_uint64, _i64, _fp]
        L.orc_fill_normal_f64.argtypes = [ctypes.c_uint64, _i64, _dp]
        L.orc_smallest_prime_factor.argtypes = [_i64]
        L.orc_smallest_prime_factor.restype = _i64
        L.orc_next_length_factorable_by.argtypes = [_i64, ctypes.POINTER(ctypes.c_int), ctypes.c_int]
        L.orc_next_length_factorable_by.restype = _i64
        L.orc_find_group_generator.argtypes = [_i64]
        L.orc_find_group_generator.restype = _i64
        L.orc_rader_permutations.argtypes = [_i64, _ip, _ip]
        L.orc_bluestein_phase.argtypes = [_i64, _ip]
        L.orc_bluestein_length.argtypes = [_i64]
        L.orc_bluestein_length.restype = _i64
        L.orc_dft.argtypes = [_i64, ctypes.c_int, _dp, _dp]
        L.orc_fft.argtypes = [_i64, ctypes.c_int, _dp, _dp]
        L.orc_fft_f32.argtypes = [_i64, ctypes.c_int, _fp, _fp]
        L.orc_fft_describe.argtypes = [_i64, ctypes.c_char_p, ctypes.c_int]
        L.orc_c2c_batched.argtypes = [_i64, ctypes.c_int, _i64, _i64, _i64, _i64, _i64, _dp, _dp]
        L.orc_c2c_batched_f32.argtypes = [_i64, ctypes.c_int, _i64, _i64, _i64, _i64, _i64, _fp, _fp, ctypes.c_int]
        L.orc_r2c_batched.argtypes = [_i64, _i64, _i64, _i64, _i64, _i64, _dp, _dp]
        L.orc_c2r_batched.argtypes = [_i64, _i64, _i64, _i64, _i64, _i64, _dp, _dp]
        L.orc_r2r_direct.argtypes = [ctypes.c_int, _i64, _dp, _dp]
        L.orc_r2r_direct.restype = ctypes.c_int
        L.orc_r2r_fast.argtypes = [ctypes.c_int, _i64, _dp, _dp]
        L.orc_r2r_fast.restype = ctypes.c_int
        L.orc_r2r_batched.argtypes = [ctypes.c_int, _i64, _i64, _i64, _i64, _i64, _i64, _dp, _dp]
        L.orc_q_convert.argtypes = [ctypes.c_double, ctypes.c_int]
        L.orc_q_convert.restype = ctypes.c_int32
        L.orc_fixed_chain.argtypes = [_i64, ctypes.POINTER(ctypes.c_int)]
        L.orc_fixed_chain.restype = ctypes.c_int
        L.orc_fixed_c2c.argtypes = [_i64, ctypes.c_int, ctypes.c_int,
                                    ctypes.POINTER(ctypes.c_int32), ctypes.POINTER(ctypes.c_int32)]
        L.orc_fixed_c2c.restype = ctypes.c_int
        L.orc_fixed_codelet.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int32),
                                        ctypes.POINTER(ctypes.c_int32)]
        L.orc_fixed_codelet.restype = ctypes.c_int
        L.orc_fixed_set_plant.argtypes = [ctypes.c_int]
        L.orc_fixed_set_plant.restype = None
        _lib = L
    return _lib


def _d(a):
    return a.ctypes.data_as(_dp)


def _as_c128(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, dtype=np.complex128))


# ---- seeded inputs ----------------------------------------------------------
def uniform_f64(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    lib().orc_fill_uniform_f64(seed, count, _d(out))
    return out


def uniform_f32(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float32)
    lib().orc_fill_uniform_f32(seed, count, out.ctypes.data_as(_fp))
    return out


def normal_f64(seed: int, count: int) -> np.ndarray:
    out = np.empty(count, np.float64)
    lib().orc_fill_normal_f64(seed, count, _d(out))
    return out


# ---- integer helpers --------------------------------------------------------
def smallest_prime_factor(n: int) -> int:
    return int(lib().orc_smallest_prime_factor(n))


def find_group_generator(n: int) -> int:
    return int(lib().orc_find_group_generator(n))


def rader_permutations(n: int):
    perm = np.zeros(n - 1, np.int64)
    rev = np.zeros(n, np.int64)
    lib().orc_rader_permutations(n, perm.ctypes.data_as(_ip), rev.ctypes.data_as(_ip))
    return perm, rev


def bluestein_phase(n: int) -> np.ndarray:
    out = np.zeros(n, np.int64)
    lib().orc_bluestein_phase(n, out.ctypes.data_as(_ip))
    return out


def bluestein_length(n: int) -> int:
    return int(lib().orc_bluestein_length(n))


def next_length_factorable_by(n: int, factors) -> int:
    arr = (ctypes.c_int * len(factors))(*factors)
    return int(lib().orc_next_length_factorable_by(n, arr, len(factors)))


# ---- transforms ---------------------------------------------------------------
def dft(x, direction: int = -1) -> np.ndarray:
    x = _as_c128(x)
    y = np.empty_like(x)
    lib().orc_dft(x.size, direction, _d(x.view(np.float64)), _d(y.view(np.float64)))
    return y


def fft(x, direction: int = -1) -> np.ndarray:
    x = _as_c128(x)
    y = np.empty_like(x)
    lib().orc_fft(x.size, direction, _d(x.view(np.float64)), _d(y.view(np.float64)))
    return y


def fft_describe(n: int) -> str:
    buf = ctypes.create_string_buffer(2048)
    lib().orc_fft_describe(n, buf, 2048)
    return buf.value.decode()


def c2c_batched(x, n, direction, howmany, istride, idist, ostride, odist, out_size) -> np.ndarray:
    """x: complex128 buffer (flat). Returns a complex128 buffer of out_size."""
    x = _as_c128(x).ravel()
    y = np.zeros(out_size, np.complex128)
    lib().orc_c2c_batched(n, direction, howmany, istride, idist, ostride, odist,
                          _d(x.view(np.float64)), _d(y.view(np.float64)))
    return y


def c2c_batched_f32(x: np.ndarray, n, direction, howmany, istride, idist, ostride, odist,
                    out: np.ndarray, threads: int = 0) -> None:
    """FP32 port of the planned transform over a batch (CPU baseline)."""
    lib().orc_c2c_batched_f32(n, direction, howmany, istride, idist, ostride, odist,
                              x.view(np.float32).ctypes.data_as(_fp),
                              out.view(np.float32).ctypes.data_as(_fp), threads)


def r2c_batched(x, n, howmany, istride, idist, ostride, odist, out_size) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64)).ravel()
    y = np.zeros(out_size, np.complex128)
    lib().orc_r2c_batched(n, howmany, istride, idist, ostride, odist, _d(x), _d(y.view(np.float64)))
    return y


def c2r_batched(x, n, howmany, istride, idist, ostride, odist, out_size) -> np.ndarray:
    x = _as_c128(x).ravel()
    y = np.zeros(out_size, np.float64)
    lib().orc_c2r_batched(n, howmany, istride, idist, ostride, odist, _d(x.view(np.float64)), _d(y))
    return y


def r2r_direct(kind: int, x) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64))
    y = np.empty_like(x)
    if not lib().orc_r2r_direct(kind, x.size, _d(x), _d(y)):
        raise ValueError("invalid r2r kind/size")
    return y


def r2r_fast(kind: int, x) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64))
    y = np.empty_like(x)
    if not lib().orc_r2r_fast(kind, x.size, _d(x), _d(y)):
        raise ValueError("invalid r2r kind/size")
    return y


def r2r_batched(kind, x, n, howmany, istride, idist, ostride, odist, out_size) -> np.ndarray:
    x = np.ascontiguousarray(np.asarray(x, np.float64)).ravel()
    y = np.zeros(out_size, np.float64)
    lib().orc_r2r_batched(kind, n, howmany, istride, idist, ostride, odist, _d(x), _d(y))
    return y


def q_convert(v: float, bits: int) -> int:
    return int(lib().orc_q_convert(v, bits))


def fixed_chain(n: int):
    buf = (ctypes.c_int * 64)()
    ns = lib().orc_fixed_chain(n, buf)
    return None if ns < 0 else list(buf[:ns])


def fixed_c2c(x, direction: int = -1, bits: int = 15) -> np.ndarray:
    """Fixed-point C2C of one member. x: int array of shape (n, 2) or (2n,)
    (re, im interleaved) within the storage range; returns int32 (n, 2)."""
    a = np.ascontiguousarray(np.asarray(x, np.int32).reshape(-1))
    n = a.size // 2
    y = np.empty_like(a)
    ip = ctypes.POINTER(ctypes.c_int32)
    st = lib().orc_fixed_c2c(n, direction, bits, a.ctypes.data_as(ip), y.ctypes.data_as(ip))
    if st != 0:
        raise ValueError(f"fixed-point C2C needs a power-of-two n and bits 15 or 7 (n={n})")
    return y.reshape(n, 2)


def fixed_codelet(x, bits: int = 15) -> np.ndarray:
    """One forward fixed-point codelet (oracle.c orc_fixed_codelet) on R
    (re, im) pairs, R in (2, 4, 8, 16)."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.int32).reshape(-1, 2))
    y = np.zeros_like(a)
    ip = ctypes.POINTER(ctypes.c_int32)
    if lib().orc_fixed_codelet(a.shape[0], bits, a.ctypes.data_as(ip), y.ctypes.data_as(ip)):
        raise ValueError(f"no fixed-point codelet of size {a.shape[0]} for {bits} bits")
    return y


def fixed_set_plant(mode: int) -> None:
    """Test hook: plant a known mistake in fixed_c2c (0 = none)."""
    lib().orc_fixed_set_plant(mode)
