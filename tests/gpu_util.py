"""Helpers shared by the GPU parity tests."""
import math

import numpy as np

REAL_NP = {"f32": np.float32, "f64": np.float64, "f16": np.float16}
CPLX_NP = {"f32": np.complex64, "f64": np.complex128}

# rel-L2 bound vs the FP64 oracle; f32 is the 2e-6 of the target metric.
TOL = {"f32": 2e-6, "f64": 1e-13, "f16": 4e-3, "bf16": 3e-2}


def span(count, stride, howmany, dist):
    return (howmany - 1) * dist + (count - 1) * stride + 1


def relerr(a, b):
    a = np.asarray(a, np.complex128).ravel()
    b = np.asarray(b, np.complex128).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def default_tol(prec, n):
    # test/test_utils.hpp:120-129: eps * 4 * n * log2(n)
    eps = {"f32": 2.0 ** -23, "f64": 2.0 ** -52, "f16": 2.0 ** -10, "bf16": 2.0 ** -7}[prec]
    return eps * 4 * n * max(1.0, math.log2(n))


def to_dev(torch, arr):
    t = torch.from_numpy(np.ascontiguousarray(arr)).to("cuda")
    torch.cuda.synchronize()
    return t


def from_dev(torch, t):
    torch.cuda.synchronize()
    return t.cpu().numpy()


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """float32 -> bf16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = u + 0x7FFF + ((u >> 16) & 1)
    return (u >> 16).astype(np.uint16)


def bf16_to_f32(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32)
