"""Adjoint kernels and the torch.autograd harness.

The transforms are linear, so backward = adjoint. Checked three ways:
the dot-product identity <A x, y> = <x, A^T y> in FP64; torch.autograd
gradcheck (FP64 finite differences) through the C ABI; and the
differentiable-transform configuration's loss sum(w_k |X_k|^2) against
its closed-form gradient 2 Re(sum_k w_k X_k exp(+2 pi i j k / n)) and FP64
central differences.
"""
import numpy as np
import pytest

import plfft_amd as P
from plfft_amd import rng
from gpu_util import from_dev, relerr, span, to_dev

pytestmark = pytest.mark.gpu


def _rdot(a, b):
    return float(np.sum(np.real(np.asarray(a).ravel() * np.conj(np.asarray(b).ravel()))))


CASES = [("c2c", 1024, -1, None), ("c2c", 97, 1, None), ("c2c", 65536, -1, None),
         ("r2c", 4096, -1, None), ("r2c", 1001, -1, None), ("r2c", 31, -1, None),
         ("c2r", 4096, 1, None), ("c2r", 999, 1, None),
         ("r2c", 65536, -1, None), ("c2r", 65536, 1, None),
         ("r2r", 4096, -1, "dct2"), ("r2r", 4096, -1, "dct3"), ("r2r", 4096, -1, "dst2"),
         ("r2r", 4096, -1, "dst3"), ("r2r", 99, -1, "dct2"), ("r2r", 99, -1, "dst3"),
         ("r2r", 64, -1, "dct1"), ("r2r", 99, -1, "dct1"), ("r2r", 64, -1, "dct4"),
         ("r2r", 99, -1, "dst1"), ("r2r", 100, -1, "dst4"), ("r2r", 99, -1, "dht"),
         ("r2r", 64, -1, "r2hc"), ("r2r", 99, -1, "r2hc"), ("r2r", 64, -1, "hc2r"),
         ("r2r", 99, -1, "hc2r")]


@pytest.mark.parametrize("kind,n,direction,r2r", CASES)
def test_dot_product_identity(torch_cuda, kind, n, direction, r2r):
    torch = torch_cuda
    hm, is_, id_ = 3, 2, None
    cin, cout = (n, n) if kind in ("c2c", "r2r") else ((n, n // 2 + 1) if kind == "r2c" else (n // 2 + 1, n))
    id_ = cin * is_ + 5
    plan = P.Plan(kind, n, direction, "f64", hm, is_, id_, 1, cout, r2r=r2r or 0)
    isz, osz = span(cin, is_, hm, id_), span(cout, 1, hm, cout)
    cplx_in = kind in ("c2c", "c2r")
    cplx_out = kind in ("c2c", "r2c")
    x = rng.normal_f64(1, 2 * isz).view(np.complex128) if cplx_in else rng.normal_f64(1, isz)
    y = rng.normal_f64(2, 2 * osz).view(np.complex128) if cplx_out else rng.normal_f64(2, osz)
    # only logical elements carry values (the gaps are not part of the map)
    mask = np.zeros(isz, bool)
    for b in range(hm):
        mask[b * id_ + is_ * np.arange(cin)] = True
    x = np.where(mask, x, 0)
    if kind == "c2r":  # the C2R map only sees Re of bins 0 and n/2
        for b in range(hm):
            x[b * id_] = x[b * id_].real
            if n % 2 == 0:
                x[b * id_ + is_ * (n // 2)] = x[b * id_ + is_ * (n // 2)].real
    xd, yd = to_dev(torch, x), to_dev(torch, y)
    Ax = torch.zeros_like(yd)
    Aty = torch.zeros_like(xd)
    plan.execute(xd, Ax)
    plan.adjoint(yd, Aty)
    lhs = _rdot(from_dev(torch, Ax), y)
    rhs = _rdot(x, from_dev(torch, Aty))
    assert abs(lhs - rhs) <= 1e-10 * (abs(lhs) + abs(rhs) + 1), (lhs, rhs, plan.describe())
    if kind == "r2c":  # the adjoint writes nothing into the gaps
        assert not np.any(from_dev(torch, Aty)[~mask])


@pytest.mark.parametrize("kind,r2r", [("r2c", 0), ("r2r", "dct2"), ("r2r", "dst2"),
                                      ("c2c", 0), ("c2r", 0)])
def test_gradcheck_fp64(torch_cuda, kind, r2r):
    torch = torch_cuda
    from plfft_amd.autograd import PlanFunction, input_size
    n = 24
    plan = P.Plan(kind, n, 1 if kind == "c2r" else -1, "f64", howmany=2, r2r=r2r)
    cplx = kind in ("c2c", "c2r")
    dt = torch.complex128 if cplx else torch.float64
    x = torch.randn(input_size(plan), dtype=dt, device="cuda", requires_grad=True)
    f = lambda v: PlanFunction.apply(v, plan)  # noqa: E731
    assert torch.autograd.gradcheck(f, (x,), eps=1e-6, atol=1e-7)


def _loss_fp64(x, w, n, hm, is_, id_):
    L = 0.0
    for b in range(hm):
        X = np.fft.rfft(x[b * id_ + is_ * np.arange(n)])
        L += float(np.sum(w[b] * np.abs(X) ** 2))
    return L


def test_differentiable_r2c_loss(torch_cuda, orc):
    """loss = sum w_k |X_k|^2, X = R2C(x), n=4096, input stride 2, distance
    2n+64 (batch reduced); f32 autograd vs closed form and FP64 FD."""
    torch = torch_cuda
    from plfft_amd.autograd import PlanFunction
    n, hm, is_, id_ = 4096, 8, 2, 2 * 4096 + 64
    m = n // 2 + 1
    plan = P.Plan("r2c", n, -1, "f32", hm, is_, id_, 1, m)
    isz = span(n, is_, hm, id_)
    x = rng.normal_f64(4, isz)
    w = rng.uniform_f64(5, hm * m).reshape(hm, m) * 0.5 + 0.5  # U(0,1)
    xt = torch.tensor(x.astype(np.float32), device="cuda", requires_grad=True)
    wt = torch.tensor(w.ravel().astype(np.float32), device="cuda")
    X = PlanFunction.apply(xt, plan)
    loss = (wt * X.abs() ** 2).sum()
    loss.backward()
    g = xt.grad.cpu().numpy().astype(np.float64)
    # closed form in FP64
    ref = np.zeros(isz)
    for b in range(hm):
        idx = b * id_ + is_ * np.arange(n)
        Xb = np.fft.rfft(x[idx])
        full = np.zeros(n, complex)
        full[:m] = w[b] * Xb
        # 2 Re sum_{k<=n/2} w_k X_k e^{+2 pi i j k / n}
        ref[idx] = 2 * np.real(np.fft.ifft(full) * n)
    assert relerr(g, ref) < 2e-5
    # FP64 central differences at a few coordinates
    for j in (0, 1, 4097, 3 * id_ + 2 * 17):
        h = 1e-4
        xp, xm = x.copy(), x.copy()
        xp[j] += h
        xm[j] -= h
        fd = (_loss_fp64(xp, w, n, hm, is_, id_) - _loss_fp64(xm, w, n, hm, is_, id_)) / (2 * h)
        assert abs(fd - ref[j]) <= 1e-4 * max(1.0, abs(ref[j]))
        assert abs(g[j] - ref[j]) <= 2e-3 * max(1.0, abs(ref[j]))


def _r2r_matrix(kind, n):
    """FFTW's REDFT10 (DCT-II) / RODFT10 (DST-II) as explicit FP64 matrices:
    Y_k = 2 sum_j x_j cos(pi (2j+1) k / 2n), Y_k = 2 sum_j x_j sin(pi (2j+1) (k+1) / 2n)."""
    j = np.arange(n)[None, :]
    k = np.arange(n)[:, None]
    if kind == "dct2":
        return 2.0 * np.cos(np.pi * (2 * j + 1) * k / (2.0 * n))
    return 2.0 * np.sin(np.pi * (2 * j + 1) * (k + 1) / (2.0 * n))


@pytest.mark.parametrize("r2r", ["dct2", "dst2"])
def test_differentiable_r2r_loss(torch_cuda, r2r):
    """loss = sum w_k Y_k^2, Y = DCT-II / DST-II of x at the differentiable
    layout (n = 4096, input stride 2, distance 2n+64, batch reduced); f32
    autograd against the closed form 2 A^T (w Y) from the explicit FP64
    matrix A, and the closed form against FP64 central differences."""
    torch = torch_cuda
    from plfft_amd.autograd import PlanFunction
    n, hm, is_, id_ = 4096, 4, 2, 2 * 4096 + 64
    plan = P.Plan("r2r", n, -1, "f32", hm, is_, id_, 1, n, r2r=r2r)
    isz = span(n, is_, hm, id_)
    x = rng.normal_f64(6, isz)
    w = rng.uniform_f64(7, hm * n).reshape(hm, n) * 0.5 + 0.5
    xt = torch.tensor(x.astype(np.float32), device="cuda", requires_grad=True)
    wt = torch.tensor(w.ravel().astype(np.float32), device="cuda")
    Y = PlanFunction.apply(xt, plan)
    loss = (wt * Y ** 2).sum()
    loss.backward()
    g = xt.grad.cpu().numpy().astype(np.float64)
    A = _r2r_matrix(r2r, n)
    ref = np.zeros(isz)
    rows = []
    for b in range(hm):
        idx = b * id_ + is_ * np.arange(n)
        Yb = A @ x[idx]
        rows.append(Yb)
        ref[idx] = 2.0 * (A.T @ (w[b] * Yb))
    # the forward values too, against the same matrix
    Yg = Y.detach().cpu().numpy().astype(np.float64).reshape(hm, n)
    assert relerr(Yg, np.stack(rows)) < 2e-6
    assert relerr(g, ref) < 2e-5

    def loss64(xx):
        return sum(float(np.sum(w[b] * (A @ xx[b * id_ + is_ * np.arange(n)]) ** 2))
                   for b in range(hm))

    for jj in (0, 1, 4095 * 2, 2 * id_ + 2 * 33):
        h = 1e-4
        xp, xm = x.copy(), x.copy()
        xp[jj] += h
        xm[jj] -= h
        fd = (loss64(xp) - loss64(xm)) / (2 * h)
        assert abs(fd - ref[jj]) <= 1e-4 * max(1.0, abs(ref[jj]))
        assert abs(g[jj] - ref[jj]) <= 2e-3 * max(1.0, abs(ref[jj]))
