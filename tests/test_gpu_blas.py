"""GPU tests of the BLAS boundary (blaslapack.cpp:5-37: any_gemm -> dgemm_,
any_axpy -> daxpy_, any_nrm2 -> dnrm2_, any_dot -> ddot_) through the C ABI
against numpy: every transpose pair, alpha / beta, leading dimensions larger
than the row count, empty and one-element shapes."""
import ctypes as C

import numpy as np
import pytest

import oracle
import strassen_amd as sa
from gpu_util import ctx

pytestmark = pytest.mark.gpu


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def dgemm(ta, tb, m, n, k, alpha, A, lda, B, ldb, beta, Cm, ldc):
    rc = sa.lib().sa_dgemm(ctx().h, ta, tb, m, n, k, alpha, _p(A), lda, _p(B), ldb, beta, _p(Cm), ldc)
    return rc


def padded(rng, rows, cols, ld):
    """Column-major rows x cols matrix stored with leading dimension ld (the
    rows beyond `rows` hold values that must be ignored)."""
    store = np.asfortranarray(rng.standard_normal((ld, max(cols, 1))))
    return store, store[:rows, :cols]


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("m,n,k,alpha,beta", [(37, 53, 29, 1.0, 0.0), (128, 64, 256, -0.5, 1.0),
                                              (1, 1, 1, 2.0, 0.5), (200, 3, 130, 1.5, -2.0)])
def test_dgemm_vs_numpy(ta, tb, m, n, k, alpha, beta):
    rng = np.random.default_rng(m * 7 + n * 3 + k + 10 * ta + tb)
    ar, ac = (k, m) if ta else (m, k)
    br, bc = (n, k) if tb else (k, n)
    As, A = padded(rng, ar, ac, ar + 5)
    Bs, Bm = padded(rng, br, bc, br + 3)
    Cs, C0 = padded(rng, m, n, m + 2)
    c_before = Cs.copy()
    want = alpha * ((A.T if ta else A) @ (Bm.T if tb else Bm)) + beta * C0
    assert dgemm(ta, tb, m, n, k, alpha, As, ar + 5, Bs, br + 3, beta, Cs, m + 2) == sa.OK
    got = Cs[:m, :n]
    assert np.linalg.norm(got - want) <= 1e-14 * max(1.0, np.linalg.norm(want)) * np.sqrt(k)
    # rows beyond m of each column are left alone
    assert np.array_equal(Cs[m:, :n], c_before[m:, :n])


def test_dgemm_edge_and_errors():
    A = np.asfortranarray(np.ones((4, 4))); B = A.copy(); Cm = np.asfortranarray(np.full((4, 4), 3.0))
    # k = 0: C = beta C
    assert dgemm(0, 0, 4, 4, 0, 1.0, A, 4, B, 4, 0.5, Cm, 4) == sa.OK
    assert np.array_equal(Cm, np.full((4, 4), 1.5))
    # bad leading dimension and negative sizes are rejected before any device work
    assert dgemm(0, 0, 4, 4, 4, 1.0, A, 3, B, 4, 0.0, Cm, 4) == sa.ERR_INVALID
    assert dgemm(0, 0, -1, 4, 4, 1.0, A, 4, B, 4, 0.0, Cm, 4) == sa.ERR_INVALID


@pytest.mark.parametrize("n", [0, 1, 63, 64, 65, 1000, 100003])
def test_dnrm2_ddot_daxpy_vs_numpy(n):
    rng = np.random.default_rng(n + 1)
    x = rng.standard_normal(n) * np.exp(rng.uniform(-20, 20, n))
    y = rng.standard_normal(n)
    out = np.zeros(1)
    assert sa.lib().sa_dnrm2(ctx().h, n, _p(x), _p(out)) == sa.OK
    # the canonical block-norm order: bit-identical to the oracle's norm
    assert out[0] == oracle.block_norm(x)
    assert abs(out[0] - np.linalg.norm(x)) <= 1.2e-16 * (n + 1) * np.linalg.norm(x)
    assert sa.lib().sa_ddot(ctx().h, n, _p(x), _p(y), _p(out)) == sa.OK
    want = float(np.dot(x, y))
    assert abs(out[0] - want) <= 1.2e-16 * (n + 1) * float(np.abs(x) @ np.abs(y))
    yy = y.copy()
    assert sa.lib().sa_daxpy(ctx().h, n, -1.75, _p(x), _p(yy)) == sa.OK
    # one fused multiply-add per element: exact up to the final rounding of x*a+y
    assert np.all(np.abs(yy - (y - 1.75 * x)) <= 2.3e-16 * (np.abs(1.75 * x) + np.abs(y)))
