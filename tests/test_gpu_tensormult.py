"""GPU test of the drop-in tensormult overloads (strassen_amd/cpp/
blocksparse3d_tensor.h; reference blocksparse3d_tensor.h:10-43) compiled into a
C++ program (tests/cpp/test_tensormult.cpp) against the oracle's screened
multiplies on the same inputs."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from conftest import GOLDEN, rel_fro

pytestmark = pytest.mark.gpu
BIN = os.path.join(os.path.dirname(GOLDEN), "..", "strassen_amd", "bin", "test_tensormult")


def oracle_mm(c0, a, ta, b, tb, bs, thresh, alpha, beta):
    n = c0.shape[0]
    oC = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(c0)
    oA = oracle.OBsm.uniform(n, n, bs, bs, thresh).from_dense(a)
    oB = oracle.OBsm.uniform(n, n, bs, bs, thresh).from_dense(b)
    oracle.matmult(oC, oA, ta, oB, tb, thresh, alpha, beta)
    return oC.to_dense(), oC


@pytest.mark.parametrize("ta,tb,alpha,beta", [(0, 0, 1.0, 0.0), (1, 0, -0.5, 1.0), (0, 1, 2.0, 0.25),
                                              (1, 1, 1.0, 1.0)])
def test_tensormult_forms_vs_oracle(tmp_path, ta, tb, alpha, beta):
    rng = np.random.default_rng(10 + 4 * ta + 2 * tb)
    L, n, bs, thresh = 3, 50, 16, 1e-3
    I = np.repeat(np.arange(4), [16, 16, 16, 2])
    decay = np.exp(-1.5 * np.abs(I[:, None] - I[None, :]))  # block decay: some products screened out
    A = [np.asfortranarray(rng.standard_normal((n, n)) * decay) for _ in range(L)]
    B = [np.asfortranarray(rng.standard_normal((n, n)) * decay) for _ in range(L)]
    c0 = np.asfortranarray(rng.standard_normal((n, n)))
    col_major = lambda x, name: x.ravel(order="F").tofile(tmp_path / name)  # tofile alone writes row-major
    for l in range(L):
        col_major(A[l], f"A{l}.data")
        col_major(B[l], f"B{l}.data")
    col_major(c0, "C0.data")
    r = subprocess.run([BIN, str(tmp_path), str(L), str(n), str(bs), str(ta), str(tb), repr(thresh), repr(alpha),
                        repr(beta)], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    assert r.returncode == 0, r.stdout
    read = lambda name: np.fromfile(tmp_path / name).reshape(n, n, order="F")
    # form 1: beta applied once, then every slice accumulated in order
    oC = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(c0)
    for l in range(L):
        oA = oracle.OBsm.uniform(n, n, bs, bs, thresh).from_dense(A[l])
        oB = oracle.OBsm.uniform(n, n, bs, bs, thresh).from_dense(B[l])
        oracle.matmult(oC, oA, ta, oB, tb, thresh, alpha, beta if l == 0 else 1.0)
    want1 = oC.to_dense()
    got1 = read("form1.data")
    assert rel_fro(got1, want1) < 1e-12
    # block(rb, cb) of the drop-in: the ragged last block (n = 50 = 3 * 16 + 2)
    last = np.fromfile(tmp_path / "block_last.data")
    if beta != 0.0:  # C0's blocks are all present
        assert last.size == 4
    if last.size == 0:  # absent: screened out and no C0 block kept (beta = 0)
        assert not np.any(got1[48:, 48:])
    else:
        assert last.size == 4 and np.array_equal(last.reshape(2, 2, order="F"), got1[48:, 48:])
    # the screened sum differs from the unscreened one (screening is exercised)
    full = sum((A[l].T if ta else A[l]) @ (B[l].T if tb else B[l]) for l in range(L)) * alpha + beta * c0
    assert rel_fro(want1, full) > 1e-8
    for l in range(L):
        want2, _ = oracle_mm(c0, A[0], ta, B[l], tb, bs, thresh, alpha, beta)
        assert rel_fro(read(f"form2_{l}.data"), want2) < 1e-12
        # transC: C^l = alpha op(B^l)^T op(A)^T + beta C0
        want2t, _ = oracle_mm(c0, B[l], 1 - tb, A[0], 1 - ta, bs, thresh, alpha, beta)
        assert rel_fro(read(f"form2t_{l}.data"), want2t) < 1e-12
        # fourth overload: C^l = alpha op(B^l) op(A) + beta C0
        want4, _ = oracle_mm(c0, B[l], tb, A[0], ta, bs, thresh, alpha, beta)
        assert rel_fro(read(f"form4_{l}.data"), want4) < 1e-12
