"""GPU test of the allocation-free multiply (sa_matmult_noalloc) through the C
ABI: against sa_matmult on the same inputs -- task list bit-exact, C within
the FP64 bar (identical when both ran the same tile) -- on two calls in a row,
the second smaller, with transposes and an existing C; and its errors when a
table would exceed the scratch region or C's arena is too small."""
import ctypes as C

import numpy as np
import pytest

import strassen_amd as sa
from strassen_amd import synth
from gpu_util import ctx, rand_dense, rel_fro

pytestmark = pytest.mark.gpu


def noalloc(Cm, A, ta, B, tb, thresh, alpha, beta):
    return sa.lib().sa_matmult_noalloc(Cm.h, A.h, ta, B.h, tb, thresh, alpha, beta)


def reserve(m, elements):
    assert sa.lib().sa_bsm_reserve(m.h, elements) == sa.OK


def case(n, seed, ta, tb, beta):
    w = synth.small_uneven(n=n, fill=0.3, seed=seed)
    rng = np.random.default_rng(seed)
    c0 = rand_dense(rng, w.rsz, w.csz, 0.5)
    c = ctx()
    A = sa.BSM(c, w.ksz if ta else w.rsz, w.rsz if ta else w.ksz, w.thresh).generate(1, w.a_scale.T if ta else w.a_scale)
    B = sa.BSM(c, w.csz if tb else w.ksz, w.ksz if tb else w.csz, w.thresh).generate(2, w.b_scale.T if tb else w.b_scale)
    C1 = sa.BSM(c, w.rsz, w.csz, 0.0).from_dense(c0)
    sa.matmult(C1, A, ta, B, tb, w.thresh, 0.75, beta)
    t1 = c.tasklist()
    cfg1 = c.gemm_config()[0]
    C2 = sa.BSM(c, w.rsz, w.csz, 0.0).from_dense(c0)
    reserve(C2, int(w.rsz.sum()) * int(w.csz.sum()))
    assert noalloc(C2, A, ta, B, tb, w.thresh, 0.75, beta) == sa.OK, sa.lib().sa_last_error()
    t2 = c.tasklist()
    cfg2 = c.gemm_config()[0]
    assert len(t1) > 100
    assert np.array_equal(t1, t2)
    d1, d2 = C1.to_dense(), C2.to_dense()
    if cfg1 == cfg2:
        assert np.array_equal(d1, d2)
    assert rel_fro(d2, d1) < 1e-12
    p1, n1 = C1.meta()
    p2, n2 = C2.meta()
    assert np.array_equal(p1, p2) and np.array_equal(n1, n2)


@pytest.mark.parametrize("ta,tb,beta", [(0, 0, 0.0), (1, 0, 1.0), (0, 1, 0.5), (1, 1, 0.0)])
def test_noalloc_matches_matmult_two_calls(ta, tb, beta):
    case(900, 3 + ta + 2 * tb, ta, tb, beta)   # first call
    case(400, 11 + ta + 2 * tb, ta, tb, beta)  # second call, smaller input


def test_noalloc_uniform_band_repeated():
    """The band workload's shape (64-blocks) three times on the same matrices:
    the region keeps the previous call's tables, and nothing needs clearing."""
    cfg = synth.band_workload(n=1024)
    c = ctx()
    A = sa.BSM.uniform(c, 1024, 1024, 64, 64, cfg.thresh).generate(1, cfg.a_scale)
    B = sa.BSM.uniform(c, 1024, 1024, 64, 64, cfg.thresh).generate(2, cfg.b_scale)
    C1 = sa.BSM.uniform(c, 1024, 1024, 64, 64, 0.0)
    sa.matmult(C1, A, 0, B, 0, cfg.thresh)
    ref = C1.to_dense()
    C2 = sa.BSM.uniform(c, 1024, 1024, 64, 64, 0.0)
    reserve(C2, 1024 * 1024)
    for _ in range(3):
        assert noalloc(C2, A, 0, B, 0, cfg.thresh, 1.0, 0.0) == sa.OK, sa.lib().sa_last_error()
        assert rel_fro(C2.to_dense(), ref) < 1e-12


def test_noalloc_errors():
    c = ctx()
    n = 64
    a = np.asfortranarray(np.ones((n, n)))
    A = sa.BSM.uniform(c, n, n, 8, 8, 0.0).from_dense(a)
    # C's arena cannot hold the new blocks: refused, C unchanged
    Cm = sa.BSM.uniform(c, n, n, 8, 8, 0.0)
    assert noalloc(Cm, A, 0, A, 0, 0.0, 1.0, 0.0) == sa.ERR_INVALID
    assert b"sa_bsm_reserve" in sa.lib().sa_last_error()
    assert Cm.nalloc() == 0
    # a block grid whose task tables exceed the scratch region: refused
    # before any kernel reads or writes the tables
    cap = C.c_int64(0)
    assert sa.lib().sa_matmult_scratch_bytes(C.byref(cap)) == sa.OK
    nb = int(np.sqrt(cap.value / 8)) + 16  # nb^2 C blocks of 8-byte counts alone exceed it
    big = sa.BSM.uniform(c, nb, nb, 1, 1, 0.0)
    Cb = sa.BSM.uniform(c, nb, nb, 1, 1, 0.0)
    assert noalloc(Cb, big, 0, big, 0, 0.0, 1.0, 0.0) == sa.ERR_INVALID
    assert b"scratch region" in sa.lib().sa_last_error()
