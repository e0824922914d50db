"""GPU parity of the block-sparse container (blocksparsematrix.hpp) against the
oracle: block presence and cached norms bit-exact, values exact."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from strassen_amd import synth
from gpu_util import (assert_same_meta, ctx, empty_pair_blocked, empty_pair_uniform, pair_blocked, pair_uniform,
                      rand_dense, rel_fro)

pytestmark = pytest.mark.gpu


def test_device_sqrt_and_norm_bit_exact():
    """Canonical block norms (fma partial sums, halving tree, IEEE sqrt) are
    bit-identical between the device kernel and the oracle."""
    rng = np.random.default_rng(0)
    for shape, bs in [((300, 257), 17), ((128, 128), 64), ((65, 63), 8), ((40, 7), 5)]:
        a = rng.standard_normal(shape) * np.exp(rng.uniform(-30, 30, shape))
        o, d = pair_uniform(a, bs, bs, 0.0)
        assert_same_meta(o, d)
        assert np.array_equal(o.to_dense(), d.to_dense())


def test_copy_in_drop_rule_uniform_ragged():
    rng = np.random.default_rng(1)
    a = rand_dense(rng, [16] * 9 + [7], [16] * 6 + [3], 0.5, decay=0.8)
    for thresh in [0.0, 1e-10, 1e-3, 1e-1]:
        o, d = pair_uniform(a, 16, 16, thresh)
        assert_same_meta(o, d)
        assert np.array_equal(o.to_dense(), d.to_dense())
        assert o.nalloc() == d.nalloc()


def test_copy_in_uneven_and_empty():
    w = synth.small_uneven(n=300, fill=0.3, seed=11)
    a = synth.dense_from_scales(5, w.rsz, w.ksz, w.a_scale)
    o, d = pair_blocked(a, w.rsz, w.ksz, 1e-10)
    assert_same_meta(o, d)
    assert np.array_equal(o.to_dense(), d.to_dense())
    z = np.zeros((50, 40))
    o, d = pair_uniform(z, 8, 8, 1e-12)  # every block dropped
    assert d.nalloc() == 0 and o.nalloc() == 0
    o, d = pair_uniform(z, 8, 8, 0.0)    # thresh 0 keeps zero blocks (0 < 0 is false)
    assert d.nalloc() == o.nalloc() == 7 * 5


def test_generator_matches_host_definition():
    w = synth.small_uneven(n=400, fill=0.25, seed=3)
    d = sa.BSM(ctx(), w.rsz, w.ksz, 1e-10).generate(7, w.a_scale)
    a = synth.dense_from_scales(7, w.rsz, w.ksz, w.a_scale)
    o = oracle.OBsm(w.rsz, w.ksz, 1e-10).from_dense(a)
    assert np.array_equal(d.to_dense(), a)
    assert_same_meta(o, d)
    cfg = synth.band_workload(n=512, bs=64)
    d = sa.BSM.uniform(ctx(), 512, 512, 64, 64, 1e-10).generate(1, cfg.a_scale)
    assert np.array_equal(d.to_dense(), synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale))


def test_fill_zero_scale_recompress_calc_norms():
    rng = np.random.default_rng(2)
    a = rand_dense(rng, [6] * 5 + [4], [6] * 4 + [1], 0.6)
    o, d = pair_uniform(a, 6, 6, 1e-3)
    for op, args in [("scale", (-1.5,)), ("recompress", ()), ("calc_norms", ()), ("scale", (0.25,)),
                     ("fill", (0.5,)), ("zero", ()), ("fill", (-2.0,))]:
        getattr(o, op)(*args)
        getattr(d, op)(*args)
        assert_same_meta(o, d)
        assert np.array_equal(o.to_dense(), d.to_dense()), op


@pytest.mark.parametrize("sign", [1.0, -1.0])
def test_add_sub_semantics(sign):
    rng = np.random.default_rng(3)
    rsz = [5, 5, 5, 3]
    a = rand_dense(rng, rsz, rsz, 0.5)
    b = rand_dense(rng, rsz, rsz, 0.5)
    oa, da = pair_uniform(a, 5, 5, 1e-12)
    ob, db = pair_uniform(b, 5, 5, 1e-12)
    oa.add(ob, sign)
    da.add(db, sign)
    assert_same_meta(oa, da)
    assert np.array_equal(oa.to_dense(), da.to_dense())


def test_add_transpose_copy_dot_norm():
    rng = np.random.default_rng(4)
    w = synth.small_uneven(n=120, fill=0.5, seed=2)
    a = rand_dense(rng, w.rsz, w.csz, 0.5)
    b = rand_dense(rng, w.csz, w.rsz, 0.5)
    oa, da = pair_blocked(a, w.rsz, w.csz, 1e-12)
    ob, db = pair_blocked(b, w.csz, w.rsz, 1e-12)
    oc, dc = oa.clone(), sa.BSM(ctx(), w.rsz, w.csz, 1e-12).copy_from(da)
    assert_same_meta(oc, dc)
    oc.add_transpose(ob)
    dc.add_transpose(db)
    assert_same_meta(oc, dc)
    assert np.array_equal(oc.to_dense(), dc.to_dense())
    assert abs(da.frobenius_norm() - oa.frobenius_norm()) <= 1e-14 * oa.frobenius_norm()
    a2 = rand_dense(rng, w.rsz, w.csz, 0.5)
    oa2, da2 = pair_blocked(a2, w.rsz, w.csz, 1e-12)
    for thr in [0.0, 0.1]:
        ref = oa.dot(oa2, thr)
        assert abs(da.dot(da2, thr) - ref) <= 1e-13 * max(abs(ref), 1.0)


def test_error_paths_on_device():
    _, d = empty_pair_uniform(10, 10, 5, 5)
    _, e = empty_pair_uniform(12, 10, 5, 5)
    with pytest.raises(sa.SAError):
        d.add(e)
    with pytest.raises(sa.SAError):
        sa.matmult(d, d, 0, d, 0, 0.0)  # aliasing C and A
    with pytest.raises(sa.SAError):
        sa.BSM(ctx(), [3, 0], [2], 0.0)  # zero-sized block row


def test_scale_recompress_designed_ties_on_device():
    """Device scale + recompress on blocks exactly at the target-size threshold
    and ragged blocks between the target- and actual-size thresholds: the kept
    set and cached norms are fixed by construction (tests/designed.py)."""
    import designed

    a, thr, s, pres, norms = designed.recompress_case()
    d = sa.BSM.uniform(ctx(), designed.NR, designed.NC, designed.BS, designed.BS, thr).from_dense(a)
    d.scale(s)
    d.recompress()
    p, n = d.meta()
    assert np.array_equal(p, pres)
    assert np.array_equal(n[p], norms[p])


def test_dot_thresholded_designed_ties_on_device():
    import designed

    a, b, thr, expect, expect_all = designed.dot_case()
    A = sa.BSM.uniform(ctx(), designed.NR, designed.NC, designed.BS, designed.BS, 0.0).from_dense(a)
    B = sa.BSM.uniform(ctx(), designed.NR, designed.NC, designed.BS, designed.BS, 0.0).from_dense(b)
    assert A.dot(B, thr) == expect
    assert A.dot(B, 0.0) == expect_all


def test_scale_with_recomputed_norms():
    """The opt-in corrected norm rule: after scale(s, recompute_norms) the cached
    norms are the canonical norms of the scaled blocks (equal to a later
    calc_norms, and to |s| times the old norm up to rounding); the default
    keeps the reference's s*s."""
    rng = np.random.default_rng(23)
    a = rand_dense(rng, [7] * 5 + [3], [9] * 4 + [2], 0.6)
    for s in [-2.5, 0.3, 1e-3]:
        o, d = pair_uniform(a, 7, 9, 0.0)
        _, n0 = d.meta()
        d.scale(s, recompute_norms=True)
        p1, n1 = d.meta()
        d.calc_norms()
        _, n2 = d.meta()
        assert np.array_equal(n1, n2)
        np.testing.assert_allclose(n1[p1], abs(s) * n0[p1], rtol=1e-14)
        o.scale(s)
        assert np.array_equal(d.to_dense(), o.to_dense())
        d2 = sa.BSM.uniform(ctx(), a.shape[0], a.shape[1], 7, 9, 0.0).from_dense(a)
        d2.scale(s)
        assert_same_meta(o, d2)  # the reference rule, bit-exact with the oracle


def test_get_blocks_sub_grid():
    """sa_bsm_get_blocks (block(rb, cb) read access) against the dense copy, on
    uneven blocks with absent ones, single blocks and ranges."""
    rng = np.random.default_rng(31)
    rsz, csz = [5, 9, 14, 3], [7, 2, 11]
    a = rand_dense(rng, rsz, csz, 0.5)
    o, d = pair_blocked(a, rsz, csz, 0.0)
    full = d.to_dense()
    ro, co = np.concatenate([[0], np.cumsum(rsz)]), np.concatenate([[0], np.cumsum(csz)])
    for rb0, rb1, cb0, cb1 in [(0, 4, 0, 3), (1, 2, 2, 3), (3, 4, 0, 1), (1, 3, 1, 3), (2, 2, 0, 3)]:
        got = d.blocks(rb0, rb1, cb0, cb1)
        assert np.array_equal(got, full[ro[rb0]:ro[rb1], co[cb0]:co[cb1]])
