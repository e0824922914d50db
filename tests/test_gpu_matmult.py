"""GPU parity of the screened block GEMM (blocksparsematrix.hpp:292-414):
task lists bit-exact against the oracle, C within 1e-12 relative Frobenius
error (FP64; tolerance stated per assert), C block pattern and cached norms
identical."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from strassen_amd import synth
from gpu_util import (assert_same_meta, ctx, empty_pair_blocked, empty_pair_uniform, pair_blocked, pair_uniform,
                      rand_dense, rel_fro)

pytestmark = pytest.mark.gpu
TOL = 1e-12  # FP64 mode bar: relative Frobenius error vs the oracle


def check_mm(oC, dC, oA, dA, ta, oB, dB, tb, thresh, alpha, beta, tile=-1):
    t_o = oracle.matmult(oC, oA, ta, oB, tb, thresh, alpha, beta)
    sa.matmult(dC, dA, ta, dB, tb, thresh, alpha, beta, sa.default_opts(tile=tile))
    t_d = ctx().tasklist()
    assert np.array_equal(t_o, t_d), f"task lists differ ({len(t_o)} vs {len(t_d)})"
    assert_same_meta(oC, dC)
    err = rel_fro(dC.to_dense(), oC.to_dense())
    assert err < TOL, err
    return len(t_o), err


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
@pytest.mark.parametrize("bs", [64, 17, 5])
def test_uniform_ragged_transposes(ta, tb, bs):
    rng = np.random.default_rng(100 + bs + 2 * ta + tb)
    ni, nk, nj = 3 * bs + 7, 2 * bs + 3, 4 * bs - 1
    a_shape = (nk, ni) if ta else (ni, nk)
    b_shape = (nj, nk) if tb else (nk, nj)
    a = rand_dense(rng, synth.uniform_sizes(a_shape[0], bs), synth.uniform_sizes(a_shape[1], bs), 0.7, 0.2)
    b = rand_dense(rng, synth.uniform_sizes(b_shape[0], bs), synth.uniform_sizes(b_shape[1], bs), 0.7, 0.2)
    oA, dA = pair_uniform(a, bs, bs, 1e-12)
    oB, dB = pair_uniform(b, bs, bs, 1e-12)
    oC, dC = empty_pair_uniform(ni, nj, bs, bs)
    check_mm(oC, dC, oA, dA, ta, oB, dB, tb, 1e-6, 1.0, 0.0)


@pytest.mark.parametrize("beta", [0.0, 1.0, 0.5])
def test_alpha_beta_existing_c(beta):
    rng = np.random.default_rng(7)
    a = rand_dense(rng, [32] * 5, [32] * 4, 0.6)
    b = rand_dense(rng, [32] * 4, [32] * 6, 0.6)
    c0 = rand_dense(rng, [32] * 5, [32] * 6, 0.5)
    oA, dA = pair_uniform(a, 32, 32, 1e-12)
    oB, dB = pair_uniform(b, 32, 32, 1e-12)
    oC, dC = pair_uniform(c0, 32, 32, 1e-12)
    check_mm(oC, dC, oA, dA, 0, oB, dB, 0, 1e-4, -1.25, beta)


@pytest.mark.parametrize("tile", list(range(28)))
def test_every_tile_config(tile):
    rng = np.random.default_rng(11)
    w = synth.small_uneven(n=500, fill=0.3, seed=4)
    a = synth.dense_from_scales(1, w.rsz, w.ksz, w.a_scale)
    b = synth.dense_from_scales(2, w.ksz, w.csz, w.b_scale)
    oA, dA = pair_blocked(a, w.rsz, w.ksz, 1e-10)
    oB, dB = pair_blocked(b, w.ksz, w.csz, 1e-10)
    oC, dC = empty_pair_blocked(w.rsz, w.csz)
    n, err = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, 1e-10, 1.0, 0.0, tile=tile)
    assert n > 100


@pytest.mark.parametrize("tile", [14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26, 27])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_pipelined_tiles_transposes(tile, ta, tb):
    """Large ragged blocks (150, with 7/3/149-wide edges) span several pipeline
    tiles and k-stages, including partial ones, for every transpose variant."""
    bs = 150
    rng = np.random.default_rng(300 + tile + 2 * ta + tb)
    ni, nk, nj = 2 * bs + 7, 2 * bs + 3, 3 * bs - 1
    a_shape = (nk, ni) if ta else (ni, nk)
    b_shape = (nj, nk) if tb else (nk, nj)
    a = rand_dense(rng, synth.uniform_sizes(a_shape[0], bs), synth.uniform_sizes(a_shape[1], bs), 0.8, 0.2)
    b = rand_dense(rng, synth.uniform_sizes(b_shape[0], bs), synth.uniform_sizes(b_shape[1], bs), 0.8, 0.2)
    oA, dA = pair_uniform(a, bs, bs, 1e-12)
    oB, dB = pair_uniform(b, bs, bs, 1e-12)
    oC, dC = empty_pair_uniform(ni, nj, bs, bs)
    check_mm(oC, dC, oA, dA, ta, oB, dB, tb, 1e-6, 0.75, 0.0, tile=tile)


@pytest.mark.parametrize("ta", [0, 1])
def test_many_small_blocks(ta):
    """A 100 x 100 grid of 3-blocks with 70 inner blocks: more than one scan
    tile for the screening's segment, arena and output scans (10,000 C
    blocks, 10,000 padded output slots), and more than one 64-wide step in
    the per-row lists of op(A)'s blocks."""
    rng = np.random.default_rng(77 + ta)
    ni, nk, nj, bs = 300, 210, 300, 3
    a_shape = (nk, ni) if ta else (ni, nk)
    a = rand_dense(rng, synth.uniform_sizes(a_shape[0], bs), synth.uniform_sizes(a_shape[1], bs), 0.5, 0.3)
    b = rand_dense(rng, synth.uniform_sizes(nk, bs), synth.uniform_sizes(nj, bs), 0.5, 0.3)
    oA, dA = pair_uniform(a, bs, bs, 1e-12)
    oB, dB = pair_uniform(b, bs, bs, 1e-12)
    oC, dC = empty_pair_uniform(ni, nj, bs, bs)
    n, _ = check_mm(oC, dC, oA, dA, ta, oB, dB, 0, 1e-3, 1.0, 0.0)
    assert n > 0


def test_binding_threshold_screens_tasks():
    cfg = synth.band_workload(n=1024, bs=64, band=6, decay=2.5)
    a = synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale)
    b = synth.dense_from_scales(2, cfg.ksz, cfg.csz, cfg.b_scale)
    oA, dA = pair_uniform(a, 64, 64, 1e-12)
    oB, dB = pair_uniform(b, 64, 64, 1e-12)
    oC, dC = empty_pair_uniform(1024, 1024, 64, 64)
    dense_tasks = (oA.meta()[0].astype(int) @ oB.meta()[0].astype(int)).sum()
    n, _ = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, 1e-7, 1.0, 0.0)
    assert 0 < n < dense_tasks  # the threshold removes some products


def test_empty_operands():
    oA, dA = empty_pair_uniform(64, 64, 16, 16)
    oB, dB = empty_pair_uniform(64, 64, 16, 16)
    oC, dC = empty_pair_uniform(64, 64, 16, 16)
    n, _ = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, 1e-10, 1.0, 0.0)
    assert n == 0 and dC.nalloc() == 0


def test_band_full_size():
    """n=2048, 64x64 blocks, band |I-J|<=4, exp(-0.5|I-J|) scales, thresh 1e-10."""
    cfg = synth.band_workload()
    oA = oracle.OBsm.uniform(2048, 2048, 64, 64, cfg.thresh).from_dense(
        synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale))
    oB = oracle.OBsm.uniform(2048, 2048, 64, 64, cfg.thresh).from_dense(
        synth.dense_from_scales(2, cfg.ksz, cfg.csz, cfg.b_scale))
    dA = sa.BSM.uniform(ctx(), 2048, 2048, 64, 64, cfg.thresh).generate(1, cfg.a_scale)
    dB = sa.BSM.uniform(ctx(), 2048, 2048, 64, 64, cfg.thresh).generate(2, cfg.b_scale)
    assert_same_meta(oA, dA)
    oC, dC = empty_pair_uniform(2048, 2048, 64, 64)
    n, err = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, cfg.thresh, 1.0, 0.0)
    assert n == 2292


def test_uneven_atoms_moderate():
    w = synth.atoms_workload(n=2500, fill=0.1, seed=9)
    a = synth.dense_from_scales(1, w.rsz, w.ksz, w.a_scale)
    b = synth.dense_from_scales(2, w.ksz, w.csz, w.b_scale)
    oA, dA = pair_blocked(a, w.rsz, w.ksz, w.thresh)
    oB, dB = pair_blocked(b, w.ksz, w.csz, w.thresh)
    oC, dC = empty_pair_blocked(w.rsz, w.csz)
    n, err = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, w.thresh, 1.0, 0.0)
    assert n > 1000


def test_chained_products_use_stale_norms():
    """A product's C blocks carry zero cached norms (blocksparsematrix.hpp:398-404):
    feeding C into a second screened product without calc_norms screens
    everything out, exactly as the reference."""
    rng = np.random.default_rng(5)
    a = rand_dense(rng, [8] * 4, [8] * 4, 1.0)
    oA, dA = pair_uniform(a, 8, 8, 0.0)
    oC, dC = empty_pair_uniform(32, 32, 8, 8)
    check_mm(oC, dC, oA, dA, 0, oA, dA, 0, 1e-20, 1.0, 0.0)
    oD, dD = empty_pair_uniform(32, 32, 8, 8)
    n, _ = check_mm(oD, dD, oC, dC, 0, oA, dA, 0, 1e-20, 1.0, 0.0)
    assert n == 0
    oC.calc_norms(); dC.calc_norms()
    n, _ = check_mm(oD, dD, oC, dC, 0, oA, dA, 0, 1e-20, 1.0, 0.0)
    assert n == 64


@pytest.mark.parametrize("beta", [0.0, 1.0])
def test_kwindows_match_single_launch(monkeypatch, beta):
    """A 1 MB operand budget splits the band product into many k-windows; the
    task list is unchanged and C stays within the FP64 bar, with existing C
    blocks (beta = 1) accumulated only in their first window."""
    cfg = synth.band_workload(n=1024, bs=64, band=6)
    a = synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale)
    b = synth.dense_from_scales(2, cfg.ksz, cfg.csz, cfg.b_scale)
    c0 = synth.dense_from_scales(3, cfg.rsz, cfg.csz, cfg.a_scale)
    oA, dA = pair_uniform(a, 64, 64, cfg.thresh)
    oB, dB = pair_uniform(b, 64, 64, cfg.thresh)
    oC, dC = pair_uniform(c0, 64, 64, 0.0)
    monkeypatch.setenv("SA_KWINDOW_MB", "1")
    n, err = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, cfg.thresh, 1.0, beta)
    assert ctx().kwindows() > 4
    monkeypatch.setenv("SA_KWINDOW_MB", "0")
    oD, dD = empty_pair_uniform(1024, 1024, 64, 64)
    check_mm(oD, dD, oA, dA, 0, oB, dB, 0, cfg.thresh, 1.0, 0.0)
    assert ctx().kwindows() == 1



def test_near_tie_task_list_bit_exact():
    """Screening at near-ties (tests/designed.py: all block norms equal in exact
    arithmetic, threshold at the median estimate): the device's task list is
    the oracle's, triple for triple, because both screen with the same
    canonical block norms and the same FP64 product and comparison."""
    import designed

    a, b, bs = designed.near_tie_case()
    n = a.shape[0]
    oA, dA = pair_uniform(a, bs, bs, 0.0)
    oB, dB = pair_uniform(b, bs, bs, 0.0)
    _, na = oA.meta(); _, nbm = oB.meta()
    est = (na[:, :, None] * nbm[None, :, :]).ravel()
    thresh = float(np.median(np.sort(est)[: est.size // 2 * 2 - 1])) / float(bs ** 3)
    oC = oracle.OBsm.uniform(n, n, bs, bs, 0.0)
    t_or = oracle.matmult(oC, oA, 0, oB, 0, thresh)
    dC = sa.BSM.uniform(ctx(), n, n, bs, bs, 0.0)
    sa.matmult(dC, dA, 0, dB, 0, thresh)
    t_dev = ctx().tasklist()
    assert 0 < len(t_or) < (n // bs) ** 3
    assert np.array_equal(t_or, t_dev)
    assert rel_fro(dC.to_dense(), oC.to_dense()) < 1e-13


def test_uneven_atoms_full_size():
    """The uneven-atoms workload at full size (n = 12000, 5-40 atom blocks,
    10 % fill, ~2.4e6 screened products): operands generated in HBM and read
    back for the oracle (the generator is pinned to its host definition in
    test_gpu_bsm), task list bit-exact, C within 1e-12."""
    w = synth.atoms_workload()
    dA = sa.BSM(ctx(), w.rsz, w.ksz, w.thresh).generate(1, w.a_scale)
    dB = sa.BSM(ctx(), w.ksz, w.csz, w.thresh).generate(2, w.b_scale)
    oA = oracle.OBsm(w.rsz, w.ksz, w.thresh).from_dense(dA.to_dense())
    oB = oracle.OBsm(w.ksz, w.csz, w.thresh).from_dense(dB.to_dense())
    assert_same_meta(oA, dA)
    assert_same_meta(oB, dB)
    oC, dC = empty_pair_blocked(w.rsz, w.csz)
    n, err = check_mm(oC, dC, oA, dA, 0, oB, dB, 0, w.thresh, 1.0, 0.0)
    assert n > 2_000_000
