"""GPU split-precision (Ozaki) screened multiply: the screened task list is
the FP64 one (bit-exact vs the oracle) and C is within 1e-10 relative
Frobenius error of the FP64 oracle, overall and per block row (bar stated per
assert).  Integer-valued inputs whose bits fit the leading slice are exact."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from strassen_amd import synth
from gpu_util import assert_same_meta, ctx, empty_pair_blocked, empty_pair_uniform, pair_blocked, pair_uniform, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-10  # split-precision bar vs the FP64 oracle


def block_row_errors(got, ref, rsz):
    edges = np.concatenate([[0], np.cumsum(rsz)])
    out = []
    for i in range(len(rsz)):
        r = ref[edges[i]:edges[i + 1]]
        n = np.linalg.norm(r)
        if n > 0:
            out.append(np.linalg.norm(got[edges[i]:edges[i + 1]] - r) / n)
    return np.array(out)


def run(oA, dA, ta, oB, dB, tb, oC, dC, thresh, opts, alpha=1.0, beta=0.0):
    t_o = oracle.matmult(oC, oA, ta, oB, tb, thresh, alpha, beta)
    sa.matmult(dC, dA, ta, dB, tb, thresh, alpha, beta, opts)
    assert np.array_equal(t_o, ctx().tasklist())
    assert_same_meta(oC, dC)
    return oC.to_dense(), dC.to_dense()


@pytest.mark.parametrize("slices,stype", [(4, sa.SLICE_FP16), (5, sa.SLICE_FP16), (5, sa.SLICE_BF16),
                                          (6, sa.SLICE_BF16), (4, sa.SLICE_INT8), (5, sa.SLICE_INT8),
                                          (6, sa.SLICE_INT8)])
def test_band_error_per_block_row(slices, stype):
    cfg = synth.band_workload(n=1024, bs=64)
    a = synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale)
    b = synth.dense_from_scales(2, cfg.ksz, cfg.csz, cfg.b_scale)
    oA, dA = pair_uniform(a, 64, 64, cfg.thresh)
    oB, dB = pair_uniform(b, 64, 64, cfg.thresh)
    oC, dC = empty_pair_uniform(1024, 1024, 64, 64)
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, cfg.thresh, sa.default_opts(mode=sa.MODE_OZAKI, slices=slices,
                                                                               slice_type=stype))
    err = rel_fro(got, ref)
    rows = block_row_errors(got, ref, cfg.rsz)
    bound = TOL if (stype == sa.SLICE_FP16 or slices >= 5) else 1e-8  # 4 BF16 / INT8 slices: 30-31 bits
    assert err < bound, err
    assert rows.max() < bound, rows.max()


@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_uneven_transposes(ta, tb):
    w = synth.small_uneven(n=400, fill=0.4, seed=8)
    a = synth.dense_from_scales(1, w.rsz, w.ksz, w.a_scale)
    b = synth.dense_from_scales(2, w.ksz, w.csz, w.b_scale)
    if ta:
        a = np.asfortranarray(a.T)
    if tb:
        b = np.asfortranarray(b.T)
    oA, dA = pair_blocked(a, w.ksz if ta else w.rsz, w.rsz if ta else w.ksz, w.thresh)
    oB, dB = pair_blocked(b, w.csz if tb else w.ksz, w.ksz if tb else w.csz, w.thresh)
    oC, dC = empty_pair_blocked(w.rsz, w.csz)
    ref, got = run(oA, dA, ta, oB, dB, tb, oC, dC, w.thresh, sa.default_opts(mode=sa.MODE_OZAKI, slices=4))
    assert rel_fro(got, ref) < TOL


@pytest.mark.parametrize("stype,slices", [(sa.SLICE_FP16, 4), (sa.SLICE_BF16, 5), (sa.SLICE_INT8, 5)])
def test_deep_k_fold_schedule(stype, slices):
    """Products with k = 1000 and 1024 (32 k-steps, one ragged) chain over
    several fold intervals and product boundaries of one output block."""
    rng = np.random.default_rng(21)
    ksz = [1024, 1000, 1024, 37]
    a = rng.standard_normal((80, sum(ksz))) * np.exp(rng.standard_normal((80, 1)))
    b = rng.standard_normal((sum(ksz), 72)) * np.exp(rng.standard_normal((1, 72)))
    oA, dA = pair_blocked(a, [80], ksz, 0.0)
    oB, dB = pair_blocked(b, ksz, [40, 32], 0.0)
    oC, dC = empty_pair_blocked([80], [40, 32])
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, 0.0, sa.default_opts(mode=sa.MODE_OZAKI, slices=slices,
                                                                       slice_type=stype))
    assert rel_fro(got, ref) < TOL


@pytest.mark.parametrize("slices", [4, 5, 6])
def test_int8_extreme_digits_short_products(slices):
    """Entries whose 5-slice INT8 digits sit at the ends of their ranges
    (leading digit near +-127, the others -128 or 127) in products of 1 to 4
    k-steps, against a host restatement of the split product (exact integer
    slice sums, FP64 only in the fold)."""
    rng = np.random.default_rng(50 + slices)
    pats = np.array([272722001792, 272722001791, -272738844800, -272738844801], dtype=np.float64) * 2.0 ** -38
    ksz = [192, 130, 64, 200, 17]
    rsz, csz = [48, 80], [64, 33]
    a = rng.choice(pats, (sum(rsz), sum(ksz)))
    b = rng.choice(pats, (sum(ksz), sum(csz)))
    a[:, :192] = pats[0]  # one product of 3 k-steps with every digit pair at its largest
    b[:192, :] = pats[0]
    oA, dA = pair_blocked(a, rsz, ksz, 0.0)
    oB, dB = pair_blocked(b, ksz, csz, 0.0)
    oC, dC = empty_pair_blocked(rsz, csz)
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=slices, slice_type=sa.SLICE_INT8))
    # the scheme itself: exact integer slice products, FP64 only in the fold
    assert rel_fro(got, int8_slice_model(a, rsz, ksz, b, csz, slices)) < 1e-14
    if slices >= 5:  # 4 slices round these coherent inputs to 2^-31: ~2e-9 against FP64
        assert rel_fro(got, ref) < TOL


def int8_slice_model(a, rsz, ksz, b, csz, S):
    """Split product with INT8 digits restated on the host: per row of an
    A block / column of a B block the scale 2^e above its largest entry mx,
    lowered to 2^(e-1) when rint(mx 2^(8S-1-e)) + H < 2^(8S-1) (the leading
    digit then still fits int8), n = rint(x 2^(8S-2-e)), digits of n + H
    (leading one m >> 8(S-1), the others byte - 128), slice pairs
    p + q <= S - 1 summed exactly in int64, folded into FP64 per product in k
    order."""
    H = sum(128 << (8 * j) for j in range(S - 1))

    def digits(x):  # vectors along axis 0
        mx = np.abs(x).max(axis=1)
        e = np.where(mx > 0, np.frexp(mx)[1], 0)
        lower = (mx > 0) & (np.rint(np.ldexp(mx, 8 * S - 1 - e)).astype(np.int64) + H < (1 << (8 * S - 1)))
        e = e - lower.astype(e.dtype)
        m = np.rint(np.ldexp(x, (8 * S - 2 - e)[:, None])).astype(np.int64) + H
        lead = m >> (8 * (S - 1))
        assert lead.min() >= -128 and lead.max() <= 127
        return [m >> (8 * (S - 1))] + [((m >> (8 * j)) & 255) - 128 for j in range(S - 2, -1, -1)], e

    re_, ke, ce = np.cumsum([0] + rsz), np.cumsum([0] + ksz), np.cumsum([0] + csz)
    C = np.zeros((re_[-1], ce[-1]))
    for i in range(len(rsz)):
        for j in range(len(csz)):
            acc = np.zeros((rsz[i], csz[j]))
            for k in range(len(ksz)):
                da, er = digits(a[re_[i]:re_[i + 1], ke[k]:ke[k + 1]])
                db, ec = digits(b[ke[k]:ke[k + 1], ce[j]:ce[j + 1]].T)
                t = np.zeros((rsz[i], csz[j]))
                for w in range(S):
                    sw = sum(da[p] @ db[w - p].T for p in range(w + 1))
                    t += np.ldexp(sw.astype(np.float64), -8 * w)
                acc += np.ldexp(t, er[:, None] + ec[None, :] + 4 - 16)
            C[re_[i]:re_[i + 1], ce[j]:ce[j + 1]] = acc
    return C


@pytest.mark.parametrize("slices,stype", [(5, sa.SLICE_INT8), (4, sa.SLICE_FP16), (5, sa.SLICE_BF16)])
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0), (0, 1)])
def test_shared_exponents(slices, stype, ta, tb):
    """Exponents shared along op(A) rows / op(B) columns across inner blocks:
    the products of an output chain in one accumulation (many short products
    of ragged blocks with decaying norms, so the shared scale is not every
    block's own; vectors that are zero in one block)."""
    shared_case(slices, stype, ta, tb, -1)


@pytest.mark.parametrize("tile", [23, 26, 29, 31, 32, 33])
def test_shared_exponents_every_int8_kernel(tile):
    """The same with each 5-slice INT8 kernel forced: 64x64 and 32x32 chain
    the products, the deep and wide kernels fold each one."""
    shared_case(5, sa.SLICE_INT8, 0, 0, tile)


def shared_case(slices, stype, ta, tb, tile):
    rng = np.random.default_rng(40 + slices + 3 * ta + 5 * tb)
    rsz, ksz, csz = [24, 40, 33], [17, 64, 9, 48, 30, 21, 64, 5], [36, 28]
    ni, nk, nj = sum(rsz), sum(ksz), sum(csz)
    kb = np.repeat(np.arange(len(ksz)), ksz)
    a = rng.standard_normal((ni, nk)) * np.exp(-0.7 * kb)[None, :]
    b = rng.standard_normal((nk, nj)) * np.exp(-0.4 * kb)[:, None]
    a *= 1e-6  # below 1: a zero vector must not lift its line's exponent to 0
    a[:8, :17] = 0.0
    b[81:90, :5] = 0.0
    a_, b_ = (a.T.copy(), ksz, rsz) if ta else (a, rsz, ksz), (b.T.copy(), csz, ksz) if tb else (b, ksz, csz)
    oA, dA = pair_blocked(np.asfortranarray(a_[0]), a_[1], a_[2], 0.0)
    oB, dB = pair_blocked(np.asfortranarray(b_[0]), b_[1], b_[2], 0.0)
    oC, dC = empty_pair_blocked(rsz, csz)
    ref, got = run(oA, dA, ta, oB, dB, tb, oC, dC, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=slices, slice_type=stype, shared_exp=1, tile=tile))
    assert rel_fro(got, ref) < TOL


DEEP = [(12, 3, sa.SLICE_FP16), (13, 4, sa.SLICE_FP16), (14, 5, sa.SLICE_FP16), (15, 3, sa.SLICE_FP16),
        (16, 4, sa.SLICE_FP16), (17, 5, sa.SLICE_FP16), (18, 4, sa.SLICE_BF16), (19, 5, sa.SLICE_BF16),
        (20, 4, sa.SLICE_BF16), (21, 5, sa.SLICE_BF16), (22, 4, sa.SLICE_INT8), (23, 5, sa.SLICE_INT8),
        (24, 6, sa.SLICE_INT8), (25, 4, sa.SLICE_INT8), (26, 5, sa.SLICE_INT8), (28, 4, sa.SLICE_INT8),
        (29, 5, sa.SLICE_INT8), (30, 4, sa.SLICE_INT8), (31, 5, sa.SLICE_INT8), (32, 5, sa.SLICE_INT8),
        (33, 5, sa.SLICE_INT8), (34, 4, sa.SLICE_INT8), (35, 4, sa.SLICE_INT8)]


@pytest.mark.parametrize("tile,slices,stype", DEEP)
@pytest.mark.parametrize("ta,tb", [(0, 0), (1, 0)])
def test_deep_pipelined_kernels(tile, slices, stype, ta, tb):
    """Every kernel table entry by id (64x64, 32x32, 128x64 / 64x128 deep and
    wide) on ragged 150-blocks (several tiles, partial fragments and waves
    without live fragments, multi-product outputs, several k-steps)."""
    bs = 150
    rng = np.random.default_rng(500 + tile + ta)
    ni, nk, nj = 2 * bs + 7, 3 * bs + 3, 2 * bs - 1
    a_shape = (nk, ni) if ta else (ni, nk)
    b_shape = (nj, nk) if tb else (nk, nj)
    a = rng.standard_normal(a_shape)
    b = rng.standard_normal(b_shape)
    oA, dA = pair_uniform(a, bs, bs, 0.0)
    oB, dB = pair_uniform(b, bs, bs, 0.0)
    oC, dC = empty_pair_uniform(ni, nj, bs, bs)
    ref, got = run(oA, dA, ta, oB, dB, tb, oC, dC, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=slices, slice_type=stype, tile=tile))
    # bars: 2^-(beta S - TS) truncation; 3 FP16 slices (28 bits) and 4 BF16 slices (31 bits) are coarser
    bound = {(3, sa.SLICE_FP16): 1e-7, (4, sa.SLICE_BF16): 1e-8, (4, sa.SLICE_INT8): 1e-8}.get((slices, stype), TOL)
    assert rel_fro(got, ref) < bound


@pytest.mark.parametrize("stype,slices,tile", [(sa.SLICE_INT8, 5, -1), (sa.SLICE_INT8, 5, 26), (sa.SLICE_INT8, 5, 32),
                                               (sa.SLICE_INT8, 4, -1), (sa.SLICE_FP16, 4, -1), (sa.SLICE_BF16, 5, -1)])
@pytest.mark.parametrize("shared", [0, 1])
@pytest.mark.parametrize("ta", [0, 1])
@pytest.mark.parametrize("zmin", ["3", "0"])
def test_zero_slice_skip_bit_identical(monkeypatch, stype, slices, tile, shared, ta, zmin):
    """Inner blocks 2^-9 .. 2^-45 below the largest of their lines: with
    shared exponents their leading slices are all zero, and outputs with
    enough zero slice pairs (SA_OZ_ZERO_MIN sixteenths; 0: any) take the path
    that skips them (neither loaded nor multiplied).  C is bit-identical to
    the run that multiplies every pair (SA_OZ_ZERO_SKIP=0), on the 64x64,
    32x32 and wide kernels, and within the bar of the FP64 oracle."""
    monkeypatch.setenv("SA_OZ_ZERO_MIN", zmin)
    rng = np.random.default_rng(70 + slices + 7 * shared + 3 * ta)
    rsz, ksz, csz = [100, 37, 64], [64, 50, 64, 30, 64, 17], [80, 48]
    kb = np.repeat(np.arange(len(ksz)), ksz)
    a = rng.standard_normal((sum(rsz), sum(ksz))) * np.exp2(-9.0 * kb)[None, :]
    b = rng.standard_normal((sum(ksz), sum(csz)))
    a_, ar, ac = (a.T.copy(), ksz, rsz) if ta else (a, rsz, ksz)
    outs = []
    for skip in ("0", "1"):
        monkeypatch.setenv("SA_OZ_ZERO_SKIP", skip)
        oA, dA = pair_blocked(np.asfortranarray(a_), ar, ac, 0.0)
        oB, dB = pair_blocked(b, ksz, csz, 0.0)
        oC, dC = empty_pair_blocked(rsz, csz)
        ref, got = run(oA, dA, ta, oB, dB, 0, oC, dC, 0.0,
                       sa.default_opts(mode=sa.MODE_OZAKI, slices=slices, slice_type=stype, shared_exp=shared,
                                       tile=tile))
        outs.append(got)
    assert np.array_equal(outs[0], outs[1])
    bound = TOL if (stype == sa.SLICE_FP16 or slices >= 5) else 1e-8  # 4 INT8 slices: 31 bits
    assert rel_fro(outs[1], ref) < bound


def test_zero_slice_pair_counts(monkeypatch):
    """The context's slice-pair counters: no pair is skipped without shared
    exponents or with skipping switched off, some are with shared exponents
    on decaying blocks, and multiplied + skipped is the same work either way."""
    rng = np.random.default_rng(77)
    rsz, ksz, csz = [100, 64], [64, 50, 64, 30, 64], [80, 48]
    kb = np.repeat(np.arange(len(ksz)), ksz)
    a = rng.standard_normal((sum(rsz), sum(ksz))) * np.exp2(-9.0 * kb)[None, :]
    b = rng.standard_normal((sum(ksz), sum(csz)))
    counts = {}
    for shared, skip in ((0, "1"), (1, "0"), (1, "1")):
        monkeypatch.setenv("SA_OZ_ZERO_SKIP", skip)
        monkeypatch.setenv("SA_OZ_ZERO_MIN", "0")
        _, dA = pair_blocked(a, rsz, ksz, 0.0)
        _, dB = pair_blocked(b, ksz, csz, 0.0)
        _, dC = empty_pair_blocked(rsz, csz)
        ctx().slice_pairs()
        sa.matmult(dC, dA, 0, dB, 0, 0.0, 1.0, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=5, slice_type=sa.SLICE_INT8, shared_exp=shared))
        counts[(shared, skip)] = ctx().slice_pairs()
    assert counts[(0, "1")][1] == 0 and counts[(1, "0")][1] == 0
    assert counts[(1, "1")][1] > 0
    assert sum(counts[(1, "1")]) == sum(counts[(1, "0")]) == sum(counts[(0, "1")]) > 0
    assert ctx().slice_pairs() == (0, 0)


@pytest.mark.parametrize("stype", [sa.SLICE_FP16, sa.SLICE_BF16, sa.SLICE_INT8])
def test_integer_inputs_exact(stype):
    """Inputs with at most 3 significant bits per row max live in the leading
    slice, so every dropped slice product is zero and the result is exact."""
    rng = np.random.default_rng(3)
    a = rng.integers(-4, 5, (96, 80)).astype(float)
    b = rng.integers(-4, 5, (80, 112)).astype(float)
    oA, dA = pair_uniform(a, 16, 16, 0.0)
    oB, dB = pair_uniform(b, 16, 16, 0.0)
    oC, dC = empty_pair_uniform(96, 112, 16, 16)
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=4, slice_type=stype), alpha=1.0, beta=0.0)
    assert np.array_equal(got, a @ b)


def test_existing_c_and_alpha():
    rng = np.random.default_rng(4)
    a = rng.standard_normal((100, 70)); b = rng.standard_normal((70, 90)); c0 = rng.standard_normal((100, 90))
    oA, dA = pair_uniform(a, 20, 20, 0.0)
    oB, dB = pair_uniform(b, 20, 20, 0.0)
    oC, dC = pair_uniform(c0, 20, 20, 0.0)
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, 0.0, sa.default_opts(mode=sa.MODE_OZAKI, slices=4),
                   alpha=-0.75, beta=1.0)
    assert rel_fro(got, ref) < TOL


def test_unsupported_slice_count_rejected():
    _, dA = pair_uniform(np.ones((32, 32)), 16, 16, 0.0)
    _, dB = pair_uniform(np.ones((32, 32)), 16, 16, 0.0)
    _, dC = empty_pair_uniform(32, 32, 16, 16)
    with pytest.raises(sa.SAError):
        sa.matmult(dC, dA, 0, dB, 0, 0.0, 1.0, 0.0, sa.default_opts(mode=sa.MODE_OZAKI, slices=9))


def test_oversized_blocks_rejected():
    """The split GEMM addresses a block's slices with 32-bit offsets in 16-byte
    units: a 90000 x 90000 block at 5 INT8 slices (5 * 5625 * 1407 * 64 units,
    above 2^31) is refused before any work; the operands are empty."""
    big = [90000]
    dA, dB, dC = (sa.BSM(ctx(), big, big, 0.0) for _ in range(3))
    with pytest.raises(sa.SAError, match="too large") as err:
        sa.matmult(dC, dA, 0, dB, 0, 0.0, 1.0, 0.0,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=5, slice_type=sa.SLICE_INT8))
    assert err.value.code == sa.ERR_UNSUPPORTED


def test_kwindows_split_precision(monkeypatch):
    """k-windows under the split-precision kernel: same task list, same bar."""
    cfg = synth.band_workload(n=1024, bs=64)
    a = synth.dense_from_scales(1, cfg.rsz, cfg.ksz, cfg.a_scale)
    b = synth.dense_from_scales(2, cfg.ksz, cfg.csz, cfg.b_scale)
    oA, dA = pair_uniform(a, 64, 64, cfg.thresh)
    oB, dB = pair_uniform(b, 64, 64, cfg.thresh)
    oC, dC = empty_pair_uniform(1024, 1024, 64, 64)
    monkeypatch.setenv("SA_KWINDOW_MB", "1")
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, cfg.thresh, sa.default_opts(mode=sa.MODE_OZAKI, slices=4))
    assert ctx().kwindows() > 4
    assert rel_fro(got, ref) < TOL



@pytest.mark.parametrize("slices,stype", [(4, sa.SLICE_FP16), (3, sa.SLICE_FP16), (4, sa.SLICE_BF16),
                                          (5, sa.SLICE_INT8)])
def test_band_full_size_error_per_block_row(slices, stype):
    """The split-precision band workload at full size (n = 2048, 64-blocks,
    band 4, thresh 1e-10) with 3-4 FP16 / 4 BF16 slices as the workload names
    and the INT8 5-slice form: task list bit-exact against the FP64 oracle and
    the relative error per block row stated against a bound set by the digit
    count (2^-(bits) growth: 3 FP16 slices carry 28 bits, 4 BF16 31)."""
    cfg = synth.band_workload()
    dA = sa.BSM.uniform(ctx(), 2048, 2048, 64, 64, cfg.thresh).generate(1, cfg.a_scale)
    dB = sa.BSM.uniform(ctx(), 2048, 2048, 64, 64, cfg.thresh).generate(2, cfg.b_scale)
    oA = oracle.OBsm.uniform(2048, 2048, 64, 64, cfg.thresh).from_dense(dA.to_dense())
    oB = oracle.OBsm.uniform(2048, 2048, 64, 64, cfg.thresh).from_dense(dB.to_dense())
    oC, dC = empty_pair_uniform(2048, 2048, 64, 64)
    ref, got = run(oA, dA, 0, oB, dB, 0, oC, dC, cfg.thresh,
                   sa.default_opts(mode=sa.MODE_OZAKI, slices=slices, slice_type=stype))
    rows = block_row_errors(got, ref, cfg.rsz)
    bound = {(4, sa.SLICE_FP16): TOL, (3, sa.SLICE_FP16): 1e-7, (4, sa.SLICE_BF16): 1e-8,
             (5, sa.SLICE_INT8): TOL}[(slices, stype)]
    print(f"band n=2048 {slices} slices type {stype}: rel err {rel_fro(got, ref):.3e}, "
          f"per block row max {rows.max():.3e} median {np.median(rows):.3e}")
    assert rel_fro(got, ref) < bound
    assert rows.max() < bound
    assert len(ctx().tasklist()) == 2292
