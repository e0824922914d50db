"""GPU parity of the hybrid Strassen-sparse path (strassen.hpp): dry-run
decision trees bit-exact against the oracle, screened 8-path leaf task lists
bit-exact, integer-valued inputs reproduce the dense product exactly (every
FP64 operation exact), random inputs within 1e-12 relative Frobenius error of
the oracle run on the same tree."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from strassen_amd import synth
from gpu_util import assert_same_meta, ctx, pair_uniform, rel_fro

pytestmark = pytest.mark.gpu
TOL = 1e-12


def inputs(rng, nb, bs, kind, integer=False, ragged=0):
    n = nb * bs - ragged
    sz = synth.uniform_sizes(n, bs)
    if integer:
        a = rng.integers(-4, 5, (n, n)).astype(float)
        b = rng.integers(-4, 5, (n, n)).astype(float)
    else:
        a = rng.standard_normal((n, n))
        b = rng.standard_normal((n, n))
    if kind != "dense":
        if kind == "band":
            s = synth.band_scales(len(sz), 3, 0.5) > 0
        else:
            s = rng.random((len(sz), len(sz))) < 0.5
        rb = np.repeat(np.arange(len(sz)), sz)
        a[~s[rb][:, rb]] = 0.0
        b[~s.T[rb][:, rb]] = 0.0
    return n, np.asfortranarray(a), np.asfortranarray(b)


@pytest.mark.parametrize("nb,bs,kind", [(8, 32, "dense"), (16, 64, "dense"), (16, 64, "band"), (13, 32, "random"),
                                        (32, 16, "random"), (9, 64, "dense")])
def test_dryrun_bits_exact(nb, bs, kind):
    rng = np.random.default_rng(nb * bs)
    n, a, b = inputs(rng, nb, bs, kind, ragged=(5 if nb == 9 else 0))
    oA, dA = pair_uniform(a, bs, bs, 1e-12)
    oB, dB = pair_uniform(b, bs, bs, 1e-12)
    thr = 1e-10 * bs ** 3
    lo, fo = oracle.strassen_dryrun(oA, oB, thr)
    ld, fd = sa.strassen_dryrun(dA, dB, thr)
    assert len(lo) == len(ld)
    for l, (x, y) in enumerate(zip(lo, ld)):
        assert np.array_equal(x, y), f"level {l} differs"
    assert np.array_equal(fo, fd)


def run_both(a, b, n, bs, thresh, cutoff, tree=None, cfill=0.0):
    oA, dA = pair_uniform(a, bs, bs, 0.0)
    oB, dB = pair_uniform(b, bs, bs, 0.0)
    oC = oracle.OBsm.uniform(n, n, bs, bs, 0.0); oC.fill(cfill)
    dC = sa.BSM.uniform(ctx(), n, n, bs, bs, 0.0); dC.fill(cfill)
    leaf_o, nroots_o = oracle.strassen_sparse(oC, oA, oB, thresh, cutoff, tree=tree)
    sa.matmult_strassen_sparse(dC, dA, dB, thresh, cutoff, tree=tree)
    st = sa.strassen_last_stats()
    leaf_d = ctx().tasklist()
    assert np.array_equal(leaf_o, leaf_d), f"leaf task lists differ ({len(leaf_o)} vs {len(leaf_d)})"
    assert st["nroots"] == nroots_o
    return oC.to_dense(), dC.to_dense(), st


@pytest.mark.parametrize("nb,bs,cutoff,kind,levels", [
    (8, 4, 0, "dense", {1}), (16, 4, 8, "dense", {2}), (16, 2, 0, "random", {0, 3}),
    (32, 2, 4, "dense", {4}), (16, 4, 0, "random", {1}), (16, 8, 0, "band", {0, 2}),
    (8, 32, 0, "dense", None), (16, 64, 512, "dense", None)])
def test_integer_inputs_exact(nb, bs, cutoff, kind, levels):
    rng = np.random.default_rng(nb + bs + cutoff)
    n, a, b = inputs(rng, nb, bs, kind, integer=True)
    tree = None if levels is None else oracle.forced_tree(nb, levels)
    co, cd, st = run_both(a, b, n, bs, 0.0, cutoff, tree)
    assert np.array_equal(co, a @ b)
    assert np.array_equal(cd, a @ b)
    if levels is None or levels:
        assert st["nroots"] > 0


@pytest.mark.parametrize("nb,bs,cutoff,kind,levels", [
    (16, 64, 512, "dense", None), (16, 32, 0, "random", {1, 2}), (32, 16, 64, "band", {2}),
    (12, 32, 0, "random", {1})])
def test_random_inputs_vs_oracle(nb, bs, cutoff, kind, levels):
    rng = np.random.default_rng(3 * nb + bs)
    n, a, b = inputs(rng, nb, bs, kind)
    tree = None if levels is None else oracle.forced_tree(nb, levels)
    co, cd, st = run_both(a, b, n, bs, 1e-10, cutoff, tree, cfill=0.25)
    assert rel_fro(cd, co) < TOL
    assert rel_fro(cd, a @ b + 0.25) < 1e-12


def test_levels_run_as_one_graph_repeatedly():
    """The execution phase (pre-adds, leaf GEMM, post-adds) runs as one HIP
    graph per call: three calls in a row on new integer inputs (and eager
    screened multiplies between them) stay exact, one graph launch each."""
    c = ctx()
    c.set_profiling(True)
    c.profile_reset()
    try:
        for seed in range(3):
            rng = np.random.default_rng(90 + seed)
            n, a, b = inputs(rng, 16, 4, "dense", integer=True)
            co, cd, st = run_both(a, b, n, 4, 0.0, 0, oracle.forced_tree(16, {1, 2}))
            assert st["nroots"] > 0
            assert np.array_equal(cd, a @ b)
            oA, dA = pair_uniform(a, 4, 4, 0.0)
            oB, dB = pair_uniform(b, 4, 4, 0.0)
            dC = sa.BSM.uniform(c, n, n, 4, 4, 0.0)
            sa.matmult(dC, dA, 0, dB, 0, 0.0, 1.0, 0.0)
            assert np.array_equal(dC.to_dense(), a @ b)
        c.synchronize()
        assert c.profile("graph")[1] == 3
    finally:
        c.set_profiling(False)


def test_graph_executable_replayed():
    """One executable graph is kept per context: a call of the same shape
    updates it in place from the new capture and launches it again, a call of
    another shape instantiates it anew; every replay is exact on new integer
    inputs, with eager screened multiplies between the replays."""
    c = ctx()
    s0 = c.graph_stats()
    shapes = [(16, 4, {1, 2})] * 3 + [(8, 8, {1})] + [(16, 4, {1, 2})] * 2
    for i, (nb, bs, levels) in enumerate(shapes):
        rng = np.random.default_rng(200 + i)
        n, a, b = inputs(rng, nb, bs, "dense", integer=True)
        co, cd, st = run_both(a, b, n, bs, 0.0, 0, oracle.forced_tree(nb, levels))
        assert st["nroots"] > 0
        assert np.array_equal(cd, a @ b), f"call {i}"
        _, dA = pair_uniform(a, bs, bs, 0.0)
        _, dB = pair_uniform(b, bs, bs, 0.0)
        dC = sa.BSM.uniform(c, n, n, bs, bs, 0.0)
        sa.matmult(dC, dA, 0, dB, 0, 0.0, 1.0, 0.0)
        assert np.array_equal(dC.to_dense(), a @ b)
    s1 = c.graph_stats()
    d = {k: s1[k] - s0[k] for k in s1}
    assert d["launched"] == len(shapes)
    assert d["instantiated"] + d["updated"] == len(shapes)
    assert d["updated"] >= 3, d       # calls 2, 3 and 6 replay the kept executable
    assert d["instantiated"] >= 2, d  # the other shape, and the way back


def random_plan(rng, nb, p_s, p_d):
    sizes = oracle.tree_level_sizes(nb)
    return ([(rng.random(s) < p_s).astype(np.uint8) for s in sizes],
            [(rng.random(s) < p_d).astype(np.uint8) for s in sizes])


def interior(nb, bits):
    """Clear bits of nodes that reach past the block grid (edge super-blocks)."""
    out = []
    for lev, b in enumerate(bits):
        w = 2 ** (lev + 1)
        ns = -(-nb // w)
        idx = np.arange(b.size)
        si, sj, sk = idx % ns, (idx // ns) % ns, idx // (ns * ns)
        ok = ((si + 1) * w <= nb) & ((sj + 1) * w <= nb) & ((sk + 1) * w <= nb)
        out.append((b * ok).astype(np.uint8))
    return out


@pytest.mark.parametrize("seed,nb,bs,integer", [(1, 8, 16, True), (2, 16, 8, True), (3, 16, 32, False),
                                                (4, 12, 16, False), (5, 32, 8, True)])
def test_dense_roots_vs_oracle(seed, nb, bs, integer):
    """Plans mixing Strassen and dense roots (random bits at every level):
    screened leaf task lists bit-exact, integer inputs exact, random inputs
    within the FP64 bar of the oracle run on the same plan."""
    rng = np.random.default_rng(seed)
    n, a, b = inputs(rng, nb, bs, "random", integer=integer)
    tree, dense = random_plan(rng, nb, 0.1, 0.2)
    tree, dense = interior(nb, tree), interior(nb, dense)
    oA, dA = pair_uniform(a, bs, bs, 0.0)
    oB, dB = pair_uniform(b, bs, bs, 0.0)
    oC = oracle.OBsm.uniform(n, n, bs, bs, 0.0); oC.fill(0.5)
    dC = sa.BSM.uniform(ctx(), n, n, bs, bs, 0.0); dC.fill(0.5)
    leaf_o, nr, nd = oracle.strassen_sparse(oC, oA, oB, 1e-10, 0, tree=tree, dense=dense, counts=True)
    sa.matmult_strassen_sparse(dC, dA, dB, 1e-10, 0, tree=tree, dense=dense)
    st = sa.strassen_last_stats()
    assert np.array_equal(leaf_o, ctx().tasklist())
    assert (st["nroots"], st["dense_roots"]) == (nr, nd)
    assert nd > 0
    if integer:
        assert np.array_equal(dC.to_dense(), oC.to_dense())
    else:
        assert rel_fro(dC.to_dense(), oC.to_dense()) < TOL


def test_empty_c_blocks_drop_products():
    """C must be pre-filled; products into empty C blocks are dropped."""
    rng = np.random.default_rng(9)
    n, a, b = inputs(rng, 8, 32, "dense", integer=True)
    oA, dA = pair_uniform(a, 32, 32, 0.0)
    oB, dB = pair_uniform(b, 32, 32, 0.0)
    c0 = np.zeros((n, n)); c0[:64, :] = 1.0
    oC, dC = pair_uniform(c0, 32, 32, 1e-12)  # only the first two block rows exist
    oracle.strassen_sparse(oC, oA, oB, 0.0, 0)
    sa.matmult_strassen_sparse(dC, dA, dB, 0.0, 0)
    assert_same_meta(oC, dC)
    assert np.array_equal(oC.to_dense(), dC.to_dense())
    assert np.array_equal(dC.to_dense()[:64], (a @ b + 1.0)[:64])


def test_preconditions_rejected():
    _, dA = pair_uniform(np.ones((64, 64)), 16, 16, 0.0)
    _, dB = pair_uniform(np.ones((64, 64)), 16, 16, 0.0)
    _, dC = pair_uniform(np.ones((64, 64)), 16, 16, 0.0)
    L = sa.lib()
    assert L.sa_matmult_strassen_sparse(dC.h, dA.h, 1, dB.h, 0, 0.0, 1.0, 0.0, None) == -1  # transpose
    assert L.sa_matmult_strassen_sparse(dC.h, dA.h, 0, dB.h, 0, 0.0, 2.0, 0.0, None) == -1  # alpha != 1
    _, dR = pair_uniform(np.ones((64, 48)), 16, 16, 0.0)
    assert L.sa_matmult_strassen_sparse(dC.h, dR.h, 0, dB.h, 0, 0.0, 1.0, 0.0, None) == -1  # not square
    # a caller tree that selects an edge super-block is rejected before any launch
    _, dE = pair_uniform(np.ones((60, 60)), 16, 16, 0.0)
    _, dF = pair_uniform(np.ones((60, 60)), 16, 16, 0.0)
    _, dG = pair_uniform(np.ones((60, 60)), 16, 16, 0.0)
    with pytest.raises(sa.SAError):
        sa.matmult_strassen_sparse(dG, dE, dF, 0.0, 0, tree=oracle.forced_tree(4, {1}))


def hybrid_shape_integer(nb, bs, band, seed):
    """Band matrices of the large hybrid workload's shape (128-blocks, band
    |I-J| <= band) with integer entries in [-4, 4] times power-of-two block
    scales 2^-(|I-J| // 8): every FP64 operation of every path (screened
    leaves, Winograd pre/post-adds, dense roots) is exact, so C must equal the
    dense product bit for bit."""
    rng = np.random.default_rng(seed)
    n = nb * bs
    sz = synth.uniform_sizes(n, bs)
    I = np.arange(nb)
    d = np.abs(I[:, None] - I[None, :])
    sc = np.where(d <= band, 2.0 ** -(d // 8), 0.0)
    rb = np.repeat(np.arange(nb), sz)
    a = rng.integers(-4, 5, (n, n)).astype(float) * sc[rb][:, rb]
    b = rng.integers(-4, 5, (n, n)).astype(float) * sc[rb][:, rb]
    return n, np.asfortranarray(a), np.asfortranarray(b)


@pytest.mark.parametrize("cutoff", [2048, 512])
def test_hybrid_workload_shape_exact_batched(cutoff, monkeypatch):
    """The large hybrid recursion at n = 8192 in its own blocking (128-blocks,
    band 24 of 64 block rows ~ the n = 32768 band 95 of 256), threshold 1e-10:
    the cost-model tree, the learned plan (committed model) and a forced
    level-2 tree, each with the Strassen scratch budget capped (SA_SCRATCH_MB)
    so that every root runs in a batch of its own.  cutoff 2048 is the
    workload's (min_size_for_strassen); 512 lets the recursion reach depth 4
    at this size."""
    monkeypatch.setenv("SA_SCRATCH_MB", "1")
    nb, bs = 64, 128
    n, a, b = hybrid_shape_integer(nb, bs, 24, 7 + cutoff)
    ref = a @ b
    c = ctx()
    sa.mlp_load(c, sa.DEFAULT_PATH_MODEL)
    dA = sa.BSM.uniform(c, n, n, bs, bs, 1e-10).from_dense(a)
    dB = sa.BSM.uniform(c, n, n, bs, bs, 1e-10).from_dense(b)
    dC = sa.BSM.uniform(c, n, n, bs, bs, 0.0)
    most_batches = 0
    for name, opts, tree in [("heuristic", sa.default_opts(scheduler=sa.SCHED_HEURISTIC), None),
                             ("learned", sa.default_opts(scheduler=sa.SCHED_MLP), None),
                             ("forced", None, oracle.forced_tree(nb, {2}))]:
        dC.fill(0.0)
        sa.matmult_strassen_sparse(dC, dA, dB, 1e-10, cutoff, opts=opts, tree=tree)
        st = sa.strassen_last_stats()
        print(name, cutoff, st, flush=True)
        assert np.array_equal(dC.to_dense(), ref), f"{name}: hybrid result differs from the exact product"
        if st["nroots"] > 0:
            assert st["batches"] == st["nroots"], (name, st)
        most_batches = max(most_batches, st["batches"])
    assert most_batches >= 3
