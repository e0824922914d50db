"""Pins the CPU oracle (oracle/oracle.c) independently of its own text:
known-answer vectors from the reference's h2o fixtures, closed forms (dense
products, exact integer arithmetic), brute-force task lists on small inputs,
and -- where oracle/_ref was built from the reference sources -- the
reference's own templates on the same inputs."""
import numpy as np
import pytest

import oracle
from strassen_amd import synth
from conftest import rel_fro

needs_ref = pytest.mark.skipif(not oracle.ref_available(), reason="oracle/_ref (reference build) not present")


def rand_bsm_dense(rng, rsz, csz, fill, decay=0.0):
    rsz = np.asarray(rsz); csz = np.asarray(csz)
    mask = rng.random((len(rsz), len(csz))) < fill
    a = rng.standard_normal((rsz.sum(), csz.sum()))
    rb = np.repeat(np.arange(len(rsz)), rsz)
    cb = np.repeat(np.arange(len(csz)), csz)
    a[~mask[rb][:, cb]] = 0.0
    if decay:
        a *= np.exp(-decay * np.abs(rb[:, None] - cb[None, :]))
    return np.asfortranarray(a)


# ------------------------------------------------------------------ norms --
def test_block_norm_closed_form():
    rng = np.random.default_rng(0)
    for n in [1, 5, 63, 64, 65, 4096, 1600]:
        x = rng.standard_normal(n)
        assert abs(oracle.block_norm(x) - np.linalg.norm(x)) <= 4e-16 * np.linalg.norm(x) * np.sqrt(n)
    assert oracle.block_norm(np.zeros(17)) == 0.0
    assert oracle.block_norm(np.array([-3.0])) == 3.0
    assert oracle.block_norm(np.array([3.0, 4.0])) == 5.0
    # order: the 64 strided partial sums (exact fma via rationals), then the halving tree
    from fractions import Fraction
    x = rng.standard_normal(200)
    s = np.zeros(64)
    for lane in range(64):
        a = 0.0
        for e in range(lane, 200, 64):
            a = float(Fraction(x[e]) * Fraction(x[e]) + Fraction(a))
        s[lane] = a
    off = 32
    while off:
        s[:off] = s[:off] + s[off:2 * off]
        off //= 2
    assert oracle.block_norm(x) == np.sqrt(s[0])


@needs_ref
def test_block_norm_vs_reference_dnrm2():
    rng = np.random.default_rng(1)
    for n in [9, 64, 4096, 16384]:
        x = rng.standard_normal(n)
        r = oracle.ref().ref_norm(oracle._ptr(x), n)
        assert abs(oracle.block_norm(x) - r) <= 8e-16 * r


# ----------------------------------------------------------------- copy-in --
def test_copy_in_drop_rule_bruteforce():
    rng = np.random.default_rng(2)
    a = rand_bsm_dense(rng, [4] * 6 + [3], [4] * 5 + [2], 0.6)
    a[:4, :4] *= 1e-9  # a block just below / above threshold
    thresh = 1e-10 * 2
    m = oracle.OBsm.uniform(a.shape[0], a.shape[1], 4, 4, thresh).from_dense(a)
    pres, norms = m.meta()
    for rb in range(m.nrb):
        for cb in range(m.ncb):
            blk = a[rb * 4:(rb + 1) * 4, cb * 4:(cb + 1) * 4]
            nrm = oracle.block_norm(blk)
            assert pres[rb, cb] == (not nrm < thresh * 16)  # target sizes, also for ragged blocks
            assert norms[rb, cb] == (nrm if pres[rb, cb] else 0.0)
    d = m.to_dense()
    keep = np.repeat(np.repeat(pres, 4, 0), 4, 1)[: a.shape[0], : a.shape[1]]
    assert np.array_equal(d, np.where(keep, a, 0.0))


@needs_ref
def test_copy_in_vs_reference():
    rng = np.random.default_rng(3)
    a = rand_bsm_dense(rng, [8] * 9 + [5], [8] * 7 + [3], 0.5)
    n_r, n_c = a.shape
    m = oracle.OBsm.uniform(n_r, n_c, 8, 8, 1e-3).from_dense(a)
    pres, norms = m.meta()
    rp = np.zeros(m.nrb * m.ncb, np.uint8); rn = np.zeros(m.nrb * m.ncb)
    oracle.ref().ref_from_dense(oracle._ptr(a), n_r, n_c, 8, 8, 1e-3, oracle._ptr(rp), oracle._ptr(rn))
    assert np.array_equal(pres.ravel(order="F"), rp.astype(bool))
    np.testing.assert_allclose(norms.ravel(order="F"), rn, rtol=4e-16, atol=0)


# ------------------------------------------------------------------ matmult --
@pytest.mark.parametrize("tA,tB", [(0, 0), (1, 0), (0, 1), (1, 1)])
def test_matmult_dense_equivalence(tA, tB):
    rng = np.random.default_rng(10 + 2 * tA + tB)
    n_i, n_k, n_j, bs = 37, 29, 23, 6
    a = rng.standard_normal((n_k, n_i) if tA else (n_i, n_k))
    b = rng.standard_normal((n_j, n_k) if tB else (n_k, n_j))
    A = oracle.OBsm.uniform(*a.shape, bs, bs, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(*b.shape, bs, bs, 0.0).from_dense(b)
    Cm = oracle.OBsm.uniform(n_i, n_j, bs, bs, 0.0)
    tasks = oracle.matmult(Cm, A, tA, B, tB, 0.0, 1.0, 0.0)
    ref = (a.T if tA else a) @ (b.T if tB else b)
    assert rel_fro(Cm.to_dense(), ref) < 1e-15
    assert len(tasks) == A.nrb * A.ncb * (B.ncb if not tB else B.nrb)


def test_matmult_alpha_beta():
    rng = np.random.default_rng(20)
    a = rng.standard_normal((24, 18)); b = rng.standard_normal((18, 30)); c0 = rng.standard_normal((24, 30))
    A = oracle.OBsm.uniform(24, 18, 6, 6).from_dense(a)
    B = oracle.OBsm.uniform(18, 30, 6, 6).from_dense(b)
    for beta in [0.0, 1.0, -0.5]:
        Cm = oracle.OBsm.uniform(24, 30, 6, 6).from_dense(c0)
        oracle.matmult(Cm, A, 0, B, 0, 0.0, 2.5, beta)
        assert rel_fro(Cm.to_dense(), 2.5 * a @ b + beta * c0) < 1e-15
        pres, norms = Cm.meta()
        if beta == 0.0:
            assert np.all(norms == 0.0)  # fresh C blocks keep a zero cached norm


def brute_tasks(A_d, B_d, rsz, ksz, csz, thresh, a_thr, b_thr, scr=None):
    """Brute-force screened triples of the canonical order (jb, ib, kb)."""
    ro = np.concatenate([[0], np.cumsum(rsz)]); ko = np.concatenate([[0], np.cumsum(ksz)]); co = np.concatenate([[0], np.cumsum(csz)])
    def blk(M, r0, r1, c0, c1):
        return M[r0:r1, c0:c1]
    out = []
    for jb in range(len(csz)):
        for ib in range(len(rsz)):
            for kb in range(len(ksz)):
                a = blk(A_d, ro[ib], ro[ib + 1], ko[kb], ko[kb + 1])
                b = blk(B_d, ko[kb], ko[kb + 1], co[jb], co[jb + 1])
                na, nb = oracle.block_norm(a), oracle.block_norm(b)
                if na < a_thr * a.size or nb < b_thr * b.size:
                    continue
                s = (rsz[ib] * csz[jb] * ksz[kb]) if scr is None else scr
                if na * nb >= thresh * s:
                    out.append((ib, jb, kb))
    return np.array(out, dtype=np.int32).reshape(-1, 3)


def test_matmult_tasklist_bruteforce_uneven():
    rng = np.random.default_rng(30)
    w = synth.small_uneven(n=150, fill=0.4, seed=5)
    a = synth.dense_from_scales(1, w.rsz, w.ksz, w.a_scale)
    b = synth.dense_from_scales(2, w.ksz, w.csz, w.b_scale)
    thresh = 1e-3
    A = oracle.OBsm(w.rsz, w.ksz, 1e-12).from_dense(a)
    B = oracle.OBsm(w.ksz, w.csz, 1e-12).from_dense(b)
    Cm = oracle.OBsm(w.rsz, w.csz, 0.0)
    tasks = oracle.matmult(Cm, A, 0, B, 0, thresh)
    bt = brute_tasks(a, b, w.rsz, w.ksz, w.csz, thresh, 1e-12, 1e-12)
    assert len(bt) > 10 and np.array_equal(tasks, bt)
    # C equals the sum of the surviving products
    ro = np.concatenate([[0], np.cumsum(w.rsz)]); ko = np.concatenate([[0], np.cumsum(w.ksz)]); co = np.concatenate([[0], np.cumsum(w.csz)])
    ref = np.zeros((w.n, w.n))
    for ib, jb, kb in bt:
        ref[ro[ib]:ro[ib + 1], co[jb]:co[jb + 1]] += a[ro[ib]:ro[ib + 1], ko[kb]:ko[kb + 1]] @ b[ko[kb]:ko[kb + 1], co[jb]:co[jb + 1]]
    assert rel_fro(Cm.to_dense(), ref) < 1e-14


def test_matmult_uniform_threshold_uses_target_sizes():
    # ragged edge blocks are screened with the target block size (blocksparsematrix.hpp:349-350)
    rng = np.random.default_rng(31)
    a = rand_bsm_dense(rng, [5, 5, 2], [5, 5, 2], 1.0) * 0.1
    b = rand_bsm_dense(rng, [5, 5, 2], [5, 5, 2], 1.0) * 0.1
    A = oracle.OBsm.uniform(12, 12, 5, 5).from_dense(a)
    B = oracle.OBsm.uniform(12, 12, 5, 5).from_dense(b)
    Cm = oracle.OBsm.uniform(12, 12, 5, 5)
    thresh = 2e-4
    tasks = oracle.matmult(Cm, A, 0, B, 0, thresh)
    bt = brute_tasks(a, b, [5, 5, 2], [5, 5, 2], [5, 5, 2], thresh, 0.0, 0.0, scr=125)
    assert np.array_equal(tasks, bt)


@needs_ref
@pytest.mark.parametrize("tA,tB,beta", [(0, 0, 0.0), (1, 0, 0.0), (0, 1, 1.0), (1, 1, 0.5), (0, 0, 1.0)])
def test_matmult_vs_reference(tA, tB, beta):
    rng = np.random.default_rng(40 + tA + 2 * tB)
    bs = 8
    n_i, n_k, n_j = 70, 61, 53
    a = rand_bsm_dense(rng, synth.uniform_sizes(n_k if tA else n_i, bs), synth.uniform_sizes(n_i if tA else n_k, bs), 0.6, 0.3)
    b = rand_bsm_dense(rng, synth.uniform_sizes(n_j if tB else n_k, bs), synth.uniform_sizes(n_k if tB else n_j, bs), 0.6, 0.3)
    c0 = rand_bsm_dense(rng, synth.uniform_sizes(n_i, bs), synth.uniform_sizes(n_j, bs), 0.5)
    thresh = 3e-3
    A = oracle.OBsm.uniform(*a.shape, bs, bs, 1e-9).from_dense(a)
    B = oracle.OBsm.uniform(*b.shape, bs, bs, 1e-9).from_dense(b)
    Cm = oracle.OBsm.uniform(n_i, n_j, bs, bs, 1e-9).from_dense(c0)
    oracle.matmult(Cm, A, tA, B, tB, thresh, 1.5, beta)
    pres, _ = Cm.meta()
    cio = c0.copy(order="F")
    rp = np.zeros(pres.size, np.uint8)
    oracle.ref().ref_matmult(oracle._ptr(a), a.shape[0], a.shape[1], 1e-9, oracle._ptr(b), b.shape[0], b.shape[1], 1e-9,
                             oracle._ptr(cio), n_i, n_j, 1e-9, 0, bs, tA, tB, thresh, 1.5, beta,
                             oracle._ptr(rp), None, None)
    assert np.array_equal(pres.ravel(order="F"), rp.astype(bool))
    assert rel_fro(Cm.to_dense(), cio) < 1e-14


# ------------------------------------------------------------------ dry run --
def dryrun_inputs(rng, nb, bs, kind):
    n = nb * bs - (3 if kind == "ragged" else 0)
    sz = synth.uniform_sizes(n, bs)
    if kind == "band":
        s = synth.band_scales(len(sz), 3, 0.5)
    elif kind == "random":
        s = (rng.random((len(sz), len(sz))) < 0.5).astype(float)
    else:
        s = np.ones((len(sz), len(sz)))
    a = synth.dense_from_scales(1, sz, sz, s)
    b = synth.dense_from_scales(2, sz, sz, s)
    return n, a, b


@needs_ref
@pytest.mark.parametrize("nb,bs,kind", [(8, 4, "dense"), (16, 4, "dense"), (12, 4, "random"), (16, 8, "band"),
                                        (13, 4, "ragged"), (32, 2, "dense"), (20, 2, "random")])
def test_dryrun_vs_reference(nb, bs, kind):
    rng = np.random.default_rng(nb * 100 + bs)
    n, a, b = dryrun_inputs(rng, nb, bs, kind)
    A = oracle.OBsm.uniform(n, n, bs, bs, 1e-12).from_dense(a)
    B = oracle.OBsm.uniform(n, n, bs, bs, 1e-12).from_dense(b)
    thr = 1e-10 * bs ** 3
    lv, _ = oracle.strassen_dryrun(A, B, thr)
    flat = np.concatenate(lv) if lv else np.zeros(0, np.uint8)
    rb = np.zeros(max(flat.size, 1), np.uint8)
    ml = oracle.ref().ref_dryrun(oracle._ptr(a), oracle._ptr(b), n, bs, 1e-12, thr, oracle._ptr(rb), rb.size)
    assert ml == len(lv)
    assert np.array_equal(flat, rb[: flat.size])


def test_dryrun_cost_model_closed_form():
    # dense 16x16 grid of 64^3 blocks: level-l node flops are 8^(l+1) * 2*64^3 when all
    # children are sparse, Strassen once that reaches 7^(l+1)*2*64^3 + 15*fac*4^(l+1)*64^2
    bs, nb = 64, 16
    sz = synth.uniform_sizes(nb * bs, bs)
    A = oracle.OBsm(sz, sz); A.fill(1.0)
    B = oracle.OBsm(sz, sz); B.fill(1.0)
    lv, fl = oracle.strassen_dryrun(A, B, 0.0)
    fac = [10, 10, 10, 25]
    child = 2 * bs ** 3
    for level, bits in enumerate(lv):
        cost_s = 7 ** (level + 1) * 2 * bs ** 3 + int(15 * fac[level] * 4 ** (level + 1) * bs * bs)
        f = 8 * child
        expect = f >= cost_s
        assert np.all(bits == expect), (level, f, cost_s)
        child = cost_s if expect else f


# ----------------------------------------------------------------- Strassen --
@pytest.mark.parametrize("nb,bs,cutoff,kind,levels", [
    (8, 4, 0, "dense", {1}), (16, 4, 8, "dense", {2}), (16, 2, 0, "random", {0, 3}),
    (13, 4, 0, "ragged", set()), (32, 2, 4, "dense", {4}), (16, 4, 0, "random", {1}),
    (8, 32, 0, "dense", None)])
def test_strassen_integer_exact(nb, bs, cutoff, kind, levels):
    """Integer-valued inputs make every FP64 operation exact, so any Strassen
    schedule must reproduce the dense product bit for bit.  levels=None uses the
    reference cost model (which picks Strassen at level 2 for 8x8 blocks of 32)."""
    rng = np.random.default_rng(nb + bs + cutoff)
    n = nb * bs - (3 if kind == "ragged" else 0)
    sz = synth.uniform_sizes(n, bs)
    a = rng.integers(-4, 5, (n, n)).astype(float)
    b = rng.integers(-4, 5, (n, n)).astype(float)
    if kind == "random":
        m = rng.random((len(sz), len(sz))) < 0.5
        rbi = np.repeat(np.arange(len(sz)), sz)
        a[~m[rbi][:, rbi]] = 0.0
    a = np.asfortranarray(a); b = np.asfortranarray(b)
    A = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(b)
    Cm = oracle.OBsm.uniform(n, n, bs, bs, 0.0); Cm.fill(0.0)
    tree = None if levels is None else oracle.forced_tree(len(sz), levels)
    leaf, nroots = oracle.strassen_sparse(Cm, A, B, 0.0, cutoff, tree=tree)
    assert np.array_equal(Cm.to_dense(), a @ b)
    if levels is None or levels:
        assert nroots > 0


def _brute_plan(a, b, sz, thr, roots, cmask):
    """Product of integer-valued a, b by brute force over block triples: a triple
    inside a Strassen or dense root region is always counted, any other one
    only when it passes the screen; products into absent C blocks are dropped."""
    nb = len(sz)
    e = np.concatenate([[0], np.cumsum(sz)])
    blk = lambda x, i, j: x[e[i]:e[i + 1], e[j]:e[j + 1]]
    c = np.zeros_like(a)
    for i in range(nb):
        for j in range(nb):
            if not cmask[i, j]:
                continue
            for k in range(nb):
                ab, bb = blk(a, i, k), blk(b, k, j)
                covered = any(r0 <= i < r0 + w and c0 <= j < c0 + w and k0 <= k < k0 + w
                              for (r0, c0, k0, w) in roots)
                na, nbm = np.linalg.norm(ab), np.linalg.norm(bb)
                if covered or (na > 0 and nbm > 0 and na * nbm >= thr):
                    c[e[i]:e[i + 1], e[j]:e[j + 1]] += ab @ bb
    return c


def _plan_roots(tree, dense, nb):
    """Roots of a plan as the planner finds them: from the top level down, a
    node with its Strassen bit (else its dense bit) set is a root covering its
    region; otherwise its children are visited."""
    roots = []
    top = len(tree) - 1

    def visit(lev, si, sj, sk):
        w = 2 ** (lev + 1)
        ns = -(-nb // w)
        idx = si + sj * ns + sk * ns * ns
        if tree[lev][idx] or dense[lev][idx]:
            roots.append((si * w, sj * w, sk * w, w, "S" if tree[lev][idx] else "D"))
            return
        if lev == 0:
            return
        h = w // 2
        for dj in range(2):
            for di in range(2):
                for dk in range(2):
                    ci, cj, ck = 2 * si + di, 2 * sj + dj, 2 * sk + dk
                    if ci * h < nb and cj * h < nb and ck * h < nb:
                        visit(lev - 1, ci, cj, ck)

    nt = -(-nb // 2 ** (top + 1))
    for sk in range(nt):
        for sj in range(nt):
            for si in range(nt):
                visit(top, si, sj, sk)
    return roots


@pytest.mark.parametrize("seed", [1, 2, 3, 4, 5])
def test_strassen_plan_dense_roots_integer_exact(seed):
    """Random plans mixing Strassen and dense roots at every level (Strassen bit
    wins where both are set) on integer inputs with absent A, B and C blocks and
    a threshold that screens some leaf triples: bit-exact against a brute-force
    sum over block triples."""
    rng = np.random.default_rng(seed)
    nb, bs = 8, 3
    n = nb * bs
    sz = synth.uniform_sizes(n, bs)
    rbi = np.repeat(np.arange(nb), sz)
    a = rng.integers(-3, 4, (n, n)).astype(float)
    b = rng.integers(-3, 4, (n, n)).astype(float)
    a[~(rng.random((nb, nb)) < 0.7)[rbi][:, rbi]] = 0.0
    b[~(rng.random((nb, nb)) < 0.7)[rbi][:, rbi]] = 0.0
    a *= np.exp2(-np.abs(rbi[:, None] - rbi[None, :]))  # decaying blocks so the screen binds; still exact
    cmask = rng.random((nb, nb)) < 0.85
    a = np.asfortranarray(a); b = np.asfortranarray(b)
    A = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(b)
    c0 = np.asfortranarray(np.where(cmask[rbi][:, rbi], 1.0, 0.0))
    Cm = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(c0)
    pres, _ = Cm.meta()
    thresh = 2e-3
    sizes = oracle.tree_level_sizes(nb)
    tree = [(rng.random(s) < 0.15).astype(np.uint8) for s in sizes]
    dense = [(rng.random(s) < 0.25).astype(np.uint8) for s in sizes]
    _, nr, nd = oracle.strassen_sparse(Cm, A, B, thresh, 0, tree=tree, dense=dense, counts=True)
    roots = _plan_roots(tree, dense, nb)
    want = _brute_plan(a, b, sz, thresh * bs ** 3, [r[:4] for r in roots], pres.astype(bool)) + c0
    assert np.array_equal(Cm.to_dense(), want)
    assert nr == sum(r[4] == "S" for r in roots) and nd == sum(r[4] == "D" for r in roots)
    assert nr + nd > 0


@needs_ref
@pytest.mark.parametrize("nb,bs,cutoff,kind", [(16, 4, 8, "dense"), (16, 4, 0, "band"), (12, 4, 0, "random")])
def test_strassen_vs_reference(nb, bs, cutoff, kind):
    rng = np.random.default_rng(7 * nb + bs)
    n, a, b = dryrun_inputs(rng, nb, bs, kind)
    A = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(b)
    Cm = oracle.OBsm.uniform(n, n, bs, bs, 0.0); Cm.fill(0.0)
    oracle.strassen_sparse(Cm, A, B, 1e-10, cutoff)
    cio = np.zeros((n, n), order="F")
    oracle.ref().ref_strassen(oracle._ptr(a), oracle._ptr(b), oracle._ptr(cio), n, bs, 0.0, 1e-10, cutoff, None)
    assert rel_fro(Cm.to_dense(), cio) < 1e-14
    assert rel_fro(cio, a @ b) < 1e-12


# ------------------------------------------------------------- RI-K drivers --
def test_occ_ri_k_h2o_kat(h2o):
    K = oracle.occ_ri_k(24, 113, 5, 322, h2o["pqm12"], h2o["mo"], h2o["keys"], 3, 1e-20)
    assert abs(np.linalg.norm(K) - 10.415071322740426) < 1e-12
    Kd = oracle.occ_ri_k_dense(24, 113, 5, 322, h2o["pqm12"], h2o["mo"], h2o["keys"])
    assert rel_fro(K, Kd) < 1e-14
    # closed form K = sum_P U_P C C^T (U_P + U_P^T) C
    U = np.zeros((113, 24 * 24))
    U[:, h2o["keys"].astype(np.int64)] = h2o["pqm12"].T
    C = h2o["mo"]
    ref = np.zeros((24, 5))
    for p in range(113):
        Up = U[p].reshape(24, 24, order="F")
        ref += Up @ C @ (C.T @ (Up + Up.T) @ C)
    assert rel_fro(K, ref) < 1e-14


@needs_ref
def test_occ_ri_k_vs_reference(h2o):
    K = oracle.occ_ri_k(24, 113, 5, 322, h2o["pqm12"], h2o["mo"], h2o["keys"], 3, 1e-20)
    Kr = np.zeros((24, 5), order="F")
    oracle.ref().ref_occ_ri_k(24, 113, 5, 322, oracle._ptr(np.asfortranarray(h2o["pqm12"])),
                              oracle._ptr(np.asfortranarray(h2o["mo"])), oracle._ptr(h2o["keys"]), 3, 1e-20, oracle._ptr(Kr))
    assert rel_fro(K, Kr) < 1e-14


def test_3c_2c_transform_kat(h2o):
    """3c_ints_h2o x 2c_ints_inv05_h2o reproduces the shipped 3c_ints_PQm12_h2o."""
    A = oracle.OBsm.uniform(322, 113, 3, 3, 0.0).from_dense(h2o["ints3c"])
    B = oracle.OBsm.uniform(113, 113, 3, 3, 0.0).from_dense(h2o["inv05"])
    Cm = oracle.OBsm.uniform(322, 113, 3, 3, 0.0); Cm.fill(0.0)
    oracle.matmult(Cm, A, 0, B, 0, 1e-20, 1.0, 1.0)
    assert rel_fro(Cm.to_dense(), h2o["pqm12"]) < 1e-13


# --------------------------------------------------------- BSM helpers -----
def test_bsm_helpers_semantics():
    rng = np.random.default_rng(50)
    a = rand_bsm_dense(rng, [3] * 4, [3] * 4, 0.6)
    b = rand_bsm_dense(rng, [3] * 4, [3] * 4, 0.6)
    A = oracle.OBsm.uniform(12, 12, 3, 3).from_dense(a)
    B = oracle.OBsm.uniform(12, 12, 3, 3).from_dense(b)
    pa, na = A.meta(); pb, nb = B.meta()
    X = A.clone(); X.add(B, 1.0)
    assert np.array_equal(X.to_dense(), a + b)
    px, nx = X.meta()
    # += keeps the cached norm of an existing block, copies the rhs norm into an empty one
    assert np.array_equal(nx[pa], na[pa])
    assert np.array_equal(nx[~pa & pb], nb[~pa & pb])
    Y = A.clone(); Y.add(B, -1.0)  # -= copies (not negates) rhs into an empty block
    expect = np.where(np.repeat(np.repeat(pa, 3, 0), 3, 1), a - b, b)
    assert np.array_equal(Y.to_dense(), expect)
    T = A.clone(); T.add_transpose(B)
    assert np.array_equal(T.to_dense(), a + b.T)
    assert abs(A.dot(B) - np.sum(a * b)) < 1e-13
    assert abs(A.frobenius_norm() - np.linalg.norm(a)) < 1e-13
    S = A.clone(); S.scale(-2.0)
    _, ns = S.meta()
    assert np.array_equal(ns, na * 4.0)  # the reference multiplies the cached norm by s*s
    F = A.clone(); F.fill(0.5)
    _, nf = F.meta()
    assert np.allclose(nf, 1.5)


# ----------------------------------------------------- path selector network --
def test_mlp_oracle_gradient_matches_finite_differences():
    """The numpy network's analytic gradient (the device kernels' check) is
    pinned by central differences of the loss, masks included."""
    from oracle import mlp
    rng = np.random.default_rng(0)
    F, O, n = 10, 2, 37
    prm = mlp.init(F, O, seed=1) + 0.05 * rng.standard_normal(mlp.n_params(F, O))
    X = rng.random((n, F)); T = rng.standard_normal((n, O)); M = (rng.random((n, O)) < 0.8).astype(float)
    mu, sd = X.mean(0), X.std(0) + 0.1
    g = mlp.grad(prm, X, T, M, mu, sd)
    idx = rng.choice(prm.size, 60, replace=False)
    fd = mlp.fd_grad(prm, X, T, M, mu, sd, idx)
    assert np.max(np.abs(g[idx] - fd)) < 1e-8 * max(1.0, np.max(np.abs(fd)))
    # a dropped mask entry removes that target from loss and gradient
    M2 = M.copy(); M2[3, 1] = 0.0
    T2 = T.copy(); T2[3, 1] += 100.0
    assert np.allclose(mlp.grad(prm, X, T2, M2, mu, sd), mlp.grad(prm, X, T, M2, mu, sd))


def test_mlp_oracle_adam_decreases_loss():
    from oracle import mlp
    rng = np.random.default_rng(2)
    F, O, n = 10, 2, 64
    X = rng.random((n, F)); W = rng.standard_normal((F, O)); T = np.tanh(X @ W)
    M = np.ones((n, O)); mu, sd = X.mean(0), X.std(0) + 1e-3
    prm = mlp.init(F, O, seed=3)
    m1 = np.zeros_like(prm); m2 = np.zeros_like(prm)
    l0 = mlp.loss(prm, X, T, M, mu, sd)
    for t in range(1, 301):
        prm, m1, m2 = mlp.adam(prm, mlp.grad(prm, X, T, M, mu, sd), m1, m2, t, 1e-2)
    assert mlp.loss(prm, X, T, M, mu, sd) < 0.05 * l0


# ------------------------------------- recompress / thresholded dot, pinned --
def test_scale_recompress_designed_ties():
    """scale() multiplies cached norms by s*s and recompress() drops cached
    norms below thresh * (target block size) -- fixed by construction, with
    blocks exactly at the threshold and ragged blocks between the target-size
    and actual-size thresholds (tests/designed.py)."""
    import designed

    a, thr, s, pres, norms = designed.recompress_case()
    m = oracle.OBsm.uniform(designed.NR, designed.NC, designed.BS, designed.BS, thr).from_dense(a)
    m.scale(s)
    m.recompress()
    p, n = m.meta()
    assert np.array_equal(p, pres)
    assert np.array_equal(n[p], norms[p])
    assert np.array_equal(m.to_dense(), np.where(np.abs(a) * s * s >= 1.0, a * s, 0.0))


def test_dot_thresholded_designed_ties():
    import designed

    a, b, thr, expect, expect_all = designed.dot_case()
    A = oracle.OBsm.uniform(designed.NR, designed.NC, designed.BS, designed.BS, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(designed.NR, designed.NC, designed.BS, designed.BS, 0.0).from_dense(b)
    assert A.dot(B, thr) == expect
    assert A.dot(B, 0.0) == expect_all
    assert A.dot(B, 0.0) == float(np.sum(a * b))


@needs_ref
def test_scale_recompress_vs_reference():
    """The oracle's scale + recompress against the reference's templates, on
    the designed ties and on random ragged matrices at several thresholds."""
    import designed

    R = oracle.ref()
    cases = [designed.recompress_case()[:3]]
    rng = np.random.default_rng(61)
    for thr, s in [(1e-3, 0.7), (5e-2, 0.3), (0.2, -1.5)]:
        cases.append((rand_bsm_dense(rng, [5] * 7 + [3], [5] * 5 + [2], 0.7, decay=0.6), thr, s))
    for a, thr, s in cases:
        nr, nc = a.shape
        bs = 4 if (nr, nc) == (designed.NR, designed.NC) else 5
        m = oracle.OBsm.uniform(nr, nc, bs, bs, thr).from_dense(a)
        m.scale(s)
        m.recompress()
        p, n = m.meta()
        nb = m.nrb * m.ncb
        rp = np.zeros(nb, np.uint8); rn = np.zeros(nb); rd = np.zeros((nr, nc), order="F")
        R.ref_scale_recompress(oracle._ptr(np.asfortranarray(a)), nr, nc, bs, bs, thr, s, oracle._ptr(rp),
                               oracle._ptr(rn), oracle._ptr(rd))
        rp = rp.reshape(m.nrb, m.ncb, order="F").astype(bool)
        rn = rn.reshape(m.nrb, m.ncb, order="F")
        assert np.array_equal(p, rp)
        np.testing.assert_allclose(n[p], rn[rp], rtol=4e-15, atol=0)
        assert np.array_equal(m.to_dense(), rd)


@needs_ref
def test_dot_thresholded_vs_reference():
    import designed

    R = oracle.ref()
    a, b, thr, expect, _ = designed.dot_case()
    got = R.ref_dot(oracle._ptr(np.asfortranarray(a)), oracle._ptr(np.asfortranarray(b)), designed.NR, designed.NC,
                    designed.BS, designed.BS, 0.0, thr)
    assert got == expect
    rng = np.random.default_rng(62)
    for thr in [0.0, 1e-3, 0.05, 0.3]:
        a = rand_bsm_dense(rng, [6] * 6 + [4], [6] * 4 + [1], 0.8, decay=0.5)
        b = rand_bsm_dense(rng, [6] * 6 + [4], [6] * 4 + [1], 0.8, decay=0.5)
        A = oracle.OBsm.uniform(a.shape[0], a.shape[1], 6, 6, 0.0).from_dense(a)
        B = oracle.OBsm.uniform(a.shape[0], a.shape[1], 6, 6, 0.0).from_dense(b)
        want = R.ref_dot(oracle._ptr(a), oracle._ptr(b), a.shape[0], a.shape[1], 6, 6, 0.0, thr)
        got = A.dot(B, thr)
        assert abs(got - want) <= 1e-14 * max(1.0, abs(want)), (thr, got, want)


def _near_tie_setup():
    """The designed near-tie operands and a threshold at the median estimate
    (thresh * bs^3 reproduces it exactly: bs^3 is a power of two)."""
    import designed

    a, b, bs = designed.near_tie_case()
    n = a.shape[0]
    A = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(a)
    B = oracle.OBsm.uniform(n, n, bs, bs, 0.0).from_dense(b)
    _, na = A.meta(); _, nbm = B.meta()
    est = (na[:, :, None] * nbm[None, :, :]).ravel()
    thresh = float(np.median(np.sort(est)[: est.size // 2 * 2 - 1])) / float(bs ** 3)
    return a, b, bs, thresh, A, B, na, nbm


def test_near_tie_screening_follows_canonical_norms():
    """At near-ties the oracle keeps exactly the triples whose canonical-norm
    product reaches the threshold (one FP64 product and one comparison, from
    the oracle's cached norms), and both outcomes occur."""
    a, b, bs, thresh, A, B, na, nbm = _near_tie_setup()
    nb = A.nrb
    Cm = oracle.OBsm.uniform(a.shape[0], a.shape[0], bs, bs, 0.0)
    t = oracle.matmult(Cm, A, 0, B, 0, thresh)
    thr = thresh * float(bs * bs * bs)
    keep = [(i, j, k) for j in range(nb) for i in range(nb) for k in range(nb) if na[i, k] * nbm[k, j] >= thr]
    assert np.array_equal(t, np.array(keep, dtype=np.int32).reshape(-1, 3))
    assert 0 < len(t) < nb ** 3
    est = na[:, :, None] * nbm[None, :, :]
    assert np.abs(est / thr - 1).max() < 1e-14


@needs_ref
def test_near_tie_screening_vs_reference():
    """The reference's screened set at near-ties: its block norms come from
    MKL dnrm2, whose summation order differs from the canonical one, so C
    blocks may differ -- each such block holds a triple within 1e-14 of the
    threshold (all triples of this case are).  Away from ties the task lists
    match (test_matmult_vs_reference)."""
    a, b, bs, thresh, A, B, na, nbm = _near_tie_setup()
    n = a.shape[0]
    nb = n // bs
    R = oracle.ref()
    c = np.zeros((n, n), order="F")
    R.ref_matmult(oracle._ptr(a), n, n, 0.0, oracle._ptr(b), n, n, 0.0, oracle._ptr(c), n, n, 0.0, 1, bs, 0, 0,
                  thresh, 1.0, 0.0, None, None, None)
    Cm = oracle.OBsm.uniform(n, n, bs, bs, 0.0)
    t = oracle.matmult(Cm, A, 0, B, 0, thresh)
    co = Cm.to_dense()
    differ = 0
    for j in range(nb):
        for i in range(nb):
            blk = (slice(i * bs, (i + 1) * bs), slice(j * bs, (j + 1) * bs))
            if np.abs(co[blk] - c[blk]).max() > 1e-12 * max(1.0, np.abs(c[blk]).max()):
                differ += 1
    print(f"near-tie case: {len(t)} of {nb ** 3} triples kept by the canonical norms; "
          f"{differ} of {nb * nb} C blocks differ from the reference (MKL dnrm2 norms)")
    # MKL's norms of the same multiset of values: equal to the canonical ones
    # to within 2 ulp, so the reference keeps every triple whose estimate is
    # clearly above, drops every one clearly below
    rn = np.zeros(nb * nb); rp = np.zeros(nb * nb, np.uint8)
    R.ref_from_dense(oracle._ptr(a), n, n, bs, bs, 0.0, oracle._ptr(rp), oracle._ptr(rn))
    np.testing.assert_allclose(rn.reshape(nb, nb, order="F"), na, rtol=4.5e-16, atol=0)
