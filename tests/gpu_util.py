"""Helpers shared by the GPU parity tests: paired oracle / device matrices."""
import numpy as np

import oracle
import strassen_amd as sa

_CTX = None


def ctx():
    global _CTX
    if _CTX is None:
        _CTX = sa.Context(0)
    return _CTX


def pair_uniform(a, bs_r, bs_c, thresh):
    a = np.asfortranarray(a, dtype=np.float64)
    o = oracle.OBsm.uniform(a.shape[0], a.shape[1], bs_r, bs_c, thresh).from_dense(a)
    d = sa.BSM.uniform(ctx(), a.shape[0], a.shape[1], bs_r, bs_c, thresh).from_dense(a)
    return o, d


def pair_blocked(a, rsz, csz, thresh, rscr=None, cscr=None):
    a = np.asfortranarray(a, dtype=np.float64)
    o = oracle.OBsm(rsz, csz, thresh, rscr, cscr).from_dense(a)
    d = sa.BSM(ctx(), rsz, csz, thresh, rscr, cscr).from_dense(a)
    return o, d


def empty_pair_uniform(nr, nc, bs_r, bs_c, thresh=0.0):
    return (oracle.OBsm.uniform(nr, nc, bs_r, bs_c, thresh), sa.BSM.uniform(ctx(), nr, nc, bs_r, bs_c, thresh))


def empty_pair_blocked(rsz, csz, thresh=0.0):
    return oracle.OBsm(rsz, csz, thresh), sa.BSM(ctx(), rsz, csz, thresh)


def assert_same_meta(o, d, norms_exact=True):
    po, no = o.meta()
    pd, nd = d.meta()
    assert np.array_equal(po, pd), f"block presence differs: {np.argwhere(po != pd)[:5]}"
    if norms_exact:
        bad = np.argwhere(no != nd)
        assert bad.size == 0, f"cached norms differ at {bad[:5]}: {no[tuple(bad[0])]!r} vs {nd[tuple(bad[0])]!r}"
    else:
        np.testing.assert_allclose(nd, no, rtol=1e-13, atol=0)


def rel_fro(x, ref):
    n = np.linalg.norm(ref)
    return np.linalg.norm(x - ref) / (n if n > 0 else 1.0)


def rand_dense(rng, rsz, csz, fill, decay=0.0):
    rsz = np.asarray(rsz); csz = np.asarray(csz)
    mask = rng.random((len(rsz), len(csz))) < fill
    a = rng.standard_normal((int(rsz.sum()), int(csz.sum())))
    rb = np.repeat(np.arange(len(rsz)), rsz)
    cb = np.repeat(np.arange(len(csz)), csz)
    a[~mask[rb][:, cb]] = 0.0
    if decay:
        a *= np.exp(-decay * np.abs(rb[:, None] - cb[None, :]))
    return np.asfortranarray(a)
