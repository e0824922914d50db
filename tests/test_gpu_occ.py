"""Device occupied RI-K (sa_occ_ri_k): the compressed 3-centre integrals are
decompressed on the device straight into block form and the whole P loop of
occ_ri_k_step.cpp:83-131 runs there.  Checked against the oracle's restatement
of that loop (itself pinned by the h2o KAT, the closed form
K = sum_P U_P C C^T (U_P + U_P^T) C and the built reference)."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from gpu_util import ctx, rel_fro

pytestmark = pytest.mark.gpu

H2O_NORM = 10.415071322740426  # ||K||_F of test_occ_ri_k_h2o.sh (24 113 5 322 PQm12 ... 3 1e-20)


def team_oracle(n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams):
    """The reference's team: P split as OpenMP's static schedule, one oracle
    run (one thread of occ_ri_k_step.cpp:90-120) per member, K_sub added in
    member order."""
    T = max(1, min(teams, max(n_aux, 1)))
    q, r = divmod(n_aux, T)
    K = np.zeros((n_bf, n_occ), order="F")
    for t in range(T):
        p0, n = t * q + min(t, r), q + (1 if t < r else 0)
        K = K + oracle.occ_ri_k(n_bf, n, n_occ, n_vec2, np.asfortranarray(ints[:, p0:p0 + n]), mo, keys, bs, thresh)
    return K


@pytest.mark.parametrize("teams", [1, 8])
def test_h2o_kat(h2o, teams):
    K = sa.occ_ri_k(ctx(), 24, 113, 5, 322, h2o["pqm12"], h2o["mo"], h2o["keys"], 3, 1e-20, teams=teams)
    Ko = team_oracle(24, 113, 5, 322, h2o["pqm12"], h2o["mo"], h2o["keys"], 3, 1e-20, teams)
    assert rel_fro(K, Ko) < 1e-14
    assert abs(np.linalg.norm(K) - H2O_NORM) < 1e-12 * H2O_NORM


def synthetic(seed, n_bf, n_aux, n_occ, frac=0.4, dup=0):
    """Pairs (r <= c) of a random subset of the upper triangle with decaying
    values, keys r + n_bf c as in the v2m_key format; dup extra keys repeat
    earlier ones (a later value overwrites)."""
    rng = np.random.default_rng(seed)
    r, c = np.triu_indices(n_bf)
    keep = rng.random(r.size) < frac
    r, c = r[keep], c[keep]
    keys = (r + n_bf * c).astype(np.uint64)
    if dup:
        keys = np.concatenate([keys, rng.choice(keys, dup)])
    n_vec2 = keys.size
    dist = np.abs(r - c).astype(float)
    if dup:
        dist = np.concatenate([dist, rng.random(dup) * n_bf])
    ints = rng.standard_normal((n_vec2, n_aux)) * np.exp(-0.15 * dist)[:, None]
    mo = rng.standard_normal((n_bf, n_occ))
    return n_vec2, np.asfortranarray(ints), np.asfortranarray(mo), keys


@pytest.mark.parametrize("n_bf,n_aux,n_occ,bs,thresh,dup,teams",
                         [(70, 40, 12, 8, 1e-10, 0, 1), (70, 40, 12, 8, 1e-10, 0, 8), (45, 17, 9, 5, 1e-6, 0, 3),
                          (33, 12, 6, 4, 1e-10, 25, 5), (33, 3, 6, 4, 1e-10, 0, 8)])
def test_synthetic_vs_oracle(n_bf, n_aux, n_occ, bs, thresh, dup, teams):
    """Ragged last blocks, screening that drops products, repeated keys, team
    sizes that do and do not divide n_aux (and exceed it)."""
    n_vec2, ints, mo, keys = synthetic(n_bf + n_aux, n_bf, n_aux, n_occ, dup=dup)
    K = sa.occ_ri_k(ctx(), n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams=teams)
    Ko = team_oracle(n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams)
    assert np.linalg.norm(Ko) > 0
    assert rel_fro(K, Ko) < 1e-12


@pytest.mark.parametrize("batch", ["1", "3", "7", "0"])
@pytest.mark.parametrize("n_bf,n_aux,n_occ,bs,thresh,teams", [(70, 40, 12, 8, 1e-10, 1), (45, 17, 9, 5, 1e-6, 3)])
def test_aux_batches(monkeypatch, batch, n_bf, n_aux, n_occ, bs, thresh, teams):
    """Aux functions run in batches of 1, 3, 7 or as many as fit (ragged bf and
    occ blocks, a partial last batch): the same screened products as one P at
    a time, K within the FP64 bar of the oracle's team."""
    monkeypatch.setenv("SA_OCC_BATCH", batch)
    n_vec2, ints, mo, keys = synthetic(n_bf + n_aux + 7, n_bf, n_aux, n_occ)
    K = sa.occ_ri_k(ctx(), n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams=teams)
    Ko = team_oracle(n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams)
    assert np.linalg.norm(Ko) > 0
    assert rel_fro(K, Ko) < 1e-12


def test_split_precision(h2o):
    """The three products in split precision (5 INT8 slices): bar 1e-10."""
    n_vec2, ints, mo, keys = synthetic(7, 64, 30, 10)
    o = sa.default_opts(mode=sa.MODE_OZAKI, slices=5, slice_type=sa.SLICE_INT8)
    K = sa.occ_ri_k(ctx(), 64, 30, 10, n_vec2, ints, mo, keys, 8, 1e-10, teams=4, opts=o)
    Ko = team_oracle(64, 30, 10, n_vec2, ints, mo, keys, 8, 1e-10, 4)
    assert rel_fro(K, Ko) < 1e-10


def test_key_out_of_range_is_an_error(h2o):
    keys = h2o["keys"].copy()
    keys[7] = 24 * 24
    with pytest.raises(RuntimeError, match="beyond"):
        sa.occ_ri_k(ctx(), 24, 113, 5, 322, h2o["pqm12"], h2o["mo"], keys, 3, 1e-20)


def test_no_pairs():
    K = sa.occ_ri_k(ctx(), 10, 4, 3, 0, np.zeros((0, 4)), np.ones((10, 3)), np.zeros(0, np.uint64), 4, 1e-10)
    assert not K.any()
