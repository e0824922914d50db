"""GPU RI-K step: the Newton-Schulz inverse square root of the 2-centre
metric against LAPACK eigh and the reference's own h2o metric, and the full
step (transform B' = X B, K = sum_{P,i} B'^T B') against the oracle run on the
same X: task lists bit-exact, K within 1e-12 (FP64) / 1e-10 (split precision)."""
import numpy as np
import pytest

import oracle
import strassen_amd as sa
from strassen_amd import synth
from conftest import load_golden
from gpu_util import ctx, rel_fro

pytestmark = pytest.mark.gpu


def inv_sqrt_ref(M):
    w, V = np.linalg.eigh(M)
    return (V / np.sqrt(w)) @ V.T


def test_inv_sqrt_vs_eigh():
    w = synth.rik_small()
    M = w.metric()
    X, iters, res = sa.inv_sqrt(ctx(), M, tol=1e-14)
    ev = np.linalg.eigvalsh(M)
    kappa = ev[-1] / ev[0]
    assert rel_fro(X, inv_sqrt_ref(M)) < 1e-14 * kappa * 10
    assert res < 1e-12
    assert np.abs(X - X.T).max() < 1e-12 * np.abs(X).max()
    assert 5 < iters < 60


def test_inv_sqrt_h2o_round_trip():
    """M := (inv05)^{-2} from the shipped 2c_ints_inv05_h2o.data; M^{-1/2} must
    reproduce the file (cond(M) ~ 2.5e7, so agreement is ~ cond * eps)."""
    X0 = load_golden("2c_ints_inv05_h2o.data", (113, 113))
    Xi = np.linalg.inv(X0)
    M = np.asfortranarray(0.5 * ((Xi @ Xi) + (Xi @ Xi).T))
    X, iters, res = sa.inv_sqrt(ctx(), M, max_iter=200, tol=1e-13)
    assert rel_fro(X, X0) < 5e-8, rel_fro(X, X0)


def build_pair(w, X):
    """B (generated on the device and on the host from the same definition) and X."""
    bcs = w.b_col_sizes
    dB = sa.BSM(ctx(), w.p_sizes, bcs, w.thresh).generate(7, w.b_scale)
    b = synth.dense_from_scales(7, w.p_sizes, bcs, w.b_scale)
    oB = oracle.OBsm(w.p_sizes, bcs, w.thresh).from_dense(b)
    oX = oracle.OBsm(w.p_sizes, w.p_sizes, w.thresh).from_dense(X)
    return dB, oB, oX


def oracle_rik(w, oX, oB):
    bcs = w.b_col_sizes
    oBp = oracle.OBsm(w.p_sizes, bcs, 0.0)
    t1 = oracle.matmult(oBp, oX, 0, oB, 0, w.thresh)
    oBp.calc_norms()
    n_ic = len(w.i_chunks)
    vr = np.array([p * i for i in w.i_chunks for p in w.p_sizes], dtype=np.int32)
    view = oBp.view_reshape(vr, w.mu_sizes)
    oK = oracle.OBsm(w.mu_sizes, w.mu_sizes, 0.0)
    t2 = oracle.matmult(oK, view, 1, view, 0, w.thresh)
    return oK.to_dense(), oBp.to_dense(), t1, t2


@pytest.mark.parametrize("mode", ["fp64", "ozaki"])
def test_ri_k_step_vs_oracle(mode):
    """Split precision: 5 FP16 slices for the transform (X = M^{-1/2} rows mix
    large and small entries, so its cancellation amplifies slice truncation),
    4 for the Gram-type K product."""
    w = synth.rik_small()
    dX = sa.BSM(ctx(), w.p_sizes, w.p_sizes, w.thresh)
    X, iters, res = sa.inv_sqrt(ctx(), w.metric(), X_bsm=dX, tol=1e-14)
    dB, oB, oX = build_pair(w, X)
    bcs = w.b_col_sizes
    dBp = sa.BSM(ctx(), w.p_sizes, bcs, 0.0)
    dK = sa.BSM(ctx(), w.mu_sizes, w.mu_sizes, 0.0)
    ot = sa.default_opts(mode=sa.MODE_OZAKI if mode == "ozaki" else sa.MODE_FP64, slices=5)
    ok = sa.default_opts(mode=sa.MODE_OZAKI if mode == "ozaki" else sa.MODE_FP64, slices=4)
    sa.ri_k_step(dX, dB, dBp, dK, w.mu_sizes, w.i_chunks, w.thresh, ot, ok)
    t_dev = ctx().tasklist()
    K_ref, Bp_ref, t1, t2 = oracle_rik(w, oX, oB)
    assert np.array_equal(t2, t_dev), "K task lists differ"
    tol = 1e-12 if mode == "fp64" else 1e-10
    assert rel_fro(dBp.to_dense(), Bp_ref) < tol
    K = dK.to_dense()
    assert rel_fro(K, K_ref) < tol
    # the blocks below the diagonal are the transposes of those above
    assert np.array_equal(K, K.T)
    # K is symmetric and equals the dense contraction of the transformed tensor
    Bp = Bp_ref
    n_ic = len(w.i_chunks)
    # dense K from B'(P, (mu, i)) with i fastest inside each column block
    cols = []
    off = 0
    for mu in w.mu_sizes:
        for ic in w.i_chunks:
            cols.append((mu, ic, off)); off += mu * ic
    Kd = np.zeros((w.n_bf, w.n_bf))
    mu_off = np.concatenate([[0], np.cumsum(w.mu_sizes)])
    for a, (mua, ica, oa) in enumerate(cols):
        for b_, (mub, icb, ob) in enumerate(cols):
            if (a % n_ic) != (b_ % n_ic):
                continue
            ma, mb = a // n_ic, b_ // n_ic
            Ba = Bp[:, oa:oa + mua * ica].reshape(-1, ica, mua, order="F")  # (P, i, mu)
            Bb = Bp[:, ob:ob + mub * icb].reshape(-1, icb, mub, order="F")
            Kd[mu_off[ma]:mu_off[ma + 1], mu_off[mb]:mu_off[mb + 1]] += np.einsum("pim,pin->mn", Ba, Bb)
    assert rel_fro(K, Kd) < (1e-11 if mode == "fp64" else 1e-10)


@pytest.mark.parametrize("kind,s_t,s_k,shared_t,shared_k,win_mb", [
    ("int8", 5, 5, 0, 0, 1),   # the benchmarked form, K contraction in many k-windows
    ("int8", 5, 5, 0, 0, 0),   # the same in one launch
    ("int8", 5, 5, 1, 0, 0),   # shared exponents in the transform
    ("int8", 5, 5, 0, 1, 1),   # shared exponents on the K contraction's one shared slicing
    ("int8", 5, 5, 1, 1, 1),   # shared exponents in both stages (the fastest benchmarked form)
    ("bf16", 5, 5, 0, 0, 1),
    ("fp16", 4, 4, 0, 0, 1),
])
def test_ri_k_step_split_forms_vs_oracle(kind, s_t, s_k, shared_t, shared_k, win_mb, monkeypatch):
    """The split-precision forms of the RI-K step against the FP64 oracle on
    the same X and B: K task list bit-exact, B' and K within 1e-10.  The K
    contraction B'^T B' reads one slicing of B' in both roles (a reshape view
    used transposed and plain); SA_KWINDOW_MB=1 splits it into k-windows."""
    monkeypatch.setenv("SA_KWINDOW_MB", str(win_mb))
    st = {"int8": sa.SLICE_INT8, "bf16": sa.SLICE_BF16, "fp16": sa.SLICE_FP16}[kind]
    w = synth.rik_small()
    dX = sa.BSM(ctx(), w.p_sizes, w.p_sizes, w.thresh)
    X, _, _ = sa.inv_sqrt(ctx(), w.metric(), X_bsm=dX, tol=1e-14)
    dB, oB, oX = build_pair(w, X)
    dBp = sa.BSM(ctx(), w.p_sizes, w.b_col_sizes, 0.0)
    dK = sa.BSM(ctx(), w.mu_sizes, w.mu_sizes, 0.0)
    ot = sa.default_opts(mode=sa.MODE_OZAKI, slices=s_t, slice_type=st, shared_exp=shared_t)
    ok = sa.default_opts(mode=sa.MODE_OZAKI, slices=s_k, slice_type=st, shared_exp=shared_k)
    sa.ri_k_step(dX, dB, dBp, dK, w.mu_sizes, w.i_chunks, w.thresh, ot, ok)
    nwin = ctx().kwindows()
    t_dev = ctx().tasklist()
    K_ref, Bp_ref, _, t2 = oracle_rik(w, oX, oB)
    assert np.array_equal(t2, t_dev), "K task lists differ"
    if win_mb:
        assert nwin > 4, f"expected the K contraction in k-windows, got {nwin}"
    else:
        assert nwin == 1
    assert rel_fro(dBp.to_dense(), Bp_ref) < 1e-10
    K = dK.to_dense()
    err = rel_fro(K, K_ref)
    assert err < 1e-10, err
    assert np.array_equal(K, K.T)
    # per block row of K, as the split-precision error is reported
    mo = np.concatenate([[0], np.cumsum(w.mu_sizes)])
    for r in range(len(w.mu_sizes)):
        ref_r = K_ref[mo[r]:mo[r + 1]]
        if np.linalg.norm(ref_r) > 0:
            assert rel_fro(K[mo[r]:mo[r + 1]], ref_r) < 1e-9


@pytest.mark.parametrize("kind,s_k,shared_k", [("int8", 5, 1), ("int8", 5, 0), ("fp16", 4, 0), ("bf16", 5, 1)])
def test_fused_norms_and_exponents_bit_exact(kind, s_k, shared_k, monkeypatch):
    """The split K contraction takes B''s block norms and its slicing
    exponents in one pass over B' (k_norms_view_exps).  With the pass split
    (SA_RIK_FUSED_NORMS=0: the canonical block norms, then the slicer's own
    max pass) every result must be bit-identical: task list, B', K."""
    st = {"int8": sa.SLICE_INT8, "bf16": sa.SLICE_BF16, "fp16": sa.SLICE_FP16}[kind]
    w = synth.rik_small()
    dX = sa.BSM(ctx(), w.p_sizes, w.p_sizes, w.thresh)
    X, _, _ = sa.inv_sqrt(ctx(), w.metric(), X_bsm=dX, tol=1e-14)
    dB, _, _ = build_pair(w, X)
    ot = sa.default_opts(mode=sa.MODE_OZAKI, slices=5, slice_type=sa.SLICE_INT8)
    ok = sa.default_opts(mode=sa.MODE_OZAKI, slices=s_k, slice_type=st, shared_exp=shared_k)
    out = []
    for fused in ("1", "0"):
        monkeypatch.setenv("SA_RIK_FUSED_NORMS", fused)
        dBp = sa.BSM(ctx(), w.p_sizes, w.b_col_sizes, 0.0)
        dK = sa.BSM(ctx(), w.mu_sizes, w.mu_sizes, 0.0)
        sa.ri_k_step(dX, dB, dBp, dK, w.mu_sizes, w.i_chunks, w.thresh, ot, ok)
        out.append((ctx().tasklist().copy(), dBp.to_dense(), dK.to_dense(), dBp.meta()[1]))
    (t1, b1, k1, n1), (t2, b2, k2, n2) = out
    assert np.array_equal(t1, t2)
    assert np.array_equal(b1, b2)
    assert np.array_equal(n1, n2), "block norms differ between the fused and the separate pass"
    assert np.array_equal(k1, k2), "K differs between fused and separate exponent passes"
