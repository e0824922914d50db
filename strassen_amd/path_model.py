"""Training of the learned path selector on timings the method generates.

Each sample is a homogeneous problem: a square uniform block-sparse product of
m 2^(l+1) x m 2^(l+1) blocks (levels 1..5, block sizes 64 and 128, random block
fill with norm decay, seeded) whose level-l recursion nodes all take one path,
run through the hybrid multiply with a caller plan:

  screened-sparse  no root (the screened grouped GEMM of all triples),
  Strassen         every level-l node a Strassen-Winograd root (children
                   recurse down to the cutoff),
  dense            every level-l node a dense root (packed regions, one grouped
                   FP64 MFMA GEMM launch per chunk of roots);

so each path's per-node cost is amortised over m^3 nodes as it is when a plan
runs.  The median device time of each (after a warm-up) gives the targets
log2(T_strassen / T_sparse) and log2(T_dense / T_sparse); the inputs are the
mean features of the level-l nodes (sa_strassen_features).  The network is
trained on the device by the library's own backward and Adam kernels; the model
file is written in the text format sa_mlp_load reads.

    python -m strassen_amd.path_model --out strassen_amd/data/path_mlp.txt
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

import strassen_amd as sa

CUTOFF = {64: 512, 128: 2048}  # min_size_for_strassen per block size (Strassen depth <= 4 at n = 32768, bs 128)


def _scales(rng, nb, p, decay):
    idx = np.arange(nb)
    d = np.abs(idx[:, None] - idx[None, :]).astype(float)
    return np.where(rng.random((nb, nb)) < p, np.exp(-decay * d), 0.0)


def level_problem(ctx, rng, level, bs, thresh, max_n=8192):
    w = 2 ** (level + 1)
    target = float(rng.uniform(2048, max_n))
    m = max(1, int(target // (w * bs)))
    nb = m * w
    n = nb * bs
    p = float(rng.uniform(0.08, 1.0)) if rng.random() < 0.75 else 1.0
    decay = float(rng.choice([0.0, 0.02, 0.1, 0.3]))
    seed = int(rng.integers(1, 2 ** 31))
    A = sa.BSM.uniform(ctx, n, n, bs, bs, thresh).generate(seed, _scales(rng, nb, p, decay))
    B = sa.BSM.uniform(ctx, n, n, bs, bs, thresh).generate(seed + 1, _scales(rng, nb, p, decay))
    Cm = sa.BSM.uniform(ctx, n, n, bs, bs, 0.0)
    Cm.fill(0.0)
    return A, B, Cm, {"level": level, "bs": bs, "m": m, "n": n, "p": p, "decay": decay, "seed": seed}


def _level_plan(nb, level, on):
    bits = [np.zeros(s, np.uint8) for s in sa.tree_level_sizes(nb)]
    if on:
        bits[level][:] = 1
    return bits


def time_path(ctx, Cm, A, B, thresh, level, path, reps=2):
    nb = A.nrb
    tree = _level_plan(nb, level, path == "strassen")
    dense = _level_plan(nb, level, path == "dense")
    cutoff = CUTOFF.get(A.nrow // A.nrb, 512)
    run = lambda: sa.matmult_strassen_sparse(Cm, A, B, thresh, cutoff, tree=tree, dense=dense)
    run()
    ts = []
    for _ in range(reps):
        ctx.synchronize()
        ctx.mark(0)
        run()
        ctx.mark(1)
        ctx.synchronize()
        ts.append(ctx.elapsed(0, 1))
    return float(np.median(ts))


def generate(ctx, n_problems, seed=1, thresh=1e-10, log=print):
    rng = np.random.default_rng(seed)
    X, T, meta = [], [], []
    t0 = time.time()
    for i in range(n_problems):
        level = int(rng.integers(1, 6))
        bs = int(rng.choice([64, 128]))
        A, B, Cm, m = level_problem(ctx, rng, level, bs, thresh)
        feats = sa.strassen_features(Cm, A, B, thresh)
        sizes = sa.tree_level_sizes(A.nrb)
        off = int(np.sum(sizes[:level]))
        x = feats[off:off + sizes[level]].mean(0)
        tm = {p: time_path(ctx, Cm, A, B, thresh, level, p) for p in ("sparse", "strassen", "dense")}
        X.append(x)
        T.append([np.log2(tm["strassen"] / tm["sparse"]), np.log2(tm["dense"] / tm["sparse"])])
        m.update({k: round(v, 4) for k, v in tm.items()})
        meta.append(m)
        if (i + 1) % 10 == 0:
            log(f"[path_model] {i + 1}/{n_problems} level problems timed ({time.time() - t0:.0f} s)")
        del A, B, Cm
    return np.array(X), np.array(T), meta


def glorot(F, O, seed):
    rng = np.random.default_rng(seed)
    parts = []
    for shp in [(32, F), (32,), (32, 32), (32,), (O, 32), (O,)]:
        if len(shp) == 2:
            lim = np.sqrt(6.0 / (shp[0] + shp[1]))
            parts.append(rng.uniform(-lim, lim, shp).ravel())
        else:
            parts.append(np.zeros(shp))
    return np.concatenate(parts)


def fit(ctx, X, T, epochs=4000, lr=3e-3, seed=0, holdout=0.2, chunk=200, clip=3.0):
    """Train on the device with early stopping on a holdout set; targets are
    clipped to +-clip (a path 8x slower than screened is simply bad).  Returns
    fit statistics on the training and holdout sets (rmse of the log2 ratios,
    how often the predicted best of {sparse, Strassen, dense} is the measured
    best, and the mean time of the predicted best relative to the true best)."""
    F, O, _ = sa.mlp_info()
    T = np.clip(T, -clip, clip)
    rng = np.random.default_rng(seed)
    perm = rng.permutation(len(X))
    nh = int(round(holdout * len(X)))
    te, tr = perm[:nh], perm[nh:]
    mu = X[tr].mean(0)
    sd = X[tr].std(0)
    sd[sd < 1e-6] = 1.0
    sa.mlp_set(ctx, glorot(F, O, seed), mu, sd)

    def stats(idx):
        if len(idx) == 0:
            return {}
        Y = sa.mlp_forward(ctx, X[idx])
        full = np.c_[np.zeros(len(idx)), T[idx]]
        best_m = np.argmin(full, 1)
        best_p = np.argmin(np.c_[np.zeros(len(idx)), Y], 1)
        rel = np.exp2(full[np.arange(len(idx)), best_p] - full[np.arange(len(idx)), best_m])
        return {"rmse_log2": float(np.sqrt(np.mean((Y - T[idx]) ** 2))), "best_path_hit": float(np.mean(best_m == best_p)),
                "mean_time_vs_best": float(np.mean(rel)), "n": int(len(idx))}

    best, best_prm, done, loss = np.inf, None, 0, None
    while done < epochs:
        loss = sa.mlp_train(ctx, X[tr], T[tr], epochs=chunk, lr=lr)
        done += chunk
        h = stats(te).get("rmse_log2", 0.0) if len(te) else loss
        if h < best:
            best, best_prm, best_at = h, sa.mlp_get(ctx), done
    sa.mlp_set(ctx, *best_prm)
    return {"loss": loss, "epochs_kept": best_at, "train": stats(tr), "holdout": stats(te)}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--problems", type=int, default=240)
    ap.add_argument("--epochs", type=int, default=4000)
    ap.add_argument("--lr", type=float, default=3e-3)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", default=sa.DEFAULT_PATH_MODEL)
    ap.add_argument("--samples", default=None, help="also write the timed samples (JSON)")
    args = ap.parse_args(argv)
    ctx = sa.Context(0)
    log = lambda s: print(s, file=sys.stderr, flush=True)
    X, T, meta = generate(ctx, args.problems, args.seed, log=log)
    st = fit(ctx, X, T, args.epochs, args.lr, args.seed)
    sa.mlp_save(ctx, args.out)
    if args.samples:
        with open(args.samples, "w") as fh:
            json.dump({"X": X.tolist(), "T": T.tolist(), "meta": meta, "fit": st}, fh)
    print(json.dumps({"model": args.out, "samples": len(X), **st}))


if __name__ == "__main__":
    main()
