"""Benchmark of the block-sparse multiply hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]

Default workload (the headline metric): the RI-K step (n_bf=4000, n_aux=12000,
n_occ=200, 300 random atoms).  A step is B' = X B followed by
K = sum_{P,i} B'^T B' -- screening, task lists, block allocation and the
grouped matrix-core GEMMs -- on inputs resident in HBM.  It is timed in split
precision (INT8- or FP16-slice Ozaki products, FP64 screening / norms /
accumulation) and in plain FP64; value = dense-equivalent TFLOP/s of the
fastest split-precision form within 1e-10 of FP64, with both step times, their
ratio and the relative error of K against the FP64 result in the same JSON
line.  The metric's inverse square root is set-up, timed separately.

Other workloads (--workload NAME): band (n 2048, 64-blocks, band |I-J| <= 4),
atoms (n 12000, uneven 5-40 atom blocks, 10 % fill), band_split (the band
inputs in split precision): one screened block-sparse multiply C = A B per
step; hybrid (n 32768, 128-blocks, band 95): screened vs cost-model Strassen
tree vs learned plan; occ: the occupied RI-K loop.

Timing: HIP events on the library's stream around the timed steps; the
dominant kernel's average launch time comes from the library's per-launch
HIP events (same stream).  The CPU baseline is the oracle (single-threaded C
restatement) on a bounded sample, scaled by algorithmic flops; where the
reference's own build (oracle/_ref: its templates, MKL, OpenMP) is present,
cpu_reference times it on the host's cores as well.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import strassen_amd as sa  # noqa: E402
from strassen_amd import synth  # noqa: E402

PEAK_FP64 = 78.6     # TFLOP/s, MI355X dense FP64 (matrix)
PEAK_FP16 = 2500.0   # TFLOP/s, MI355X dense FP16/BF16 matrix
PEAK_INT8 = 5000.0   # TOP/s, MI355X dense int8 matrix
METRIC = "dense-equiv TFLOP/s & RI-K step time: mixed vs FP64 GPU at <=1e-10 rel err"


def log(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def rel(x, ref):
    n = np.linalg.norm(ref)
    return float(np.linalg.norm(x - ref) / (n if n > 0 else 1.0))


def tasks_flops(t, rsz, ksz, csz):
    r, k, c = (np.asarray(x, dtype=np.float64) for x in (rsz, ksz, csz))
    return float((2.0 * r[t[:, 0]] * c[t[:, 1]] * k[t[:, 2]]).sum())


def tasks_issued_flops(t, rsz, ksz, csz, tm, tn, kq=4):
    """MFMA flops the FP64 kernels issue for the task list: each output block
    is covered by whole tm x tn tiles and each product's k is padded to the
    MFMA k (4)."""
    r, k, c = (np.asarray(x, dtype=np.float64) for x in (rsz, ksz, csz))
    rp = np.ceil(r / tm) * tm
    cp = np.ceil(c / tn) * tn
    kp = np.ceil(k / kq) * kq
    return float((2.0 * rp[t[:, 0]] * cp[t[:, 1]] * kp[t[:, 2]]).sum())


def tile_dims(name):
    """(tm, tn) of an FP64 tile config name such as 'stage16x32_k16_n4'."""
    import re

    mt = re.search(r"(\d+)x(\d+)", name or "")
    return (int(mt.group(1)), int(mt.group(2))) if mt else (16, 16)


def timed(ctx, fn, steps):
    ctx.profile_reset()
    ctx.mark(0)
    for _ in range(steps):
        fn()
    ctx.mark(1)
    ctx.synchronize()
    return ctx.elapsed(0, 1) / steps


# ----------------------------------------------------------------- RI-K --
# rocprofv3 summaries of the RI-K step's split-precision kernel, per form
RIK_PROFILES = {"int8ss_5_5": "profiles/gemm_ozaki_int8_shared_ri_k_step.md",
                "int8_5_5": "profiles/gemm_ozaki_int8_ri_k_step.md"}


def profiled_traffic(path):
    """Bytes below L2 per launch of the dominant kernel, as measured by
    rocprofv3 (FETCH_SIZE and WRITE_SIZE, scaled by a copy-kernel calibration)
    for this exact workload and recorded in the committed profile summary;
    null for any other workload."""
    if path:
        try:
            with open(os.path.join(ROOT, path)) as fh:
                for line in fh:
                    if line.startswith("traffic_bytes_per_launch:"):
                        return {"traffic": int(line.split(":")[1]), "traffic_source": path}
        except OSError:
            pass
    return {"traffic": None}


def cpu_port_rik(w, X, budget_s=6.0):
    """Oracle transform on a bounded sample: P-groups 0..g-1 against all Q for
    the first column block of B, grown until about budget_s of work (one core)."""
    import oracle

    bcs = w.b_col_sizes
    b0 = synth.dense_from_scales(7, w.p_sizes, bcs[:1], w.b_scale[:, :1])
    oB = oracle.OBsm(w.p_sizes, bcs[:1], w.thresh).from_dense(b0)
    g = 1
    while True:
        rows = int(np.sum(w.p_sizes[:g]))
        oX = oracle.OBsm(w.p_sizes[:g], w.p_sizes, w.thresh).from_dense(np.asfortranarray(X[:rows]))
        oC = oracle.OBsm(w.p_sizes[:g], bcs[:1], 0.0)
        t0 = time.perf_counter()
        t = oracle.matmult(oC, oX, 0, oB, 0, w.thresh)
        dt = time.perf_counter() - t0
        fl = tasks_flops(t, w.p_sizes[:g], w.p_sizes, bcs[:1])
        if dt > budget_s / 3 or g >= len(w.p_sizes):
            return dt, fl, g, rows
        g = min(len(w.p_sizes), max(g + 1, int(g * min(4.0, budget_s / 3 / max(dt, 1e-3)))))


REF_THREADS = int(os.environ.get("OMP_NUM_THREADS", "16"))  # the host-core share of one GPU


def ref_cores():
    return max(1, min(REF_THREADS, os.cpu_count() or 1))


def load_ref():
    """oracle/_ref (the reference's own templates, Intel MKL, OpenMP), or None.
    MKL stays sequential inside the reference's OpenMP loops over blocks."""
    os.environ.setdefault("MKL_NUM_THREADS", "1")
    import oracle

    return oracle.ref() if oracle.ref_available() else None


def cpu_reference_rik(ctx, w, X, n_mu=4):
    """The reference's screened multiply (blocksparsematrix.hpp:292-414: MKL
    dgemm per block, OpenMP over C super-blocks on ref_cores() threads) on the
    transform B' = X B restricted to the column blocks of the first n_mu
    mu-groups (all rows of X, all i-chunks) -- about n_mu / n_mu_groups of the
    transform -- in uniform blocks of the mean P-group size, since the
    reference takes uniform blocks only.  B's columns are generated in HBM and
    read back (the generator is pinned to its host definition by the tests).
    The task list of that product (for its algorithmic flops) comes from this
    library, whose screening matches the reference's bit for bit away from
    near-ties.  Best of two runs."""
    R = load_ref()
    if R is None:
        return None
    import oracle

    cores = R.ref_set_threads(ref_cores())
    bs = int(round(float(np.mean(w.p_sizes))))
    bcs = w.b_col_sizes
    ncb = n_mu * len(w.i_chunks)
    dBs = sa.BSM(ctx, w.p_sizes, bcs[:ncb], w.thresh).generate(7, w.b_scale[:, :ncb])
    b0 = dBs.to_dense()
    dBs.close()
    a = np.asfortranarray(X)
    rows, nk = a.shape
    nc = b0.shape[1]
    usz = lambda n: [bs] * (n // bs) + ([n % bs] if n % bs else [])
    c = np.zeros((rows, nc), order="F")
    sec = np.zeros(1)
    best = None
    for _ in range(2):
        R.ref_matmult(oracle._ptr(a), rows, nk, w.thresh, oracle._ptr(b0), nk, nc, w.thresh, oracle._ptr(c), rows, nc,
                      0.0, 1, bs, 0, 0, w.thresh, 1.0, 0.0, None, None, oracle._ptr(sec))
        best = float(sec[0]) if best is None else min(best, float(sec[0]))
    dA = sa.BSM.uniform(ctx, rows, nk, bs, bs, w.thresh).from_dense(a)
    dB = sa.BSM.uniform(ctx, nk, nc, bs, bs, w.thresh).from_dense(b0)
    dC = sa.BSM.uniform(ctx, rows, nc, bs, bs, 0.0)
    sa.matmult(dC, dA, 0, dB, 0, w.thresh, 1.0, 0.0)
    fl = tasks_flops(ctx.tasklist(), usz(rows), usz(nk), usz(nc))
    diff = rel(dC.to_dense(), c)
    for m in (dA, dB, dC):
        m.close()
    return best, fl, ncb, bs, cores, diff


def near_tie_count(na, pa, nb, pb, iscr, jscr, kscr, thresh, tol=1e-12):
    """Candidate triples (op(A) block (i, k) and op(B) block (k, j) present)
    whose screening estimate lies within a relative tol of the threshold
    (|est / thr - 1| < tol), and the number of candidates."""
    iscr, jscr, kscr = (np.asarray(x, dtype=np.float64) for x in (iscr, jscr, kscr))
    total = near = 0
    for k in range(na.shape[1]):
        ai = np.nonzero(pa[:, k])[0]
        bj = np.nonzero(pb[k, :])[0]
        if ai.size == 0 or bj.size == 0:
            continue
        est = na[ai, k][:, None] * nb[k, bj][None, :]
        thr = thresh * (iscr[ai][:, None] * jscr[bj][None, :]) * kscr[k]
        near += int(np.count_nonzero(np.abs(est / thr - 1.0) < tol))
        total += est.size
    return near, total


def bench_rik(args):
    w = synth.rik_config(n_occ=args.n_occ) if args.n_occ == 200 else \
        synth.rik_config(n_occ=args.n_occ, name=f"ri_k_step_n_occ{args.n_occ}")
    ctx = sa.Context(0)
    bcs = w.b_col_sizes
    n_ic, n_mu, npb = len(w.i_chunks), len(w.mu_sizes), len(w.p_sizes)
    # set-up: metric, X = M^{-1/2} (Newton-Schulz on the device), timed separately
    log(f"RI-K geometry: {npb} P-groups, {n_mu} mu-groups, i-chunks {[int(x) for x in w.i_chunks]}")
    t0 = time.perf_counter()
    M = w.metric()
    t_metric = time.perf_counter() - t0
    log(f"metric built on the host ({t_metric:.1f} s)")
    dX = sa.BSM(ctx, w.p_sizes, w.p_sizes, w.thresh)
    t0 = time.perf_counter()
    X, ns_it, ns_res = sa.inv_sqrt(ctx, M, X_bsm=dX, tol=1e-14, want_x=not args.no_cpu)
    t_ns = time.perf_counter() - t0
    log(f"X = M^-1/2: {ns_it} Newton-Schulz iterations, residual {ns_res:.2e} ({t_ns:.1f} s)")
    del M
    dB = sa.BSM(ctx, w.p_sizes, bcs, w.thresh).generate(7, w.b_scale)
    dBp = sa.BSM(ctx, w.p_sizes, bcs, 0.0)
    dK = sa.BSM(ctx, w.mu_sizes, w.mu_sizes, 0.0)
    o64 = sa.default_opts()
    # kind suffix "s": the transform's operands sliced with shared exponents (one
    # int32 accumulation and one FP64 fold per output instead of one per product)
    # suffix "ss": shared exponents in the K contraction as well
    kinds = {"fp16": sa.SLICE_FP16, "bf16": sa.SLICE_BF16, "int8": sa.SLICE_INT8,
             "fp16s": sa.SLICE_FP16, "bf16s": sa.SLICE_BF16, "int8s": sa.SLICE_INT8,
             "fp16ss": sa.SLICE_FP16, "bf16ss": sa.SLICE_BF16, "int8ss": sa.SLICE_INT8}
    forms = []
    for f in args.forms.split(","):
        parts = f.split(":")
        kind, s_t, s_k = parts[0], int(parts[1]), int(parts[2])
        tiles = (int(parts[3]), int(parts[4])) if len(parts) == 5 else (-1, -1)
        forms.append((kind, s_t, s_k, tiles))

    # algorithmic flops from the task lists (transform, then K)
    log(f"B generated in HBM ({dB.info()[5] / 1e9:.1f} GB)")
    sa.matmult(dBp, dX, 0, dB, 0, w.thresh, 1.0, 0.0)
    t_tr = ctx.tasklist()
    log(f"transform task list: {len(t_tr)} products; B' {dBp.info()[5] / 1e9:.1f} GB")
    f_tr = tasks_flops(t_tr, w.p_sizes, w.p_sizes, bcs)

    def step64():
        sa.ri_k_step(dX, dB, dBp, dK, w.mu_sizes, w.i_chunks, w.thresh, o64, o64)

    step64()
    ctx.synchronize()
    t_k = ctx.tasklist()
    kr = np.array([p * i for i in w.i_chunks for p in w.p_sizes])
    f_k = tasks_flops(t_k, w.mu_sizes, kr, w.mu_sizes)
    # K = B'^T B' is symmetric: the products of blocks above the diagonal run,
    # those below are their mirror images (written as transposes)
    up = t_k[:, 0] <= t_k[:, 1]
    f_k_exec = tasks_flops(t_k[up], w.mu_sizes, kr, w.mu_sizes)
    log(f"FP64 warm-up step done; K task list: {len(t_k)} products ({int(up.sum())} computed, the rest mirrored)")
    K64 = dK.to_dense()
    # screening near-ties of both products (|est/thr - 1| < 1e-12), from the cached block norms
    px, nx = dX.meta()
    pb_, nb_ = dB.meta()
    pbp, nbp = dBp.meta()
    ties_tr = near_tie_count(nx, px, nb_, pb_, w.p_sizes, bcs, w.p_sizes, w.thresh)
    nv = nbp.reshape(npb * n_ic, n_mu, order="F")
    pv = pbp.reshape(npb * n_ic, n_mu, order="F")
    vscr = np.array([p * i for i in w.i_chunks for p in w.p_sizes])
    ties_k = near_tie_count(nv.T, pv.T, nv, pv, w.mu_sizes, w.mu_sizes, vscr, w.thresh)
    samp_pg = [0, npb // 2]
    bp_sample = {"fp64": [dBp.blocks(g, g + 1, 0, n_ic) for g in samp_pg]}
    for _ in range(max(args.warmup - 1, 0)):
        step64()
    ctx.set_profiling(True)
    ms64 = timed(ctx, step64, args.steps)
    log(f"FP64 step: {ms64:.1f} ms")
    g64_ms, g64_n = ctx.profile("gemm_fp64")
    cfg64 = ctx.gemm_config()[1]
    st64 = {k: round(ctx.profile(k)[0] / args.steps, 3) for k in ("rik_transform", "rik_k")}
    ctx.profile_reset()
    dense_fl = 2.0 * w.n_aux ** 2 * w.n_bf * w.n_occ + 2.0 * w.n_bf ** 2 * w.n_aux * w.n_occ
    fp64_achieved = (f_tr + f_k_exec) / (g64_ms / args.steps * 1e-3) / 1e12
    mo = np.concatenate([[0], np.cumsum(w.mu_sizes)])

    results = {}
    for kind, s_t, s_k, tiles in forms:
        name = f"{kind}_{s_t}_{s_k}" + ("" if tiles == (-1, -1) else f"_t{tiles[0]}_{tiles[1]}")
        oz_t = sa.default_opts(mode=sa.MODE_OZAKI, slices=s_t, slice_type=kinds[kind], tile=tiles[0],
                               shared_exp=1 if kind.endswith("s") else 0)
        oz_k = sa.default_opts(mode=sa.MODE_OZAKI, slices=s_k, slice_type=kinds[kind], tile=tiles[1],
                               shared_exp=1 if kind.endswith("ss") else 0)

        def stepmix():
            sa.ri_k_step(dX, dB, dBp, dK, w.mu_sizes, w.i_chunks, w.thresh, oz_t, oz_k)

        ctx.slice_pairs()  # reset the slice-pair counters
        for _ in range(args.warmup):
            stepmix()
        ms = timed(ctx, stepmix, args.steps)
        log(f"split-precision step ({name}): {ms:.1f} ms")
        n_mult, n_skip = ctx.slice_pairs()
        skip_frac = n_skip / max(1, n_mult + n_skip)
        goz_ms, goz_n = ctx.profile("gemm_ozaki")
        prof = {k: round(ctx.profile(k)[0] / args.steps, 3) for k in ("gemm_ozaki", "ozaki_slice", "screen", "norms",
                                                                       "rik_transform", "rik_k")}
        K = dK.to_dense()
        err = rel(K, K64)
        rows_err = [rel(K[mo[r]:mo[r + 1]], K64[mo[r]:mo[r + 1]]) for r in range(n_mu)]
        bp_sample[name] = [dBp.blocks(g, g + 1, 0, n_ic) for g in samp_pg]
        # slice-pair products of the step, less those the kernel skips as zero
        # (an all-zero op(A) slice in the tile's rows), i.e. executed MFMA work
        pairs = (s_t * (s_t + 1) // 2 * f_tr + s_k * (s_k + 1) // 2 * f_k_exec) * (1.0 - skip_frac)
        peak = PEAK_INT8 if kind.startswith("int8") else PEAK_FP16
        ach = pairs / (goz_ms / args.steps * 1e-3) / 1e12
        results[name] = {"slice_type": kind.rstrip("s"), "slices_transform": s_t, "slices_k": s_k,
                         "shared_exponents": {"transform": kind.endswith("s"), "k": kind.endswith("ss")},
                         "step_ms": round(ms, 3),
                         "speedup_vs_fp64_gpu": round(ms64 / ms, 3), "rel_err_K_vs_fp64": err,
                         "rel_err_K_per_block_row": {"max": max(rows_err), "median": float(np.median(rows_err))},
                         "meets_1e-10": bool(err <= 1e-10), "ms_per_step": prof,
                         "lowp_tops": round(ach, 3), "lowp_peak": peak, "lowp_frac": round(ach / peak, 4),
                         "slice_pairs_skipped_as_zero": round(skip_frac, 4),
                         "kernel_avg_ms": round(goz_ms / max(goz_n, 1), 3), "kernel_launches_per_step": goz_n // args.steps}
    ctx.set_profiling(False)
    ok = [k for k, v in results.items() if v["meets_1e-10"]]
    if not ok:
        log("no split-precision form reached 1e-10 against FP64: the headline reports the fastest form, flagged")
    head = min(ok or results, key=lambda k: results[k]["step_ms"])
    H = results[head]
    msmix = H["step_ms"]

    rec = {
        "metric": METRIC,
        "value": round(dense_fl / (msmix * 1e-3) / 1e12, 3),
        "unit": "dense-equivalent TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(msmix, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "meets_1e-10": bool(ok),
        "dtype": f"f64 with {H['slice_type']}-slice products",
        "data": "synthetic (seeded geometry; B generated in HBM from a counter-based hash)",
        "config": {"workload": w.name, "natoms": w.natoms, "n_bf": w.n_bf, "n_aux": w.n_aux, "n_occ": w.n_occ,
                   "p_groups": int(npb), "mu_groups": int(n_mu),
                   "i_chunks": [int(x) for x in w.i_chunks], "b_block_fill": round(float((w.b_scale > 0).mean()), 4),
                   "thresh": w.thresh, "form": head, "parallelism": "single GPU"},
        "ri_k": {"step_ms_split_precision": msmix, "step_ms_fp64": round(ms64, 3),
                 "speedup_vs_fp64_gpu": H["speedup_vs_fp64_gpu"], "rel_err_K_vs_fp64": H["rel_err_K_vs_fp64"],
                 "dense_equiv_tflops_fp64": round(dense_fl / (ms64 * 1e-3) / 1e12, 3),
                 "transform_tasks": int(len(t_tr)), "k_tasks": int(len(t_k)), "k_tasks_computed": int(up.sum()),
                 "algorithmic_flop": f_tr + f_k, "algorithmic_flop_computed": f_tr + f_k_exec,
                 "dense_equiv_flop": dense_fl,
                 "fp64_gemm_tflops": round(fp64_achieved, 3), "fp64_gemm_frac": round(fp64_achieved / PEAK_FP64, 4),
                 "fp64_gemm_tile": cfg64, "stage_ms_fp64": st64, "forms": results,
                 "screening_near_ties": {"transform": {"within_1e-12": ties_tr[0], "candidates": ties_tr[1]},
                                         "k": {"within_1e-12": ties_k[0], "candidates": ties_k[1]}},
                 "setup": {"metric_host_s": round(t_metric, 3), "inv_sqrt_s": round(t_ns, 3),
                           "newton_schulz_iterations": ns_it, "newton_schulz_residual": ns_res}},
        "roofline": {"bound": "mfma", "achieved": H["lowp_tops"], "peak": H["lowp_peak"],
                     "unit": "TOP/s" if H["slice_type"].startswith("int8") else "TFLOP/s",
                     "frac": H["lowp_frac"], "kernel": "k_gemm_ozaki",
                     "slice_pairs_skipped_as_zero": H["slice_pairs_skipped_as_zero"],
                     "frac_counting_skipped_pairs": round(H["lowp_frac"] / max(1e-9, 1.0 - H["slice_pairs_skipped_as_zero"]), 4),
                     **profiled_traffic(RIK_PROFILES.get(head) if args.n_occ == 200 else None),
                     "kernel_avg_ms": H["kernel_avg_ms"],
                     "note": "achieved = executed low-precision slice-product operations (slice pairs x algorithmic "
                             "flop of the computed products: the transform's and K's upper triangle, less the "
                             "fraction of slice pairs the kernel skips as zero, slice_pairs_skipped_as_zero) / "
                             "k_gemm_ozaki time over the timed steps; the FP64 path's k_gemm_fp64 is in "
                             "ri_k.fp64_gemm_tflops (peak 78.6)"},
    }
    if not args.no_cpu:
        import oracle

        log("B' sample against the CPU oracle ...")
        rows = np.concatenate([[0], np.cumsum(w.p_sizes)])
        ps = [int(w.p_sizes[g]) for g in samp_pg]
        xs = np.asfortranarray(np.vstack([X[rows[g]:rows[g + 1]] for g in samp_pg]))
        dBs = sa.BSM(ctx, w.p_sizes, bcs[:n_ic], w.thresh).generate(7, w.b_scale[:, :n_ic])
        oBs = oracle.OBsm(w.p_sizes, bcs[:n_ic], w.thresh).from_dense(dBs.to_dense())
        dBs.close()
        oXs = oracle.OBsm(ps, w.p_sizes, w.thresh).from_dense(xs)
        oC = oracle.OBsm(ps, bcs[:n_ic], 0.0)
        oracle.matmult(oC, oXs, 0, oBs, 0, w.thresh)
        ref = oC.to_dense()
        rec["ri_k"]["bp_sample_vs_cpu_oracle"] = {
            "sample": f"B' rows of P-groups {samp_pg} ({sum(ps)} rows) x the {n_ic} column blocks of mu-group 0",
            **{k: rel(np.vstack(v), ref) for k, v in bp_sample.items()}}
        del oBs, oXs, oC
        log("CPU reference (oracle/_ref, OpenMP + MKL) ...")
        rr = cpu_reference_rik(ctx, w, X)
        log("CPU port (oracle, one core) ...")
        dt, fl, g, prow = cpu_port_rik(w, X)
        rate_p = fl / dt
        est_p = (f_tr + f_k) / rate_p
        port = {"value": round(dense_fl / est_p / 1e12, 6), "unit": "dense-equivalent TFLOP/s", "cores": 1,
                "kind": "port",
                "sample": f"oracle screened transform of P-groups 0..{g - 1} ({prow} rows) against all of X for "
                          f"B's first column block: {fl:.3g} algorithmic flop in {dt:.1f} s; full step estimated "
                          f"from the step's algorithmic flops",
                "est_step_s": round(est_p, 1), "gpu_speedup_split_precision": round(est_p / (msmix * 1e-3), 1)}
        if rr is not None:
            dt, fl, ncb, ubs, cores, diff = rr
            rate = fl / dt
            est = (f_tr + f_k) / rate
            rec["cpu_baseline"] = {
                "value": round(dense_fl / est / 1e12, 6), "unit": "dense-equivalent TFLOP/s", "cores": cores,
                "kind": "reference", "host_cpus": os.cpu_count(),
                "sample": f"the reference's matmult (MKL dgemm per block, OpenMP on {cores} threads: the host-core "
                          f"share of one GPU, of {os.cpu_count()} host CPUs) on the transform X B restricted to B's "
                          f"first {ncb} column blocks ({ncb} of {len(bcs)}: {fl / f_tr:.1%} of the transform's "
                          f"algorithmic flop), all 12000 rows, in uniform {ubs}-blocks (the reference takes uniform "
                          f"blocks only): {fl:.3g} algorithmic flop in {dt:.2f} s (best of 2); the reference has no "
                          f"K contraction (tensormult form 1 is declared only), so the whole step "
                          f"({f_tr + f_k:.3g} flop, K without symmetry) is estimated at the measured rate",
                "measured_tflops": round(rate / 1e12, 3),
                "rel_diff_gpu_vs_reference_on_sample": diff,
                "est_step_s": round(est, 2),
                "gpu_speedup_split_precision": round(est / (msmix * 1e-3), 1),
                "gpu_speedup_fp64": round(est / (ms64 * 1e-3), 1)}
            rec["cpu_port"] = port
        else:
            rec["cpu_baseline"] = port
    print(json.dumps(rec))


# ------------------------------------------------------ hybrid Strassen --
def occ_problem(n_bf=600, n_aux=1200, n_occ=60, frac=0.35, seed=11):
    """Synthetic occupied RI-K input in the reference's formats: pairs r <= c of
    the basis kept with probability frac, values decaying with |r - c|, keys
    r + n_bf c (v2m_key layout), MO coefficients N(0, 1)."""
    rng = np.random.default_rng(seed)
    r, c = np.triu_indices(n_bf)
    keep = rng.random(r.size) < frac
    r, c = r[keep], c[keep]
    keys = (r + n_bf * c).astype(np.uint64)
    ints = np.asfortranarray(rng.standard_normal((keys.size, n_aux)) * np.exp(-0.01 * np.abs(r - c))[:, None])
    mo = np.asfortranarray(rng.standard_normal((n_bf, n_occ)))
    return keys.size, ints, mo, keys


def bench_occ(args):
    """Device occupied RI-K (sa_occ_ri_k, occ_ri_k_step.cpp:83-131) on a synthetic
    problem; the oracle's restatement timed on a bounded subset of aux functions
    for the CPU baseline.  The call includes the upload of the compressed
    integrals (host arrays in the reference's formats)."""
    n_bf, n_aux, n_occ, bs, thresh = 600, 1200, 60, 12, 1e-10
    n_vec2, ints, mo, keys = occ_problem(n_bf, n_aux, n_occ)
    ctx = sa.Context(0)
    teams = args.teams
    run = lambda: sa.occ_ri_k(ctx, n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh, teams=teams)
    for _ in range(args.warmup):
        K = run()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        K = run()
    ms = (time.perf_counter() - t0) / args.steps * 1e3
    fl = 2.0 * (n_bf * n_bf * n_occ * n_aux + 2.0 * n_bf * n_occ * n_occ * n_aux)  # occ_ri_k_step.cpp:75
    rec = {"metric": METRIC, "value": round(fl / (ms * 1e-3) / 1e12, 4), "unit": "dense-equivalent TFLOP/s",
           "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 3),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
           "data": "synthetic (seeded compressed pairs and MO coefficients, host arrays)",
           "config": {"workload": "occ_ri_k_device", "n_bf": n_bf, "n_aux": n_aux, "n_occ": n_occ, "n_vec2": n_vec2,
                      "bs": bs, "thresh": thresh, "teams": teams, "parallelism": f"single GPU, {teams} streams",
                      "timing": "wall clock of the call, including the upload of the integrals"}}
    if not args.no_cpu:
        import oracle

        g = 8
        while True:
            t0 = time.perf_counter()
            oracle.occ_ri_k(n_bf, g, n_occ, n_vec2, np.asfortranarray(ints[:, :g]), mo, keys, bs, thresh)
            dt = time.perf_counter() - t0
            if dt > 3.0 or g >= n_aux:
                break
            g = min(n_aux, g * 4)
        K_dev = sa.occ_ri_k(ctx, n_bf, g, n_occ, n_vec2, np.asfortranarray(ints[:, :g]), mo, keys, bs, thresh,
                            teams=1)
        K_or = oracle.occ_ri_k(n_bf, g, n_occ, n_vec2, np.asfortranarray(ints[:, :g]), mo, keys, bs, thresh)
        rate = fl * g / n_aux / dt
        rec["cpu_baseline"] = {"value": round(rate / 1e12, 6), "unit": "dense-equivalent TFLOP/s", "cores": 1,
                               "kind": "port", "sample": f"oracle P loop over the first {g} of {n_aux} aux functions "
                                                         f"({dt:.1f} s), scaled by aux count",
                               "rel_diff_device_vs_oracle_on_sample": rel(K_dev, K_or)}
        log("CPU reference (oracle/_ref team) ...")
        rr = cpu_reference_occ(n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh)
        if rr is not None:
            dt, cores, K_ref = rr
            rec["cpu_reference"] = {
                "value": round(fl / dt / 1e12, 6), "unit": "dense-equivalent TFLOP/s", "cores": cores,
                "kind": "reference",
                "sample": f"the whole problem: the reference's per-thread chain (occ_ri_k_step.cpp:90-120 with "
                          f"MKL) on a team of {cores} threads, aux functions split statically, in {dt:.2f} s",
                "rel_diff_device_vs_reference": rel(K, K_ref),
                "gpu_speedup": round(dt / (ms * 1e-3), 2)}
    print(json.dumps(rec))


def cpu_reference_occ(n_bf, n_aux, n_occ, n_vec2, ints, mo, keys, bs, thresh):
    """The reference's occupied RI-K team (occ_ri_k_step.cpp:83-131): each of
    ref_cores() host threads runs the reference's own chain (oracle/_ref) over
    its static share of the aux functions, K_sub summed.  Returns (seconds,
    threads, K) or None."""
    import threading

    import oracle

    R = load_ref()
    if R is None:
        return None
    T = min(ref_cores(), n_aux)
    q, r = divmod(n_aux, T)
    parts = []
    for t in range(T):
        p0, n = t * q + min(t, r), q + (1 if t < r else 0)
        parts.append((np.asfortranarray(ints[:, p0:p0 + n]), n, np.zeros((n_bf, n_occ), order="F")))
    mo_f = np.asfortranarray(mo)
    keys_c = np.ascontiguousarray(keys, dtype=np.uint64)

    def member(x, n, Kt):
        R.ref_set_threads(1)  # nested regions serial, as inside the reference's team
        R.ref_occ_ri_k(n_bf, n, n_occ, n_vec2, oracle._ptr(x), oracle._ptr(mo_f), oracle._ptr(keys_c), bs, thresh,
                       oracle._ptr(Kt))

    th = [threading.Thread(target=member, args=pt) for pt in parts]
    t0 = time.perf_counter()
    for t in th:
        t.start()
    for t in th:
        t.join()
    dt = time.perf_counter() - t0
    K = np.zeros((n_bf, n_occ), order="F")
    for _, _, Kt in parts:
        K = K + Kt
    return dt, T, K


def bench_hybrid(args):
    """Large hybrid recursion: n=32768, 128x128 blocks, band |I-J|<=95 under
    exp(-0.05|I-J|), cutoff 2048 (Strassen depth <= 4).  One step = one
    hybrid multiply into a pre-filled C, planning (dry run, features, network,
    plan) included; variants: plain screened multiply, the reference cost-model
    tree, the learned plan."""
    w = synth.hybrid_workload()
    n, bs, cutoff = w.n, w.uniform_bs, w.extra["cutoff"]
    ctx = sa.Context(0)
    A = sa.BSM.uniform(ctx, n, n, bs, bs, w.thresh).generate(1, w.a_scale)
    B = sa.BSM.uniform(ctx, n, n, bs, bs, w.thresh).generate(2, w.b_scale)
    Cm = sa.BSM.uniform(ctx, n, n, bs, bs, 0.0)
    Cm.fill(0.0)
    log(f"hybrid: A, B generated ({A.info()[5] / 1e9:.1f} GB each), C pre-filled ({Cm.info()[5] / 1e9:.1f} GB)")
    model = os.path.exists(sa.DEFAULT_PATH_MODEL)
    if model:
        sa.mlp_load(ctx, sa.DEFAULT_PATH_MODEL)

    variants = {
        "screened": lambda: sa.matmult(Cm, A, 0, B, 0, w.thresh, 1.0, 1.0),
        "heuristic": lambda: sa.matmult_strassen_sparse(Cm, A, B, w.thresh, cutoff,
                                                        opts=sa.default_opts(scheduler=sa.SCHED_HEURISTIC)),
        "learned": lambda: sa.matmult_strassen_sparse(Cm, A, B, w.thresh, cutoff,
                                                      opts=sa.default_opts(scheduler=sa.SCHED_MLP)),
    }
    res, stats, outs = {}, {}, {}
    ref = None
    for name, fn in variants.items():
        Cm.fill(0.0)
        fn()
        ctx.synchronize()
        stats[name] = {**sa.strassen_last_stats(), **sa.strassen_last_plan()} if name != "screened" else \
            {"tasks": int(len(ctx.tasklist()))}
        outs[name] = Cm.frobenius_norm()
        if name == "heuristic":
            ref = Cm.to_dense() if not args.no_check else None
        if name == "learned" and ref is not None:
            stats[name]["rel_diff_vs_heuristic"] = rel(Cm.to_dense(), ref)
        for _ in range(max(args.warmup - 1, 0)):
            fn()
        ms = timed(ctx, fn, args.steps)
        res[name] = ms
        log(f"hybrid {name}: {ms:.1f} ms per step")
    dense = 2.0 * n ** 3
    best = "learned" if model else "heuristic"
    rec = {
        "metric": METRIC, "value": round(dense / (res[best] * 1e-3) / 1e12, 3), "unit": "dense-equivalent TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(res[best], 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": w.name, "n": n, "bs": bs, "band": w.extra["band"], "decay": w.extra["decay"],
                   "thresh": w.thresh, "cutoff": cutoff, "scheduler": best, "parallelism": "single GPU"},
        "hybrid": {"ms_per_step": {k: round(v, 3) for k, v in res.items()},
                   "learned_speedup_vs_heuristic": round(res["heuristic"] / res["learned"], 3),
                   "model_loaded": model, "stats": stats},
    }
    print(json.dumps(rec))


# -------------------------------------------------------- screened GEMM --
def bench_gemm(args):
    if args.workload == "band_split":
        w = synth.band_workload()
        w.name += "_split_fp16"
        opts = sa.default_opts(mode=sa.MODE_OZAKI, slices=args.slices_k, tile=args.tile)
    else:
        w = {"band": synth.band_workload, "atoms": synth.atoms_workload}[args.workload]()
        opts = sa.default_opts(tile=args.tile)
    ozaki = opts.mode == sa.MODE_OZAKI
    n = w.n
    ctx = sa.Context(0)
    A = sa.BSM(ctx, w.rsz, w.ksz, w.thresh).generate(1, w.a_scale)
    B = sa.BSM(ctx, w.ksz, w.csz, w.thresh).generate(2, w.b_scale)
    Cm = sa.BSM(ctx, w.rsz, w.csz, 0.0)

    def step():
        sa.matmult(Cm, A, 0, B, 0, w.thresh, 1.0, 0.0, opts)

    for _ in range(args.warmup):
        step()
    ctx.synchronize()
    t = ctx.tasklist()
    flops_alg = tasks_flops(t, w.rsz, w.ksz, w.csz)
    err = None
    if ozaki:
        ref = sa.BSM(ctx, w.rsz, w.csz, 0.0)
        sa.matmult(ref, A, 0, B, 0, w.thresh, 1.0, 0.0)
        err = rel(Cm.to_dense(), ref.to_dense())
        del ref
    ctx.set_profiling(True)
    ms = timed(ctx, step, args.steps)
    cls = "gemm_ozaki" if ozaki else "gemm_fp64"
    g_ms, g_n = ctx.profile(cls)
    scr_ms, _ = ctx.profile("screen")
    ctx.set_profiling(False)
    cfg_name = ctx.gemm_config()[1] if not ozaki else f"ozaki_fp16_s{args.slices_k}"
    dense = 2.0 * n ** 3
    g_avg = g_ms / max(g_n, 1)
    issued = None
    if ozaki:
        pairs = args.slices_k * (args.slices_k + 1) // 2
        achieved, peak = pairs * flops_alg / (g_avg * 1e-3) / 1e12, PEAK_FP16
    else:
        achieved, peak = flops_alg / (g_avg * 1e-3) / 1e12, PEAK_FP64
        issued = tasks_issued_flops(t, w.rsz, w.ksz, w.csz, *tile_dims(cfg_name))
    rec = {
        "metric": METRIC, "value": round(dense / (ms * 1e-3) / 1e12, 3), "unit": "dense-equivalent TFLOP/s",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(ms, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64 with fp16-slice products" if ozaki else "f64",
        "data": "synthetic (seeded counter-based generator, generated in HBM)",
        "config": {"workload": w.name, "n": n, "nblocks": int(len(w.rsz)), "thresh": w.thresh, "tasks": int(len(t)),
                   "algorithmic_flop": flops_alg, "gemm_tile": cfg_name,
                   "gemm_tile_choice": "forced" if args.tile >= 0 else "autotuned", "parallelism": "single GPU",
                   "rel_err_vs_fp64": err, "issued_mfma_flop": issued,
                   "useful_frac_of_issued": round(flops_alg / issued, 4) if issued else None,
                   "issued_tflops": round(issued / (g_avg * 1e-3) / 1e12, 3) if issued else None},
        "roofline": {"bound": "mfma", "achieved": round(achieved, 3), "peak": peak, "unit": "TFLOP/s",
                     "frac": round(achieved / peak, 4), "traffic": None,
                     "kernel": "k_gemm_ozaki" if ozaki else "k_gemm_fp64", "kernel_avg_ms": round(g_avg, 4),
                     "screen_ms_per_step": round(scr_ms / args.steps, 4)},
    }
    if not args.no_cpu:
        import oracle

        ncols, dt, fl = 1, 0.0, 0.0
        a = synth.dense_from_scales(1, w.rsz, w.ksz, w.a_scale)
        oA = oracle.OBsm(w.rsz, w.ksz, w.thresh).from_dense(a)
        while True:
            csz = w.csz[:ncols]
            b = synth.dense_from_scales(2, w.ksz, w.csz, w.b_scale)[:, : int(np.sum(csz))]
            oB = oracle.OBsm(w.ksz, csz, w.thresh).from_dense(b)
            oC = oracle.OBsm(w.rsz, csz, 0.0)
            t0 = time.perf_counter()
            tt = oracle.matmult(oC, oA, 0, oB, 0, w.thresh)
            dt = time.perf_counter() - t0
            fl = tasks_flops(tt, w.rsz, w.ksz, csz)
            if dt > 5.0 or ncols >= len(w.csz):
                break
            ncols = min(len(w.csz), max(ncols + 1, int(ncols * min(8.0, 5.0 / max(dt, 1e-3)))))
        est = flops_alg / (fl / dt)
        rec["cpu_baseline"] = {"value": round(dense / est / 1e12, 6), "unit": "dense-equivalent TFLOP/s", "cores": 1,
                               "kind": "port",
                               "sample": f"oracle screened multiply over the first {ncols} of {len(w.csz)} C block "
                                         f"columns ({dt:.1f} s, {fl:.3g} algorithmic flop), scaled by flops"}
    print(json.dumps(rec))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="rik", choices=["rik", "band", "atoms", "band_split", "hybrid", "occ"])
    ap.add_argument("--tile", type=int, default=-1)
    ap.add_argument("--slices-k", type=int, default=4, help="split-precision slices (band_split workload)")
    ap.add_argument("--forms", default="int8ss:5:5,int8:5:5,fp16:4:4",
                    help="RI-K split-precision forms kind:slices_transform:slices_k[:tile_transform:tile_k] "
                         "(kind fp16|bf16|int8, suffix s for shared exponents in the transform, ss in both "
                         "stages, e.g. int8ss; tiles = split-precision kernel table ids, default chosen)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--teams", type=int, default=8, help="occ: team members (streams, host threads)")
    ap.add_argument("--n-occ", type=int, default=200, help="RI-K occupied orbitals (200 = the headline workload)")
    ap.add_argument("--no-check", action="store_true", help="hybrid: skip the learned-vs-heuristic comparison")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1", "n_gpus": args.gpus}))
        sys.exit(2)
    if args.workload == "rik":
        bench_rik(args)
    elif args.workload == "hybrid":
        bench_hybrid(args)
    elif args.workload == "occ":
        bench_occ(args)
    else:
        bench_gemm(args)


if __name__ == "__main__":
    main()
