/*
 * strassen_amd.h -- C ABI of the MI355X (gfx950) block-sparse multiply library.
 *
 * The reference (diersam/Strassen) is a header-only C++ template library with
 * no FFI; its callers include blocksparsematrix.h / strassen.hpp directly.
 * This ABI is what those headers' bodies call in the drop-in C++ layer
 * (strassen_amd/cpp/): every entry point below names the reference interface
 * it replaces.  Plain pointers and sizes only; no C++ types cross it.
 *
 * Conventions
 *  - return value: SA_OK (0) or a negative SA_ERR_* code; sa_last_error()
 *    gives the message of the calling thread's last failure;
 *  - host buffers are caller-owned and column-major; device memory is owned by
 *    the context / matrices;
 *  - a context owns one HIP stream; all calls on a context are ordered on it
 *    and calls that return host data synchronise it;
 *  - a block-sparse matrix (sa_bsm) is a grid of column-major blocks, grid
 *    index rb + cb*nrb (blocksparsematrix.h:71-72), each block with a cached
 *    Frobenius norm that follows the reference's staleness rules.
 */
#ifndef STRASSEN_AMD_H
#define STRASSEN_AMD_H
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SA_OK 0
#define SA_ERR_INVALID (-1)     /* bad argument, shape or block geometry */
#define SA_ERR_HIP (-2)         /* HIP runtime failure */
#define SA_ERR_OOM (-3)         /* device allocation failed */
#define SA_ERR_UNSUPPORTED (-4) /* valid request outside what this build supports */

typedef struct sa_context* sa_ctx_t;
typedef struct sa_bsm* sa_bsm_t;

/* Multiply options (no counterpart in the reference; defaults = reference
 * semantics in FP64). */
#define SA_MODE_FP64 0
#define SA_MODE_OZAKI 1
#define SA_SLICE_FP16 0 /* 10-bit digits in FP16, FP32 MFMA accumulation (3..5 slices) */
#define SA_SLICE_BF16 1 /* 8-bit digits in BF16, FP32 accumulation (4..6 slices) */
#define SA_SLICE_INT8 2 /* 8-bit digits in int8, exact int32 accumulation (4..6 slices) */
#define SA_SCHED_HEURISTIC 0 /* dry-run cost model, strassen.hpp:579-673 */
#define SA_SCHED_MLP 1       /* learned per-node selector, heuristic as fallback */
#define SA_SCHED_TREE 2      /* caller-supplied decision tree bits */
typedef struct {
  int32_t mode;        /* SA_MODE_* */
  int32_t slices;      /* Ozaki: slices per operand (FP16: 3..5, BF16 and INT8: 4..6) */
  int32_t slice_type;  /* SA_SLICE_* */
  int32_t tile;        /* FP64: GEMM tile config id (sa_ctx_gemm_config), -1 = autotuned;
                          Ozaki: kernel table id, ignored unless it runs this slice count/type */
  int32_t scheduler;   /* SA_SCHED_* (hybrid Strassen only) */
  int32_t shared_exp;  /* Ozaki: 0 (default) = power-of-two scale per row / column of each
                          block; 1 = per row of op(A) and column of op(B) across all inner
                          blocks, so an output's products share one int32 accumulation and
                          fold into FP64 once (fewer folds for short inner blocks; small
                          blocks then carry fewer significant bits) */
  int64_t cutoff;      /* min_size_for_strassen (main.cpp:9), elements */
  const uint8_t* tree; /* SA_SCHED_TREE: Strassen-root bits in dry-run layout */
  const uint8_t* tree_dense; /* SA_SCHED_TREE (optional): dense-root bits, same layout; a
                                dense root adds all present block products of its region,
                                unscreened, as one library GEMM (no reference counterpart:
                                the learned path selector's third path) */
} sa_opts_t;

void sa_opts_default(sa_opts_t* o);

/* ------------------------------------------------------------ context -- */
/* Last error message of the calling thread ("" if none). */
const char* sa_last_error(void);
int sa_device_count(int* n);
/* Context on one device: stream, workspace, profiling counters. */
int sa_ctx_create(int device, sa_ctx_t* out);
int sa_ctx_destroy(sa_ctx_t ctx);
int sa_ctx_synchronize(sa_ctx_t ctx);
/* Profiling: when on, every launch of the dominant kernels is bracketed by
 * HIP events on the context stream; sa_ctx_profile returns the accumulated
 * kernel time (ms) and launch count of one kernel class
 * ("gemm_fp64", "gemm_ozaki", "screen", "norms", "winograd", ...). */
int sa_ctx_set_profiling(sa_ctx_t ctx, int on);
int sa_ctx_profile(sa_ctx_t ctx, const char* kernel_class, double* ms, int64_t* launches);
int sa_ctx_profile_reset(sa_ctx_t ctx);
/* Tile configuration id and name used by the context's last FP64 GEMM launch
 * (chosen by the autotuner unless opts->tile forced one). */
int sa_ctx_gemm_config(sa_ctx_t ctx, int32_t* cfg, char* name, int32_t name_len);
/* Number of k-windows (grouped launches in ascending inner-block order) the
 * context's last screened multiply ran in; 1 = one launch.  Windows are used
 * when the operands far exceed the 256 MB Infinity Cache while the written C
 * blocks fit in it; env SA_KWINDOW_MB sets the operand budget per window
 * (default 160, 0 = never).  No counterpart in the reference, which runs
 * per-block dgemm calls (blocksparsematrix.hpp:363-405). */
int sa_ctx_kwindows(sa_ctx_t ctx, int32_t* nwin);
/* HIP-graph counters of the context: executables instantiated, executables
 * updated in place from a new capture (replays of the same topology) and
 * graph launches.  The Strassen execution phase of sa_matmult_strassen_sparse is
 * captured on every call and replayed through one kept executable.  No
 * counterpart in the reference, which has no device graphs. */
int sa_ctx_graph_stats(sa_ctx_t ctx, int64_t* instantiated, int64_t* updated, int64_t* launched);
/* Split-precision slice-pair work of the context's multiplies since the last
 * call (then reset): slice-pair MFMA k-steps times GEMM tiles multiplied, and
 * those skipped because the op(A) slice of the pair is all zero in the tile's
 * rows (a block far below its line's shared exponent; the skipped pairs add
 * nothing, results are unchanged).  No counterpart in the reference, which
 * multiplies in FP64 (blocksparsematrix.hpp:391-404). */
int sa_ctx_slice_pairs(sa_ctx_t ctx, int64_t* multiplied, int64_t* skipped);
/* Wall time of the context stream between two sa_ctx_mark calls (ms). */
int sa_ctx_mark(sa_ctx_t ctx, int slot);
int sa_ctx_elapsed(sa_ctx_t ctx, int slot0, int slot1, float* ms);

/* ---------------------------------------------------- block-sparse matrix -- */
/* BlockSparseMatrix(nr, nc, bs_r, bs_c, thresh) -- blocksparsematrix.hpp:6-16
 * (uniform target block size, ragged last block row/column). */
int sa_bsm_create(sa_ctx_t ctx, int64_t nrow, int64_t ncol, int32_t bs_row, int32_t bs_col,
                  double thresh, sa_bsm_t* out);
/* Extension: uneven block rows/columns (actual sizes rsz/csz; screening sizes
 * rscr/cscr, NULL = actual sizes). */
int sa_bsm_create_blocked(sa_ctx_t ctx, int32_t nrb, const int32_t* rsz, const int32_t* rscr,
                          int32_t ncb, const int32_t* csz, const int32_t* cscr, double thresh,
                          sa_bsm_t* out);
int sa_bsm_destroy(sa_bsm_t m);
/* nrow(), ncol(), nrowblocks(), ncolblocks(), no_of_alloc_blocks()
 * (blocksparsematrix.h:49-58); arena bytes in use. */
int sa_bsm_info(sa_bsm_t m, int64_t* nrow, int64_t* ncol, int32_t* nrb, int32_t* ncb,
                int64_t* nalloc, int64_t* arena_bytes);
/* Copy assignment / copy ctor (blocks and cached norms). */
int sa_bsm_copy(sa_bsm_t dst, sa_bsm_t src);
/* copy_from_input_array (blocksparsematrix.hpp:35-61): blocks with
 * norm < thresh*bs_r*bs_c are dropped. */
int sa_bsm_from_dense(sa_bsm_t m, const double* host, int64_t ld);
/* to_pointer (blocksparsematrix.hpp:103-124): writes present blocks only. */
int sa_bsm_to_dense(sa_bsm_t m, double* host, int64_t ld);
/* Dense values of the block sub-grid rows [rb0, rb1) x columns [cb0, cb1)
 * (absent blocks as zeros) into host (column-major, leading dimension ld >=
 * the sub-grid's rows): the reference's per-block read access block(rb, cb)
 * (blocksparsematrix.h:71-72) for one block or a range. */
int sa_bsm_get_blocks(sa_bsm_t m, int32_t rb0, int32_t rb1, int32_t cb0, int32_t cb1, double* host, int64_t ld);
/* Block presence (1/0) and cached norms, grid order. Either may be NULL. */
int sa_bsm_get_meta(sa_bsm_t m, uint8_t* present, double* norms);
/* Extension: synthetic input generated in HBM.  Block b gets
 * block_scale[b] * z(seed, global element index) (z defined in
 * strassen_amd/synth.py), then the copy-in drop rule applies, exactly as
 * sa_bsm_from_dense of the equivalent dense matrix. */
int sa_bsm_generate(sa_bsm_t m, uint64_t seed, const double* block_scale);
/* Extension: reshape view.  A matrix over the same block payloads, cached
 * norms and grid index b as `base`, with new block geometry (each block keeps
 * its element count and its column-major storage, e.g. a block of nP x (ni*nmu)
 * read as (nP*ni) x nmu).  Views are read-only (matmult inputs, to_dense,
 * meta, dot); they stay valid while `base` lives; destroy with sa_bsm_destroy. */
int sa_bsm_view_reshape(sa_bsm_t base, int32_t nrb, const int32_t* rsz, const int32_t* rscr, int32_t ncb,
                        const int32_t* csz, const int32_t* cscr, sa_bsm_t* out);
/* fill_with_values (blocksparsematrix.hpp:84-101): all blocks, norm sqrt(r*c)*v. */
int sa_bsm_fill(sa_bsm_t m, double v);
/* zero() (blocksparsematrix.hpp:79-82): all blocks empty. */
int sa_bsm_zero(sa_bsm_t m);
/* scale / *= (blocksparsematrix.hpp:71-77, 250-254): cached norm *= s*s. */
int sa_bsm_scale(sa_bsm_t m, double s);
/* scale with a choice of cached-norm update: SA_NORM_RULE_REFERENCE is
 * sa_bsm_scale (the reference's norm *= s*s, blocksparsematrix.hpp:75-76);
 * SA_NORM_RULE_RECOMPUTE recomputes each present block's norm from its scaled
 * values in the same pass (bit-identical to sa_bsm_calc_norms afterwards), the
 * corrected rule |s| * norm up to rounding.  No counterpart in the reference. */
#define SA_NORM_RULE_REFERENCE 0
#define SA_NORM_RULE_RECOMPUTE 1
int sa_bsm_scale_ex(sa_bsm_t m, double s, int32_t norm_rule);
/* calc_frobenius_norms (blocksparsematrix.hpp:139-147). */
int sa_bsm_calc_norms(sa_bsm_t m);
/* recompress (blocksparsematrix.hpp:149-156). */
int sa_bsm_recompress(sa_bsm_t m);
/* operator+= (sign>0, :170-194) / operator-= (sign<0, :196-221). */
int sa_bsm_add(sa_bsm_t m, sa_bsm_t rhs, double sign);
/* add_transpose (blocksparsematrix.hpp:224-248). */
int sa_bsm_add_transpose(sa_bsm_t m, sa_bsm_t rhs);
/* frobenius_norm (blocksparsematrix.hpp:256-261). */
int sa_bsm_frobenius_norm(sa_bsm_t m, double* out);
/* dot(lhs, rhs, thresh) (blocksparsematrix.hpp:263-290). */
int sa_bsm_dot(sa_bsm_t a, sa_bsm_t b, double thresh, double* out);

/* ------------------------------------------------------------- multiply -- */
/* matmult(C, A, transA, B, transB, thresh, alpha, beta)
 * -- blocksparsematrix.h:95-99 / blocksparsematrix.hpp:292-414.
 * opts may be NULL (FP64).  opts->mode = SA_MODE_OZAKI evaluates the same
 * screened block products in split precision (opts->slices FP16 or BF16
 * slices per operand on the low-precision matrix cores, FP32 accumulation per
 * k-step, FP64 recombination). */
int sa_matmult(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh,
               double alpha, double beta, const sa_opts_t* opts);
/* Allocation-free matmult (FP64).  No counterpart in the reference: for
 * callers that cannot allocate device memory (code under HIP graph capture,
 * tight serving loops, harnesses that forbid hipMalloc).  Same semantics,
 * screening, task list and result as sa_matmult in FP64 mode with an
 * autotuned-or-heuristic tile and a single GEMM launch; it calls no hipMalloc,
 * hipFree or allocating library routine and records no profiling events.
 * Its tables (task counts and lists, GEMM records, work items, scan partials)
 * live in one static per-device scratch region of sa_matmult_scratch_bytes()
 * (64 MiB; held by every device the library runs on until the process ends);
 * a multiply whose tables exceed it fails with SA_ERR_INVALID before its
 * GEMM runs.  Output blocks that C gains must fit C's arena: reserve it
 * beforehand with sa_bsm_reserve (else SA_ERR_INVALID).  Every call on one
 * device shares that device's region, so calls that overlap on one device
 * (several host threads or streams) race; the region needs no clearing
 * between calls.  sa_tasklist_get reads the last call's list until the next
 * multiply on the device. */
int sa_matmult_noalloc(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh, double alpha,
                       double beta);
int sa_matmult_scratch_bytes(int64_t* bytes);
/* Extension: grow a matrix's block arena to hold `elements` doubles (contents
 * kept); lets a later sa_matmult_noalloc create blocks without allocating. */
int sa_bsm_reserve(sa_bsm_t m, int64_t elements);

/* Screened task list of the context's last sa_matmult / hybrid leaf set:
 * (ib, jb, kb) int32 triples in canonical order (C blocks column-major, kb
 * ascending).  *n receives the full count (the reference's nsig,
 * blocksparsematrix.hpp:353,402). */
int sa_tasklist_get(sa_ctx_t ctx, int32_t* triples, int64_t cap, int64_t* n);

/* ------------------------------------------------------- BLAS boundary -- */
/* any_gemm<double> -> dgemm_ (blaslapack.cpp:24-27, matrix.hpp:229-278):
 * C = alpha op(A) op(B) + beta C on host column-major arrays, computed on the
 * device (rocBLAS). */
int sa_dgemm(sa_ctx_t ctx, int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
             const double* A, int64_t lda, const double* B, int64_t ldb, double beta, double* C, int64_t ldc);
/* any_nrm2<double> -> dnrm2_ (blaslapack.cpp:34-37): the canonical norm order
 * of the block-norm kernel. */
int sa_dnrm2(sa_ctx_t ctx, int64_t n, const double* x, double* out);
/* any_dot<double> -> ddot_ (blaslapack.cpp:29-32). */
int sa_ddot(sa_ctx_t ctx, int64_t n, const double* x, const double* y, double* out);
/* any_axpy<double> -> daxpy_ (blaslapack.cpp:12-15): y += alpha x. */
int sa_daxpy(sa_ctx_t ctx, int64_t n, double alpha, const double* x, double* y);

/* ------------------------------------------------------------- RI-K step -- */
/* Extension (the reference reads M^{-1/2} as data, 3c_2c_transform.cpp:23-26):
 * X = M^{-1/2} of a symmetric positive definite n x n host matrix M by the
 * coupled Newton-Schulz iteration on device DGEMM, stopped when
 * ||Z Y - I||_F/sqrt(n) < tol.  Outputs (each optional): X as a host array,
 * X copied into the block-sparse X_bsm (its copy-in threshold applies),
 * iteration count and the residual ||X M X - I||_F/sqrt(n). */
int sa_inv_sqrt(sa_ctx_t ctx, int64_t n, const double* M, double* X, sa_bsm_t X_bsm, int32_t max_iter, double tol,
                int32_t* iters, double* residual);
/* RI-K step (tensormult form 2 then form 1, blocksparse3d_tensor.h:10-24):
 *   B' = X B           (screened; rows P, column blocks (mu-block, i-chunk)
 *                       with column block index ic + n_ic*mu and i fastest
 *                       inside a block),
 *   K  = sum_{P,i} B'(P,mu,i) B'(P,nu,i)  (screened B'^T B' over a reshape view
 *                       of B' whose rows are (P, i) pairs).
 * Bp receives B', K receives K (n_mu x n_mu blocks of mu_sizes).  Each
 * product runs in the mode of its options (FP64 or split precision);
 * screening, norms and accumulation are FP64. */
int sa_ri_k_step(sa_bsm_t X, sa_bsm_t B, sa_bsm_t Bp, sa_bsm_t K, int32_t n_mu, const int32_t* mu_sizes,
                 int32_t n_ic, const int32_t* i_chunks, double thresh, const sa_opts_t* opts_transform,
                 const sa_opts_t* opts_k);
/* Occupied RI-K exchange step, sparse path of occ_ri_k_step.cpp:83-131,
 * entirely on the device: K_{mu j} = sum_{P,i} (mu i|P)(ij|P).  The OpenMP
 * team of the reference becomes `teams` members (<= 0: 8), each with its own
 * stream and host thread: P is split as OpenMP's static schedule (contiguous
 * ranges, the first n_aux % teams members one more), each member runs its
 * range in batches of aux functions (stacked U_P, mui_P, ij_P, ijs_P: the same
 * screened block products as one P at a time, since the reference empties mui
 * and ij for every P; environment variable SA_OCC_BATCH caps the batch), and
 * the K_sub are added into K in member order (the reference adds them in a critical section in arrival
 * order; teams = 1 is its one-thread run).  ints3c is the compressed
 * 3-centre tensor (n_vec2 x n_aux, column-major: column P holds the pairs of
 * aux function P), keys[n] = r + n_bf c places pair n in the n_bf x n_bf
 * matrix (v2m_key format, occ_ri_k_step.cpp:133-146; a key named twice keeps
 * its later value), mo the n_bf x n_occ occupied coefficients, K the n_bf x
 * n_occ result (all host, column-major).  The integrals are uploaded once and
 * decompressed per P on the device straight into the block arena of U_P (no
 * dense n_bf x n_bf matrix), with U's copy-in drop rule at 1e-20 as in the
 * reference; then mui = U C (norms refreshed), ij = C^T mui, ijs = ij + ij^T
 * (norms refreshed where the transposed block exists, as add_transpose does),
 * Ks += mui ijs, each screened at thresh and run in the mode of opts (NULL:
 * FP64).  Errors: SA_ERR_INVALID for a key beyond n_bf^2 (the reference's
 * assert, occ_ri_k_step.cpp:139) or bad sizes. */
int sa_occ_ri_k(sa_ctx_t ctx, int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2, const double* ints3c,
                const double* mo, const uint64_t* keys, int32_t bs, double thresh, int32_t teams,
                const sa_opts_t* opts, double* K);

/* ------------------------------------------------------ hybrid Strassen -- */
/* matmult_strassen_dryrun(A, B, thresh_per_block) -- strassen.hpp:510-679.
 * Level bits concatenated (level 0 first, node index si + sj*ns + sk*ns^2),
 * level_size[l] = ns^3, level_flops[0..nlevels] = per-level flop totals
 * (index 0: sum over block triples; l+1: what strassen.hpp:676 prints).
 * Any output pointer may be NULL except nlevels. */
int sa_strassen_dryrun(sa_bsm_t A, sa_bsm_t B, double thresh_per_block, uint8_t* bits, int64_t cap,
                       int64_t* level_size, uint64_t* level_flops, int32_t* nlevels);
/* matmult_strassen_sparse(C, A, transA, B, transB, thresh, alpha, beta)
 * -- strassen.hpp:687-753.  Square, uniform, no transposes, alpha == 1, beta
 * ignored, C accumulated into (pre-filled by the caller; products into empty C
 * blocks are dropped as in the reference).  opts->cutoff is
 * min_size_for_strassen (main.cpp:9); opts->scheduler selects the decision
 * tree (dry-run heuristic, learned selector, or caller bits in opts->tree).
 * sa_tasklist_get then returns the screened 8-path leaf triples. */
int sa_matmult_strassen_sparse(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh,
                               double alpha, double beta, const sa_opts_t* opts);
/* Statistics of the context-independent last hybrid call: executed Strassen
 * roots, Winograd nodes, dense leaf block products, peak scratch bytes of a
 * batch, batches, screened 8-path leaf products, dense roots. */
int sa_strassen_last_stats(int64_t* nroots, int64_t* nnodes, int64_t* nleaf_products, int64_t* scratch_bytes,
                           int64_t* batches, int64_t* screened, int64_t* dense_roots);

/* ---------------------------------------------------- learned path selector --
 * No counterpart in the reference, whose recursion is steered by the fixed
 * cost model of matmult_strassen_dryrun (strassen.hpp:510-679).  A small MLP
 * (kMlpFeatures node features -> 32 -> 32 (tanh) -> 2) predicts, per dry-run
 * node, log2 of the cost ratio of the Strassen-Winograd and of the dense path
 * against the screened-sparse path; a bottom-up pass over the tree picks the
 * cheapest of descend / Strassen / dense per node (interior nodes only).  The
 * heuristic is the fallback (no model loaded) and the judge: a learned choice
 * is admissible only where the reference cost model prices it at no more than
 * 1.5x the node's screened flops (Strassen: its node cost; dense: the region's
 * dense flops), and a whole plan costing more than 4x the heuristic plan's
 * model flops is overruled.  The screening threshold is never touched. */
int sa_mlp_info(int32_t* n_features, int32_t* n_outputs, int64_t* n_params);
/* Install a model: n_params parameters, feature means and scales (x' = (x-mu)/sd). */
int sa_mlp_set(sa_ctx_t ctx, const double* params, int64_t n_params, const double* feat_mu, const double* feat_sd);
int sa_mlp_get(sa_ctx_t ctx, double* params, double* feat_mu, double* feat_sd);
/* Text model file: "path_mlp F 32 O P", P parameters, F means, F scales. */
int sa_mlp_load(sa_ctx_t ctx, const char* path);
/* Y (n x O, row-major) = model(X (n x F, row-major)) on the device. */
int sa_mlp_forward(sa_ctx_t ctx, const double* X, int64_t n, double* Y);
/* Gradient of L = 1/(2n) sum m (y - t)^2 w.r.t. the parameters (device
 * backward kernels); M (n x O) masks missing targets, NULL = all present. */
int sa_mlp_grad(sa_ctx_t ctx, const double* X, const double* T, const double* M, int64_t n, double* grad,
                double* loss);
/* epochs full-batch Adam steps on the device; loss_out = loss before the last step. */
int sa_mlp_train(sa_ctx_t ctx, const double* X, const double* T, const double* M, int64_t n, int32_t epochs,
                 double lr, double* loss_out);
/* Features of every dry-run node of A*B (nnodes x F, dry-run order); C may be
 * NULL (all C blocks present).  feat NULL: *nnodes only. */
int sa_strassen_features(sa_bsm_t C, sa_bsm_t A, sa_bsm_t B, double thresh, double* feat, int64_t cap,
                         int64_t* nnodes);
/* The learned plan (Strassen-root and dense-root bits, dry-run layout) and the
 * judge's verdict, without executing it. */
int sa_strassen_plan(sa_bsm_t C, sa_bsm_t A, sa_bsm_t B, double thresh, uint8_t* sbits, uint8_t* dbits, int64_t cap,
                     int32_t* overruled, double* model_cost, double* heuristic_cost);
/* Schedule used by the last hybrid call (SA_SCHED_*) and the judge's figures. */
int sa_strassen_last_plan(int32_t* scheduler, int32_t* overruled, double* model_cost, double* heuristic_cost);

#ifdef __cplusplus
}
#endif
#endif
