/*
 * oracle.h -- CPU restatement of the diersam/Strassen block-sparse multiply path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP path is
 * compared against.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg may load it; the product library never links or calls it.
 *
 * Every function cites the reference file:line it restates.  Pinning of each
 * function (what fixes it independently of this text) is listed in
 * tests/test_oracle_*.py; a function without such a check would say
 * "parity unpinned" here -- none currently does.
 *
 * Conventions (identical to the reference):
 *   - a block-sparse matrix is a grid of dense column-major blocks; the grid is
 *     itself column-major: block(rb, cb) = blocks[rb + cb * nrb]
 *     (blocksparsematrix.h:71-72);
 *   - an empty block has no storage (size()==0 in the reference);
 *   - every block carries a cached Frobenius norm with the reference's
 *     staleness rules (see or_* comments).
 *
 * Extension over the reference: block rows/columns may have uneven sizes.  Each
 * block row/column then has an actual size and a "screening size"; for a
 * uniform matrix the screening size is the target block size for every row
 * (also the ragged last one, as blocksparsematrix.hpp:349-350 uses the max
 * sizes), for an uneven matrix it equals the actual size.
 */
#ifndef STRASSEN_ORACLE_H
#define STRASSEN_ORACLE_H
#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct or_bsm or_bsm;

/* Canonical block norm: 64 strided partial sums s_l = fma-accumulated x[e]^2
 * over e = l, l+64, ..., then a halving tree s_l += s_{l+off} for off=32..1,
 * then IEEE sqrt.  The device kernel evaluates exactly this order.  The
 * reference calls MKL dnrm2 (matrix.hpp:189-193, :281-286). */
double or_block_norm(const double* x, int64_t n);

/* Construction (blocksparsematrix.hpp:6-16).  rsz/csz: actual sizes, rscr/cscr:
 * screening sizes (NULL = same as actual). */
or_bsm* or_bsm_create(int32_t nrb, const int32_t* rsz, const int32_t* rscr,
                      int32_t ncb, const int32_t* csz, const int32_t* cscr,
                      double thresh);
/* Uniform constructor with ragged last block (blocksparsematrix.hpp:6-16). */
or_bsm* or_bsm_create_uniform(int64_t nrow, int64_t ncol, int32_t bs_r, int32_t bs_c, double thresh);
or_bsm* or_bsm_clone(const or_bsm* m);           /* copy ctor: blocks + norms */
void    or_bsm_destroy(or_bsm* m);
int64_t or_bsm_nrow(const or_bsm* m);
int64_t or_bsm_ncol(const or_bsm* m);
int32_t or_bsm_nrb(const or_bsm* m);
int32_t or_bsm_ncb(const or_bsm* m);

/* Reshape view: copy of s with new block geometry (equal element count per
 * block, same grid index); NULL if the geometry does not fit. */
or_bsm* or_bsm_view_reshape(const or_bsm* s, int32_t nrb, const int32_t* rsz, const int32_t* rscr, int32_t ncb,
                            const int32_t* csz, const int32_t* cscr);

/* copy_from_input_array (blocksparsematrix.hpp:35-61): copy every block,
 * compute its norm, drop it if norm < thresh * scr_r * scr_c. */
int  or_bsm_from_dense(or_bsm* m, const double* dense, int64_t ld);
/* to_pointer (blocksparsematrix.hpp:103-124): writes present blocks only. */
int  or_bsm_to_dense(const or_bsm* m, double* dense, int64_t ld);
/* present[b] = 1/0, norms[b] = cached norm, for b = rb + cb*nrb. */
int  or_bsm_get_meta(const or_bsm* m, uint8_t* present, double* norms);
/* Upload one block directly (rows x cols col-major), norm recomputed. */
int  or_bsm_set_block(or_bsm* m, int32_t rb, int32_t cb, const double* data);

void or_bsm_fill(or_bsm* m, double val);              /* blocksparsematrix.hpp:84-101, matrix.hpp:26-33 */
void or_bsm_zero(or_bsm* m);                          /* :79-82 */
void or_bsm_scale(or_bsm* m, double s);               /* :71-77 (norm *= s*s, as the reference) */
void or_bsm_calc_norms(or_bsm* m);                    /* :139-147 */
void or_bsm_recompress(or_bsm* m);                    /* :149-156 */
int  or_bsm_add(or_bsm* m, const or_bsm* rhs, double sign); /* += (:170-194, no norm refresh) / -= (:196-221, refresh) */
int  or_bsm_add_transpose(or_bsm* m, const or_bsm* rhs);    /* :224-248 */
double or_bsm_frobenius_norm(const or_bsm* m);        /* :256-261 */
double or_bsm_dot(const or_bsm* a, const or_bsm* b, double thresh); /* :263-290 */
int64_t or_bsm_nalloc(const or_bsm* m);               /* no_of_alloc_blocks :158-162 */

/* Screened block GEMM (blocksparsematrix.hpp:292-414).
 * tasks (optional): receives up to cap triples (ib, jb, kb) in canonical order
 * (C blocks column-major, kb ascending); *ntasks gets the full count (nsig). */
int or_matmult(or_bsm* C, const or_bsm* A, int transA, const or_bsm* B, int transB,
               double thresh, double alpha, double beta,
               int32_t* tasks, int64_t cap, int64_t* ntasks);

/* Dry-run decision tree (strassen.hpp:510-679).  Writes levels' bits
 * concatenated into bits (cap entries), level_size[l] = n_super^3 at level l,
 * returns the number of levels (max_level) or <0 on error.  level_flops[l]
 * (optional, max_level+1 entries) receives the per-level flop totals the
 * reference prints at strassen.hpp:675-676 (index 0 = level 0 triple sum). */
int or_strassen_dryrun(const or_bsm* A, const or_bsm* B, double thresh_per_block,
                       uint8_t* bits, int64_t cap, int64_t* level_size,
                       uint64_t* level_flops);

/* Hybrid Strassen-sparse multiply (strassen.hpp:687-753) with the reference's
 * schedule; min_size_for_strassen is the global cutoff (main.cpp:9).
 * Divergences: empty blocks read as zero in the Strassen path; scratch is
 * sized exactly.  leaf_tasks (optional) receives the screened 8-path leaf
 * triples in canonical order; *nleaf gets their count.  nroots (optional)
 * gets the number of executed Strassen roots. */
int or_matmult_strassen_sparse(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                               int64_t min_size_for_strassen,
                               int32_t* leaf_tasks, int64_t cap, int64_t* nleaf,
                               int64_t* nroots);

/* As or_matmult_strassen_sparse, but with the decision tree given by the
 * caller (same layout as or_strassen_dryrun's bits; NULL = run the dry run).
 * Used to check Winograd schedules on trees the cost model would not pick. */
int or_matmult_strassen_tree(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                             int64_t min_size_for_strassen, const uint8_t* tree_bits,
                             int32_t* leaf_tasks, int64_t cap, int64_t* nleaf, int64_t* nroots);

/* As or_matmult_strassen_tree, plus dense roots: a node whose dense_bits
 * entry is set (and whose Strassen bit is not) adds every product of present
 * blocks inside its region, without screening, into the present C blocks of
 * the region (C blocks visited column-major, inner blocks ascending).  This is
 * the "dense" path of the learned path selector; it has no counterpart in the
 * reference, whose dense fallback (strassen.hpp:296-353) exists only inside
 * Strassen nodes.  ndense (optional) gets the number of dense roots. */
int or_matmult_strassen_plan(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                             int64_t min_size_for_strassen, const uint8_t* tree_bits,
                             const uint8_t* dense_bits, int32_t* leaf_tasks, int64_t cap,
                             int64_t* nleaf, int64_t* nroots, int64_t* ndense);

/* Dense column-major GEMM used by the oracle's block products:
 * C = alpha op(A) op(B) + beta C (beta==0: C not read). */
void or_dgemm(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
              const double* A, int64_t lda, const double* B, int64_t ldb,
              double beta, double* C, int64_t ldc);

/* decompress_integrals (occ_ri_k_step.cpp:133-146): out[key[n]] = comp[n + P*nvec2]. */
void or_decompress(double* out, int64_t P, const double* comp, int64_t nvec2, const uint64_t* keys);

/* Sparse occ-RI-K path of occ_ri_k_step.cpp:83-123 (one run): K (N_bf x N_occ
 * dense, column-major) = sum over P of the BSM chain with block size bs and
 * threshold thresh. */
int or_occ_ri_k(int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2,
                const double* ints3c, const double* mo, const uint64_t* keys,
                int32_t bs, double thresh, double* K);
/* Dense occ-RI-K path of occ_ri_k_step.cpp:38-72. */
int or_occ_ri_k_dense(int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2,
                      const double* ints3c, const double* mo, const uint64_t* keys, double* K);

#ifdef __cplusplus
}
#endif
#endif
