/*
 * oracle.c -- CPU restatement of the diersam/Strassen block-sparse multiply path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker).  Loaded by tests/, by
 * __graft_entry__.smoke() and by bench.py's cpu_baseline leg; never by the
 * product library.  Single-threaded, deterministic, written from the
 * behaviour of the reference files cited at each function.
 *
 * Built with -O2 -ffp-contract=off (oracle/Makefile) so that the order of
 * every floating-point operation is the one written here.
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

struct or_bsm {
  int32_t nrb, ncb;
  int32_t *rsz, *csz;   /* actual block sizes */
  int32_t *rscr, *cscr; /* screening sizes */
  int64_t nrow, ncol;
  double thresh;
  double** blk;  /* nrb*ncb pointers, NULL = empty block */
  double* norm;  /* cached norms (0 for empty blocks) */
};

#define BIDX(m, rb, cb) ((int64_t)(rb) + (int64_t)(cb) * (m)->nrb)

/* ---------------------------------------------------------------- norms -- */
double or_block_norm(const double* x, int64_t n) {
  double s[64];
  for (int l = 0; l < 64; ++l) {
    double a = 0.0;
    for (int64_t e = l; e < n; e += 64) a = fma(x[e], x[e], a);
    s[l] = a;
  }
  for (int off = 32; off >= 1; off >>= 1)
    for (int l = 0; l < off; ++l) s[l] = s[l] + s[l + off];
  return sqrt(s[0]);
}

/* ----------------------------------------------------------- containers -- */
static int32_t* dup32(const int32_t* src, int32_t n) {
  int32_t* d = (int32_t*)malloc(sizeof(int32_t) * (size_t)(n > 0 ? n : 1));
  if (n > 0) memcpy(d, src, sizeof(int32_t) * (size_t)n);
  return d;
}

or_bsm* or_bsm_create(int32_t nrb, const int32_t* rsz, const int32_t* rscr,
                      int32_t ncb, const int32_t* csz, const int32_t* cscr, double thresh) {
  or_bsm* m = (or_bsm*)calloc(1, sizeof(or_bsm));
  m->nrb = nrb;
  m->ncb = ncb;
  m->rsz = dup32(rsz, nrb);
  m->csz = dup32(csz, ncb);
  m->rscr = dup32(rscr ? rscr : rsz, nrb);
  m->cscr = dup32(cscr ? cscr : csz, ncb);
  m->nrow = 0;
  for (int32_t i = 0; i < nrb; ++i) m->nrow += rsz[i];
  m->ncol = 0;
  for (int32_t j = 0; j < ncb; ++j) m->ncol += csz[j];
  m->thresh = thresh;
  int64_t nb = (int64_t)nrb * ncb;
  m->blk = (double**)calloc((size_t)(nb > 0 ? nb : 1), sizeof(double*));
  m->norm = (double*)calloc((size_t)(nb > 0 ? nb : 1), sizeof(double));
  return m;
}

/* blocksparsematrix.hpp:6-16: nrb = ceil(nr/bs), ragged last block, the
 * screening size is the target size for every block row/column. */
or_bsm* or_bsm_create_uniform(int64_t nrow, int64_t ncol, int32_t bs_r, int32_t bs_c, double thresh) {
  int32_t nrb = (int32_t)(nrow == 0 ? 0 : 1 + (nrow - 1) / bs_r);
  int32_t ncb = (int32_t)(ncol == 0 ? 0 : 1 + (ncol - 1) / bs_c);
  int32_t* rsz = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nrb + 1));
  int32_t* rsc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(nrb + 1));
  int32_t* csz = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncb + 1));
  int32_t* csc = (int32_t*)malloc(sizeof(int32_t) * (size_t)(ncb + 1));
  for (int32_t i = 0; i < nrb; ++i) {
    int64_t end = (int64_t)(i + 1) * bs_r < nrow ? (int64_t)(i + 1) * bs_r : nrow;
    rsz[i] = (int32_t)(end - (int64_t)i * bs_r);
    rsc[i] = bs_r;
  }
  for (int32_t j = 0; j < ncb; ++j) {
    int64_t end = (int64_t)(j + 1) * bs_c < ncol ? (int64_t)(j + 1) * bs_c : ncol;
    csz[j] = (int32_t)(end - (int64_t)j * bs_c);
    csc[j] = bs_c;
  }
  or_bsm* m = or_bsm_create(nrb, rsz, rsc, ncb, csz, csc, thresh);
  free(rsz); free(rsc); free(csz); free(csc);
  return m;
}

static double* blk_alloc(const or_bsm* m, int32_t rb, int32_t cb) {
  size_t n = (size_t)m->rsz[rb] * (size_t)m->csz[cb];
  return (double*)malloc(sizeof(double) * (n > 0 ? n : 1));
}

static void blk_free(or_bsm* m, int64_t b) {
  free(m->blk[b]);
  m->blk[b] = NULL;
  m->norm[b] = 0.0;
}

or_bsm* or_bsm_clone(const or_bsm* s) {
  or_bsm* m = or_bsm_create(s->nrb, s->rsz, s->rscr, s->ncb, s->csz, s->cscr, s->thresh);
  for (int32_t cb = 0; cb < s->ncb; ++cb)
    for (int32_t rb = 0; rb < s->nrb; ++rb) {
      int64_t b = BIDX(s, rb, cb);
      if (s->blk[b]) {
        size_t n = (size_t)s->rsz[rb] * (size_t)s->csz[cb];
        m->blk[b] = blk_alloc(m, rb, cb);
        memcpy(m->blk[b], s->blk[b], n * sizeof(double));
      }
      m->norm[b] = s->norm[b];
    }
  return m;
}

void or_bsm_destroy(or_bsm* m) {
  if (!m) return;
  int64_t nb = (int64_t)m->nrb * m->ncb;
  for (int64_t b = 0; b < nb; ++b) free(m->blk[b]);
  free(m->blk); free(m->norm);
  free(m->rsz); free(m->csz); free(m->rscr); free(m->cscr);
  free(m);
}

int64_t or_bsm_nrow(const or_bsm* m) { return m->nrow; }
int64_t or_bsm_ncol(const or_bsm* m) { return m->ncol; }
int32_t or_bsm_nrb(const or_bsm* m) { return m->nrb; }
int32_t or_bsm_ncb(const or_bsm* m) { return m->ncb; }

/* Reshape view (the device library's sa_bsm_view_reshape): same block
 * payloads, norms and grid index, new block geometry.  Returns a copy. */
or_bsm* or_bsm_view_reshape(const or_bsm* s, int32_t nrb, const int32_t* rsz, const int32_t* rscr, int32_t ncb,
                            const int32_t* csz, const int32_t* cscr) {
  if ((int64_t)nrb * ncb != (int64_t)s->nrb * s->ncb) return NULL;
  or_bsm* m = or_bsm_create(nrb, rsz, rscr, ncb, csz, cscr, s->thresh);
  for (int64_t b = 0; b < (int64_t)nrb * ncb; ++b) {
    const int64_t n = (int64_t)rsz[b % nrb] * csz[b / nrb];
    if (n != (int64_t)s->rsz[b % s->nrb] * s->csz[b / s->nrb]) {
      or_bsm_destroy(m);
      return NULL;
    }
    if (s->blk[b]) {
      m->blk[b] = (double*)malloc(sizeof(double) * (size_t)n);
      memcpy(m->blk[b], s->blk[b], sizeof(double) * (size_t)n);
    }
    m->norm[b] = s->norm[b];
  }
  return m;
}

/* blocksparsematrix.hpp:35-61 */
int or_bsm_from_dense(or_bsm* m, const double* dense, int64_t ld) {
  if (ld < m->nrow) return -1;
  int64_t c0 = 0;
  for (int32_t cb = 0; cb < m->ncb; ++cb) {
    int64_t r0 = 0;
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      free(m->blk[b]);
      int32_t nr = m->rsz[rb], nc = m->csz[cb];
      double* d = blk_alloc(m, rb, cb);
      for (int32_t c = 0; c < nc; ++c)
        memcpy(d + (int64_t)c * nr, dense + r0 + (c0 + c) * ld, sizeof(double) * (size_t)nr);
      double est = or_block_norm(d, (int64_t)nr * nc);
      double thr = m->thresh * (double)((int64_t)m->cscr[cb] * (int64_t)m->rscr[rb]);
      if (est < thr) {
        free(d);
        m->blk[b] = NULL;
        m->norm[b] = 0.0;
      } else {
        m->blk[b] = d;
        m->norm[b] = est;
      }
      r0 += nr;
    }
    c0 += m->csz[cb];
  }
  return 0;
}

/* blocksparsematrix.hpp:103-124: only present blocks are written. */
int or_bsm_to_dense(const or_bsm* m, double* dense, int64_t ld) {
  if (ld < m->nrow) return -1;
  int64_t c0 = 0;
  for (int32_t cb = 0; cb < m->ncb; ++cb) {
    int64_t r0 = 0;
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      const double* d = m->blk[BIDX(m, rb, cb)];
      int32_t nr = m->rsz[rb], nc = m->csz[cb];
      if (d)
        for (int32_t c = 0; c < nc; ++c)
          memcpy(dense + r0 + (c0 + c) * ld, d + (int64_t)c * nr, sizeof(double) * (size_t)nr);
      r0 += nr;
    }
    c0 += m->csz[cb];
  }
  return 0;
}

int or_bsm_get_meta(const or_bsm* m, uint8_t* present, double* norms) {
  int64_t nb = (int64_t)m->nrb * m->ncb;
  for (int64_t b = 0; b < nb; ++b) {
    if (present) present[b] = m->blk[b] ? 1 : 0;
    if (norms) norms[b] = m->norm[b];
  }
  return 0;
}

int or_bsm_set_block(or_bsm* m, int32_t rb, int32_t cb, const double* data) {
  if (rb < 0 || rb >= m->nrb || cb < 0 || cb >= m->ncb) return -1;
  int64_t b = BIDX(m, rb, cb);
  if (!m->blk[b]) m->blk[b] = blk_alloc(m, rb, cb);
  int64_t n = (int64_t)m->rsz[rb] * m->csz[cb];
  memcpy(m->blk[b], data, sizeof(double) * (size_t)n);
  m->norm[b] = or_block_norm(m->blk[b], n);
  return 0;
}

/* blocksparsematrix.hpp:84-101 with Matrix(val, nr, nc) at matrix.hpp:26-33:
 * every block allocated, cached norm sqrt(nr*nc)*val. */
void or_bsm_fill(or_bsm* m, double val) {
  for (int32_t cb = 0; cb < m->ncb; ++cb)
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      int64_t n = (int64_t)m->rsz[rb] * m->csz[cb];
      if (!m->blk[b]) m->blk[b] = blk_alloc(m, rb, cb);
      for (int64_t e = 0; e < n; ++e) m->blk[b][e] = val;
      m->norm[b] = sqrt((double)n) * val;
    }
}

/* blocksparsematrix.hpp:79-82 */
void or_bsm_zero(or_bsm* m) {
  int64_t nb = (int64_t)m->nrb * m->ncb;
  for (int64_t b = 0; b < nb; ++b) blk_free(m, b);
}

/* blocksparsematrix.hpp:71-77: data *= s, cached norm *= s*s (as written). */
void or_bsm_scale(or_bsm* m, double s) {
  for (int32_t cb = 0; cb < m->ncb; ++cb)
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      if (m->blk[b]) {
        int64_t n = (int64_t)m->rsz[rb] * m->csz[cb];
        for (int64_t e = 0; e < n; ++e) m->blk[b][e] = m->blk[b][e] * s;
      }
      m->norm[b] = m->norm[b] * (s * s);
    }
}

/* blocksparsematrix.hpp:139-147 (empty blocks: norm of a 0-element block = 0) */
void or_bsm_calc_norms(or_bsm* m) {
  for (int32_t cb = 0; cb < m->ncb; ++cb)
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      m->norm[b] = m->blk[b] ? or_block_norm(m->blk[b], (int64_t)m->rsz[rb] * m->csz[cb]) : 0.0;
    }
}

/* blocksparsematrix.hpp:149-156 */
void or_bsm_recompress(or_bsm* m) {
  for (int32_t cb = 0; cb < m->ncb; ++cb)
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      double thr = m->thresh * (double)((int64_t)m->cscr[cb] * (int64_t)m->rscr[rb]);
      if (m->norm[b] < thr) blk_free(m, b);
    }
}

static int same_geometry(const or_bsm* a, const or_bsm* b) {
  if (a->nrb != b->nrb || a->ncb != b->ncb || a->nrow != b->nrow || a->ncol != b->ncol) return 0;
  for (int32_t i = 0; i < a->nrb; ++i) if (a->rsz[i] != b->rsz[i]) return 0;
  for (int32_t j = 0; j < a->ncb; ++j) if (a->csz[j] != b->csz[j]) return 0;
  return 1;
}

/* += : blocksparsematrix.hpp:170-194 (copy rhs block incl. its cached norm if
 * this block is empty, else axpy; norm not refreshed).
 * -= : :196-221 (copy or axpy(-1), then norm refreshed). */
int or_bsm_add(or_bsm* m, const or_bsm* rhs, double sign) {
  if (!same_geometry(m, rhs)) return -1;
  for (int32_t cb = 0; cb < m->ncb; ++cb)
    for (int32_t rb = 0; rb < m->nrb; ++rb) {
      int64_t b = BIDX(m, rb, cb);
      if (!rhs->blk[b]) continue;
      int64_t n = (int64_t)m->rsz[rb] * m->csz[cb];
      if (!m->blk[b]) {
        m->blk[b] = blk_alloc(m, rb, cb);
        memcpy(m->blk[b], rhs->blk[b], sizeof(double) * (size_t)n);
        m->norm[b] = rhs->norm[b];
      } else if (sign > 0) {
        for (int64_t e = 0; e < n; ++e) m->blk[b][e] = m->blk[b][e] + rhs->blk[b][e];
      } else {
        for (int64_t e = 0; e < n; ++e) m->blk[b][e] = m->blk[b][e] - rhs->blk[b][e];
      }
      if (sign < 0) m->norm[b] = or_block_norm(m->blk[b], n);
    }
  return 0;
}

/* blocksparsematrix.hpp:224-248 with Matrix::add_transpose matrix.hpp:179-187 */
int or_bsm_add_transpose(or_bsm* m, const or_bsm* rhs) {
  if (m->nrb != rhs->ncb || m->ncb != rhs->nrb || m->nrow != rhs->ncol || m->ncol != rhs->nrow) return -1;
  for (int32_t ib = 0; ib < m->nrb; ++ib)
    for (int32_t jb = 0; jb < m->ncb; ++jb) {
      const double* r = rhs->blk[BIDX(rhs, jb, ib)];
      if (!r) continue;
      int64_t b = BIDX(m, ib, jb);
      int32_t nr = m->rsz[ib], nc = m->csz[jb];
      if (!m->blk[b]) {
        m->blk[b] = blk_alloc(m, ib, jb);
        memset(m->blk[b], 0, sizeof(double) * (size_t)nr * (size_t)nc);
      }
      /* rhs block is nc x nr */
      for (int32_t c = 0; c < nc; ++c)
        for (int32_t rr = 0; rr < nr; ++rr)
          m->blk[b][rr + (int64_t)c * nr] = m->blk[b][rr + (int64_t)c * nr] + r[c + (int64_t)rr * nc];
      m->norm[b] = or_block_norm(m->blk[b], (int64_t)nr * nc);
    }
  return 0;
}

/* blocksparsematrix.hpp:256-261: sqrt of the sum over blocks (grid order) of
 * cached norm squared. */
double or_bsm_frobenius_norm(const or_bsm* m) {
  double acc = 0.0;
  int64_t nb = (int64_t)m->nrb * m->ncb;
  for (int64_t b = 0; b < nb; ++b) acc = m->norm[b] * m->norm[b] + acc;
  return sqrt(acc);
}

/* blocksparsematrix.hpp:263-290 (serial order over blocks) */
double or_bsm_dot(const or_bsm* a, const or_bsm* b, double thresh) {
  double acc = 0.0;
  for (int32_t cb = 0; cb < a->ncb; ++cb)
    for (int32_t rb = 0; rb < a->nrb; ++rb) {
      int64_t i = BIDX(a, rb, cb);
      if (!a->blk[i] || !b->blk[i]) continue;
      double thr = thresh * (double)((int64_t)a->rscr[rb] * (int64_t)a->cscr[cb]);
      if (a->norm[i] * b->norm[i] >= thr) {
        int64_t n = (int64_t)a->rsz[rb] * a->csz[cb];
        double d = 0.0;
        for (int64_t e = 0; e < n; ++e) d += a->blk[i][e] * b->blk[i][e];
        acc += d;
      }
    }
  return acc;
}

int64_t or_bsm_nalloc(const or_bsm* m) {
  int64_t nb = (int64_t)m->nrb * m->ncb, c = 0;
  for (int64_t b = 0; b < nb; ++b) c += m->blk[b] ? 1 : 0;
  return c;
}

/* ----------------------------------------------------------------- gemm -- */
void or_dgemm(int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha,
              const double* A, int64_t lda, const double* B, int64_t ldb,
              double beta, double* C, int64_t ldc) {
  double* acc = (double*)malloc(sizeof(double) * (size_t)(m > 0 ? m : 1));
  for (int64_t j = 0; j < n; ++j) {
    for (int64_t i = 0; i < m; ++i) acc[i] = 0.0;
    if (!transA) {
      for (int64_t l = 0; l < k; ++l) {
        double b = transB ? B[j + l * ldb] : B[l + j * ldb];
        const double* a = A + l * lda;
        for (int64_t i = 0; i < m; ++i) acc[i] += a[i] * b;
      }
    } else {
      for (int64_t i = 0; i < m; ++i) {
        const double* a = A + i * lda;
        double s = 0.0;
        for (int64_t l = 0; l < k; ++l) s += a[l] * (transB ? B[j + l * ldb] : B[l + j * ldb]);
        acc[i] = s;
      }
    }
    double* c = C + j * ldc;
    if (beta == 0.0)
      for (int64_t i = 0; i < m; ++i) c[i] = alpha * acc[i];
    else
      for (int64_t i = 0; i < m; ++i) c[i] = alpha * acc[i] + beta * c[i];
  }
  free(acc);
}

/* ------------------------------------------------------- screened matmult -- */
/* blocksparsematrix.hpp:292-414 */
int or_matmult(or_bsm* C, const or_bsm* A, int transA, const or_bsm* B, int transB,
               double thresh, double alpha, double beta,
               int32_t* tasks, int64_t cap, int64_t* ntasks) {
  /* geometry of op(A) (nib x nkb) and op(B) (nkb x njb) */
  const int32_t nib = transA ? A->ncb : A->nrb;
  const int32_t nkb = transA ? A->nrb : A->ncb;
  const int32_t nkb2 = transB ? B->ncb : B->nrb;
  const int32_t njb = transB ? B->nrb : B->ncb;
  const int32_t* a_isz = transA ? A->csz : A->rsz;
  const int32_t* a_iscr = transA ? A->cscr : A->rscr;
  const int32_t* a_ksz = transA ? A->rsz : A->csz;
  const int32_t* a_kscr = transA ? A->rscr : A->cscr;
  const int32_t* b_ksz = transB ? B->csz : B->rsz;
  const int32_t* b_jsz = transB ? B->rsz : B->csz;
  const int32_t* b_jscr = transB ? B->rscr : B->cscr;
  if (nkb != nkb2 || nib != C->nrb || njb != C->ncb) return -2;
  for (int32_t i = 0; i < nib; ++i) if (a_isz[i] != C->rsz[i] || a_iscr[i] != C->rscr[i]) return -2;
  for (int32_t j = 0; j < njb; ++j) if (b_jsz[j] != C->csz[j] || b_jscr[j] != C->cscr[j]) return -2;
  for (int32_t k = 0; k < nkb; ++k) if (a_ksz[k] != b_ksz[k]) return -2;

  /* beta pre-pass (:299-305) */
  if (beta == 1.0) {
  } else if (beta == 0.0) {
    or_bsm_zero(C);
  } else {
    or_bsm_scale(C, beta);
  }

  int64_t nsig = 0;
  for (int32_t jb = 0; jb < njb; ++jb)
    for (int32_t ib = 0; ib < nib; ++ib)
      for (int32_t kb = 0; kb < nkb; ++kb) {
        int64_t ai = transA ? BIDX(A, kb, ib) : BIDX(A, ib, kb);
        int64_t bi = transB ? BIDX(B, jb, kb) : BIDX(B, kb, jb);
        const double* a = A->blk[ai];
        const double* b = B->blk[bi];
        if (!a || !b) continue;
        const double est = A->norm[ai] * B->norm[bi];
        const double thr = thresh * (double)((int64_t)a_iscr[ib] * (int64_t)b_jscr[jb] * (int64_t)a_kscr[kb]);
        if (!(est >= thr)) continue;
        int64_t ci = BIDX(C, ib, jb);
        double nb_ = 1.0;
        if (!C->blk[ci]) {
          C->blk[ci] = blk_alloc(C, ib, jb);
          C->norm[ci] = 0.0;
          nb_ = 0.0;
        }
        const int64_t m = C->rsz[ib], n = C->csz[jb], k = a_ksz[kb];
        or_dgemm(transA, transB, m, n, k, alpha, a, transA ? k : m, b, transB ? n : k, nb_, C->blk[ci], m);
        if (tasks && nsig < cap) {
          tasks[3 * nsig + 0] = ib;
          tasks[3 * nsig + 1] = jb;
          tasks[3 * nsig + 2] = kb;
        }
        ++nsig;
      }
  if (ntasks) *ntasks = nsig;
  return 0;
}

/* ------------------------------------------------------------ dry run -- */
static int64_t ceil_div(int64_t a, int64_t b) { return a == 0 ? 0 : 1 + (a - 1) / b; }
#define IJK(i, j, k, di, dj) ((int64_t)(i) + (int64_t)(j) * (di) + (int64_t)(k) * (di) * (dj))

static const double k_strassen_fac[20] = {10, 10, 10, 25, 35, 35, 35, 35, 35, 35,
                                          35, 35, 35, 35, 35, 35, 35, 35, 35, 35};

static int uniform_square(const or_bsm* A, const or_bsm* B) {
  if (A->nrb != A->ncb || B->nrb != B->ncb || A->nrb != B->nrb) return 0;
  if (A->nrow != A->ncol || B->nrow != B->ncol || A->nrow != B->nrow) return 0;
  if (A->nrb == 0) return 0;
  int32_t bs = A->rscr[0];
  for (int32_t i = 0; i < A->nrb; ++i)
    if (A->rscr[i] != bs || A->cscr[i] != bs || B->rscr[i] != bs || B->cscr[i] != bs) return 0;
  return 1;
}

/* strassen.hpp:510-679 */
int or_strassen_dryrun(const or_bsm* A, const or_bsm* B, double thresh_per_block,
                       uint8_t* bits, int64_t cap, int64_t* level_size, uint64_t* level_flops) {
  if (!uniform_square(A, B)) return -2;
  const int64_t nib = A->nrb, njb = B->ncb, nkb = A->ncb;
  const int64_t bs = A->rscr[0];
  const int64_t ni = A->nrow, nj = B->ncol, nk = A->ncol;
  const int max_level = (int)(size_t)log2((double)nib + 0.5);

  uint64_t* prev = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nib * njb * nkb));
  uint64_t tot0 = 0;
  for (int64_t kb = 0; kb < nkb; ++kb)
    for (int64_t jb = 0; jb < njb; ++jb)
      for (int64_t ib = 0; ib < nib; ++ib) {
        int64_t ai = BIDX(A, ib, kb), bi = BIDX(B, kb, jb);
        const double est = A->norm[ai] * B->norm[bi];
        uint64_t mm = A->blk[ai] ? (uint64_t)A->rsz[ib] : 0;
        uint64_t kk = A->blk[ai] ? (uint64_t)A->csz[kb] : 0;
        uint64_t nn = B->blk[bi] ? (uint64_t)B->csz[jb] : 0;
        uint64_t f = 2u * mm * nn * kk;
        prev[IJK(ib, jb, kb, nib, njb)] = (est >= thresh_per_block ? 1u : 0u) * f;
        tot0 += prev[IJK(ib, jb, kb, nib, njb)];
      }
  if (level_flops) level_flops[0] = tot0;

  const uint64_t max_mult = 2u * (uint64_t)bs * (uint64_t)bs * (uint64_t)bs;
  const uint64_t max_add = (uint64_t)bs * (uint64_t)bs;
  int64_t written = 0;
  for (int level = 0; level < max_level; ++level) {
    const int64_t prev_sb = (int64_t)exp2((double)level);
    const int64_t sb = (int64_t)exp2((double)(level + 1));
    const uint64_t cost_mult = (uint64_t)round(pow(7.0, (double)(level + 1))) * max_mult;
    const uint64_t cost_add = (uint64_t)(15e0 * k_strassen_fac[level] * pow(4.0, (double)(level + 1)) * (double)max_add);
    const uint64_t cost_s = cost_mult + cost_add;
    const int64_t nsi = ceil_div(nib, sb), nsj = ceil_div(njb, sb), nsk = ceil_div(nkb, sb);
    const int64_t pnsi = ceil_div(nib, prev_sb), pnsj = ceil_div(njb, prev_sb);
    const int64_t nnodes = nsi * nsj * nsk;
    uint64_t* cur = (uint64_t*)malloc(sizeof(uint64_t) * (size_t)(nnodes > 0 ? nnodes : 1));
    if (level_size) level_size[level] = nnodes;
    uint64_t tot = 0;
    for (int64_t s = 0; s < nnodes; ++s) {
      const int64_t si = s % nsi, sj = (s / nsi) % nsj, sk = s / (nsi * nsj);
      const int64_t ib0 = si * sb, jb0 = sj * sb, kb0 = sk * sb;
      const int64_t ib1 = ib0 + sb < nib ? ib0 + sb : nib;
      const int64_t jb1 = jb0 + sb < njb ? jb0 + sb : njb;
      const int64_t kb1 = kb0 + sb < nkb ? kb0 + sb : nkb;
      const int64_t i_end = ib1 * bs < ni ? ib1 * bs : ni;
      const int64_t j_end = jb1 * bs < nj ? jb1 * bs : nj;
      const int64_t k_end = kb1 * bs < nk ? kb1 * bs : nk;
      const int edge = (i_end - ib0 * bs != bs * sb) || (j_end - jb0 * bs != bs * sb) || (k_end - kb0 * bs != bs * sb);
      const int64_t isub = ceil_div(ib1 - ib0, prev_sb), jsub = ceil_div(jb1 - jb0, prev_sb), ksub = ceil_div(kb1 - kb0, prev_sb);
      uint64_t f = 0;
      for (int64_t a = 0; a < isub; ++a)
        for (int64_t b = 0; b < jsub; ++b)
          for (int64_t c = 0; c < ksub; ++c) f += prev[IJK(2 * si + a, 2 * sj + b, 2 * sk + c, pnsi, pnsj)];
      const int do_s = (f >= cost_s) && !edge;
      if (bits) {
        if (written + s >= cap) { free(cur); free(prev); return -3; }
        bits[written + s] = (uint8_t)do_s;
      }
      cur[s] = do_s ? cost_s : f;
      tot += cur[s];
    }
    if (level_flops) level_flops[level + 1] = tot;
    written += nnodes;
    free(prev);
    prev = cur;
  }
  free(prev);
  return max_level;
}

/* -------------------------------------------------------------- Strassen -- */
/* grid of bs x bs blocks; p[rb + cb*stride], NULL reads as a zero block. */
typedef struct {
  double** p;
  int64_t nrb, ncb, stride;
} grid_t;

typedef struct {
  double* mem;
  double** ptrs;
  int64_t nrb, ncb;
} owned_t;

static owned_t owned_new(int64_t nrb, int64_t ncb, int64_t bs2, int zero) {
  owned_t o;
  o.nrb = nrb;
  o.ncb = ncb;
  o.mem = (double*)malloc(sizeof(double) * (size_t)(nrb * ncb * bs2 > 0 ? nrb * ncb * bs2 : 1));
  if (zero) memset(o.mem, 0, sizeof(double) * (size_t)(nrb * ncb * bs2));
  o.ptrs = (double**)malloc(sizeof(double*) * (size_t)(nrb * ncb > 0 ? nrb * ncb : 1));
  for (int64_t i = 0; i < nrb * ncb; ++i) o.ptrs[i] = o.mem + i * bs2;
  return o;
}
static void owned_free(owned_t* o) { free(o->mem); free(o->ptrs); }
static grid_t owned_view(const owned_t* o) { grid_t g = {o->ptrs, o->nrb, o->ncb, o->nrb}; return g; }
static grid_t quad(grid_t g, int64_t rb0, int64_t cb0, int64_t hr, int64_t hc) {
  grid_t q = {g.p + rb0 + cb0 * g.stride, hr, hc, g.stride};
  return q;
}
#define GB(g, rb, cb) ((g).p[(rb) + (cb) * (g).stride])

typedef struct {
  int64_t bs;
  int64_t cutoff;
} sctx_t;

/* C += A B for single blocks (gemm_block with beta=1, strassen.hpp:42-50) */
static void blk_gemm_acc(double* c, const double* a, const double* b, int64_t bs) {
  if (!a || !b || !c) return;
  or_dgemm(0, 0, bs, bs, bs, 1.0, a, bs, b, bs, 1.0, c, bs);
}

static void s_recurse(owned_t* C00, owned_t* C01, owned_t* C10, owned_t* C11,
                      grid_t a, grid_t b, grid_t c, grid_t d, grid_t A, grid_t Cq, grid_t B, grid_t D,
                      const sctx_t* sc);

/* strassen.hpp:296-353 */
static void s_matmult_acc(owned_t* C, grid_t A, grid_t B, const sctx_t* sc) {
  const int64_t nrb = C->nrb, nkb = A.ncb, ncb = C->ncb, bs = sc->bs, bs2 = bs * bs;
  grid_t Cg = owned_view(C);
  if (nrb == 1 && nkb == 1 && ncb == 1) {
    blk_gemm_acc(GB(Cg, 0, 0), GB(A, 0, 0), GB(B, 0, 0), bs);
    return;
  }
  const int too_small = nrb * bs <= sc->cutoff;
  if (nrb % 2 || nkb % 2 || ncb % 2 || too_small) {
    for (int64_t cb = 0; cb < ncb; ++cb)
      for (int64_t rb = 0; rb < nrb; ++rb)
        for (int64_t kb = 0; kb < nkb; ++kb) blk_gemm_acc(GB(Cg, rb, cb), GB(A, rb, kb), GB(B, kb, cb), bs);
    return;
  }
  const int64_t hr = nrb / 2, hk = nkb / 2, hc = ncb / 2;
  owned_t S00 = owned_new(hr, hc, bs2, 1), S01 = owned_new(hr, hc, bs2, 1);
  owned_t S10 = owned_new(hr, hc, bs2, 1), S11 = owned_new(hr, hc, bs2, 1);
  s_recurse(&S00, &S01, &S10, &S11, quad(A, 0, 0, hr, hk), quad(A, 0, hk, hr, hk), quad(A, hr, 0, hr, hk),
            quad(A, hr, hk, hr, hk), quad(B, 0, 0, hk, hc), quad(B, 0, hc, hk, hc), quad(B, hk, 0, hk, hc),
            quad(B, hk, hc, hk, hc), sc);
  for (int64_t cb = 0; cb < hc; ++cb)
    for (int64_t rb = 0; rb < hr; ++rb) {
      double* t[4] = {GB(Cg, rb, cb), GB(Cg, rb, cb + hc), GB(Cg, rb + hr, cb), GB(Cg, rb + hr, cb + hc)};
      const double* s[4] = {S00.ptrs[rb + cb * hr], S01.ptrs[rb + cb * hr], S10.ptrs[rb + cb * hr], S11.ptrs[rb + cb * hr]};
      for (int q = 0; q < 4; ++q)
        for (int64_t e = 0; e < bs2; ++e) t[q][e] = t[q][e] + s[q][e];
    }
  owned_free(&S00); owned_free(&S01); owned_free(&S10); owned_free(&S11);
}

static inline double rd(const double* p, int64_t e) { return p ? p[e] : 0.0; }

/* strassen.hpp:359-405 with the fused adds of :153-237 */
static void s_recurse(owned_t* C00, owned_t* C01, owned_t* C10, owned_t* C11,
                      grid_t a, grid_t b, grid_t c, grid_t d, grid_t A, grid_t Cq, grid_t B, grid_t D,
                      const sctx_t* sc) {
  const int64_t nrb = C00->nrb, nkb = a.ncb, ncb = C00->ncb, bs2 = sc->bs * sc->bs;
  owned_t na_pc = owned_new(nrb, nkb, bs2, 0), pc_pd = owned_new(nrb, nkb, bs2, 0);
  owned_t na_pc_pd = owned_new(nrb, nkb, bs2, 0), pa_pb_nc_nd = owned_new(nrb, nkb, bs2, 0);
  for (int64_t cb = 0; cb < nkb; ++cb)
    for (int64_t rb = 0; rb < nrb; ++rb) {
      const int64_t i = rb + cb * nrb;
      const double *ap = GB(a, rb, cb), *bp = GB(b, rb, cb), *cp = GB(c, rb, cb), *dp = GB(d, rb, cb);
      for (int64_t e = 0; e < bs2; ++e) {
        na_pc.ptrs[i][e] = rd(cp, e) - rd(ap, e);
        const double v = rd(cp, e) + rd(dp, e);
        pc_pd.ptrs[i][e] = v;
        const double w = v - rd(ap, e);
        na_pc_pd.ptrs[i][e] = w;
        pa_pb_nc_nd.ptrs[i][e] = rd(bp, e) - w;
      }
    }
  owned_t pC_nD = owned_new(nkb, ncb, bs2, 0), nA_pC = owned_new(nkb, ncb, bs2, 0);
  owned_t pA_nC_pD = owned_new(nkb, ncb, bs2, 0), nA_pB_pC_nD = owned_new(nkb, ncb, bs2, 0);
  for (int64_t cb = 0; cb < ncb; ++cb)
    for (int64_t rb = 0; rb < nkb; ++rb) {
      const int64_t i = rb + cb * nkb;
      const double *Ap = GB(A, rb, cb), *Bp = GB(B, rb, cb), *Cp = GB(Cq, rb, cb), *Dp = GB(D, rb, cb);
      for (int64_t e = 0; e < bs2; ++e) {
        pC_nD.ptrs[i][e] = rd(Cp, e) - rd(Dp, e);
        const double v = rd(Cp, e) - rd(Ap, e);
        nA_pC.ptrs[i][e] = v;
        const double w = rd(Dp, e) - v;
        pA_nC_pD.ptrs[i][e] = w;
        nA_pB_pC_nD.ptrs[i][e] = rd(Bp, e) - w;
      }
    }
  s_matmult_acc(C00, a, A, sc);                                             /* M1 */
  s_matmult_acc(C10, owned_view(&na_pc), owned_view(&pC_nD), sc);          /* M2 */
  s_matmult_acc(C01, owned_view(&pc_pd), owned_view(&nA_pC), sc);          /* M3 */
  s_matmult_acc(C11, owned_view(&na_pc_pd), owned_view(&pA_nC_pD), sc);    /* M4 */
  for (int64_t i = 0; i < C11->nrb * C11->ncb; ++i) {                       /* :215-237 */
    double *nw = C11->ptrs[i], *c00 = C00->ptrs[i], *c01 = C01->ptrs[i], *c10 = C10->ptrs[i];
    for (int64_t e = 0; e < bs2; ++e) {
      const double t = nw[e] + c00[e];
      c10[e] += t;
      nw[e] = c01[e] + c10[e];
      c01[e] += t;
    }
  }
  s_matmult_acc(C01, owned_view(&pa_pb_nc_nd), D, sc);                     /* M5 */
  s_matmult_acc(C10, d, owned_view(&nA_pB_pC_nD), sc);                     /* M6 */
  s_matmult_acc(C00, b, B, sc);                                             /* M7 */
  owned_free(&na_pc); owned_free(&pc_pd); owned_free(&na_pc_pd); owned_free(&pa_pb_nc_nd);
  owned_free(&pC_nD); owned_free(&nA_pC); owned_free(&pA_nC_pD); owned_free(&nA_pB_pC_nD);
}

typedef struct {
  or_bsm* C;
  const or_bsm *A, *B;
  const uint8_t* bits;
  const uint8_t* dense;
  int64_t ndense;
  const int64_t* lvl_off;
  double thr;
  sctx_t sc;
  int32_t* leaf;
  int64_t cap, nleaf, nroots;
  grid_t Ag, Bg, Cg; /* whole-matrix grids over the BSM block pointers */
} hctx_t;

/* strassen.hpp:418-508 */
static void h_recurse(hctx_t* h, int64_t ib0, int64_t jb0, int64_t kb0, int64_t sb, int level) {
  const int64_t nib = h->A->nrb, njb = h->B->ncb, nkb = h->A->ncb;
  if (level < 0) {
    if (ib0 >= nib || jb0 >= njb || kb0 >= nkb) return;
    const int64_t ai = BIDX(h->A, ib0, kb0), bi = BIDX(h->B, kb0, jb0);
    const double *ab = h->A->blk[ai], *bb = h->B->blk[bi];
    if (!ab || !bb) return;
    if (h->A->norm[ai] * h->B->norm[bi] < h->thr) return;
    double* cb = h->C->blk[BIDX(h->C, ib0, jb0)];
    if (!cb) return; /* empty C block: the reference's dgemm runs with m=n=0 */
    const int64_t m = h->C->rsz[ib0], n = h->C->csz[jb0], k = h->A->csz[kb0];
    or_dgemm(0, 0, m, n, k, 1.0, ab, m, bb, k, 1.0, cb, m);
    if (h->leaf && h->nleaf < h->cap) {
      h->leaf[3 * h->nleaf + 0] = (int32_t)ib0;
      h->leaf[3 * h->nleaf + 1] = (int32_t)jb0;
      h->leaf[3 * h->nleaf + 2] = (int32_t)kb0;
    }
    h->nleaf++;
    return;
  }
  const int64_t nsi = ceil_div(nib, sb), nsj = ceil_div(njb, sb);
  const int64_t si = ib0 / sb, sj = jb0 / sb, sk = kb0 / sb;
  const int64_t node = h->lvl_off[level] + IJK(si, sj, sk, nsi, nsj);
  const int go = h->bits[node];
  if (!go && h->dense && h->dense[node]) {
    /* dense root: every present triple of the region, unscreened */
    h->ndense++;
    const int64_t ie = ib0 + sb < nib ? ib0 + sb : nib, je = jb0 + sb < njb ? jb0 + sb : njb;
    const int64_t ke = kb0 + sb < nkb ? kb0 + sb : nkb;
    for (int64_t jb = jb0; jb < je; ++jb)
      for (int64_t ib = ib0; ib < ie; ++ib) {
        double* cb = h->C->blk[BIDX(h->C, ib, jb)];
        if (!cb) continue;
        for (int64_t kb = kb0; kb < ke; ++kb) {
          const double *ab = h->A->blk[BIDX(h->A, ib, kb)], *bb = h->B->blk[BIDX(h->B, kb, jb)];
          if (!ab || !bb) continue;
          const int64_t m = h->C->rsz[ib], n = h->C->csz[jb], k = h->A->csz[kb];
          or_dgemm(0, 0, m, n, k, 1.0, ab, m, bb, k, 1.0, cb, m);
        }
      }
    return;
  }
  if (go) {
    const int64_t hh = sb / 2, bs2 = h->sc.bs * h->sc.bs;
    h->nroots++;
    owned_t S00 = owned_new(hh, hh, bs2, 1), S01 = owned_new(hh, hh, bs2, 1);
    owned_t S10 = owned_new(hh, hh, bs2, 1), S11 = owned_new(hh, hh, bs2, 1);
    s_recurse(&S00, &S01, &S10, &S11,
              quad(h->Ag, ib0, kb0, hh, hh), quad(h->Ag, ib0, kb0 + hh, hh, hh),
              quad(h->Ag, ib0 + hh, kb0, hh, hh), quad(h->Ag, ib0 + hh, kb0 + hh, hh, hh),
              quad(h->Bg, kb0, jb0, hh, hh), quad(h->Bg, kb0, jb0 + hh, hh, hh),
              quad(h->Bg, kb0 + hh, jb0, hh, hh), quad(h->Bg, kb0 + hh, jb0 + hh, hh, hh), &h->sc);
    owned_t* S[4] = {&S00, &S01, &S10, &S11};
    const int64_t oi[4] = {0, 0, hh, hh}, oj[4] = {0, hh, 0, hh};
    for (int q = 0; q < 4; ++q)
      for (int64_t cb = 0; cb < hh; ++cb)
        for (int64_t rb = 0; rb < hh; ++rb) {
          double* t = GB(h->Cg, ib0 + oi[q] + rb, jb0 + oj[q] + cb);
          if (!t) continue; /* += into an empty block is a no-op in the reference */
          const double* s = S[q]->ptrs[rb + cb * hh];
          for (int64_t e = 0; e < bs2; ++e) t[e] = t[e] + s[e];
        }
    owned_free(&S00); owned_free(&S01); owned_free(&S10); owned_free(&S11);
  } else {
    const int64_t hh = sb / 2;
    for (int64_t dj = 0; dj < 2; ++dj) {
      const int64_t cj = jb0 + dj * hh;
      if (cj >= njb) continue;
      for (int64_t di = 0; di < 2; ++di) {
        const int64_t ci = ib0 + di * hh;
        if (ci >= nib) continue;
        for (int64_t dk = 0; dk < 2; ++dk) {
          const int64_t ck = kb0 + dk * hh;
          if (ck >= nkb) continue;
          h_recurse(h, ci, cj, ck, hh, level - 1);
        }
      }
    }
  }
}

static int cmp_task(const void* x, const void* y) {
  const int32_t* a = (const int32_t*)x;
  const int32_t* b = (const int32_t*)y;
  if (a[1] != b[1]) return a[1] < b[1] ? -1 : 1;
  if (a[0] != b[0]) return a[0] < b[0] ? -1 : 1;
  if (a[2] != b[2]) return a[2] < b[2] ? -1 : 1;
  return 0;
}

/* strassen.hpp:687-753 */
int or_matmult_strassen_sparse(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                               int64_t min_size_for_strassen,
                               int32_t* leaf_tasks, int64_t cap, int64_t* nleaf, int64_t* nroots) {
  return or_matmult_strassen_tree(C, A, B, thresh, min_size_for_strassen, NULL, leaf_tasks, cap, nleaf, nroots);
}

int or_matmult_strassen_tree(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                             int64_t min_size_for_strassen, const uint8_t* tree_bits,
                             int32_t* leaf_tasks, int64_t cap, int64_t* nleaf, int64_t* nroots) {
  return or_matmult_strassen_plan(C, A, B, thresh, min_size_for_strassen, tree_bits, NULL, leaf_tasks, cap, nleaf,
                                  nroots, NULL);
}

int or_matmult_strassen_plan(or_bsm* C, const or_bsm* A, const or_bsm* B, double thresh,
                             int64_t min_size_for_strassen, const uint8_t* tree_bits,
                             const uint8_t* dense_bits, int32_t* leaf_tasks, int64_t cap,
                             int64_t* nleaf, int64_t* nroots, int64_t* ndense) {
  if (!uniform_square(A, B) || !same_geometry(C, A)) return -2;
  const int64_t nib = A->nrb;
  const int64_t bs = A->rscr[0];
  const double thr = thresh * (double)(bs * bs * bs);
  /* tree sizes */
  const int max_level = (int)(size_t)log2((double)nib + 0.5);
  int64_t total = 0;
  int64_t* lvl_off = (int64_t*)malloc(sizeof(int64_t) * (size_t)(max_level + 1));
  for (int l = 0; l < max_level; ++l) {
    int64_t sb = (int64_t)1 << (l + 1);
    int64_t n = ceil_div(nib, sb);
    lvl_off[l] = total;
    total += n * n * n;
  }
  uint8_t* bits = (uint8_t*)malloc((size_t)(total > 0 ? total : 1));
  int64_t* lsz = (int64_t*)malloc(sizeof(int64_t) * (size_t)(max_level + 1));
  if (tree_bits) {
    memcpy(bits, tree_bits, (size_t)total);
  } else {
    int rc = or_strassen_dryrun(A, B, thr, bits, total, lsz, NULL);
    if (rc < 0) { free(lvl_off); free(bits); free(lsz); return rc; }
  }

  hctx_t h;
  memset(&h, 0, sizeof(h));
  h.C = C; h.A = A; h.B = B; h.bits = bits; h.dense = dense_bits; h.lvl_off = lvl_off; h.thr = thr;
  h.sc.bs = bs; h.sc.cutoff = min_size_for_strassen;
  h.leaf = leaf_tasks; h.cap = cap;
  grid_t ag = {A->blk, A->nrb, A->ncb, A->nrb}, bg = {B->blk, B->nrb, B->ncb, B->nrb}, cg = {C->blk, C->nrb, C->ncb, C->nrb};
  h.Ag = ag; h.Bg = bg; h.Cg = cg;
  const int top_level = max_level - 1;
  const int64_t top_sb = (int64_t)exp2((double)max_level);
  const int64_t n_top = ceil_div(nib, top_sb);
  for (int64_t sk = 0; sk < n_top; ++sk)
    for (int64_t sj = 0; sj < n_top; ++sj)
      for (int64_t si = 0; si < n_top; ++si) h_recurse(&h, si * top_sb, sj * top_sb, sk * top_sb, top_sb, top_level);
  if (leaf_tasks) qsort(leaf_tasks, (size_t)(h.nleaf < cap ? h.nleaf : cap), 3 * sizeof(int32_t), cmp_task);
  if (nleaf) *nleaf = h.nleaf;
  if (nroots) *nroots = h.nroots;
  if (ndense) *ndense = h.ndense;
  free(lvl_off); free(bits); free(lsz);
  return 0;
}

/* ---------------------------------------------------------------- RI-K -- */
void or_decompress(double* out, int64_t P, const double* comp, int64_t nvec2, const uint64_t* keys) {
  const double* src = comp + P * nvec2;
  for (int64_t n = 0; n < nvec2; ++n) out[keys[n]] = src[n];
}

/* occ_ri_k_step.cpp:83-123, one OpenMP thread */
int or_occ_ri_k(int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2,
                const double* ints3c, const double* mo, const uint64_t* keys,
                int32_t bs, double thresh, double* K) {
  for (int64_t n = 0; n < n_vec2; ++n) if (keys[n] >= (uint64_t)(n_bf * n_bf)) return -1;
  or_bsm* Kb = or_bsm_create_uniform(n_bf, n_occ, bs, bs, 0.0);
  or_bsm* Cb = or_bsm_create_uniform(n_bf, n_occ, bs, bs, 0.0);
  or_bsm_from_dense(Cb, mo, n_bf);
  or_bsm_fill(Kb, 0.0);
  double* dec = (double*)calloc((size_t)(n_bf * n_bf), sizeof(double));
  or_bsm* U = or_bsm_create_uniform(n_bf, n_bf, bs, bs, 1e-20);
  or_bsm* mui = or_bsm_create_uniform(n_bf, n_occ, bs, bs, 0.0);
  or_bsm* ij = or_bsm_create_uniform(n_occ, n_occ, bs, bs, 0.0);
  or_bsm* Ks = or_bsm_create_uniform(n_bf, n_occ, bs, bs, 0.0);
  or_bsm_fill(Ks, 0.0);
  int rc = 0;
  for (int64_t P = 0; P < n_aux && rc == 0; ++P) {
    or_decompress(dec, P, ints3c, n_vec2, keys);
    or_bsm_from_dense(U, dec, n_bf);
    rc |= or_matmult(mui, U, 0, Cb, 0, thresh, 1.0, 0.0, NULL, 0, NULL);
    or_bsm_calc_norms(mui);
    rc |= or_matmult(ij, Cb, 1, mui, 0, thresh, 1.0, 0.0, NULL, 0, NULL);
    or_bsm* ijs = or_bsm_clone(ij);
    rc |= or_bsm_add_transpose(ijs, ij);
    rc |= or_matmult(Ks, mui, 0, ijs, 0, thresh, 1.0, 1.0, NULL, 0, NULL);
    or_bsm_destroy(ijs);
  }
  or_bsm_add(Kb, Ks, 1.0);
  memset(K, 0, sizeof(double) * (size_t)(n_bf * n_occ));
  or_bsm_to_dense(Kb, K, n_bf);
  free(dec);
  or_bsm_destroy(Kb); or_bsm_destroy(Cb); or_bsm_destroy(U);
  or_bsm_destroy(mui); or_bsm_destroy(ij); or_bsm_destroy(Ks);
  return rc;
}

/* occ_ri_k_step.cpp:38-72, one OpenMP thread */
int or_occ_ri_k_dense(int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2,
                      const double* ints3c, const double* mo, const uint64_t* keys, double* K) {
  for (int64_t n = 0; n < n_vec2; ++n) if (keys[n] >= (uint64_t)(n_bf * n_bf)) return -1;
  double* dec = (double*)calloc((size_t)(n_bf * n_bf), sizeof(double));
  double* mui = (double*)malloc(sizeof(double) * (size_t)(n_bf * n_occ));
  double* ij = (double*)malloc(sizeof(double) * (size_t)(n_occ * n_occ));
  double* ijs = (double*)malloc(sizeof(double) * (size_t)(n_occ * n_occ));
  double* Ks = (double*)calloc((size_t)(n_bf * n_occ), sizeof(double));
  for (int64_t P = 0; P < n_aux; ++P) {
    or_decompress(dec, P, ints3c, n_vec2, keys);
    or_dgemm(0, 0, n_bf, n_occ, n_bf, 1.0, dec, n_bf, mo, n_bf, 0.0, mui, n_bf);
    or_dgemm(1, 0, n_occ, n_occ, n_bf, 1.0, mo, n_bf, mui, n_bf, 0.0, ij, n_occ);
    for (int64_t c = 0; c < n_occ; ++c)
      for (int64_t r = 0; r < n_occ; ++r) ijs[r + c * n_occ] = ij[r + c * n_occ] + ij[c + r * n_occ];
    or_dgemm(0, 0, n_bf, n_occ, n_occ, 1.0, mui, n_bf, ijs, n_occ, 1.0, Ks, n_bf);
  }
  for (int64_t e = 0; e < n_bf * n_occ; ++e) K[e] = 0.0 + Ks[e];
  free(dec); free(mui); free(ij); free(ijs); free(Ks);
  return 0;
}
