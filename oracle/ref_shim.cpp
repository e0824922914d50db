// ref_shim.cpp -- C entry points over the reference's own templates.
//
// TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile against the
// reference sources where they lie (-I <reference dir>, plus its
// blaslapack.cpp and Intel MKL from the image) into oracle/_ref/libref.so.
// Nothing here is copied from the reference: these are thin drivers that build
// the reference's BlockSparseMatrix objects from dense column-major arrays, call
// the reference functions, and hand the results back as plain arrays.  Used to
// pin the C restatement in oracle.c and as the CPU baseline.
#include "matrix.h"
#include "matrix.hpp"
#include "blocksparsematrix.h"
#include "blocksparsematrix.hpp"
#include "strassen.hpp"

#include <chrono>
#include <cstdint>
#include <cstring>
#include <omp.h>

size_t min_size_for_strassen = 0;  // the global the reference's Strassen path reads (main.cpp:9)

namespace {
using BSM = BlockSparseMatrix<double>;

void dump(const BSM& m, double* dense, uint8_t* present, double* norms) {
  if (dense) {
    std::memset(dense, 0, sizeof(double) * m.nrow() * m.ncol());
    m.to_pointer(dense);
  }
  for (size_t cb = 0; cb < m.ncolblocks(); ++cb)
    for (size_t rb = 0; rb < m.nrowblocks(); ++rb) {
      const size_t b = rb + cb * m.nrowblocks();
      if (present) present[b] = m.block(rb, cb).size() != 0;
      if (norms) norms[b] = m.block(rb, cb).frobenius_norm();
    }
}
}  // namespace

extern "C" {

int ref_set_threads(int n) {
  omp_set_num_threads(n);
  return omp_get_max_threads();
}

double ref_norm(const double* x, int64_t n) {
  Matrix<double> m(x, (size_t)n, 1);
  return m.frobenius_norm();
}

// copy_from_input_array on a uniform BSM; returns presence + norms.
int ref_from_dense(const double* a, int64_t nr, int64_t nc, int32_t bs_r, int32_t bs_c, double thresh,
                   uint8_t* present, double* norms) {
  BSM m(a, nr, nc, bs_r, bs_c, thresh);
  dump(m, nullptr, present, norms);
  return 0;
}

// C = alpha op(A) op(B) + beta C.  A is ar x ac, B is br x bc, C (cr x cc) is
// built from c_in with threshold c_thresh (so its block pattern is controlled).
int ref_matmult(const double* a, int64_t ar, int64_t ac, double a_thresh,
                const double* b, int64_t br, int64_t bc, double b_thresh,
                double* c_io, int64_t cr, int64_t cc, double c_thresh, int c_empty,
                int32_t bs, int transA, int transB, double thresh, double alpha, double beta,
                uint8_t* c_present, double* c_norms, double* seconds) {
  BSM A(a, ar, ac, bs, bs, a_thresh);
  BSM B(b, br, bc, bs, bs, b_thresh);
  BSM C(cr, cc, bs, bs, c_thresh);
  if (!c_empty) C.copy_from_input_array(c_io);
  const auto t0 = std::chrono::steady_clock::now();
  matmult(C, A, transA != 0, B, transB != 0, thresh, alpha, beta);
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  dump(C, c_io, c_present, c_norms);
  return 0;
}

// Uniform BSM built from `a` at threshold thresh (copy-in drop rule), then
// scale(s) (the cached norms follow the reference's rule), then recompress();
// returns presence, cached norms and the dense result.
int ref_scale_recompress(const double* a, int64_t nr, int64_t nc, int32_t bs_r, int32_t bs_c, double thresh,
                         double s, uint8_t* present, double* norms, double* dense) {
  BSM m(a, nr, nc, bs_r, bs_c, thresh);
  m.scale(s);
  m.recompress();
  dump(m, dense, present, norms);
  return 0;
}

// dot(A, B, thresh) of two uniform BSMs built from a and b (copy-in at
// build_thresh).
double ref_dot(const double* a, const double* b, int64_t nr, int64_t nc, int32_t bs_r, int32_t bs_c,
               double build_thresh, double thresh) {
  BSM A(a, nr, nc, bs_r, bs_c, build_thresh);
  BSM B(b, nr, nc, bs_r, bs_c, build_thresh);
  return dot(A, B, thresh);
}

// matmult_strassen_dryrun: concatenated level bits; returns max_level.
int ref_dryrun(const double* a, const double* b, int64_t n, int32_t bs, double bsm_thresh,
               double thresh_per_block, uint8_t* bits, int64_t cap) {
  BSM A(a, n, n, bs, bs, bsm_thresh);
  BSM B(b, n, n, bs, bs, bsm_thresh);
  const auto tree = matmult_strassen_dryrun(A, B, thresh_per_block);
  int64_t w = 0;
  for (const auto& lvl : tree)
    for (size_t i = 0; i < lvl.size(); ++i) {
      if (w >= cap) return -3;
      bits[w++] = lvl[i] ? 1 : 0;
    }
  return (int)tree.size();
}

// matmult_strassen_sparse with C pre-filled from c_io (all blocks present).
int ref_strassen(const double* a, const double* b, double* c_io, int64_t n, int32_t bs,
                 double bsm_thresh, double thresh, int64_t cutoff, double* seconds) {
  min_size_for_strassen = (size_t)cutoff;
  BSM A(a, n, n, bs, bs, bsm_thresh);
  BSM B(b, n, n, bs, bs, bsm_thresh);
  BSM C(c_io, n, n, bs, bs, 0.0);
  const auto t0 = std::chrono::steady_clock::now();
  matmult_strassen_sparse(C, A, false, B, false, thresh, 1.0, 0.0);
  const auto t1 = std::chrono::steady_clock::now();
  if (seconds) *seconds = std::chrono::duration<double>(t1 - t0).count();
  dump(C, c_io, nullptr, nullptr);
  return 0;
}

// The sparse occ-RI-K chain of the reference driver for one thread: same
// sequence of reference calls per aux index P, K_sub summed over P in order.
int ref_occ_ri_k(int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2,
                 const double* ints3c, const double* mo, const uint64_t* keys,
                 int32_t bs, double thresh, double* K) {
  Matrix<double> mo_m(mo, n_bf, n_occ);
  BSM K_bsm(n_bf, n_occ, bs, bs, 0.0);
  BSM mo_bsm(mo_m, bs, bs, 0.0);
  K_bsm.fill_with_values(0.0);
  Matrix<double> dec(0.0, n_bf, n_bf);
  BSM U(n_bf, n_bf, bs, bs, 1e-20);
  BSM mui(n_bf, n_occ, bs, bs, 0.0);
  BSM ij(n_occ, n_occ, bs, bs, 0.0);
  BSM ij_sym(ij);
  BSM K_sub(n_bf, n_occ, bs, bs, 0.0);
  K_sub.fill_with_values(0.0);
  for (int64_t P = 0; P < n_aux; ++P) {
    const double* src = ints3c + P * n_vec2;
    for (int64_t e = 0; e < n_vec2; ++e) dec.data_ptr()[keys[e]] = src[e];
    U.copy_from_input_matrix(dec);
    matmult(mui, U, false, mo_bsm, false, thresh, 1.0, 0.0);
    mui.calc_frobenius_norms();
    matmult(ij, mo_bsm, true, mui, false, thresh, 1.0, 0.0);
    ij_sym = ij;
    ij_sym.add_transpose(ij);
    matmult(K_sub, mui, false, ij_sym, false, thresh, 1.0, 1.0);
  }
  K_bsm += K_sub;
  dump(K_bsm, K, nullptr, nullptr);
  return 0;
}

}  // extern "C"
