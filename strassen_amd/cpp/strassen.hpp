// strassen.hpp -- drop-in for the reference's hybrid Strassen-sparse entry
// points (strassen.hpp:510-753) on the device.
//
// The global cutoff min_size_for_strassen is declared here and defined by the
// caller, as the reference's main.cpp:9 does.
#ifndef SA_STRASSEN_HPP
#define SA_STRASSEN_HPP
#include <cstdint>
#include <vector>

#include "blocksparsematrix.h"
#include "utils.hpp"

extern size_t min_size_for_strassen;

// strassen.hpp:510-679: per-level decision tree, bit [l][ijk(si,sj,sk,ns,ns)]
template <typename Num>
std::vector<std::vector<bool>> matmult_strassen_dryrun(const BlockSparseMatrix<Num>& A_mat,
                                                       const BlockSparseMatrix<Num>& B_mat,
                                                       const Num thresh_per_block) {
  const size_t nb = A_mat.nrowblocks();
  int32_t nlev = 0;
  std::vector<int64_t> lsz(64, 0);
  std::vector<uint64_t> lfl(65, 0);
  sa_cpp::check(sa_strassen_dryrun(A_mat.handle(), B_mat.handle(), thresh_per_block, nullptr, 0, lsz.data(),
                                   lfl.data(), &nlev),
                "sa_strassen_dryrun");
  int64_t total = 0;
  for (int l = 0; l < nlev; ++l) total += lsz[l];
  std::vector<uint8_t> bits((size_t)total + 1);
  sa_cpp::check(sa_strassen_dryrun(A_mat.handle(), B_mat.handle(), thresh_per_block, bits.data(), total, nullptr,
                                   nullptr, &nlev),
                "sa_strassen_dryrun");
  std::vector<std::vector<bool>> tree((size_t)nlev);
  int64_t off = 0;
  for (int l = 0; l < nlev; ++l) {
    tree[l].resize((size_t)lsz[l]);
    for (int64_t i = 0; i < lsz[l]; ++i) tree[l][i] = bits[off + i] != 0;
    off += lsz[l];
    std::printf("total flops on level %d: %llu\n", l + 1, (unsigned long long)lfl[l + 1]);
  }
  (void)nb;
  return tree;
}

// strassen.hpp:687-753: C += A*B (transposes unsupported, alpha == 1, beta
// ignored, C pre-filled by the caller)
template <typename Num>
void matmult_strassen_sparse(BlockSparseMatrix<Num>& C_mat, const BlockSparseMatrix<Num>& A_mat, const bool transA,
                             const BlockSparseMatrix<Num>& B_mat, const bool transB, const Num thresh,
                             const Num& alpha, const Num& beta) {
  sa_opts_t o;
  sa_opts_default(&o);
  o.cutoff = (int64_t)min_size_for_strassen;
  o.scheduler = sa_cpp::learned() ? SA_SCHED_MLP : SA_SCHED_HEURISTIC;
  sa_cpp::check(sa_matmult_strassen_sparse(C_mat.handle(), A_mat.handle(), transA, B_mat.handle(), transB, thresh,
                                           alpha, beta, &o),
                "sa_matmult_strassen_sparse");
}

#endif
