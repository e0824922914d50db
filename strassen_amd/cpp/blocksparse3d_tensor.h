// blocksparse3d_tensor.h -- the reference's BlockSparse3DTensor and its two
// declared tensormult forms (blocksparse3d_tensor.h:7-24), defined here.
// The reference only declares them; the semantics implemented are:
//   form 1:  C_ij = alpha * sum_{k,l} op(A^l)_ik op(B^l)_kj + beta C_ij
//   form 2:  op_C(C^l)_ij = alpha * sum_k op(A)_ik op(B^l)_kj + beta op_C(C^l)_ij
// Each term is one screened device matmult (the slices are separate matrices
// with their own block arenas, so form 1 accumulates slice by slice into C in
// l order; the RI-K step's K contraction, sa_ri_k_step, is the form-1 sum run
// as one screened multiply over a reshape view of a single arena).
#ifndef SA_BLOCKSPARSE3D_TENSOR_H
#define SA_BLOCKSPARSE3D_TENSOR_H
#include <vector>

#include "blocksparsematrix.h"

template <typename Num>
using BlockSparse3DTensor = std::vector<BlockSparseMatrix<Num>>;

// form 1 (blocksparse3d_tensor.h:10-16)
template <typename Num>
void tensormult(BlockSparseMatrix<Num>& C, const BlockSparse3DTensor<Num>& A, const bool transA,
                const BlockSparse3DTensor<Num>& B, const bool transB, const Num thresh, const Num& alpha,
                const Num& beta) {
  if (A.size() != B.size()) throw std::runtime_error("tensormult: A and B slice counts differ");
  if (beta != Num(1)) {
    if (beta == Num(0)) C.zero();
    else C.scale(beta);
  }
  for (size_t l = 0; l < A.size(); ++l) matmult(C, A[l], transA, B[l], transB, thresh, alpha, Num(1));
}

// form 2 (blocksparse3d_tensor.h:18-24); transC = true stores C^l transposed
template <typename Num>
void tensormult(BlockSparse3DTensor<Num>& C, const bool transC, const BlockSparseMatrix<Num>& A, const bool transA,
                const BlockSparse3DTensor<Num>& B, const bool transB, const Num thresh, const Num& alpha,
                const Num& beta) {
  if (C.size() != B.size()) throw std::runtime_error("tensormult: C and B slice counts differ");
  for (size_t l = 0; l < B.size(); ++l) {
    if (!transC) {
      matmult(C[l], A, transA, B[l], transB, thresh, alpha, beta);
    } else {
      // (op(A) op(B^l))^T = op(B^l)^T op(A)^T
      matmult(C[l], B[l], !transB, A, !transA, thresh, alpha, beta);
    }
  }
}

// the reference's fourth overload (blocksparse3d_tensor.h:35-43): it forwards
// to form 2 with transC = true and both operand flags flipped, i.e.
// C^l = alpha op(B^l) op(A) + beta C^l; kept with that forwarding
template <typename Num>
void tensormult(BlockSparse3DTensor<Num>& C, const BlockSparseMatrix<Num>& A, const bool transA,
                const BlockSparse3DTensor<Num>& B, const bool transB, const Num thresh, const Num& alpha,
                const Num& beta) {
  tensormult(C, true, A, !transA, B, !transB, thresh, alpha, beta);
}

#endif
