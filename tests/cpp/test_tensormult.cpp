// test_tensormult -- drives the drop-in tensormult overloads
// (strassen_amd/cpp/blocksparse3d_tensor.h, reference blocksparse3d_tensor.h:10-43)
// on the device for tests/test_gpu_tensormult.py, which checks the outputs
// against the oracle.
//   test_tensormult <dir> <L> <n> <bs> <transA> <transB> <thresh> <alpha> <beta>
// Inputs in <dir> (raw column-major n x n): A<l>.data, B<l>.data (l < L), C0.data.
// Outputs in <dir>:
//   form1.data      C = alpha sum_l op(A^l) op(B^l) + beta C0
//   form2_<l>.data  C^l = alpha op(A^0) op(B^l) + beta C0             (transC = false)
//   form2t_<l>.data (C^l)^T = alpha op(A^0) op(B^l) + beta C0^T, C^l written as stored
//   form4_<l>.data  the reference's fourth overload on C^l = C0
//   block_last.data block(last, last) of form 1's C (empty file if absent)
#include <cstdio>
#include <string>
#include <vector>

#include "blocksparse3d_tensor.h"
#include "blocksparsematrix.h"
#include "matrix.h"

int main(int argc, char** argv) {
  if (argc != 10) {
    std::puts("usage: test_tensormult <dir> <L> <n> <bs> <transA> <transB> <thresh> <alpha> <beta>");
    return 2;
  }
  const std::string dir = argv[1];
  const int L = std::stoi(argv[2]);
  const size_t n = std::stoul(argv[3]), bs = std::stoul(argv[4]);
  const bool tA = std::stoi(argv[5]) != 0, tB = std::stoi(argv[6]) != 0;
  const double thresh = std::stod(argv[7]), alpha = std::stod(argv[8]), beta = std::stod(argv[9]);
  auto path = [&](const std::string& name) { return dir + "/" + name; };
  auto load = [&](const std::string& name) {
    Matrix<double> m(n, n);
    m.read_from_file(path(name).c_str());
    return m;
  };
  try {
    BlockSparse3DTensor<double> A, B;
    for (int l = 0; l < L; ++l) {
      A.emplace_back(load("A" + std::to_string(l) + ".data"), bs, bs, thresh);
      B.emplace_back(load("B" + std::to_string(l) + ".data"), bs, bs, thresh);
    }
    const Matrix<double> c0 = load("C0.data");
    auto fresh = [&]() { return BlockSparseMatrix<double>(c0, bs, bs, 0.0); };
    auto fresh_tensor = [&]() {
      BlockSparse3DTensor<double> t;
      for (int l = 0; l < L; ++l) t.push_back(fresh());
      return t;
    };
    auto save = [&](const BlockSparseMatrix<double>& m, const std::string& name) {
      Matrix<double> out(0.0, n, n);
      m.to_pointer(out.data_ptr());
      out.write_to_file(path(name).c_str());
    };
    {
      BlockSparseMatrix<double> C = fresh();
      tensormult(C, A, tA, B, tB, thresh, alpha, beta);
      save(C, "form1.data");
      // per-block read access: the last block row / column (ragged)
      const size_t rb = C.nrowblocks() - 1, cb = C.ncolblocks() - 1;
      const auto blk = C.block(rb, cb);
      blk.write_to_file(path("block_last.data").c_str());
    }
    {
      BlockSparse3DTensor<double> C = fresh_tensor();
      tensormult(C, false, A[0], tA, B, tB, thresh, alpha, beta);
      for (int l = 0; l < L; ++l) save(C[l], "form2_" + std::to_string(l) + ".data");
    }
    {
      BlockSparse3DTensor<double> C = fresh_tensor();
      tensormult(C, true, A[0], tA, B, tB, thresh, alpha, beta);
      for (int l = 0; l < L; ++l) save(C[l], "form2t_" + std::to_string(l) + ".data");
    }
    {
      BlockSparse3DTensor<double> C = fresh_tensor();
      tensormult(C, A[0], tA, B, tB, thresh, alpha, beta);
      for (int l = 0; l < L; ++l) save(C[l], "form4_" + std::to_string(l) + ".data");
    }
  } catch (const std::exception& e) {
    std::printf("error: %s\n", e.what());
    return 1;
  }
  std::puts("ok");
  return 0;
}
