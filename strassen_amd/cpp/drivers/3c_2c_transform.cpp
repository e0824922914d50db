// 3c_2c_transform -- argv-compatible driver of the reference's 3c_2c_transform.cpp:10-64
//   <N_bf> <N_aux> <N_vec2> <3c filename> <2c filename> <blocksize> <threshold>
// (mu nu|P)' = sum_Q (mu nu|Q) (Q|P)^{-1/2}: dense device DGEMM and block-sparse
// matmult with beta=1 after a zero fill; prints GFLOP/s and the relative error.
// With a 6th argument "-o <file>" (argv[8], argv[9]) the sparse result is
// written as a raw column-major file.
#include <chrono>
#include <cstdio>
#include <cstring>
#include <string>

#include "blocksparsematrix.h"
#include "matrix.h"

int main(int argc, char** argv) {
  if (argc < 8) {
    std::puts("Usage: <program> <N_bf> <N_aux> <N_vec2> <3c filename> <2c filename> <blocksize> <threshold>");
    return 1;
  }
  const size_t N_aux = std::stoul(argv[2]), N_vec2 = std::stoul(argv[3]);
  const std::string fn_3c = argv[4], fn_2c = argv[5];
  const size_t bs = std::stoul(argv[6]);
  const double thresh = std::stod(argv[7]);
  const char* out_file = (argc >= 10 && std::strcmp(argv[8], "-o") == 0) ? argv[9] : nullptr;

  Matrix<double> i3(N_vec2, N_aux);
  i3.read_from_file(fn_3c.c_str());
  Matrix<double> i2(N_aux, N_aux);
  i2.read_from_file(fn_2c.c_str());
  Matrix<double> tr(N_vec2, N_aux);
  std::printf("--- dense matmult  ---\n");
  {
    const auto t0 = std::chrono::steady_clock::now();
    matmult(tr, i3, false, i2, false, 1.0, 0.0);
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [1] %.4f s  (%.4f GFLOPs)\n", 1e-6 * us, 2e-3 * (double)(N_vec2 * N_aux * N_aux) / us);
  }
  const double l2 = tr.calc_frobenius_norm();
  std::printf("L2 norm of output = %e\n", l2);

  BlockSparseMatrix<double> i3b(i3, bs, bs, 0.0);
  BlockSparseMatrix<double> i2b(i2, bs, bs, 0.0);
  BlockSparseMatrix<double> trb(N_vec2, N_aux, bs, bs, 0.0);
  std::printf("\n--- BSM matmult (thresh=%.2e) ---\n", thresh);
  trb.fill_with_values(0.0);
  const auto t0 = std::chrono::steady_clock::now();
  matmult(trb, i3b, false, i2b, false, thresh, 1.0, 1.0);
  const auto res = trb.to_matrix();
  const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
  std::printf("  [1] %.4f s  (%.4f GFLOPs)", 1e-6 * us, 2e-3 * (double)(N_vec2 * N_aux * N_aux) / us);
  auto diff = res - tr;
  std::printf("relative RMSD = %e\n", diff.calc_frobenius_norm() / l2);
  if (out_file) res.write_to_file(out_file);
  return 0;
}
