// strassen_bench -- argv-compatible driver of the reference's main.cpp:10-68
//   <blocksize> <strassen_min> <thresh> <n_runs> [<A file> <B file> <N>]
// Dense DGEMM, block-sparse matmult and matmult_strassen_sparse on the device,
// each n_runs times, with GFLOP/s = 2N^3/t and the relative error against the
// dense product.  The reference reads two N=11230 matrices that it does not
// ship; without file arguments this driver builds a banded synthetic pair
// (N = 2048, blocks |I-J| <= 4 filled with exp(-0.5|I-J|)-scaled values).
#include <chrono>
#include <cmath>
#include <cstdio>
#include <string>

#include "blocksparsematrix.h"
#include "matrix.h"
#include "strassen.hpp"

size_t min_size_for_strassen = 0;

static double frand(uint64_t& s) {
  s = s * 6364136223846793005ull + 1442695040888963407ull;
  return (double)(s >> 11) * (1.0 / 9007199254740992.0) - 0.5;
}

int main(int argc, char** argv) {
  if (argc < 5) {
    std::puts("Usage: <program> <blocksize> <strassen_min> <thresh> <n_runs> [<A file> <B file> <N>]");
    return 1;
  }
  const size_t bs = std::stoul(argv[1]);
  min_size_for_strassen = std::stoul(argv[2]);
  const double thresh = std::stod(argv[3]);
  const int n_runs = std::stoi(argv[4]);
  size_t N = 2048;
  Matrix<double> mat1, mat2;
  if (argc >= 8) {
    N = std::stoul(argv[7]);
    mat1 = Matrix<double>(N, N);
    mat2 = Matrix<double>(N, N);
    mat1.read_from_file(argv[5]);
    mat2.read_from_file(argv[6]);
  } else {
    mat1 = Matrix<double>(0.0, N, N);
    mat2 = Matrix<double>(0.0, N, N);
    uint64_t s = 12345;
    for (size_t c = 0; c < N; ++c)
      for (size_t r = 0; r < N; ++r) {
        const long d = std::labs((long)(r / 64) - (long)(c / 64));
        if (d <= 4) {
          mat1.elem(r, c) = std::exp(-0.5 * d) * frand(s);
          mat2.elem(r, c) = std::exp(-0.5 * d) * frand(s);
        }
      }
  }
  const double flops = 2.0 * (double)N * N * N;
  Matrix<double> dense(N, N);
  std::printf("--- Dense DGEMM (%zu x %zu) ---\n", N, N);
  for (int i = 0; i < n_runs; ++i) {
    const auto t0 = std::chrono::steady_clock::now();
    matmult(dense, mat1, false, mat2, false, 1.0, 0.0);
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [%d] %.4f s  (%.4f GFLOPs)\n", i + 1, 1e-6 * us, 1e-3 * flops / us);
  }
  const double l2 = dense.calc_frobenius_norm();
  BlockSparseMatrix<double> b1(mat1, bs, bs, 0.0), b2(mat2, bs, bs, 0.0);
  BlockSparseMatrix<double> b3(Matrix<double>(N, N), bs, bs, 0.0);
  std::printf("\n--- BSM matmult (thresh=%.2e) ---\n", thresh);
  for (int i = 0; i < n_runs; ++i) {
    b3.fill_with_values(0.0);
    const auto t0 = std::chrono::steady_clock::now();
    matmult(b3, b1, false, b2, false, thresh, 1.0, 0.0);
    const auto r = b3.to_matrix();
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [%d] %.4f s  (%.4f GFLOPs)", i + 1, 1e-6 * us, 1e-3 * flops / us);
    std::printf("relative RMSD = %e\n", (r - dense).calc_frobenius_norm() / l2);
  }
  std::printf("\n--- Strassen-Sparse (thresh=%.2e) ---\n", thresh);
  for (int i = 0; i < n_runs; ++i) {
    b3.fill_with_values(0.0);
    const auto t0 = std::chrono::steady_clock::now();
    matmult_strassen_sparse(b3, b1, false, b2, false, thresh, 1.0, 0.0);
    const auto r = b3.to_matrix();
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [%d] %.4f s  (%.4f GFLOPs)", i + 1, 1e-6 * us, 1e-3 * flops / us);
    std::printf("relative RMSD = %e\n", (r - dense).calc_frobenius_norm() / l2);
  }
  return 0;
}
