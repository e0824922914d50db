// occ_ri_k_step -- argv-compatible driver of the reference's occ_ri_k_step.cpp:14-131
//   <N_bf> <N_aux> <N_occ> <N_vec2> <3c file> <MO file> <v2m_key file> <blocksize> <threshold>
// Dense and block-sparse occupied RI-K, K_{mu j} = sum_{P,i} (mu i|P)(ij|P), both
// on the device: the dense path through the device DGEMM per aux index P, the
// sparse path as one sa_occ_ri_k call (device-side decompression straight into
// block form, then the reference's screened call sequence per P).
#include <chrono>
#include <cstdint>
#include <cstdio>
#include <string>
#include <vector>

#include "blocksparsematrix.h"
#include "matrix.h"
#include "strassen.hpp"

size_t min_size_for_strassen = 0;

// occ_ri_k_step.cpp:133-146
template <typename T>
void decompress_integrals(Matrix<T>& out, const size_t P, const Matrix<T>& comp, const Matrix<size_t>& keys) {
  T* d = out.data_ptr();
  const T* s = &comp.elem(0, P);
  for (size_t n = 0; n < keys.size(); ++n) d[keys.elem(n, 0)] = s[n];
}

int main(int argc, char** argv) {
  if (argc < 10) {
    std::puts("Usage: <program> <N_bf> <N_aux> <N_occ> <N_vec2> <3c filename> <MO filename> <v2m_key_filename> <blocksize> <threshold>");
    return 1;
  }
  const size_t N_bf = std::stoul(argv[1]), N_aux = std::stoul(argv[2]), N_occ = std::stoul(argv[3]);
  const size_t N_vec2 = std::stoul(argv[4]);
  const std::string fn_3c = argv[5], fn_mo = argv[6], fn_key = argv[7];
  const size_t bs = std::stoul(argv[8]);
  const double thresh = std::stod(argv[9]);

  Matrix<double> ints(N_vec2, N_aux);
  ints.read_from_file(fn_3c.c_str());
  Matrix<double> K(0.0, N_bf, N_occ);
  Matrix<double> C(N_bf, N_occ);
  C.read_from_file(fn_mo.c_str());
  Matrix<size_t> keys(N_vec2, 1);
  keys.read_from_file(fn_key.c_str());
  for (size_t n = 0; n < N_vec2; ++n)
    if (keys.elem(n, 0) >= N_bf * N_bf) {
      std::printf("key %zu out of range\n", n);
      return 1;
    }
  const double total_flops = 2.0 * ((double)N_bf * N_bf * N_occ * N_aux + 2.0 * N_bf * N_occ * N_occ * N_aux);

  std::printf("--- dense occ-RIK  ---\n");
  for (int it = 0; it < 3; ++it) {
    K.fill_with_values(0.0);
    const auto t0 = std::chrono::steady_clock::now();
    Matrix<double> dec(0.0, N_bf, N_bf), mui(N_bf, N_occ), ij(N_occ, N_occ), ijs(N_occ, N_occ);
    for (size_t P = 0; P < N_aux; ++P) {
      decompress_integrals(dec, P, ints, keys);
      matmult(mui, dec, false, C, false, 1.0, 0.0);
      matmult(ij, C, true, mui, false, 1.0, 0.0);
      ijs = ij;
      ijs.add_transpose(ij);
      matmult(K, mui, false, ijs, false, 1.0, 1.0);
    }
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [%d] %.4f s  (%.4f GFLOPs)\n", it + 1, 1e-6 * us, 2e-3 * total_flops / us);
  }
  const double l2 = K.calc_frobenius_norm();
  std::printf("L2 norm of output = %e\n", l2);

  std::printf("--- sparse occ-RIK  ---\n");
  // The P loop of occ_ri_k_step.cpp:89-118 as one device call: the compressed
  // integrals are uploaded once and decompressed per P on the device straight
  // into block form (no dense N_bf x N_bf matrix, no per-P host traffic).
  std::vector<uint64_t> key64(N_vec2);
  for (size_t n = 0; n < N_vec2; ++n) key64[n] = (uint64_t)keys.elem(n, 0);
  Matrix<double> Ksp(0.0, N_bf, N_occ);
  for (int it = 0; it < 3; ++it) {
    const auto t0 = std::chrono::steady_clock::now();
    sa_cpp::check(sa_occ_ri_k(sa_cpp::context(), (int64_t)N_bf, (int64_t)N_aux, (int64_t)N_occ, (int64_t)N_vec2,
                              ints.data_ptr(), C.data_ptr(), key64.data(), (int32_t)bs, thresh, 0, nullptr,
                              Ksp.data_ptr()),
                  "sa_occ_ri_k");
    const double us = (double)std::chrono::duration_cast<std::chrono::microseconds>(std::chrono::steady_clock::now() - t0).count();
    std::printf("  [%d] %.4f s  (%.4f GFLOPs)\n", it + 1, 1e-6 * us, 2e-3 * total_flops / us);
    auto diff = Ksp - K;
    std::printf("relative RMSD = %e\n", diff.calc_frobenius_norm() / l2);
  }
  // not printed by the reference: both norms to full precision, for checks
  // against known answers
  std::printf("full-precision L2 norms: dense %.15e sparse %.15e\n", l2, Ksp.calc_frobenius_norm());
  return 0;
}
