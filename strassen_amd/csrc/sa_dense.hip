// sa_dense.hip -- the reference's BLAS boundary (blaslapack.cpp:5-37) on the
// device: dense DGEMM through rocBLAS (a plain library GEMM), and the
// canonical norm / dot / axpy as small HBM-bound kernels.  Host pointers in,
// host pointers out; the copies go through the context stream.
#include <rocblas/rocblas.h>

#include <cmath>

#include "sa_internal.hpp"

namespace sa {

static rocblas_handle handle_for(sa_context* c) {
  static thread_local rocblas_handle h = nullptr;
  static thread_local hipStream_t s = nullptr;
  if (!h) {
    if (rocblas_create_handle(&h) != rocblas_status_success) SA_UNSUPPORTED("rocblas_create_handle failed");
  }
  if (s != c->stream) {
    rocblas_set_stream(h, c->stream);
    s = c->stream;
  }
  return h;
}

void device_dgemm(sa_context* c, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double beta, double* Cm, int64_t ldc) {
  rocblas_handle h = handle_for(c);
  ProfScope ps(c, "dgemm_dense");
  const rocblas_status st =
      rocblas_dgemm(h, ta ? rocblas_operation_transpose : rocblas_operation_none,
                    tb ? rocblas_operation_transpose : rocblas_operation_none, (rocblas_int)m, (rocblas_int)n,
                    (rocblas_int)k, &alpha, A, (rocblas_int)lda, B, (rocblas_int)ldb, &beta, Cm, (rocblas_int)ldc);
  if (st != rocblas_status_success) {
    set_error("rocblas_dgemm failed (%d)", (int)st);
    throw Error{SA_ERR_HIP};
  }
}

__global__ void __launch_bounds__(256) k_vec_norm_dot(const double* __restrict__ x, const double* __restrict__ y,
                                                      int64_t n, double* __restrict__ out) {
  // single wave, canonical 64-lane strided order (same as the block norm)
  const int lane = threadIdx.x;
  if (lane >= 64) return;
  double a = 0.0;
  for (int64_t e = lane; e < n; e += 64) a = fma(x[e], y ? y[e] : x[e], a);
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a = a + __shfl_down(a, o, 64);
  if (lane == 0) *out = y ? a : sqrt(a);
}

__global__ void k_axpy(int64_t n, double alpha, const double* __restrict__ x, double* __restrict__ y) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) y[i] = fma(alpha, x[i], y[i]);
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_dgemm(sa_ctx_t c, int transA, int transB, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
             int64_t lda, const double* B, int64_t ldb, double beta, double* Cm, int64_t ldc) {
  return guard([&] {
    SA_REQUIRE(c && A && B && Cm, "null argument");
    SA_REQUIRE(m >= 0 && n >= 0 && k >= 0, "negative dimension");
    SA_REQUIRE(lda >= (transA ? k : m) && ldb >= (transB ? n : k) && ldc >= m && lda > 0 && ldb > 0 && ldc > 0,
               "bad leading dimension");
    SA_REQUIRE(m < INT32_MAX && n < INT32_MAX && k < INT32_MAX, "dimension exceeds 32-bit BLAS range");
    if (m == 0 || n == 0) return;
    SA_CHECK_HIP(hipSetDevice(c->device));
    const int64_t ar = transA ? k : m, ac = transA ? m : k;
    const int64_t br = transB ? n : k, bc = transB ? k : n;
    const size_t sa_ = sizeof(double) * (size_t)(lda * (ac > 0 ? ac - 1 : 0) + ar);
    const size_t sb_ = sizeof(double) * (size_t)(ldb * (bc > 0 ? bc - 1 : 0) + br);
    const size_t sc_ = sizeof(double) * (size_t)(ldc * (n - 1) + m);
    DevBuf& bA = c->buf("dg_a");
    DevBuf& bB = c->buf("dg_b");
    DevBuf& bC = c->buf("dg_c");
    bA.reserve(sa_ + 8);
    bB.reserve(sb_ + 8);
    bC.reserve(sc_ + 8);
    if (k > 0) {
      SA_CHECK_HIP(hipMemcpyAsync(bA.p, A, sa_, hipMemcpyHostToDevice, c->stream));
      SA_CHECK_HIP(hipMemcpyAsync(bB.p, B, sb_, hipMemcpyHostToDevice, c->stream));
    }
    // C moves as its m x n part only: rows m..ldc-1 of each column belong to
    // the caller and are neither read nor written
    const size_t pitch = sizeof(double) * (size_t)ldc, width = sizeof(double) * (size_t)m;
    if (beta != 0.0)
      SA_CHECK_HIP(hipMemcpy2DAsync(bC.p, pitch, Cm, pitch, width, (size_t)n, hipMemcpyHostToDevice, c->stream));
    device_dgemm(c, transA != 0, transB != 0, m, n, k, alpha, bA.as<double>(), lda, bB.as<double>(), ldb, beta,
                 bC.as<double>(), ldc);
    SA_CHECK_HIP(hipMemcpy2DAsync(Cm, pitch, bC.p, pitch, width, (size_t)n, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_dnrm2(sa_ctx_t c, int64_t n, const double* x, double* out) {
  return guard([&] {
    SA_REQUIRE(c && out && (x || n == 0), "null argument");
    DevBuf& b = c->buf("dv_x");
    b.reserve(sizeof(double) * (size_t)(n + 1));
    if (n) SA_CHECK_HIP(hipMemcpyAsync(b.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    DevBuf& r = c->buf("dv_r");
    r.reserve(64);
    hipLaunchKernelGGL(k_vec_norm_dot, dim3(1), dim3(64), 0, c->stream, (const double*)b.p, (const double*)nullptr, n,
                       r.as<double>());
    SA_CHECK_HIP(hipMemcpyAsync(out, r.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_ddot(sa_ctx_t c, int64_t n, const double* x, const double* y, double* out) {
  return guard([&] {
    SA_REQUIRE(c && out && ((x && y) || n == 0), "null argument");
    DevBuf& b = c->buf("dv_x");
    DevBuf& b2 = c->buf("dv_y");
    b.reserve(sizeof(double) * (size_t)(n + 1));
    b2.reserve(sizeof(double) * (size_t)(n + 1));
    if (n) {
      SA_CHECK_HIP(hipMemcpyAsync(b.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
      SA_CHECK_HIP(hipMemcpyAsync(b2.p, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    }
    DevBuf& r = c->buf("dv_r");
    r.reserve(64);
    hipLaunchKernelGGL(k_vec_norm_dot, dim3(1), dim3(64), 0, c->stream, (const double*)b.p, (const double*)b2.p, n,
                       r.as<double>());
    SA_CHECK_HIP(hipMemcpyAsync(out, r.p, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_daxpy(sa_ctx_t c, int64_t n, double alpha, const double* x, double* y) {
  return guard([&] {
    SA_REQUIRE(c && ((x && y) || n == 0), "null argument");
    if (n == 0) return;
    DevBuf& b = c->buf("dv_x");
    DevBuf& b2 = c->buf("dv_y");
    b.reserve(sizeof(double) * (size_t)n);
    b2.reserve(sizeof(double) * (size_t)n);
    SA_CHECK_HIP(hipMemcpyAsync(b.p, x, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    SA_CHECK_HIP(hipMemcpyAsync(b2.p, y, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_axpy, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, n, alpha,
                       (const double*)b.p, b2.as<double>());
    SA_CHECK_HIP(hipMemcpyAsync(y, b2.p, sizeof(double) * n, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

}  // extern "C"
