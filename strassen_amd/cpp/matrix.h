// matrix.h -- drop-in for the reference's dense Matrix<Num, Allocator>
// (matrix.h:9-109) as used by the drivers: column-major host storage, cached
// Frobenius norm, raw binary file I/O.  The BLAS boundary (dgemm, dnrm2,
// daxpy; blaslapack.cpp) runs on the device through the C ABI.
#ifndef SA_MATRIX_H
#define SA_MATRIX_H
#include <cassert>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <memory>
#include <vector>

#include "sa_runtime.hpp"

template <typename Num, typename Allocator = std::allocator<Num>>
class Matrix final {
  static_assert(std::is_same<Num, double>::value || std::is_integral<Num>::value,
                "the device path computes in FP64 (integer matrices hold keys)");

 public:
  explicit Matrix(const Allocator& alloc = Allocator()) : _data(alloc) {}
  explicit Matrix(size_t nrow, size_t ncol, const Allocator& alloc = Allocator())
      : _nrow(nrow), _ncol(ncol), _data(nrow * ncol, alloc) {}
  // matrix.hpp:26-33: all elements val, cached norm sqrt(nr*nc)*val
  explicit Matrix(const Num val, size_t nrow, size_t ncol, const Allocator& alloc = Allocator())
      : _nrow(nrow), _ncol(ncol), _data(nrow * ncol, val, alloc) {
    _frobenius_norm = (Num)(std::sqrt((double)(nrow * ncol)) * (double)val);
  }
  explicit Matrix(const Num* vals, size_t nrow, size_t ncol, const Allocator& alloc = Allocator())
      : _nrow(nrow), _ncol(ncol), _data(vals, vals + nrow * ncol, alloc) {
    calc_frobenius_norm();
  }

  std::vector<Num, Allocator>& data() & { return _data; }
  const std::vector<Num, Allocator>& data() const& { return _data; }
  Num* data_ptr() & { return _data.data(); }
  const Num* data_ptr() const& { return _data.data(); }
  size_t nrow() const { return _nrow; }
  size_t ncol() const { return _ncol; }
  size_t size() const { return _nrow * _ncol; }
  Num& elem(size_t r, size_t c) & { return _data[r + c * _nrow]; }
  const Num& elem(size_t r, size_t c) const& { return _data[r + c * _nrow]; }
  Num& frobenius_norm() & { return _frobenius_norm; }
  const Num& frobenius_norm() const& { return _frobenius_norm; }

  // matrix.hpp:200-219: raw column-major binary, missing file -> exit(1)
  void read_from_file(const char* filename, const size_t start_offset_in_file = 0) {
    FILE* f = std::fopen(filename, "rb");
    if (!f) {
      std::printf("File %s not available\n", filename);
      std::exit(1);
    }
    std::fseek(f, 0, SEEK_END);
    const size_t file_elems = (size_t)std::ftell(f) / sizeof(Num);
    if (file_elems < start_offset_in_file + size()) {
      std::fclose(f);
      throw std::runtime_error(std::string("file too small: ") + filename);
    }
    std::fseek(f, (long)(start_offset_in_file * sizeof(Num)), SEEK_SET);
    const size_t got = std::fread(data_ptr(), sizeof(Num), size(), f);
    std::fclose(f);
    if (got != size()) throw std::runtime_error(std::string("short read: ") + filename);
  }
  void write_to_file(const char* filename, const size_t start_offset_in_file = 0) const {
    FILE* f = std::fopen(filename, "wb");
    if (!f) throw std::runtime_error(std::string("cannot write ") + filename);
    std::fseek(f, (long)(start_offset_in_file * sizeof(Num)), SEEK_SET);
    std::fwrite(data_ptr(), sizeof(Num), size(), f);
    std::fclose(f);
  }
  void fill_with_values(const Num val) & { std::fill(_data.begin(), _data.end(), val); }

  // matrix.hpp:189-193 (dnrm2 -> device canonical norm)
  Num calc_frobenius_norm() {
    double r = 0.0;
    sa_cpp::check(sa_dnrm2(sa_cpp::context(), (int64_t)size(), (const double*)data_ptr(), &r), "sa_dnrm2");
    _frobenius_norm = (Num)r;
    return _frobenius_norm;
  }
  void axpy(const Matrix& rhs, const Num scale) & {
    assert(rhs.size() == size());
    sa_cpp::check(sa_daxpy(sa_cpp::context(), (int64_t)size(), (double)scale, (const double*)rhs.data_ptr(),
                           (double*)data_ptr()),
                  "sa_daxpy");
  }
  Matrix& operator+=(const Matrix& rhs) & {
    axpy(rhs, Num(1));
    return *this;
  }
  Matrix& operator-=(const Matrix& rhs) & {
    axpy(rhs, Num(-1));
    return *this;
  }
  Matrix& operator*=(const Num& s) & {
    for (auto& x : _data) x *= s;
    return *this;
  }
  // matrix.hpp:179-187
  template <typename Num2, typename Allocator2>
  void add_transpose(const Matrix<Num2, Allocator2>& t) {
    for (size_t c = 0; c < _ncol; ++c)
      for (size_t r = 0; r < _nrow; ++r) elem(r, c) += t.elem(c, r);
  }

 private:
  size_t _nrow = 0, _ncol = 0;
  Num _frobenius_norm = Num();
  std::vector<Num, Allocator> _data;
};

using DMat = Matrix<double, std::allocator<double>>;

template <typename Num, typename A>
Matrix<Num, A> operator-(const Matrix<Num, A>& lhs, const Matrix<Num, A>& rhs) {
  Matrix<Num, A> r(lhs);
  r -= rhs;
  return r;
}
template <typename Num, typename A>
Matrix<Num, A> operator+(const Matrix<Num, A>& lhs, const Matrix<Num, A>& rhs) {
  Matrix<Num, A> r(lhs);
  r += rhs;
  return r;
}

// matrix.hpp:229-278: C = alpha op(A) op(B) + beta C through dgemm (device)
template <typename Num, typename A1, typename A2, typename A3>
void matmult(Matrix<Num, A1>& C, const Matrix<Num, A2>& A, const bool transA, const Matrix<Num, A3>& B,
             const bool transB, const Num& alpha = Num(1), const Num& beta = Num(0)) {
  const size_t m = C.nrow(), n = C.ncol(), k = transA ? A.nrow() : A.ncol();
  if ((transA ? A.ncol() : A.nrow()) != m || (transB ? B.nrow() : B.ncol()) != n ||
      (transB ? B.ncol() : B.nrow()) != k)
    throw std::runtime_error("matmult: dimensions do not work out");
  sa_cpp::check(sa_dgemm(sa_cpp::context(), transA, transB, (int64_t)m, (int64_t)n, (int64_t)k, alpha,
                         A.data_ptr(), (int64_t)std::max<size_t>(A.nrow(), 1), B.data_ptr(),
                         (int64_t)std::max<size_t>(B.nrow(), 1), beta, C.data_ptr(), (int64_t)std::max<size_t>(m, 1)),
                "sa_dgemm");
}

template <typename Num, typename Alloc>
Num nrm2(const Matrix<Num, Alloc>& m) {
  double r = 0.0;
  sa_cpp::check(sa_dnrm2(sa_cpp::context(), (int64_t)m.size(), (const double*)m.data_ptr(), &r), "sa_dnrm2");
  return (Num)r;
}

#endif
