// blocksparsematrix.h -- drop-in for the reference's BlockSparseMatrix<Num>
// (blocksparsematrix.h:18-102, blocksparsematrix.hpp) whose blocks live in HBM.
//
// Same constructors, members and free functions (matmult, dot); every body
// calls the C ABI in include/strassen_amd.h.  Blocks are not host objects, so
// block(rb, cb) returns a host copy (read access) and blocks() is not offered;
// block_present() / block_norm() read the metadata alone, and dense access
// goes through to_matrix() / to_pointer().  The memory pool argument is accepted for source compatibility:
// device storage is a per-matrix arena (see blockallocator.h).
#ifndef SA_BLOCKSPARSEMATRIX_H
#define SA_BLOCKSPARSEMATRIX_H
#include <algorithm>
#include <memory>
#include <vector>

#include "blockallocator.h"
#include "matrix.h"
#include "sa_runtime.hpp"

template <typename Num>
class BlockSparseMatrix final {
  static_assert(std::is_same<Num, double>::value, "the device path is FP64");

 public:
  explicit BlockSparseMatrix() = default;
  // blocksparsematrix.hpp:6-16
  explicit BlockSparseMatrix(size_t nr, size_t nc, size_t bs_r, size_t bs_c, Num thresh_in,
                             std::shared_ptr<BlockMemoryPool<Num>> pool = nullptr)
      : _bs_r(bs_r), _bs_c(bs_c), _thresh(thresh_in), _pool(pool) {
    sa_cpp::check(sa_bsm_create(sa_cpp::context(), (int64_t)nr, (int64_t)nc, (int32_t)bs_r, (int32_t)bs_c, thresh_in,
                                &_h),
                  "sa_bsm_create");
    refresh_shape();
  }
  // blocksparsematrix.hpp:19-25
  explicit BlockSparseMatrix(const Num* input_vals, size_t nr, size_t nc, size_t bs_r, size_t bs_c, Num thresh_in,
                             std::shared_ptr<BlockMemoryPool<Num>> pool = nullptr)
      : BlockSparseMatrix(nr, nc, bs_r, bs_c, thresh_in, pool) {
    copy_from_input_array(input_vals);
  }
  // blocksparsematrix.hpp:28-32
  template <class Allocator>
  explicit BlockSparseMatrix(const Matrix<Num, Allocator>& in, size_t bs_r, size_t bs_c, Num thresh_in,
                             std::shared_ptr<BlockMemoryPool<Num>> pool = nullptr)
      : BlockSparseMatrix(in.data_ptr(), in.nrow(), in.ncol(), bs_r, bs_c, thresh_in, pool) {}
  // extension: uneven block rows / columns (screening sizes = actual sizes)
  explicit BlockSparseMatrix(const std::vector<int32_t>& row_sizes, const std::vector<int32_t>& col_sizes,
                             Num thresh_in)
      : _thresh(thresh_in), _rsz(row_sizes), _csz(col_sizes) {
    sa_cpp::check(sa_bsm_create_blocked(sa_cpp::context(), (int32_t)row_sizes.size(), row_sizes.data(), nullptr,
                                        (int32_t)col_sizes.size(), col_sizes.data(), nullptr, thresh_in, &_h),
                  "sa_bsm_create_blocked");
    refresh_shape();
  }

  BlockSparseMatrix(const BlockSparseMatrix& o)
      : _bs_r(o._bs_r), _bs_c(o._bs_c), _thresh(o._thresh), _pool(o._pool), _rsz(o._rsz), _csz(o._csz) {
    if (!o._h) return;
    clone_geometry(o);
    sa_cpp::check(sa_bsm_copy(_h, o._h), "sa_bsm_copy");
  }
  BlockSparseMatrix& operator=(const BlockSparseMatrix& o) {
    if (this == &o) return *this;
    if (!_h || !same_shape(o)) {
      release();
      _bs_r = o._bs_r; _bs_c = o._bs_c; _thresh = o._thresh; _pool = o._pool; _rsz = o._rsz; _csz = o._csz;
      if (o._h) clone_geometry(o);
    }
    if (o._h) sa_cpp::check(sa_bsm_copy(_h, o._h), "sa_bsm_copy");
    return *this;
  }
  BlockSparseMatrix(BlockSparseMatrix&& o) noexcept { swap(o); }
  BlockSparseMatrix& operator=(BlockSparseMatrix&& o) noexcept {
    swap(o);
    return *this;
  }
  ~BlockSparseMatrix() { release(); }

  // blocksparsematrix.hpp:35-70
  void copy_from_input_array(const Num* input_vals) {
    sa_cpp::check(sa_bsm_from_dense(_h, input_vals, (int64_t)_nrow), "sa_bsm_from_dense");
  }
  template <class Allocator>
  void copy_from_input_matrix(const Matrix<Num, Allocator>& in) {
    if (in.nrow() != _nrow || in.ncol() != _ncol) throw std::runtime_error("copy_from_input_matrix: shape");
    copy_from_input_array(in.data_ptr());
  }
  void fill_with_values(const Num val) { sa_cpp::check(sa_bsm_fill(_h, val), "sa_bsm_fill"); }
  void scale(Num s) { sa_cpp::check(sa_bsm_scale(_h, s), "sa_bsm_scale"); }
  // extension: recompute_norms = true refreshes the cached norms from the
  // scaled blocks instead of the reference's norm *= s*s
  void scale(Num s, bool recompute_norms) {
    sa_cpp::check(sa_bsm_scale_ex(_h, s, recompute_norms ? SA_NORM_RULE_RECOMPUTE : SA_NORM_RULE_REFERENCE),
                  "sa_bsm_scale_ex");
  }
  void zero() { sa_cpp::check(sa_bsm_zero(_h), "sa_bsm_zero"); }
  // blocksparsematrix.hpp:103-124: only present blocks are written
  void to_pointer(Num* out) const { sa_cpp::check(sa_bsm_to_dense(_h, out, (int64_t)_nrow), "sa_bsm_to_dense"); }
  Matrix<Num, std::allocator<Num>> to_matrix() const {
    Matrix<Num, std::allocator<Num>> r(Num(0), _nrow, _ncol);
    to_pointer(r.data_ptr());
    return r;
  }
  void print(const char* name = "", const char* format = "%10.5f", size_t n_per_row = 6) const {
    const auto m = to_matrix();
    std::printf("%s (%zu x %zu)\n", name, _nrow, _ncol);
    for (size_t r = 0; r < _nrow; ++r) {
      for (size_t c = 0; c < _ncol; ++c) {
        std::printf(format, m.elem(r, c));
        if ((c + 1) % n_per_row == 0 || c + 1 == _ncol) std::printf("\n");
      }
    }
  }

  size_t nrow() const { return _nrow; }
  size_t ncol() const { return _ncol; }
  size_t nrowblocks() const { return _nrb; }
  size_t ncolblocks() const { return _ncb; }
  size_t max_blocksize_row() const { return _bs_r; }
  size_t max_blocksize_col() const { return _bs_c; }
  Num thresh() const { return _thresh; }
  size_t nblocks() const { return _nrb * _ncb; }
  size_t no_of_alloc_blocks() const {
    int64_t n = 0;
    sa_cpp::check(sa_bsm_info(_h, nullptr, nullptr, nullptr, nullptr, &n, nullptr), "sa_bsm_info");
    return (size_t)n;
  }
  size_t calc_alloc_dim() const {
    int64_t bytes = 0;
    sa_cpp::check(sa_bsm_info(_h, nullptr, nullptr, nullptr, nullptr, nullptr, &bytes), "sa_bsm_info");
    return (size_t)bytes / sizeof(Num);
  }

  BlockSparseMatrix& operator-=(const BlockSparseMatrix& rhs) {
    sa_cpp::check(sa_bsm_add(_h, rhs._h, -1.0), "sa_bsm_add");
    return *this;
  }
  BlockSparseMatrix& operator+=(const BlockSparseMatrix& rhs) {
    sa_cpp::check(sa_bsm_add(_h, rhs._h, 1.0), "sa_bsm_add");
    return *this;
  }
  BlockSparseMatrix& operator*=(Num rhs) {
    scale(rhs);
    return *this;
  }
  void add_transpose(const BlockSparseMatrix& rhs) {
    sa_cpp::check(sa_bsm_add_transpose(_h, rhs._h), "sa_bsm_add_transpose");
  }
  void calc_frobenius_norms() { sa_cpp::check(sa_bsm_calc_norms(_h), "sa_bsm_calc_norms"); }
  void recompress() { sa_cpp::check(sa_bsm_recompress(_h), "sa_bsm_recompress"); }
  Num frobenius_norm() const {
    double r = 0.0;
    sa_cpp::check(sa_bsm_frobenius_norm(_h, &r), "sa_bsm_frobenius_norm");
    return r;
  }
  // blocksparsematrix.h:71-72, read access: a host copy of block (rb, cb),
  // 0 x 0 when the block is absent, with the block's cached norm
  Matrix<Num, std::allocator<Num>> block(size_t rb, size_t cb) const {
    const auto m = meta();
    const size_t b = rb + cb * _nrb;
    if (!m.first[b]) return Matrix<Num, std::allocator<Num>>();
    const size_t r = block_rows(rb), c = block_cols(cb);
    Matrix<Num, std::allocator<Num>> out(r, c);
    sa_cpp::check(sa_bsm_get_blocks(_h, (int32_t)rb, (int32_t)rb + 1, (int32_t)cb, (int32_t)cb + 1, out.data_ptr(),
                                    (int64_t)r),
                  "sa_bsm_get_blocks");
    out.frobenius_norm() = m.second[b];
    return out;
  }
  // per-block metadata (replaces block(rb, cb).size() != 0 / .frobenius_norm())
  bool block_present(size_t rb, size_t cb) const { return meta().first[rb + cb * _nrb] != 0; }
  Num block_norm(size_t rb, size_t cb) const { return meta().second[rb + cb * _nrb]; }
  std::pair<std::vector<uint8_t>, std::vector<double>> meta() const {
    std::pair<std::vector<uint8_t>, std::vector<double>> r;
    r.first.resize(nblocks());
    r.second.resize(nblocks());
    sa_cpp::check(sa_bsm_get_meta(_h, r.first.data(), r.second.data()), "sa_bsm_get_meta");
    return r;
  }
  std::shared_ptr<BlockMemoryPool<Num>> mem_pool_ptr() const { return _pool; }
  sa_bsm_t handle() const { return _h; }

 private:
  sa_bsm_t _h = nullptr;
  size_t _nrow = 0, _ncol = 0, _nrb = 0, _ncb = 0, _bs_r = 0, _bs_c = 0;
  Num _thresh = Num();
  std::shared_ptr<BlockMemoryPool<Num>> _pool;
  std::vector<int32_t> _rsz, _csz;  // uneven geometry (empty for uniform)

  size_t block_rows(size_t rb) const {
    if (!_rsz.empty()) return (size_t)_rsz[rb];
    return std::min(_bs_r, _nrow - rb * _bs_r);
  }
  size_t block_cols(size_t cb) const {
    if (!_csz.empty()) return (size_t)_csz[cb];
    return std::min(_bs_c, _ncol - cb * _bs_c);
  }
  void refresh_shape() {
    int64_t nr = 0, nc = 0;
    int32_t rb = 0, cb = 0;
    sa_cpp::check(sa_bsm_info(_h, &nr, &nc, &rb, &cb, nullptr, nullptr), "sa_bsm_info");
    _nrow = (size_t)nr; _ncol = (size_t)nc; _nrb = (size_t)rb; _ncb = (size_t)cb;
  }
  bool same_shape(const BlockSparseMatrix& o) const {
    return _nrow == o._nrow && _ncol == o._ncol && _bs_r == o._bs_r && _bs_c == o._bs_c && _rsz == o._rsz &&
           _csz == o._csz;
  }
  void clone_geometry(const BlockSparseMatrix& o) {
    if (!o._rsz.empty() || o._bs_r == 0) {
      sa_cpp::check(sa_bsm_create_blocked(sa_cpp::context(), (int32_t)o._rsz.size(), o._rsz.data(), nullptr,
                                          (int32_t)o._csz.size(), o._csz.data(), nullptr, o._thresh, &_h),
                    "sa_bsm_create_blocked");
    } else {
      sa_cpp::check(sa_bsm_create(sa_cpp::context(), (int64_t)o._nrow, (int64_t)o._ncol, (int32_t)o._bs_r,
                                  (int32_t)o._bs_c, o._thresh, &_h),
                    "sa_bsm_create");
    }
    refresh_shape();
  }
  void release() {
    if (_h) sa_bsm_destroy(_h);
    _h = nullptr;
  }
  void swap(BlockSparseMatrix& o) noexcept {
    std::swap(_h, o._h); std::swap(_nrow, o._nrow); std::swap(_ncol, o._ncol); std::swap(_nrb, o._nrb);
    std::swap(_ncb, o._ncb); std::swap(_bs_r, o._bs_r); std::swap(_bs_c, o._bs_c); std::swap(_thresh, o._thresh);
    std::swap(_pool, o._pool); std::swap(_rsz, o._rsz); std::swap(_csz, o._csz);
  }
};

// blocksparsematrix.h:95-99 / blocksparsematrix.hpp:292-414
template <typename Num>
void matmult(BlockSparseMatrix<Num>& C, const BlockSparseMatrix<Num>& A, const bool transa,
             const BlockSparseMatrix<Num>& B, const bool transb, const Num thresh, const Num& alpha = Num(1),
             const Num& beta = Num(0)) {
  sa_cpp::check(sa_matmult(C.handle(), A.handle(), transa, B.handle(), transb, thresh, alpha, beta, nullptr),
                "sa_matmult");
}

// blocksparsematrix.h:101-102 / blocksparsematrix.hpp:263-290
template <typename Num>
Num dot(const BlockSparseMatrix<Num>& lhs, const BlockSparseMatrix<Num>& rhs, Num thresh = Num(0)) {
  double r = 0.0;
  sa_cpp::check(sa_bsm_dot(lhs.handle(), rhs.handle(), thresh, &r), "sa_bsm_dot");
  return r;
}

#endif
