// blockallocator.h -- source-compatibility names for the reference's
// BlockMemoryPool / BlockAllocator (blockallocator.h:8-55).
//
// On the device every block-sparse matrix owns one HBM arena: blocks are packed
// by a bump pointer (element offsets in a per-matrix table), zero()/beta=0
// restart it, growth is geometric and keeps the contents.  There is no per-block
// host allocation, so these types carry only the slab size a caller chose; the
// BlockSparseMatrix constructors accept them and keep the shared_ptr like the
// reference (blocksparsematrix.h:91).
#ifndef SA_BLOCKALLOCATOR_H
#define SA_BLOCKALLOCATOR_H
#include <cstddef>

template <typename T>
class BlockMemoryPool {
 public:
  constexpr BlockMemoryPool() noexcept = default;
  constexpr explicit BlockMemoryPool(size_t blocksize) noexcept : _blocksize(blocksize) {}
  void release_all_memory() {}
  void adjust_blocksize(size_t max_blocksize) { _blocksize = max_blocksize; }
  size_t blocksize() const { return _blocksize; }

 private:
  size_t _blocksize = 0;
};

template <typename T>
class BlockAllocator {
 public:
  using value_type = T;
  BlockAllocator() = delete;
  BlockAllocator(BlockMemoryPool<T>& p) : _pool(&p) {}
  BlockAllocator(BlockMemoryPool<T>* p) : _pool(p) {}
  constexpr bool is_equal(const BlockAllocator& o) const { return _pool == o._pool; }

 private:
  BlockMemoryPool<T>* _pool = nullptr;
};

#endif
