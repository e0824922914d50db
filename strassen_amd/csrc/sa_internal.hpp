// sa_internal.hpp -- internal types of the MI355X block-sparse multiply library.
//
// Device layout of a block-sparse matrix (HBM):
//   arena   : one device allocation per matrix holding every present block,
//             each block column-major (ld = its row count), blocks packed by a
//             bump pointer in units of doubles;
//   off[b]  : int64 element offset of block b in the arena, -1 = empty block;
//   norm[b] : cached Frobenius norm of block b (FP64);
//   b = rb + cb*nrb (+ l*nrb*ncb for a batch slice l), the reference's
//   column-major block grid (blocksparsematrix.h:71-72).
// Block geometry (sizes, screening sizes, row/column starts) lives in small
// device arrays next to host copies.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <functional>
#include <memory>
#include <vector>

#include "../../include/strassen_amd.h"

namespace sa {

// ------------------------------------------------------------------ errors --
void set_error(const char* fmt, ...);

struct Error {
  int code;
};

#define SA_CHECK_HIP(expr)                                                            \
  do {                                                                                \
    hipError_t _e = (expr);                                                           \
    if (_e != hipSuccess) {                                                           \
      ::sa::set_error("%s:%d: %s -> %s", __FILE__, __LINE__, #expr, hipGetErrorString(_e)); \
      throw ::sa::Error{_e == hipErrorOutOfMemory ? SA_ERR_OOM : SA_ERR_HIP};         \
    }                                                                                 \
  } while (0)

#define SA_REQUIRE(cond, ...)          \
  do {                                 \
    if (!(cond)) {                     \
      ::sa::set_error(__VA_ARGS__);    \
      throw ::sa::Error{SA_ERR_INVALID}; \
    }                                  \
  } while (0)

#define SA_UNSUPPORTED(...)                 \
  do {                                      \
    ::sa::set_error(__VA_ARGS__);           \
    throw ::sa::Error{SA_ERR_UNSUPPORTED};  \
  } while (0)

// Wrap an ABI body: exceptions become status codes.
template <class F>
int guard(F&& f) {
  try {
    f();
    return SA_OK;
  } catch (const Error& e) {
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return SA_ERR_OOM;
  }
}

// -------------------------------------------------------------- device mem --
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  hipStream_t stream = nullptr;      // the owning context's stream (the only one that uses the buffer)
  void reserve(size_t n);            // grow-only, contents discarded
  void reserve_keep(size_t n, hipStream_t s);  // grow-only, contents kept
  void release();
  template <class T>
  T* as() const { return static_cast<T*>(p); }
};

__host__ __device__ inline int64_t ceil_div(int64_t a, int64_t b) { return a == 0 ? 0 : 1 + (a - 1) / b; }

// Device memory for the tables of one multiply (task lists, records, work
// items, scan partials).  By default these are the context's named grow-only
// buffers; with a fixed region (the allocation-free entry point) they are
// carved from it by a bump pointer in 256-byte steps -- nothing is allocated,
// and a table that does not fit raises SA_ERR_INVALID before any kernel
// touches it.
struct TableMem {
  sa_context* c = nullptr;
  uint8_t* base = nullptr;  // fixed region; nullptr = context buffers
  size_t cap = 0, used = 0;
  explicit TableMem(sa_context* ctx) : c(ctx) {}
  TableMem(sa_context* ctx, void* region, size_t bytes) : c(ctx), base(static_cast<uint8_t*>(region)), cap(bytes) {}
  bool fixed() const { return base != nullptr; }
  void* raw(const char* name, size_t bytes);
  template <class T>
  T* get(const char* name, int64_t n) {
    return static_cast<T*>(raw(name, sizeof(T) * (size_t)(n > 0 ? n : 1)));
  }
  // n int64 values device -> host, then wait for the stream (pinned staging
  // for context buffers; straight into h for a fixed region)
  void read(const int64_t* d, int64_t* h, int n);
};

// Work-item index of workgroup `bid` such that consecutive items run on the
// same XCD: workgroups are dispatched round-robin over the 8 XCDs, so item
// ranges of about n/8 land on one XCD and share its L2.
__device__ inline int64_t xcd_remap(int64_t bid, int64_t n) {
  const int64_t q = n / 8, r = n % 8;
  const int64_t x = bid % 8, t = bid / 8;
  return x * q + (x < r ? x : r) + t;
}

// Canonical block norm: lane l accumulates fma(x,x,acc) over e = l, l+64, ...;
// a halving tree s_l += s_{l+off} (off = 32..1); IEEE sqrt in lane 0.
// The loads run 8 ahead of the fma chain (8 in flight per lane); the chain
// keeps the order e = l, l + 64, l + 128, ...
__device__ inline double wave_block_norm(const double* __restrict__ x, int64_t n, int lane) {
  double a = 0.0;
  int64_t e = lane;
  for (; e + 7 * 64 < n; e += 8 * 64) {
    double y[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) y[t] = x[e + 64 * t];
#pragma unroll
    for (int t = 0; t < 8; ++t) a = fma(y[t], y[t], a);
  }
  for (; e < n; e += 64) a = fma(x[e], x[e], a);
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) {
    double y = __shfl_down(a, off, 64);
    a = a + y;
  }
  return sqrt(a);
}

// Load through the global address space.  Pointers reached through records
// are generic, so plain dereferences compile to flat loads, which also count
// on the LDS wait counter: every wait for an LDS read would then wait for the
// in-flight operand prefetch of the next pipeline stage as well.
template <class T>
__device__ inline T gload1(const T* p) {
  static_assert(sizeof(T) % 4 == 0, "32-bit words");
  constexpr int N = (int)(sizeof(T) / 4);
  typedef const __attribute__((address_space(1))) unsigned* gptr;
  const gptr q = (gptr)__builtin_assume_aligned(p, alignof(T));
  unsigned w[N];
#pragma unroll
  for (int i = 0; i < N; ++i) w[i] = q[i];
  T v;
  __builtin_memcpy(&v, w, sizeof(T));
  return v;
}

// Position of C block (ib, jb) in the output-record order: 4 x 4 groups of C
// blocks are consecutive, so the tiles of neighbouring outputs, which share
// op(A) block rows and op(B) block columns, run together on one XCD (see
// xcd_remap) and meet in its L2.  Slots form a padded grid of
// ceil(nib/4)*4 x ceil(njb/4)*4.
constexpr int kOutGroup = 4;
__host__ __device__ inline int64_t out_slot(int ib, int jb, int nib) {
  const int64_t ngi = ceil_div(nib, kOutGroup);
  return ((int64_t)(ib / kOutGroup) + (int64_t)(jb / kOutGroup) * ngi) * (kOutGroup * kOutGroup) + ib % kOutGroup +
         (jb % kOutGroup) * kOutGroup;
}
inline int64_t out_slots(int nib, int njb) {
  return ceil_div(nib, kOutGroup) * ceil_div(njb, kOutGroup) * kOutGroup * kOutGroup;
}

// --------------------------------------------------------------- profiling --
struct ProfAcc {
  double ms = 0.0;
  int64_t launches = 0;
};

}  // namespace sa

struct sa_context {
  int device = 0;
  hipStream_t stream = nullptr;
  int num_cus = 256;
  // scratch workspaces (grow-only)
  sa::DevBuf ws_scan, ws_tmp0, ws_tmp1, ws_tmp2, ws_stage, ws_gemm_items, ws_gemm_items2;
  // named grow-only buffers of the multiply pipelines (task lists, records)
  std::map<std::string, sa::DevBuf> bufs;
  // last task list (canonical order): per C block count ("mm_cnt"), exclusive
  // start ("mm_seg") and the kb list ("mm_kb")
  int64_t last_ntasks = 0, last_nc = 0;
  int32_t last_nib = 0;
  const int64_t* last_cnt = nullptr;  // device tables of that task list
  const int64_t* last_seg = nullptr;
  const int32_t* last_kb = nullptr;
  bool last_sym = false;  // blocks ib > jb hold the mirrored task lists of (jb, ib)
  bool quiet = false;     // no profiling events (allocation-free calls)
  int32_t last_gemm_cfg = -1;
  int32_t last_kwindows = 1;  // k-windows of the last screened multiply
  bool oz_pairs_ready = false;  // buffer "oz_pairs" (split-precision slice pairs run / skipped) zeroed
  // executable graph of the Strassen execution phase, kept between calls:
  // each call's capture is applied to it (hipGraphExecUpdate) and it is
  // launched again; a capture of another shape instantiates it anew
  hipGraphExec_t graph_exec = nullptr;
  int64_t graph_instantiated = 0, graph_updated = 0, graph_launched = 0;
  // learned path selector (device copy of the network in buffer "mlp_prm")
  bool mlp_loaded = false;
  int32_t mlp_step = 0;  // Adam steps taken
  sa::DevBuf& buf(const char* name) {
    sa::DevBuf& b = bufs[name];
    b.stream = stream;
    return b;
  }
  // profiling
  bool profiling = false;
  std::map<std::string, sa::ProfAcc> prof;
  std::vector<std::pair<std::string, std::pair<hipEvent_t, hipEvent_t>>> pending;
  hipEvent_t marks[16] = {};
  // pinned host staging for small readbacks
  void* pinned = nullptr;
  size_t pinned_bytes = 0;
};

struct sa_bsm {
  sa_context* ctx = nullptr;
  int32_t nrb = 0, ncb = 0;
  int64_t nrow = 0, ncol = 0;
  double thresh = 0.0;
  bool uniform = false;
  int32_t bs_r = 0, bs_c = 0;
  std::vector<int32_t> rsz, csz, rscr, cscr;
  std::vector<int64_t> roff, coff;  // start row of block row rb, start col of block col cb
  // device geometry
  int32_t *d_rsz = nullptr, *d_csz = nullptr, *d_rscr = nullptr, *d_cscr = nullptr;
  int64_t *d_roff = nullptr, *d_coff = nullptr;
  // device block table
  int64_t* d_off = nullptr;
  double* d_norm = nullptr;
  // arena
  sa::DevBuf arena;
  int64_t used = 0;  // doubles in use (bump pointer)
  // reshape view: shares base's block table, norms and arena (read-only)
  sa_bsm* base = nullptr;
  int64_t nb() const { return (int64_t)nrb * ncb; }
  double* data() const { return base ? base->data() : arena.as<double>(); }
};

namespace sa {

// context helpers
void* pinned(sa_context* c, size_t bytes);
void prof_begin(sa_context* c, const char* cls, hipEvent_t* ev0);
void prof_end(sa_context* c, const char* cls, hipEvent_t ev0);
void prof_collect(sa_context* c);  // after a synchronize

struct ProfScope {
  sa_context* c;
  const char* cls;
  hipEvent_t ev0 = nullptr;
  ProfScope(sa_context* c_, const char* cls_) : c(c_), cls(cls_) { prof_begin(c, cls, &ev0); }
  ~ProfScope() { prof_end(c, cls, ev0); }
};

// exclusive scan of n int64 values in place (device), returns the total (sync)
int64_t scan_exclusive_i64(sa_context* c, int64_t* d, int64_t n, bool read_total = true);
// the same with its partial sums in mem
int64_t scan_exclusive_i64(TableMem& mem, int64_t* d, int64_t n);
// int64 workspace elements a scan of n values needs
int64_t scan_ws_elems(int64_t n);
// same, total written to a device slot, no synchronisation; ws (scan_ws_elems
// elements) or the context's scan workspace when null
void scan_exclusive_i64_async(sa_context* c, int64_t* d, int64_t n, int64_t* total_dev);
// the same from a separate input array (left untouched) into d
void scan_exclusive_i64_async(sa_context* c, const int64_t* src, int64_t* d, int64_t n, int64_t* total_dev,
                              int64_t* ws = nullptr);
// three independent scans (src[i] -> d[i], n[i] values, total[i]); one launch
// when each fits one scan tile (2048 values)
void scan3_exclusive_i64_async(sa_context* c, const int64_t* const src[3], int64_t* const d[3], const int64_t n[3],
                               int64_t* const total[3], int64_t* ws = nullptr);
// read one int64 from device (sync)
int64_t read_i64(sa_context* c, const int64_t* d);

// small fill kernels (used instead of memset so the work can sit in a HIP graph)
void fill_i64(sa_context* c, int64_t* p, int64_t n, int64_t v);
void fill_f64(sa_context* c, double* p, int64_t n, double v);

// arena management
void bsm_ensure_capacity(sa_bsm* m, int64_t elems);
// copy-in (blocksparsematrix.hpp:35-61) from a dense device array
void bsm_from_dense_device(sa_bsm* m, const double* d_dense, int64_t ld);
void bsm_upload_geometry(sa_bsm* m);
void bsm_alloc_tables(sa_bsm* m);

// norm kernel launcher: recompute norms of every present block (or all if mask==nullptr)
// canonical block norms; mask (may be null) selects the blocks to refresh
void launch_block_norms(sa_bsm* m, const int64_t* mask = nullptr);

// ------------------------------------------------------------ grouped GEMM --
// One output block (column-major, ld = m) and its segment [p0, p1) of products
// C = alpha * sum_p op(A_p) op(B_p) + (acc ? C : 0).
struct GemmOut {
  double* c;
  int32_t m, n;
  int32_t p0, p1;
  int32_t acc;
  int32_t group;  // 1 + group column of the C block (out_slot / 4), 0 = none: the work-item order
  int32_t hint;   // split-precision GEMM: 1 = take the zero-slice path (its A slices often all zero)
  int32_t pad_;
};
// op(A) is m x k, op(B) is k x n; stored blocks are column-major with
// ld = rows of the stored block (A: m or k, B: k or n).
struct GemmProd {
  const double* a;
  const double* b;
  int32_t k;
  int32_t pad;
};
struct GemmItem {
  int32_t out;
  int16_t tm, tn;
};
struct GemmBatch {
  const GemmOut* outs = nullptr;  // device
  int64_t nouts = 0;
  const GemmProd* prods = nullptr;  // device
  int64_t nprods = 0;
  int ta = 0, tb = 0;
  double alpha = 1.0;
  int32_t max_m = 0, max_n = 0, max_k = 0;
  double flops = 0.0;  // algorithmic flops (for autotuning statistics)
};
// Result of the shared screening stage of a screened multiply: per C block
// task counts / segment starts, the canonical kb list, new-block and
// non-empty-output scans (all device, owned by the context's "mm_*" buffers).
struct ScreenResult {
  int32_t nib = 0, njb = 0, nkb = 0;
  int64_t nc = 0, ntasks = 0, nnew = 0, nouts = 0, cbase = 0;
  const int64_t* cnt = nullptr;
  const int64_t* seg = nullptr;
  const int64_t* need_scan = nullptr;
  const int64_t* out_scan = nullptr;  // indexed by out_slot(ib, jb, nib)
  const int32_t* kb = nullptr;
  bool sym = false;
};
// sym: C = op(A) op(B) is symmetric (op(B) = op(A)^T, the RI-K K
// contraction): only C blocks ib <= jb are screened and computed, and
// mirror_lower() then writes each block (jb, ib) as the transpose of (ib, jb);
// the task list of a mirrored block is its partner's, triple for triple.
ScreenResult screen_and_allocate(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh,
                                 const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist,
                                 TableMem& mem, bool sym = false);
ScreenResult screen_and_allocate(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh,
                                 const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist);
void mirror_lower(sa_bsm* C, const ScreenResult& R);
// A split-precision operand prepared ahead of its multiply (sa_ozaki.hip).
struct OzPrep;
struct OzPrepDelete {
  void operator()(OzPrep* p) const;
};
using OzPrepHandle = std::unique_ptr<OzPrep, OzPrepDelete>;
// For op(A) = V^T, V a reshape view of M (view block b = M's block b): M's
// block norms and the slicer's exponents of V's columns in one pass over M;
// ozaki_multiply(C, V, true, ..., prep) then only cuts the digits.
OzPrepHandle ozaki_prepare_view(sa_bsm* M, sa_bsm* V, const sa_opts_t& o);
int64_t ozaki_multiply(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh, double alpha,
                       const sa_opts_t& o, bool sym = false, const OzPrep* prep = nullptr);
int64_t screened_multiply(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh, double alpha,
                          int tile_cfg, const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist,
                          bool sym = false);
// k-window batching.  When the operands of a multiply far exceed the
// memory-side cache (Infinity Cache, 256 MB) while the C blocks it writes fit
// in it, the products are split by inner block index kb into windows of
// consecutive kb whose operand bytes fit a budget, and one grouped launch runs
// per window, in ascending kb order: the outputs of window w carry the
// sub-range of their product segment that falls in w (acc = 1 after an
// output's first window), so every C block still sums its products in kb
// order, and all tiles reading a window's operands run close together in time.
struct KWindows {
  int nwin = 1;
  const GemmOut* outs = nullptr;  // device, grouped by window
  std::vector<int64_t> off;       // nwin + 1 offsets into outs
};
// nwin == 1 (and outs == full records) when windows do not pay.
KWindows plan_kwindows(sa_context* c, sa_bsm* A, bool ta, sa_bsm* B, bool tb, const ScreenResult& R,
                       const GemmOut* outs);
// per-device address of the allocation-free multiply's scratch region,
// looked up once with the device current (context creation)
void noalloc_setup(int device);
// learned path selector network (sa_mlp.hip)
constexpr int kMlpFeatures = 10, kMlpOutputs = 2;
constexpr int kMlpParams = 32 * kMlpFeatures + 32 + 32 * 32 + 32 + kMlpOutputs * 32 + kMlpOutputs;
bool mlp_loaded(sa_context* c);
// Y (n x kMlpOutputs) = network(X (n x kMlpFeatures)), device arrays, async
void mlp_forward_device(sa_context* c, const double* dX, int64_t n, double* dY);
// Tile configurations of the FP64 MFMA kernel.
int gemm_fp64_num_configs();
const char* gemm_fp64_config_name(int cfg);
// Run a batch with tile config cfg (-1: autotuned choice for this shape class).
void gemm_fp64_run(sa_context* c, const GemmBatch& b, int cfg);
// the same in two halves: the tile config (autotuned when cfg < 0) and the
// work-item list, which may synchronise; then one launch with no host
// interaction, which a stream capture can record
struct GemmPlan {
  int cfg = 0;
  const GemmItem* items = nullptr;
  int64_t nitems = 0;
};
GemmPlan gemm_fp64_plan(sa_context* c, const GemmBatch& b, int cfg);
// the same without timing runs or allocation: cfg, else the autotuner's
// cached choice for this shape class, else the size heuristic; work items in mem
GemmPlan gemm_fp64_plan_fixed(const GemmBatch& b, int cfg, TableMem& mem);
void gemm_fp64_launch(sa_context* c, const GemmBatch& b, const GemmPlan& g);
// tile config id for a batch: cfg when valid, else the autotuner's choice
int gemm_fp64_choose(sa_context* c, const GemmBatch& b, int cfg);
// output tile (rows, columns) of a tile config
void gemm_fp64_tile(int cfg, int* tm, int* tn);
// Work items of every k-window at once: the tiles of the window-ordered
// outputs (TM x TN tiles; item.out indexes outs) counted, scanned and emitted
// in one pass and split at the window boundaries off, with one host read-back
// of the nwin + 1 item offsets, so the windows' launches then follow each
// other without host synchronisation.  items == nullptr when no block exceeds
// one tile (then window w's items are its outputs off[w] .. off[w + 1]).
struct WindowItems {
  const GemmItem* items = nullptr;
  std::vector<int64_t> off;  // nwin + 1 item offsets (outputs when items == nullptr)
};
WindowItems window_items(sa_context* c, const GemmOut* outs, const std::vector<int64_t>& off, int TM, int TN,
                         int32_t max_m, int32_t max_n);
// run f's launches on c's stream as one HIP graph (captured, instantiated,
// launched once); f must not synchronise or copy from the host
void run_as_graph(sa_context* c, const std::function<void()>& f);

}  // namespace sa
