// sa_context.hip -- context, error state, device buffers, profiling and the
// device prefix scan shared by the table builders.
#include <cstdarg>
#include <cstring>

#include <cstdlib>

#include "sa_internal.hpp"

namespace sa {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

void DevBuf::reserve(size_t n) {
  if (n <= bytes) return;
  if (p) {
    // kernels queued earlier on the owning stream may still read the old
    // buffer (a buffer without an owner waits for the whole device)
    if (stream) SA_CHECK_HIP(hipStreamSynchronize(stream));
    else SA_CHECK_HIP(hipDeviceSynchronize());
    SA_CHECK_HIP(hipFree(p));
  }
  p = nullptr;
  bytes = 0;
  size_t want = n < 4096 ? 4096 : n;
  SA_CHECK_HIP(hipMalloc(&p, want));
  bytes = want;
}

void DevBuf::reserve_keep(size_t n, hipStream_t s) {
  if (n <= bytes) return;
  size_t want = bytes + bytes / 2;
  if (want < n) want = n;
  if (want < 4096) want = 4096;
  void* q = nullptr;
  SA_CHECK_HIP(hipMalloc(&q, want));
  if (p) {
    SA_CHECK_HIP(hipMemcpyAsync(q, p, bytes, hipMemcpyDeviceToDevice, s));
    SA_CHECK_HIP(hipStreamSynchronize(s));
    SA_CHECK_HIP(hipFree(p));
  }
  p = q;
  bytes = want;
}

void DevBuf::release() {
  if (p) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
}

void* pinned(sa_context* c, size_t bytes) {
  if (bytes > c->pinned_bytes) {
    if (c->pinned) SA_CHECK_HIP(hipHostFree(c->pinned));
    c->pinned = nullptr;
    c->pinned_bytes = 0;
    size_t want = bytes < 65536 ? 65536 : bytes;
    SA_CHECK_HIP(hipHostMalloc(&c->pinned, want, hipHostMallocDefault));
    c->pinned_bytes = want;
  }
  return c->pinned;
}

void* TableMem::raw(const char* name, size_t bytes) {
  if (!fixed()) {
    DevBuf& b = c->buf(name);
    b.reserve(bytes);
    return b.p;
  }
  const size_t off = (used + 255) & ~(size_t)255;
  if (off > cap || bytes > cap - off) {
    set_error("allocation-free multiply: table '%s' (%zu bytes) does not fit the %zu-byte scratch region (%zu in use)",
              name, bytes, cap, off);
    throw Error{SA_ERR_INVALID};
  }
  used = off + bytes;
  return base + off;
}

void TableMem::read(const int64_t* d, int64_t* h, int n) {
  if (fixed()) {
    SA_CHECK_HIP(hipMemcpyAsync(h, d, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    return;
  }
  int64_t* p = static_cast<int64_t*>(pinned(c, sizeof(int64_t) * (size_t)n));
  SA_CHECK_HIP(hipMemcpyAsync(p, d, sizeof(int64_t) * (size_t)n, hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (int i = 0; i < n; ++i) h[i] = p[i];
}

void prof_begin(sa_context* c, const char* cls, hipEvent_t* ev0) {
  (void)cls;
  *ev0 = nullptr;
  if (!c->profiling || c->quiet) return;
  SA_CHECK_HIP(hipEventCreate(ev0));
  SA_CHECK_HIP(hipEventRecord(*ev0, c->stream));
}

void prof_end(sa_context* c, const char* cls, hipEvent_t ev0) {
  if (!c->profiling || !ev0) return;
  hipEvent_t ev1;
  if (hipEventCreate(&ev1) != hipSuccess) return;
  if (hipEventRecord(ev1, c->stream) != hipSuccess) return;
  c->pending.push_back({std::string(cls), {ev0, ev1}});
}

void prof_collect(sa_context* c) {
  for (auto& p : c->pending) {
    float ms = 0.f;
    SA_CHECK_HIP(hipEventSynchronize(p.second.second));
    SA_CHECK_HIP(hipEventElapsedTime(&ms, p.second.first, p.second.second));
    auto& acc = c->prof[p.first];
    acc.ms += ms;
    acc.launches += 1;
    (void)hipEventDestroy(p.second.first);
    (void)hipEventDestroy(p.second.second);
  }
  c->pending.clear();
}

// ------------------------------------------------------------------- scan --
constexpr int kScanThreads = 256;
constexpr int kScanItems = 8;
constexpr int kScanTile = kScanThreads * kScanItems;

// Exclusive scan of tile `tile` of src into d (src may be d: each thread
// reads its items before writing them); *sum = the tile's total.
__device__ inline void scan_tile(const int64_t* src, int64_t* d, int64_t n, int64_t tile, int64_t* sum) {
  __shared__ int64_t wsum[kScanThreads / 64];
  const int64_t base = tile * kScanTile + (int64_t)threadIdx.x * kScanItems;
  int64_t v[kScanItems];
  int64_t run = 0;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i) {
    int64_t t = (base + i < n) ? src[base + i] : 0;
    v[i] = run;
    run += t;
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  int64_t x = run;
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) {
    int64_t y = __shfl_up(x, o, 64);
    if (lane >= o) x += y;
  }
  if (lane == 63) wsum[w] = x;
  __syncthreads();
  int64_t wpre = 0;
  for (int i = 0; i < w; ++i) wpre += wsum[i];
  const int64_t excl = wpre + x - run;
#pragma unroll
  for (int i = 0; i < kScanItems; ++i)
    if (base + i < n) d[base + i] = v[i] + excl;
  if (threadIdx.x == kScanThreads - 1) *sum = wpre + x;
}

__global__ void __launch_bounds__(kScanThreads) k_scan_tiles(const int64_t* src, int64_t* d, int64_t n,
                                                              int64_t* __restrict__ tile_sums) {
  scan_tile(src, d, n, blockIdx.x, tile_sums + blockIdx.x);
}

// up to 3 independent one-tile scans, one workgroup each
struct ScanJobs {
  const int64_t* src[3];
  int64_t* d[3];
  int64_t n[3];
  int64_t* total[3];
};
__global__ void __launch_bounds__(kScanThreads) k_scan_jobs(ScanJobs j) {
  const int b = blockIdx.x;
  scan_tile(j.src[b], j.d[b], j.n[b], 0, j.total[b]);
}

__global__ void k_scan_add(int64_t* __restrict__ d, int64_t n, const int64_t* __restrict__ tile_off) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] += tile_off[i / kScanTile];
}

// one tile: its sum is the total, written straight to total_slot
static void scan_rec(sa_context* c, const int64_t* src, int64_t* d, int64_t n, int64_t* ws, int64_t* total_slot) {
  const int64_t ntiles = ceil_div(n, kScanTile);
  int64_t* part = ntiles == 1 ? total_slot : ws;
  hipLaunchKernelGGL(k_scan_tiles, dim3((unsigned)ntiles), dim3(kScanThreads), 0, c->stream, src, d, n, part);
  if (ntiles == 1) return;
  scan_rec(c, part, part, ntiles, ws + ntiles, total_slot);
  hipLaunchKernelGGL(k_scan_add, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, d, n,
                     (const int64_t*)part);
}

int64_t scan_ws_elems(int64_t n) {
  int64_t need = 2, m = n;
  while (m > 1) {
    m = ceil_div(m, kScanTile);
    need += m;
  }
  return need;
}

void scan_exclusive_i64_async(sa_context* c, const int64_t* src, int64_t* d, int64_t n, int64_t* total_dev,
                              int64_t* ws) {
  if (!ws) {
    c->ws_scan.reserve(sizeof(int64_t) * (size_t)scan_ws_elems(n));
    ws = c->ws_scan.as<int64_t>();
  }
  if (n == 0) {
    fill_i64(c, total_dev, 1, 0);
  } else {
    scan_rec(c, src, d, n, ws, total_dev);
  }
  SA_CHECK_HIP(hipGetLastError());
}

void scan_exclusive_i64_async(sa_context* c, int64_t* d, int64_t n, int64_t* total_dev) {
  scan_exclusive_i64_async(c, d, d, n, total_dev);
}

void scan3_exclusive_i64_async(sa_context* c, const int64_t* const src[3], int64_t* const d[3], const int64_t n[3],
                               int64_t* const total[3], int64_t* ws) {
  bool small = true;
  for (int i = 0; i < 3; ++i) small = small && n[i] >= 1 && n[i] <= kScanTile;
  if (!small) {
    // one after the other on the stream: they may share one workspace
    for (int i = 0; i < 3; ++i) scan_exclusive_i64_async(c, src[i], d[i], n[i], total[i], ws);
    return;
  }
  ScanJobs j;
  for (int i = 0; i < 3; ++i) {
    j.src[i] = src[i];
    j.d[i] = d[i];
    j.n[i] = n[i];
    j.total[i] = total[i];
  }
  hipLaunchKernelGGL(k_scan_jobs, dim3(3), dim3(kScanThreads), 0, c->stream, j);
  SA_CHECK_HIP(hipGetLastError());
}

int64_t scan_exclusive_i64(TableMem& mem, int64_t* d, int64_t n) {
  const int64_t need = scan_ws_elems(n);
  int64_t* ws = mem.get<int64_t>("scan_ws", need);
  scan_exclusive_i64_async(mem.c, d, d, n, ws + need - 1, ws);
  int64_t t = 0;
  mem.read(ws + need - 1, &t, 1);
  return t;
}

int64_t scan_exclusive_i64(sa_context* c, int64_t* d, int64_t n, bool read_total) {
  // workspace: partial sums of every level + 1 total slot
  int64_t need = 2, m = n;
  while (m > 1) {
    m = ceil_div(m, kScanTile);
    need += m;
  }
  c->ws_scan.reserve(sizeof(int64_t) * (size_t)need);
  int64_t* ws = c->ws_scan.as<int64_t>();
  int64_t* total_slot = ws + need - 1;
  if (n == 0) {
    fill_i64(c, total_slot, 1, 0);
  } else {
    scan_rec(c, d, d, n, ws, total_slot);
  }
  SA_CHECK_HIP(hipGetLastError());
  if (!read_total) return 0;
  return read_i64(c, total_slot);
}

// SA_GRAPH_DEBUG=1: say on stderr why a graph fell back to eager launches
static void graph_note(const char* what, hipError_t e) {
  static const bool on = std::getenv("SA_GRAPH_DEBUG") != nullptr;
  if (on) std::fprintf(stderr, "strassen_amd: %s: %s\n", what, hipGetErrorString(e));
}

void run_as_graph(sa_context* c, const std::function<void()>& f) {
  const hipError_t be = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal);
  if (be != hipSuccess) {
    graph_note("stream capture unavailable", be);
    (void)hipGetLastError();
    f();  // capture unavailable: the same launches, eagerly
    return;
  }
  hipGraph_t g = nullptr;
  try {
    f();
  } catch (...) {
    (void)hipStreamEndCapture(c->stream, &g);
    if (g) (void)hipGraphDestroy(g);
    throw;
  }
  const hipError_t ee = hipStreamEndCapture(c->stream, &g);
  if (ee != hipSuccess || !g) {
    graph_note("stream capture failed", ee);
    (void)hipGetLastError();
    f();
    return;
  }
  // replay: the kept executable takes this capture's kernel parameters when
  // the captured topology is the same; otherwise it is instantiated anew
  bool ready = false;
  if (c->graph_exec) {
    hipGraphNode_t bad = nullptr;
    hipGraphExecUpdateResult res = hipGraphExecUpdateError;
    const hipError_t ue = hipGraphExecUpdate(c->graph_exec, g, &bad, &res);
    if (ue == hipSuccess && res == hipGraphExecUpdateSuccess) {
      ready = true;
      ++c->graph_updated;
    } else {
      (void)hipGetLastError();
      (void)hipGraphExecDestroy(c->graph_exec);
      c->graph_exec = nullptr;
    }
  }
  if (!ready) {
    const hipError_t ie = hipGraphInstantiate(&c->graph_exec, g, nullptr, nullptr, 0);
    if (ie != hipSuccess) {
      graph_note("graph instantiation failed", ie);
      (void)hipGetLastError();
      c->graph_exec = nullptr;
      (void)hipGraphDestroy(g);
      f();  // the captured launches did not run: run them eagerly
      return;
    }
    ++c->graph_instantiated;
  }
  hipError_t le;
  {
    ProfScope ps(c, "graph");
    le = hipGraphLaunch(c->graph_exec, c->stream);
  }
  const hipError_t se = hipStreamSynchronize(c->stream);
  (void)hipGraphDestroy(g);
  SA_CHECK_HIP(le);
  SA_CHECK_HIP(se);
  ++c->graph_launched;
}

int64_t read_i64(sa_context* c, const int64_t* d) {
  int64_t* h = static_cast<int64_t*>(pinned(c, sizeof(int64_t)));
  SA_CHECK_HIP(hipMemcpyAsync(h, d, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  return *h;
}

}  // namespace sa

// ===================================================================== ABI ==
using namespace sa;

extern "C" {

void sa_opts_default(sa_opts_t* o) {
  if (!o) return;
  std::memset(o, 0, sizeof(*o));
  o->mode = SA_MODE_FP64;
  o->slices = 4;
  o->slice_type = SA_SLICE_FP16;
  o->tile = -1;
  o->scheduler = SA_SCHED_HEURISTIC;
  o->cutoff = 0;
  o->tree = nullptr;
  o->tree_dense = nullptr;
}

const char* sa_last_error(void) { return g_err; }

int sa_device_count(int* n) {
  return guard([&] {
    SA_REQUIRE(n, "null output");
    SA_CHECK_HIP(hipGetDeviceCount(n));
  });
}

int sa_ctx_create(int device, sa_ctx_t* out) {
  return guard([&] {
    SA_REQUIRE(out, "null output");
    int n = 0;
    SA_CHECK_HIP(hipGetDeviceCount(&n));
    SA_REQUIRE(device >= 0 && device < n, "device %d out of range (%d devices)", device, n);
    SA_CHECK_HIP(hipSetDevice(device));
    auto* c = new sa_context();
    c->device = device;
    SA_CHECK_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    for (sa::DevBuf* b : {&c->ws_scan, &c->ws_tmp0, &c->ws_tmp1, &c->ws_tmp2, &c->ws_stage, &c->ws_gemm_items,
                          &c->ws_gemm_items2})
      b->stream = c->stream;
    hipDeviceProp_t prop;
    SA_CHECK_HIP(hipGetDeviceProperties(&prop, device));
    c->num_cus = prop.multiProcessorCount;
    for (auto& e : c->marks) SA_CHECK_HIP(hipEventCreate(&e));
    noalloc_setup(device);
    *out = c;
  });
}

int sa_ctx_destroy(sa_ctx_t c) {
  return guard([&] {
    if (!c) return;
    (void)hipSetDevice(c->device);
    (void)hipStreamSynchronize(c->stream);
    for (auto& p : c->pending) {
      (void)hipEventDestroy(p.second.first);
      (void)hipEventDestroy(p.second.second);
    }
    for (auto& e : c->marks) (void)hipEventDestroy(e);
    if (c->graph_exec) (void)hipGraphExecDestroy(c->graph_exec);
    c->ws_scan.release(); c->ws_tmp0.release(); c->ws_tmp1.release(); c->ws_tmp2.release();
    c->ws_stage.release(); c->ws_gemm_items.release(); c->ws_gemm_items2.release();
    for (auto& kv : c->bufs) kv.second.release();
    if (c->pinned) (void)hipHostFree(c->pinned);
    (void)hipStreamDestroy(c->stream);
    delete c;
  });
}

int sa_ctx_synchronize(sa_ctx_t c) {
  return guard([&] {
    SA_REQUIRE(c, "null context");
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
  });
}

int sa_ctx_set_profiling(sa_ctx_t c, int on) {
  return guard([&] {
    SA_REQUIRE(c, "null context");
    c->profiling = on != 0;
  });
}

int sa_ctx_profile(sa_ctx_t c, const char* cls, double* ms, int64_t* launches) {
  return guard([&] {
    SA_REQUIRE(c && cls, "null argument");
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    auto it = c->prof.find(cls);
    if (ms) *ms = it == c->prof.end() ? 0.0 : it->second.ms;
    if (launches) *launches = it == c->prof.end() ? 0 : it->second.launches;
  });
}

int sa_ctx_profile_reset(sa_ctx_t c) {
  return guard([&] {
    SA_REQUIRE(c, "null context");
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    prof_collect(c);
    c->prof.clear();
  });
}

int sa_ctx_gemm_config(sa_ctx_t c, int32_t* cfg, char* name, int32_t name_len) {
  return guard([&] {
    SA_REQUIRE(c, "null context");
    if (cfg) *cfg = c->last_gemm_cfg;
    if (name && name_len > 0) {
      std::snprintf(name, (size_t)name_len, "%s", c->last_gemm_cfg >= 0 ? gemm_fp64_config_name(c->last_gemm_cfg) : "");
    }
  });
}

int sa_ctx_graph_stats(sa_ctx_t c, int64_t* instantiated, int64_t* updated, int64_t* launched) {
  return guard([&] {
    SA_REQUIRE(c, "null context");
    if (instantiated) *instantiated = c->graph_instantiated;
    if (updated) *updated = c->graph_updated;
    if (launched) *launched = c->graph_launched;
  });
}

int sa_ctx_kwindows(sa_ctx_t c, int32_t* nwin) {
  return guard([&] {
    SA_REQUIRE(c && nwin, "null argument");
    *nwin = c->last_kwindows;
  });
}

int sa_ctx_slice_pairs(sa_ctx_t c, int64_t* multiplied, int64_t* skipped) {
  return guard([&] {
    SA_REQUIRE(c && multiplied && skipped, "null argument");
    int64_t h[2] = {0, 0};
    if (c->oz_pairs_ready) {
      DevBuf& b = c->buf("oz_pairs");
      SA_CHECK_HIP(hipMemcpyAsync(h, b.p, sizeof(h), hipMemcpyDeviceToHost, c->stream));
      SA_CHECK_HIP(hipStreamSynchronize(c->stream));
      fill_i64(c, b.as<int64_t>(), 2, 0);
    }
    *multiplied = h[0];
    *skipped = h[1];
  });
}

int sa_ctx_mark(sa_ctx_t c, int slot) {
  return guard([&] {
    SA_REQUIRE(c && slot >= 0 && slot < 16, "bad mark slot");
    SA_CHECK_HIP(hipEventRecord(c->marks[slot], c->stream));
  });
}

int sa_ctx_elapsed(sa_ctx_t c, int s0, int s1, float* ms) {
  return guard([&] {
    SA_REQUIRE(c && ms && s0 >= 0 && s0 < 16 && s1 >= 0 && s1 < 16, "bad mark slot");
    SA_CHECK_HIP(hipEventSynchronize(c->marks[s1]));
    SA_CHECK_HIP(hipEventElapsedTime(ms, c->marks[s0], c->marks[s1]));
  });
}

}  // extern "C"
