// sa_bsm.hip -- device block-sparse matrix: geometry, arena, block table and the
// container operations of blocksparsematrix.hpp (copy-in/out, norms, fill,
// zero, scale, recompress, +=, -=, add_transpose, frobenius_norm, dot) plus the
// HBM-resident synthetic-input generator.
//
// Kernels here are HBM-bound streaming passes: one wave per block for the
// norm reductions (the canonical 64-lane strided order, see block_norm below),
// one workgroup per block for copies and elementwise updates, coalesced along
// the block's contiguous column-major storage.
#include <cmath>
#include <cstring>

#include "sa_internal.hpp"

namespace sa {

// ------------------------------------------------------------- primitives --
__global__ void k_fill_i64(int64_t* __restrict__ p, int64_t n, int64_t v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}
__global__ void k_fill_f64(double* __restrict__ p, int64_t n, double v) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

void fill_i64(sa_context* c, int64_t* p, int64_t n, int64_t v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_i64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, p, n, v);
}
void fill_f64(sa_context* c, double* p, int64_t n, double v) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_fill_f64, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, p, n, v);
}

// Block-table views passed to kernels by value.
struct Geo {
  const int32_t* rsz;
  const int32_t* csz;
  const int32_t* rscr;
  const int32_t* cscr;
  const int64_t* roff;
  const int64_t* coff;
  int32_t nrb, ncb;
  int64_t nrow;
};

static Geo geo(const sa_bsm* m) {
  return Geo{m->d_rsz, m->d_csz, m->d_rscr, m->d_cscr, m->d_roff, m->d_coff, m->nrb, m->ncb, m->nrow};
}

// ---------------------------------------------------------------- kernels --
// norms of present blocks; mask (optional) restricts to blocks with mask[b]!=0
__global__ void __launch_bounds__(256) k_norms_arena(const double* __restrict__ arena,
                                                     const int64_t* __restrict__ off, Geo g, int64_t nb,
                                                     const int64_t* __restrict__ mask,
                                                     double* __restrict__ norm) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  if (mask && mask[b] == 0) return;
  const int64_t o = off[b];
  if (o < 0) {
    if (lane == 0) norm[b] = 0.0;
    return;
  }
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int64_t n = (int64_t)g.rsz[rb] * g.csz[cb];
  const double v = wave_block_norm(arena + o, n, lane);
  if (lane == 0) norm[b] = v;
}

// norms of blocks of a dense device matrix; need[b] = block size if the block
// survives the copy-in drop rule (blocksparsematrix.hpp:58-59), else 0
__global__ void __launch_bounds__(256) k_dense_norms(const double* __restrict__ dense, int64_t ld, Geo g,
                                                     int64_t nb, double thresh,
                                                     double* __restrict__ norm, int64_t* __restrict__ need) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  const double* base = dense + g.roff[rb] + g.coff[cb] * ld;
  double a = 0.0;
  for (int64_t e = lane; e < n; e += 64) {
    const double x = base[(e % nr) + (e / nr) * ld];
    a = fma(x, x, a);
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a = a + __shfl_down(a, o, 64);
  if (lane == 0) {
    const double est = sqrt(a);
    const double thr = thresh * (double)((int64_t)g.cscr[cb] * (int64_t)g.rscr[rb]);
    const bool drop = est < thr;
    norm[b] = drop ? 0.0 : est;
    need[b] = drop ? 0 : n;
  }
}

// need_scan holds exclusive offsets; keep[] marks survivors -> offsets or -1
__global__ void k_finalize_offsets(const int64_t* __restrict__ scan, const int64_t* __restrict__ keep,
                                   int64_t nb, int64_t base, int64_t* __restrict__ off) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  off[b] = keep[b] ? base + scan[b] : -1;
}

__global__ void __launch_bounds__(256) k_copy_dense_blocks(const double* __restrict__ dense, int64_t ld, Geo g,
                                                           const int64_t* __restrict__ off,
                                                           double* __restrict__ arena) {
  const int64_t b = blockIdx.x;
  const int64_t o = off[b];
  if (o < 0) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  const double* src = dense + g.roff[rb] + g.coff[cb] * ld;
  double* dst = arena + o;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[(e % nr) + (e / nr) * ld];
}

__global__ void __launch_bounds__(256) k_arena_to_dense(const double* __restrict__ arena,
                                                        const int64_t* __restrict__ off, Geo g,
                                                        double* __restrict__ dense, int64_t ld) {
  const int64_t b = blockIdx.x;
  const int64_t o = off[b];
  if (o < 0) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  double* dst = dense + g.roff[rb] + g.coff[cb] * ld;
  const double* src = arena + o;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[(e % nr) + (e / nr) * ld] = src[e];
}

// blocks of rows [rb0, rb0 + nrb) x columns [cb0, cb0 + ncb) into a dense
// sub-matrix whose (0, 0) is the first element of block (rb0, cb0)
__global__ void __launch_bounds__(256) k_arena_to_dense_sub(const double* __restrict__ arena,
                                                            const int64_t* __restrict__ off, Geo g, int32_t rb0,
                                                            int32_t nrb, int32_t cb0, double* __restrict__ dense,
                                                            int64_t ld) {
  const int rb = rb0 + (int)(blockIdx.x % nrb), cb = cb0 + (int)(blockIdx.x / nrb);
  const int64_t o = off[rb + (int64_t)cb * g.nrb];
  if (o < 0) return;
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  double* dst = dense + (g.roff[rb] - g.roff[rb0]) + (g.coff[cb] - g.coff[cb0]) * ld;
  const double* src = arena + o;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[(e % nr) + (e / nr) * ld] = src[e];
}

// --- synthetic generator (definition shared with strassen_amd/synth.py) ----
__device__ inline uint64_t splitmix64(uint64_t x) {
  uint64_t z = x + 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__device__ inline double normal_like(uint64_t base, uint64_t idx) {
  double acc = 0.0;
  const double sc = 1.7763568394002505e-15;  // 2^-49
#pragma unroll
  for (int t = 0; t < 12; ++t) {
    const uint64_t h = splitmix64(base + idx * 12ull + (uint64_t)t);
    acc = acc + (double)(h >> 15) * sc;
  }
  return acc - 6.0;
}

__device__ inline double gen_value(uint64_t base, double s, int64_t r, int64_t c, int64_t nrow) {
  if (s == 0.0) return 0.0;
  return s * normal_like(base, (uint64_t)r + (uint64_t)c * (uint64_t)nrow);
}

__global__ void __launch_bounds__(256) k_gen_norms(uint64_t base, const double* __restrict__ scale, Geo g,
                                                   int64_t nb, double thresh, double* __restrict__ norm,
                                                   int64_t* __restrict__ keep) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  const double s = scale[b];
  double a = 0.0;
  if (s != 0.0)
    for (int64_t e = lane; e < n; e += 64) {
      const double x = gen_value(base, s, g.roff[rb] + e % nr, g.coff[cb] + e / nr, g.nrow);
      a = fma(x, x, a);
    }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a = a + __shfl_down(a, o, 64);
  if (lane == 0) {
    const double est = sqrt(a);
    const double thr = thresh * (double)((int64_t)g.cscr[cb] * (int64_t)g.rscr[rb]);
    const bool drop = est < thr;
    norm[b] = drop ? 0.0 : est;
    keep[b] = drop ? 0 : n;
  }
}

__global__ void __launch_bounds__(256) k_gen_fill(uint64_t base, const double* __restrict__ scale, Geo g,
                                                  const int64_t* __restrict__ off, double* __restrict__ arena) {
  const int64_t b = blockIdx.x;
  const int64_t o = off[b];
  if (o < 0) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  const double s = scale[b];
  double* dst = arena + o;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x)
    dst[e] = gen_value(base, s, g.roff[rb] + e % nr, g.coff[cb] + e / nr, g.nrow);
}

// fill_with_values: every block gets value v, cached norm sqrt(r*c)*v
__global__ void __launch_bounds__(256) k_fill_blocks(Geo g, const int64_t* __restrict__ off, double v,
                                                     double* __restrict__ arena, double* __restrict__ norm) {
  const int64_t b = blockIdx.x;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int64_t n = (int64_t)g.rsz[rb] * g.csz[cb];
  double* dst = arena + off[b];
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = v;
  if (threadIdx.x == 0) norm[b] = sqrt((double)n) * v;
}

__global__ void __launch_bounds__(256) k_scale_blocks(Geo g, const int64_t* __restrict__ off, double s,
                                                      double* __restrict__ arena, double* __restrict__ norm) {
  const int64_t b = blockIdx.x;
  if (threadIdx.x == 0) norm[b] = norm[b] * (s * s);
  const int64_t o = off[b];
  if (o < 0) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int64_t n = (int64_t)g.rsz[rb] * g.csz[cb];
  double* d = arena + o;
  for (int64_t e = threadIdx.x; e < n; e += blockDim.x) d[e] = d[e] * s;
}

// Scale every present block and recompute its cached norm from the scaled
// values in the canonical order (one wave per block, lane l over e = l, l+64,
// ...): the result equals calc_frobenius_norms() after the scale.
__global__ void __launch_bounds__(64) k_scale_blocks_renorm(Geo g, const int64_t* __restrict__ off, double s,
                                                            double* __restrict__ arena, double* __restrict__ norm) {
  const int64_t b = blockIdx.x;
  const int64_t o = off[b];
  if (o < 0) {
    if (threadIdx.x == 0) norm[b] = 0.0;
    return;
  }
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int64_t n = (int64_t)g.rsz[rb] * g.csz[cb];
  double* d = arena + o;
  double a = 0.0;
  for (int64_t e = threadIdx.x; e < n; e += 64) {
    const double y = d[e] * s;
    d[e] = y;
    a = fma(y, y, a);
  }
#pragma unroll
  for (int off2 = 32; off2 >= 1; off2 >>= 1) a = a + __shfl_down(a, off2, 64);
  if (threadIdx.x == 0) norm[b] = sqrt(a);
}

__global__ void k_recompress(Geo g, int64_t nb, double thresh, int64_t* __restrict__ off,
                             double* __restrict__ norm) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const double thr = thresh * (double)((int64_t)g.cscr[cb] * (int64_t)g.rscr[rb]);
  if (norm[b] < thr) {
    off[b] = -1;
    norm[b] = 0.0;
  }
}

// += / -= / add_transpose: blocks that must be created (rhs present, this empty)
__global__ void k_add_need(Geo g, int64_t nb, const int64_t* __restrict__ off,
                           const int64_t* __restrict__ roff_rhs, int transposed, int32_t rhs_nrb,
                           int64_t* __restrict__ need, int64_t* __restrict__ touched) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int64_t rbix = transposed ? (int64_t)cb + (int64_t)rb * rhs_nrb : b;
  const bool rp = roff_rhs[rbix] >= 0;
  need[b] = (rp && off[b] < 0) ? (int64_t)g.rsz[rb] * g.csz[cb] : 0;
  touched[b] = rp ? 1 : 0;
}

// mode 0: +=, mode 1: -=, mode 2: add_transpose
__global__ void __launch_bounds__(256) k_add_blocks(Geo g, int mode, const int64_t* __restrict__ need_scan,
                                                    const int64_t* __restrict__ touched, int64_t base,
                                                    int64_t* __restrict__ off, double* __restrict__ norm,
                                                    double* __restrict__ arena, const int64_t* __restrict__ roff,
                                                    const double* __restrict__ rnorm,
                                                    const double* __restrict__ rarena, int32_t rhs_nrb) {
  const int64_t b = blockIdx.x;
  if (!touched[b]) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  const int nr = g.rsz[rb], nc = g.csz[cb];
  const int64_t n = (int64_t)nr * nc;
  const int64_t rbix = mode == 2 ? (int64_t)cb + (int64_t)rb * rhs_nrb : b;
  const double* src = rarena + roff[rbix];
  int64_t o = off[b];
  const bool fresh = o < 0;
  if (fresh) o = base + need_scan[b];
  double* dst = arena + o;
  if (mode == 2) {
    // this(r, c) += rhs(c, r); rhs block is nc x nr
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) {
      const int64_t r = e % nr, cc = e / nr;
      const double prev = fresh ? 0.0 : dst[e];
      dst[e] = prev + src[cc + r * nc];
    }
  } else if (fresh) {
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = src[e];
  } else if (mode == 0) {
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = dst[e] + src[e];
  } else {
    for (int64_t e = threadIdx.x; e < n; e += blockDim.x) dst[e] = dst[e] - src[e];
  }
  if (threadIdx.x == 0) {
    off[b] = o;
    if (fresh && mode != 2) norm[b] = rnorm[rbix];  // += copies the block with its cached norm
  }
}

// dot: per-block partial dot products (wave-reduced), then a fixed-order sum
__global__ void __launch_bounds__(256) k_dot_blocks(Geo g, int64_t nb, double thresh,
                                                    const int64_t* __restrict__ aoff, const double* __restrict__ anorm,
                                                    const double* __restrict__ aarena,
                                                    const int64_t* __restrict__ boff, const double* __restrict__ bnorm,
                                                    const double* __restrict__ barena, double* __restrict__ part) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb) return;
  const int rb = (int)(b % g.nrb), cb = (int)(b / g.nrb);
  double a = 0.0;
  if (aoff[b] >= 0 && boff[b] >= 0) {
    const double thr = thresh * (double)((int64_t)g.rscr[rb] * (int64_t)g.cscr[cb]);
    if (anorm[b] * bnorm[b] >= thr) {
      const int64_t n = (int64_t)g.rsz[rb] * g.csz[cb];
      const double* x = aarena + aoff[b];
      const double* y = barena + boff[b];
      for (int64_t e = lane; e < n; e += 64) a = fma(x[e], y[e], a);
    }
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) a = a + __shfl_down(a, o, 64);
  if (lane == 0) part[b] = a;
}

// single-workgroup fixed-order sum of n doubles (x or x^2)
__global__ void __launch_bounds__(256) k_sum(const double* __restrict__ x, int64_t n, int squares,
                                             double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int64_t i = threadIdx.x; i < n; i += 256) a = squares ? fma(x[i], x[i], a) : a + x[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

__global__ void k_count_present(const int64_t* __restrict__ off, int64_t nb, int64_t* __restrict__ flag) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b < nb) flag[b] = off[b] >= 0 ? 1 : 0;
}

// --------------------------------------------------------------- host side --
void bsm_upload_geometry(sa_bsm* m) {
  hipStream_t s = m->ctx->stream;
  auto up32 = [&](int32_t** d, const std::vector<int32_t>& h) {
    SA_CHECK_HIP(hipMalloc(d, sizeof(int32_t) * (h.size() + 1)));
    if (!h.empty()) SA_CHECK_HIP(hipMemcpyAsync(*d, h.data(), sizeof(int32_t) * h.size(), hipMemcpyHostToDevice, s));
  };
  auto up64 = [&](int64_t** d, const std::vector<int64_t>& h) {
    SA_CHECK_HIP(hipMalloc(d, sizeof(int64_t) * (h.size() + 1)));
    if (!h.empty()) SA_CHECK_HIP(hipMemcpyAsync(*d, h.data(), sizeof(int64_t) * h.size(), hipMemcpyHostToDevice, s));
  };
  up32(&m->d_rsz, m->rsz);
  up32(&m->d_csz, m->csz);
  up32(&m->d_rscr, m->rscr);
  up32(&m->d_cscr, m->cscr);
  up64(&m->d_roff, m->roff);
  up64(&m->d_coff, m->coff);
  SA_CHECK_HIP(hipStreamSynchronize(s));  // host vectors may be reused by the caller
}

void bsm_alloc_tables(sa_bsm* m) {
  const int64_t nb = m->nb();
  SA_CHECK_HIP(hipMalloc(&m->d_off, sizeof(int64_t) * (size_t)(nb + 1)));
  SA_CHECK_HIP(hipMalloc(&m->d_norm, sizeof(double) * (size_t)(nb + 1)));
  fill_i64(m->ctx, m->d_off, nb, -1);
  fill_f64(m->ctx, m->d_norm, nb, 0.0);
}

void bsm_ensure_capacity(sa_bsm* m, int64_t elems) {
  const size_t want = sizeof(double) * (size_t)(elems > 0 ? elems : 1);
  if (want <= m->arena.bytes) return;
  if (m->used == 0) {
    // nothing to keep: release first and allocate the exact size (no transient copy)
    m->arena.release();
    m->arena.reserve(want);
    return;
  }
  m->arena.reserve_keep(want, m->ctx->stream);
}

void launch_block_norms(sa_bsm* m, const int64_t* mask) {
  const int64_t nb = m->nb();
  if (nb == 0) return;
  ProfScope ps(m->ctx, "norms");
  hipLaunchKernelGGL(k_norms_arena, dim3((unsigned)ceil_div(nb, 4)), dim3(256), 0, m->ctx->stream,
                     (const double*)m->data(), (const int64_t*)m->d_off, geo(m), nb, mask, m->d_norm);
}

static sa_bsm* make_bsm(sa_context* c, const std::vector<int32_t>& rsz, const std::vector<int32_t>& rscr,
                        const std::vector<int32_t>& csz, const std::vector<int32_t>& cscr, double thresh) {
  SA_REQUIRE(c, "null context");
  SA_CHECK_HIP(hipSetDevice(c->device));
  auto* m = new sa_bsm();
  m->ctx = c;
  m->nrb = (int32_t)rsz.size();
  m->ncb = (int32_t)csz.size();
  m->rsz = rsz; m->csz = csz; m->rscr = rscr; m->cscr = cscr;
  m->thresh = thresh;
  m->roff.resize(rsz.size());
  m->coff.resize(csz.size());
  int64_t s = 0;
  for (size_t i = 0; i < rsz.size(); ++i) { m->roff[i] = s; s += rsz[i]; }
  m->nrow = s;
  s = 0;
  for (size_t j = 0; j < csz.size(); ++j) { m->coff[j] = s; s += csz[j]; }
  m->ncol = s;
  try {
    bsm_upload_geometry(m);
    bsm_alloc_tables(m);
  } catch (...) {
    delete m;
    throw;
  }
  return m;
}

static void free_bsm(sa_bsm* m) {
  (void)hipFree(m->d_rsz); (void)hipFree(m->d_csz); (void)hipFree(m->d_rscr); (void)hipFree(m->d_cscr);
  (void)hipFree(m->d_roff); (void)hipFree(m->d_coff);
  if (!m->base) {
    (void)hipFree(m->d_off);
    (void)hipFree(m->d_norm);
    m->arena.release();
  }
  delete m;
}

#define SA_REQUIRE_WRITABLE(m) SA_REQUIRE(!(m)->base, "operation writes to a reshape view (views are read-only)")

static bool same_geometry(const sa_bsm* a, const sa_bsm* b) {
  return a->rsz == b->rsz && a->csz == b->csz && a->nrow == b->nrow && a->ncol == b->ncol;
}

// host-side compaction of the block-offset assignment: keep[] (device, sizes or
// 0) -> scanned offsets appended at m->used; returns number of new elements
static int64_t append_blocks(sa_bsm* m, int64_t* keep, int64_t* scan_tmp, int64_t* off_out) {
  sa_context* c = m->ctx;
  const int64_t nb = m->nb();
  SA_CHECK_HIP(hipMemcpyAsync(scan_tmp, keep, sizeof(int64_t) * (size_t)nb, hipMemcpyDeviceToDevice, c->stream));
  const int64_t total = scan_exclusive_i64(c, scan_tmp, nb);
  hipLaunchKernelGGL(k_finalize_offsets, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, c->stream,
                     (const int64_t*)scan_tmp, (const int64_t*)keep, nb, m->used, off_out);
  return total;
}

}  // namespace sa

// ===================================================================== ABI ==
using namespace sa;

void sa::bsm_from_dense_device(sa_bsm* m, const double* d_dense, int64_t ld) {
  sa_context* c = m->ctx;
  const int64_t nb = m->nb();
  if (nb == 0) return;
  c->ws_tmp0.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  c->ws_tmp1.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  int64_t* keep = c->ws_tmp0.as<int64_t>();
  int64_t* scan = c->ws_tmp1.as<int64_t>();
  // the copy replaces every block (blocksparsematrix.hpp:52-53): restart the arena
  m->used = 0;
  hipLaunchKernelGGL(k_dense_norms, dim3((unsigned)ceil_div(nb, 4)), dim3(256), 0, c->stream, d_dense, ld, geo(m),
                     nb, m->thresh, m->d_norm, keep);
  const int64_t total = append_blocks(m, keep, scan, m->d_off);
  bsm_ensure_capacity(m, total);
  hipLaunchKernelGGL(k_copy_dense_blocks, dim3((unsigned)nb), dim3(256), 0, c->stream, d_dense, ld, geo(m),
                     (const int64_t*)m->d_off, m->data());
  m->used = total;
  SA_CHECK_HIP(hipGetLastError());
}


extern "C" {

int sa_bsm_create(sa_ctx_t c, int64_t nrow, int64_t ncol, int32_t bs_r, int32_t bs_c, double thresh,
                  sa_bsm_t* out) {
  return guard([&] {
    SA_REQUIRE(out, "null output");
    SA_REQUIRE(nrow >= 0 && ncol >= 0 && bs_r > 0 && bs_c > 0, "bad dimensions %lld x %lld / %d x %d",
               (long long)nrow, (long long)ncol, bs_r, bs_c);
    const int64_t nrb = ceil_div(nrow, bs_r), ncb = ceil_div(ncol, bs_c);
    SA_REQUIRE(nrb < (1 << 30) && ncb < (1 << 30), "too many blocks");
    std::vector<int32_t> rsz(nrb), csz(ncb), rscr(nrb, bs_r), cscr(ncb, bs_c);
    for (int64_t i = 0; i < nrb; ++i) rsz[i] = (int32_t)std::min<int64_t>(bs_r, nrow - i * bs_r);
    for (int64_t j = 0; j < ncb; ++j) csz[j] = (int32_t)std::min<int64_t>(bs_c, ncol - j * bs_c);
    sa_bsm* m = make_bsm(c, rsz, rscr, csz, cscr, thresh);
    m->uniform = true;
    m->bs_r = bs_r;
    m->bs_c = bs_c;
    *out = m;
  });
}

int sa_bsm_create_blocked(sa_ctx_t c, int32_t nrb, const int32_t* rsz, const int32_t* rscr, int32_t ncb,
                          const int32_t* csz, const int32_t* cscr, double thresh, sa_bsm_t* out) {
  return guard([&] {
    SA_REQUIRE(out && nrb >= 0 && ncb >= 0 && (nrb == 0 || rsz) && (ncb == 0 || csz), "bad arguments");
    std::vector<int32_t> r(rsz, rsz + nrb), cc(csz, csz + ncb);
    std::vector<int32_t> rs = rscr ? std::vector<int32_t>(rscr, rscr + nrb) : r;
    std::vector<int32_t> cs = cscr ? std::vector<int32_t>(cscr, cscr + ncb) : cc;
    for (int32_t v : r) SA_REQUIRE(v > 0, "block sizes must be positive");
    for (int32_t v : cc) SA_REQUIRE(v > 0, "block sizes must be positive");
    *out = make_bsm(c, r, rs, cc, cs, thresh);
  });
}

int sa_bsm_view_reshape(sa_bsm_t base, int32_t nrb, const int32_t* rsz, const int32_t* rscr, int32_t ncb,
                        const int32_t* csz, const int32_t* cscr, sa_bsm_t* out) {
  return guard([&] {
    SA_REQUIRE(base && out && rsz && csz, "null argument");
    sa_bsm* root = base->base ? base->base : base;
    SA_REQUIRE((int64_t)nrb * ncb == base->nb(), "view: block count %lld differs from %lld",
               (long long)((int64_t)nrb * ncb), (long long)base->nb());
    for (int32_t cb = 0; cb < ncb; ++cb)
      for (int32_t rb = 0; rb < nrb; ++rb) {
        const int64_t b = rb + (int64_t)cb * nrb;
        const int64_t orb = b % base->nrb, ocb = b / base->nrb;
        SA_REQUIRE(rsz[rb] > 0 && csz[cb] > 0, "view: block sizes must be positive");
        SA_REQUIRE((int64_t)rsz[rb] * csz[cb] == (int64_t)base->rsz[orb] * base->csz[ocb],
                   "view: block %lld holds %lld elements, base block holds %lld", (long long)b,
                   (long long)((int64_t)rsz[rb] * csz[cb]), (long long)((int64_t)base->rsz[orb] * base->csz[ocb]));
      }
    std::vector<int32_t> r(rsz, rsz + nrb), cc(csz, csz + ncb);
    std::vector<int32_t> rs = rscr ? std::vector<int32_t>(rscr, rscr + nrb) : r;
    std::vector<int32_t> cs = cscr ? std::vector<int32_t>(cscr, cscr + ncb) : cc;
    sa_context* c = base->ctx;
    auto* m = new sa_bsm();
    m->ctx = c;
    m->nrb = nrb; m->ncb = ncb;
    m->rsz = r; m->csz = cc; m->rscr = rs; m->cscr = cs;
    m->thresh = base->thresh;
    m->roff.resize(nrb);
    m->coff.resize(ncb);
    int64_t s = 0;
    for (int32_t i = 0; i < nrb; ++i) { m->roff[i] = s; s += r[i]; }
    m->nrow = s;
    s = 0;
    for (int32_t j = 0; j < ncb; ++j) { m->coff[j] = s; s += cc[j]; }
    m->ncol = s;
    m->base = root;
    m->d_off = root->d_off;
    m->d_norm = root->d_norm;
    try {
      bsm_upload_geometry(m);
    } catch (...) {
      delete m;
      throw;
    }
    *out = m;
  });
}

int sa_bsm_destroy(sa_bsm_t m) {
  return guard([&] {
    if (!m) return;
    (void)hipSetDevice(m->ctx->device);
    (void)hipStreamSynchronize(m->ctx->stream);
    free_bsm(m);
  });
}

int sa_bsm_info(sa_bsm_t m, int64_t* nrow, int64_t* ncol, int32_t* nrb, int32_t* ncb, int64_t* nalloc,
                int64_t* arena_bytes) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    if (nrow) *nrow = m->nrow;
    if (ncol) *ncol = m->ncol;
    if (nrb) *nrb = m->nrb;
    if (ncb) *ncb = m->ncb;
    if (arena_bytes) *arena_bytes = m->used * (int64_t)sizeof(double);
    if (nalloc) {
      const int64_t nb = m->nb();
      if (nb == 0) {
        *nalloc = 0;
        return;
      }
      sa_context* c = m->ctx;
      c->ws_tmp0.reserve(sizeof(int64_t) * (size_t)(nb + 1));
      int64_t* f = c->ws_tmp0.as<int64_t>();
      hipLaunchKernelGGL(k_count_present, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, c->stream,
                         (const int64_t*)m->d_off, nb, f);
      *nalloc = scan_exclusive_i64(c, f, nb);
    }
  });
}

int sa_bsm_copy(sa_bsm_t dst, sa_bsm_t src) {
  return guard([&] {
    SA_REQUIRE(dst && src, "null matrix");
    SA_REQUIRE_WRITABLE(dst);
    SA_REQUIRE(same_geometry(dst, src), "copy: block geometry differs");
    if (dst == src) return;
    sa_context* c = dst->ctx;
    const int64_t used = src->base ? src->base->used : src->used;
    bsm_ensure_capacity(dst, used);
    if (used)
      SA_CHECK_HIP(hipMemcpyAsync(dst->data(), src->data(), sizeof(double) * (size_t)used,
                                  hipMemcpyDeviceToDevice, c->stream));
    dst->used = used;
    const int64_t nb = src->nb();
    SA_CHECK_HIP(hipMemcpyAsync(dst->d_off, src->d_off, sizeof(int64_t) * nb, hipMemcpyDeviceToDevice, c->stream));
    SA_CHECK_HIP(hipMemcpyAsync(dst->d_norm, src->d_norm, sizeof(double) * nb, hipMemcpyDeviceToDevice, c->stream));
  });
}


int sa_bsm_from_dense(sa_bsm_t m, const double* host, int64_t ld) {
  return guard([&] {
    SA_REQUIRE(m && (host || m->nrow * m->ncol == 0), "null argument");
    SA_REQUIRE_WRITABLE(m);
    SA_REQUIRE(ld >= m->nrow, "ld %lld < nrow %lld", (long long)ld, (long long)m->nrow);
    if (m->nb() == 0) return;
    sa_context* c = m->ctx;
    const size_t bytes = sizeof(double) * (size_t)ld * (size_t)(m->ncol - 1) + sizeof(double) * (size_t)m->nrow;
    c->ws_stage.reserve(bytes);
    SA_CHECK_HIP(hipMemcpyAsync(c->ws_stage.p, host, bytes, hipMemcpyHostToDevice, c->stream));
    bsm_from_dense_device(m, c->ws_stage.as<double>(), ld);
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_bsm_to_dense(sa_bsm_t m, double* host, int64_t ld) {
  return guard([&] {
    SA_REQUIRE(m && (host || m->nrow * m->ncol == 0), "null argument");
    SA_REQUIRE(ld >= m->nrow, "ld < nrow");
    if (m->nb() == 0) return;
    sa_context* c = m->ctx;
    const size_t bytes = sizeof(double) * (size_t)ld * (size_t)(m->ncol - 1) + sizeof(double) * (size_t)m->nrow;
    c->ws_stage.reserve(bytes);
    // to_pointer writes present blocks only: start from the caller's contents
    SA_CHECK_HIP(hipMemcpyAsync(c->ws_stage.p, host, bytes, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_arena_to_dense, dim3((unsigned)m->nb()), dim3(256), 0, c->stream, (const double*)m->data(),
                       (const int64_t*)m->d_off, geo(m), c->ws_stage.as<double>(), ld);
    SA_CHECK_HIP(hipMemcpyAsync(host, c->ws_stage.p, bytes, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_bsm_get_blocks(sa_bsm_t m, int32_t rb0, int32_t rb1, int32_t cb0, int32_t cb1, double* host, int64_t ld) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE(0 <= rb0 && rb0 <= rb1 && rb1 <= m->nrb && 0 <= cb0 && cb0 <= cb1 && cb1 <= m->ncb,
               "get_blocks: block range [%d, %d) x [%d, %d) outside the %d x %d grid", rb0, rb1, cb0, cb1, m->nrb,
               m->ncb);
    if (rb0 == rb1 || cb0 == cb1) return;
    const int64_t rows = m->roff[(size_t)rb1 - 1] + m->rsz[(size_t)rb1 - 1] - m->roff[(size_t)rb0];
    const int64_t cols = m->coff[(size_t)cb1 - 1] + m->csz[(size_t)cb1 - 1] - m->coff[(size_t)cb0];
    SA_REQUIRE(host && ld >= rows, "get_blocks: null output or ld < %lld rows", (long long)rows);
    sa_context* c = m->ctx;
    SA_CHECK_HIP(hipSetDevice(c->device));
    const int64_t elems = ld * (cols - 1) + rows;
    c->ws_stage.reserve(sizeof(double) * (size_t)elems);
    fill_f64(c, c->ws_stage.as<double>(), elems, 0.0);  // absent blocks read as zero
    const int32_t nr = rb1 - rb0, nc = cb1 - cb0;
    hipLaunchKernelGGL(k_arena_to_dense_sub, dim3((unsigned)((int64_t)nr * nc)), dim3(256), 0, c->stream,
                       (const double*)m->data(), (const int64_t*)m->d_off, geo(m), rb0, nr, cb0,
                       c->ws_stage.as<double>(), ld);
    // column by column: rows beyond the sub-matrix in each column of the
    // caller's array are left alone
    SA_CHECK_HIP(hipMemcpy2DAsync(host, sizeof(double) * (size_t)ld, c->ws_stage.p, sizeof(double) * (size_t)ld,
                                  sizeof(double) * (size_t)rows, (size_t)cols, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_bsm_get_meta(sa_bsm_t m, uint8_t* present, double* norms) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    const int64_t nb = m->nb();
    if (nb == 0) return;
    sa_context* c = m->ctx;
    std::vector<int64_t> off(nb);
    SA_CHECK_HIP(hipMemcpyAsync(off.data(), m->d_off, sizeof(int64_t) * nb, hipMemcpyDeviceToHost, c->stream));
    if (norms) SA_CHECK_HIP(hipMemcpyAsync(norms, m->d_norm, sizeof(double) * nb, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    if (present)
      for (int64_t b = 0; b < nb; ++b) present[b] = off[b] >= 0 ? 1 : 0;
  });
}

int sa_bsm_generate(sa_bsm_t m, uint64_t seed, const double* block_scale) {
  return guard([&] {
    SA_REQUIRE(m && block_scale, "null argument");
    SA_REQUIRE_WRITABLE(m);
    const int64_t nb = m->nb();
    if (nb == 0) return;
    sa_context* c = m->ctx;
    c->ws_tmp0.reserve(sizeof(int64_t) * (size_t)(nb + 1));
    c->ws_tmp1.reserve(sizeof(int64_t) * (size_t)(nb + 1));
    c->ws_tmp2.reserve(sizeof(double) * (size_t)(nb + 1));
    int64_t* keep = c->ws_tmp0.as<int64_t>();
    int64_t* scan = c->ws_tmp1.as<int64_t>();
    double* sc = c->ws_tmp2.as<double>();
    SA_CHECK_HIP(hipMemcpyAsync(sc, block_scale, sizeof(double) * nb, hipMemcpyHostToDevice, c->stream));
    const uint64_t base = seed * 0xD1B54A32D192ED03ull;
    m->used = 0;
    hipLaunchKernelGGL(k_gen_norms, dim3((unsigned)ceil_div(nb, 4)), dim3(256), 0, c->stream, base,
                       (const double*)sc, geo(m), nb, m->thresh, m->d_norm, keep);
    const int64_t total = append_blocks(m, keep, scan, m->d_off);
    bsm_ensure_capacity(m, total);
    hipLaunchKernelGGL(k_gen_fill, dim3((unsigned)nb), dim3(256), 0, c->stream, base, (const double*)sc, geo(m),
                       (const int64_t*)m->d_off, m->data());
    m->used = total;
    SA_CHECK_HIP(hipGetLastError());
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_bsm_fill(sa_bsm_t m, double v) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    const int64_t nb = m->nb();
    if (nb == 0) return;
    sa_context* c = m->ctx;
    std::vector<int64_t> off(nb);
    int64_t s = 0;
    for (int32_t cb = 0; cb < m->ncb; ++cb)
      for (int32_t rb = 0; rb < m->nrb; ++rb) {
        off[rb + (int64_t)cb * m->nrb] = s;
        s += (int64_t)m->rsz[rb] * m->csz[cb];
      }
    bsm_ensure_capacity(m, s);
    SA_CHECK_HIP(hipMemcpyAsync(m->d_off, off.data(), sizeof(int64_t) * nb, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_fill_blocks, dim3((unsigned)nb), dim3(256), 0, c->stream, geo(m), (const int64_t*)m->d_off, v,
                       m->data(), m->d_norm);
    m->used = s;
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  });
}

int sa_bsm_zero(sa_bsm_t m) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    fill_i64(m->ctx, m->d_off, m->nb(), -1);
    fill_f64(m->ctx, m->d_norm, m->nb(), 0.0);
    m->used = 0;
  });
}

int sa_bsm_scale(sa_bsm_t m, double s) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    if (m->nb() == 0) return;
    hipLaunchKernelGGL(k_scale_blocks, dim3((unsigned)m->nb()), dim3(256), 0, m->ctx->stream, geo(m),
                       (const int64_t*)m->d_off, s, m->data(), m->d_norm);
  });
}

int sa_bsm_reserve(sa_bsm_t m, int64_t elements) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    SA_REQUIRE(elements >= 0, "reserve: negative size");
    SA_CHECK_HIP(hipSetDevice(m->ctx->device));
    bsm_ensure_capacity(m, elements);
    SA_CHECK_HIP(hipStreamSynchronize(m->ctx->stream));
  });
}

int sa_bsm_scale_ex(sa_bsm_t m, double s, int32_t norm_rule) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE(norm_rule == SA_NORM_RULE_REFERENCE || norm_rule == SA_NORM_RULE_RECOMPUTE,
               "scale: unknown norm rule %d", norm_rule);
    if (norm_rule == SA_NORM_RULE_REFERENCE) {
      const int rc = sa_bsm_scale(m, s);
      if (rc != SA_OK) throw Error{rc};
      return;
    }
    SA_REQUIRE_WRITABLE(m);
    if (m->nb() == 0) return;
    hipLaunchKernelGGL(k_scale_blocks_renorm, dim3((unsigned)m->nb()), dim3(64), 0, m->ctx->stream, geo(m),
                       (const int64_t*)m->d_off, s, m->data(), m->d_norm);
    SA_CHECK_HIP(hipGetLastError());
  });
}

int sa_bsm_calc_norms(sa_bsm_t m) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    launch_block_norms(m);
  });
}

int sa_bsm_recompress(sa_bsm_t m) {
  return guard([&] {
    SA_REQUIRE(m, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    const int64_t nb = m->nb();
    if (nb == 0) return;
    hipLaunchKernelGGL(k_recompress, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, m->ctx->stream, geo(m), nb,
                       m->thresh, m->d_off, m->d_norm);
  });
}

static void add_impl(sa_bsm* m, sa_bsm* rhs, int mode) {
  sa_context* c = m->ctx;
  const int64_t nb = m->nb();
  if (nb == 0) return;
  c->ws_tmp0.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  c->ws_tmp1.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  c->ws_tmp2.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  int64_t* need = c->ws_tmp0.as<int64_t>();
  int64_t* scan = need;  // scanned in place
  int64_t* touched = c->ws_tmp1.as<int64_t>();
  hipLaunchKernelGGL(k_add_need, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, c->stream, geo(m), nb,
                     (const int64_t*)m->d_off, (const int64_t*)rhs->d_off, mode == 2 ? 1 : 0, rhs->nrb, need, touched);
  const int64_t total = scan_exclusive_i64(c, scan, nb);
  const int64_t base = m->used;
  bsm_ensure_capacity(m, base + total);
  hipLaunchKernelGGL(k_add_blocks, dim3((unsigned)nb), dim3(256), 0, c->stream, geo(m), mode, (const int64_t*)scan,
                     (const int64_t*)touched, base, m->d_off, m->d_norm, m->data(), (const int64_t*)rhs->d_off,
                     (const double*)rhs->d_norm, (const double*)rhs->data(), rhs->nrb);
  m->used = base + total;
  if (mode != 0) {
    ProfScope ps(c, "norms");
    hipLaunchKernelGGL(k_norms_arena, dim3((unsigned)ceil_div(nb, 4)), dim3(256), 0, c->stream,
                       (const double*)m->data(), (const int64_t*)m->d_off, geo(m), nb, (const int64_t*)touched,
                       m->d_norm);
  }
  SA_CHECK_HIP(hipGetLastError());
}

int sa_bsm_add(sa_bsm_t m, sa_bsm_t rhs, double sign) {
  return guard([&] {
    SA_REQUIRE(m && rhs, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    SA_REQUIRE(same_geometry(m, rhs), "add: block geometry differs");
    SA_REQUIRE(m != rhs, "add: aliasing not supported");
    add_impl(m, rhs, sign < 0 ? 1 : 0);
  });
}

int sa_bsm_add_transpose(sa_bsm_t m, sa_bsm_t rhs) {
  return guard([&] {
    SA_REQUIRE(m && rhs, "null matrix");
    SA_REQUIRE_WRITABLE(m);
    SA_REQUIRE(m->rsz == rhs->csz && m->csz == rhs->rsz, "add_transpose: geometry is not transposed");
    SA_REQUIRE(m != rhs, "add_transpose: aliasing not supported");
    add_impl(m, rhs, 2);
  });
}

int sa_bsm_frobenius_norm(sa_bsm_t m, double* out) {
  return guard([&] {
    SA_REQUIRE(m && out, "null argument");
    sa_context* c = m->ctx;
    c->ws_tmp2.reserve(sizeof(double) * 2);
    double* r = c->ws_tmp2.as<double>();
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, c->stream, (const double*)m->d_norm, m->nb(), 1, r);
    double* h = static_cast<double*>(pinned(c, sizeof(double)));
    SA_CHECK_HIP(hipMemcpyAsync(h, r, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    *out = std::sqrt(*h);
  });
}

int sa_bsm_dot(sa_bsm_t a, sa_bsm_t b, double thresh, double* out) {
  return guard([&] {
    SA_REQUIRE(a && b && out, "null argument");
    SA_REQUIRE(same_geometry(a, b), "dot: block geometry differs");
    sa_context* c = a->ctx;
    const int64_t nb = a->nb();
    c->ws_tmp2.reserve(sizeof(double) * (size_t)(nb + 2));
    double* part = c->ws_tmp2.as<double>();
    if (nb)
      hipLaunchKernelGGL(k_dot_blocks, dim3((unsigned)ceil_div(nb, 4)), dim3(256), 0, c->stream, geo(a), nb, thresh,
                         (const int64_t*)a->d_off, (const double*)a->d_norm, (const double*)a->data(),
                         (const int64_t*)b->d_off, (const double*)b->d_norm, (const double*)b->data(), part);
    hipLaunchKernelGGL(k_sum, dim3(1), dim3(256), 0, c->stream, (const double*)part, nb, 0, part + nb);
    double* h = static_cast<double*>(pinned(c, sizeof(double)));
    SA_CHECK_HIP(hipMemcpyAsync(h, part + nb, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    *out = *h;
  });
}

}  // extern "C"
