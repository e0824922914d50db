// sa_rik.hip -- the RI-K step on the device.
//
//  * 2-centre metric inverse square root X = M^{-1/2} (the reference reads it
//    as input data, 2c_ints_inv05_h2o.data / 3c_2c_transform.cpp:23-26): coupled
//    Newton-Schulz iteration (N. J. Higham, Functions of Matrices, SIAM 2008,
//    ch. 6): A = M/c with c = ||M||_F >= lambda_max, Y_0 = A, Z_0 = I,
//        T_k = (3I - Z_k Y_k)/2,  Y_{k+1} = Y_k T_k,  Z_{k+1} = T_k Z_k,
//    Z_k -> A^{-1/2}, X = Z/sqrt(c).  Three dense FP64 GEMMs per iteration
//    (rocBLAS, plain library GEMMs), stopped when ||Z Y - I||_F/sqrt(n) < tol.
//  * RI-K step: B' = X B (screened block multiply, 3c_2c_transform.cpp:56 /
//    tensormult form 2), cached norms of B' refreshed, then
//    K = sum_{P,i} B'(P,mu,i) B'(P,nu,i) (tensormult form 1) as one transposed
//    screened multiply over a reshape view of B' whose rows are (P, i) pairs.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "sa_internal.hpp"

namespace sa {

void device_dgemm(sa_context* c, bool ta, bool tb, int64_t m, int64_t n, int64_t k, double alpha, const double* A,
                  int64_t lda, const double* B, int64_t ldb, double beta, double* Cm, int64_t ldc);
void bsm_from_dense_device(sa_bsm* m, const double* d_dense, int64_t ld);
void check_mm_geometry(const sa_bsm* C, const sa_bsm* A, bool ta, const sa_bsm* B, bool tb);
void beta_prepass(sa_bsm* C, double beta);

__global__ void k_scale_copy(const double* __restrict__ src, double* __restrict__ dst, int64_t n, double s) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i] * s;
}

__global__ void k_identity(double* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n * n) d[i] = (i % n == i / n) ? 1.0 : 0.0;
}

// T <- (3I - T)/2
__global__ void k_ns_update(double* __restrict__ t, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * n) return;
  const double d = (i % n == i / n) ? 3.0 : 0.0;
  t[i] = 0.5 * (d - t[i]);
}

// per-block partial sums of squares of (T - I) or of T
__global__ void __launch_bounds__(256) k_fro_parts(const double* __restrict__ t, int64_t n, int minus_identity,
                                                   double* __restrict__ part) {
  __shared__ double s[256];
  double a = 0.0;
  const int64_t nn = n * n;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < nn; i += (int64_t)gridDim.x * 256) {
    const double v = t[i] - ((minus_identity && i % n == i / n) ? 1.0 : 0.0);
    a = fma(v, v, a);
  }
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}

__global__ void __launch_bounds__(256) k_sum_parts(const double* __restrict__ part, int n, double* __restrict__ out) {
  __shared__ double s[256];
  double a = 0.0;
  for (int i = threadIdx.x; i < n; i += 256) a = a + part[i];
  s[threadIdx.x] = a;
  __syncthreads();
  for (int o = 128; o >= 1; o >>= 1) {
    if ((int)threadIdx.x < o) s[threadIdx.x] = s[threadIdx.x] + s[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) *out = s[0];
}

static double dev_fro(sa_context* c, const double* t, int64_t n, bool minus_identity) {
  const int nparts = 1024;
  DevBuf& bp = c->buf("ns_parts");
  bp.reserve(sizeof(double) * (nparts + 2));
  hipLaunchKernelGGL(k_fro_parts, dim3(nparts), dim3(256), 0, c->stream, t, n, minus_identity ? 1 : 0,
                     bp.as<double>());
  hipLaunchKernelGGL(k_sum_parts, dim3(1), dim3(256), 0, c->stream, (const double*)bp.as<double>(), nparts,
                     bp.as<double>() + nparts);
  double* h = static_cast<double*>(pinned(c, sizeof(double)));
  SA_CHECK_HIP(hipMemcpyAsync(h, bp.as<double>() + nparts, sizeof(double), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  return std::sqrt(*h);
}

// X = M^{-1/2}; dM, dX device n x n (ld n).  Returns iterations, residual.
static void inv_sqrt_device(sa_context* c, int64_t n, const double* dM, double* dX, int max_iter, double tol,
                            int* iters, double* resid) {
  const size_t nn = (size_t)n * (size_t)n;
  DevBuf& bY = c->buf("ns_y");
  DevBuf& bZ = c->buf("ns_z");
  DevBuf& bT = c->buf("ns_t");
  DevBuf& bY2 = c->buf("ns_y2");
  DevBuf& bZ2 = c->buf("ns_z2");
  for (DevBuf* b : {&bY, &bZ, &bT, &bY2, &bZ2}) b->reserve(sizeof(double) * nn);
  double *Y = bY.as<double>(), *Z = bZ.as<double>(), *T = bT.as<double>(), *Y2 = bY2.as<double>(),
         *Z2 = bZ2.as<double>();
  const unsigned g = (unsigned)ceil_div((int64_t)nn, 256);
  const double cn = dev_fro(c, dM, n, false);
  SA_REQUIRE(cn > 0 && std::isfinite(cn), "inv_sqrt: metric norm is %g", cn);
  hipLaunchKernelGGL(k_scale_copy, dim3(g), dim3(256), 0, c->stream, dM, Y, (int64_t)nn, 1.0 / cn);
  hipLaunchKernelGGL(k_identity, dim3(g), dim3(256), 0, c->stream, Z, n);
  int it = 0;
  double r = 1e300;
  {
    ProfScope ps(c, "newton_schulz");
    for (it = 0; it < max_iter; ++it) {
      device_dgemm(c, false, false, n, n, n, 1.0, Z, n, Y, n, 0.0, T, n);
      r = dev_fro(c, T, n, true) / std::sqrt((double)n);
      if (r < tol) break;
      hipLaunchKernelGGL(k_ns_update, dim3(g), dim3(256), 0, c->stream, T, n);
      device_dgemm(c, false, false, n, n, n, 1.0, Y, n, T, n, 0.0, Y2, n);
      device_dgemm(c, false, false, n, n, n, 1.0, T, n, Z, n, 0.0, Z2, n);
      std::swap(Y, Y2);
      std::swap(Z, Z2);
    }
  }
  hipLaunchKernelGGL(k_scale_copy, dim3(g), dim3(256), 0, c->stream, (const double*)Z, dX, (int64_t)nn,
                     1.0 / std::sqrt(cn));
  // residual ||X M X - I||_F / sqrt(n)
  device_dgemm(c, false, false, n, n, n, 1.0, dX, n, dM, n, 0.0, T, n);
  device_dgemm(c, false, false, n, n, n, 1.0, T, n, dX, n, 0.0, Y2, n);
  *resid = dev_fro(c, Y2, n, true) / std::sqrt((double)n);
  *iters = it;
  SA_CHECK_HIP(hipGetLastError());
}

// ------------------------------------------------------------ occ-RI-K --
// Device-side decompression of the compressed 3-centre integrals
// (occ_ri_k_step.cpp:133-146) straight into the block arena of U_P: every key
// (r + n_bf c of the n_bf x n_bf matrix) is resolved once on the host to its
// position in a fixed slot layout of the blocks any key touches; per aux
// function P the N_vec2 values are scattered there (positions no key names stay
// zero for every P, as in the reference's reused dense buffer), then each
// touched block gets the canonical norm and the copy-in drop rule
// (blocksparsematrix.hpp:55-59) of ints_3c_decompressed_bsm.copy_from_input_matrix.
// Host -> device copy of a caller's (pageable) array through two pinned 8 MB
// halves of the context's staging buffer: the host fills one half while the
// other is in flight (stream-ordered; a half is refilled only after the copy
// that read it has completed; later users of the staging buffer on this
// stream are ordered behind the last copies).
static void upload_staged(sa_context* x, void* dst, const void* src, size_t bytes) {
  constexpr size_t kChunk = (size_t)8 << 20;
  char* stage = static_cast<char*>(pinned(x, 2 * kChunk));
  hipEvent_t ev[2] = {nullptr, nullptr};
  SA_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  SA_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  try {
    size_t i = 0;
    for (size_t off = 0; off < bytes; off += kChunk, ++i) {
      const size_t n = std::min(kChunk, bytes - off);
      char* half = stage + (i & 1) * kChunk;
      if (i >= 2) SA_CHECK_HIP(hipEventSynchronize(ev[i & 1]));
      std::memcpy(half, static_cast<const char*>(src) + off, n);
      SA_CHECK_HIP(hipMemcpyAsync(static_cast<char*>(dst) + off, half, n, hipMemcpyHostToDevice, x->stream));
      SA_CHECK_HIP(hipEventRecord(ev[i & 1], x->stream));
    }
  } catch (...) {
    (void)hipEventDestroy(ev[0]);
    (void)hipEventDestroy(ev[1]);
    throw;
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
}

// aux functions per batch of the occupied RI-K: SA_OCC_BATCH when set (> 0),
// else as many as 256 MB of stacked U hold
static int64_t occ_batch_cap() {
  const char* v = std::getenv("SA_OCC_BATCH");
  return v ? std::atoll(v) : 0;
}

// Batched occupied RI-K: a batch of nP aux functions of one team member runs
// as three screened multiplies over stacked matrices (U_P side by side, mui_P
// side by side, ijs_P stacked), each P's blocks in its own segment.

// U_P values of the batch into segment P of the stacked arena (P-major input)
__global__ void k_occ_scatter_batch(const double* __restrict__ vals, const int64_t* __restrict__ dest, int64_t n_vec2,
                                    int64_t nP, int64_t total, double* __restrict__ arena) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n_vec2 * nP) return;
  const int64_t P = i / n_vec2, d = dest[i % n_vec2];
  if (d >= 0) arena[P * total + d] = vals[i];
}

// block table and norms of the stacked U (segments P >= nP stay empty)
__global__ void __launch_bounds__(256) k_occ_block_norms_batch(const double* __restrict__ arena,
                                                               const int64_t* __restrict__ slot,
                                                               const int32_t* __restrict__ rsz,
                                                               const int32_t* __restrict__ csz,
                                                               const int32_t* __restrict__ rscr,
                                                               const int32_t* __restrict__ cscr, int32_t nrb,
                                                               int64_t nb, int64_t nseg, int64_t nP, int64_t total,
                                                               double thresh, int64_t* __restrict__ off,
                                                               double* __restrict__ norm) {
  const int lane = threadIdx.x & 63;
  const int64_t b = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (b >= nb * nseg) return;
  const int64_t P = b / nb, sl = P < nP ? slot[b % nb] : -1;
  if (sl < 0) {  // all-zero block (or unused segment): dropped by the copy-in rule
    if (lane == 0) {
      off[b] = -1;
      norm[b] = 0.0;
    }
    return;
  }
  const int rb = (int)(b % nrb), cb = (int)(b / nrb);
  const double est = wave_block_norm(arena + P * total + sl, (int64_t)rsz[rb] * csz[cb], lane);
  if (lane == 0) {
    const bool drop = est < thresh * (double)((int64_t)cscr[cb] * (int64_t)rscr[rb]);
    off[b] = drop ? -1 : P * total + sl;
    norm[b] = drop ? 0.0 : est;
  }
}

// block-diagonal repeat of C's block table (all segments share C's arena copy)
__global__ void k_occ_blockdiag(const int64_t* __restrict__ coff, const double* __restrict__ cnorm, int32_t nrb,
                                int32_t nocb, int32_t nseg, int64_t* __restrict__ off, double* __restrict__ norm) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t R = (int64_t)nrb * nseg;
  if (b >= R * nocb * nseg) return;
  const int64_t row = b % R, col = b / R;
  const int64_t P = row / nrb, Q = col / nocb;
  const int64_t src = row % nrb + (col % nocb) * nrb;
  off[b] = P == Q ? coff[src] : -1;
  norm[b] = P == Q ? cnorm[src] : 0.0;
}

// ijs_P = ij_P + ij_P^T as occ_ri_k_step.cpp:109-110 does it (ij_sym = ij;
// ij_sym.add_transpose(ij)): ij is n_occ x (nseg n_occ) with segment P = ij_P,
// ijs is (nseg n_occ) x n_occ with segment P = ijs_P.  A block exists where
// ij_P has it or its transposed partner; its norm is refreshed (touched) only
// where the partner exists, else it keeps ij_P's fresh-block norm 0
// (blocksparsematrix.hpp:225-248).  Fixed arena slots: segment P, block (a, b)
// at P n_occ^2 + start_b n_occ + start_a size_b.
__global__ void __launch_bounds__(256) k_occ_sym(const double* __restrict__ ij, const int64_t* __restrict__ ij_off,
                                                 const int32_t* __restrict__ osz, int32_t nocb, int64_t n_occ,
                                                 int32_t nseg, double* __restrict__ out,
                                                 int64_t* __restrict__ out_off, double* __restrict__ out_norm,
                                                 int64_t* __restrict__ touched) {
  const int64_t ob = blockIdx.x;  // (P, a, b), a fastest
  const int64_t per = (int64_t)nocb * nocb;
  const int64_t P = ob / per;
  const int a = (int)(ob % per % nocb), b = (int)(ob % per / nocb);
  const int64_t ia = a + (P * nocb + b) * nocb, ib = b + (P * nocb + a) * nocb;  // ij blocks (a, b), (b, a) of P
  const int64_t o = (P * nocb + a) + (int64_t)b * ((int64_t)nseg * nocb);     // ijs block (P, a), b
  const int64_t oa = ij_off[ia], ob2 = ij_off[ib];
  if (oa < 0 && ob2 < 0) {
    if (threadIdx.x == 0) {
      out_off[o] = -1;
      out_norm[o] = 0.0;
      touched[o] = 0;
    }
    return;
  }
  int64_t sa = 0, sb = 0;
  for (int i = 0; i < a; ++i) sa += osz[i];
  for (int i = 0; i < b; ++i) sb += osz[i];
  const int ma = osz[a], mb = osz[b];
  const int64_t dst = P * n_occ * n_occ + sb * n_occ + sa * mb;
  for (int e = threadIdx.x; e < ma * mb; e += blockDim.x) {
    const int r = e % ma, c = e / ma;
    double v = oa >= 0 ? ij[oa + e] : 0.0;
    if (ob2 >= 0) v += ij[ob2 + c + (int64_t)r * mb];
    out[dst + e] = v;
  }
  if (threadIdx.x == 0) {
    out_off[o] = dst;
    out_norm[o] = 0.0;
    touched[o] = ob2 >= 0 ? 1 : 0;
  }
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_inv_sqrt(sa_ctx_t c, int64_t n, const double* M, double* X, sa_bsm_t X_bsm, int32_t max_iter, double tol,
                int32_t* iters, double* residual) {
  return guard([&] {
    SA_REQUIRE(c && M && n > 0, "null argument");
    SA_REQUIRE(max_iter > 0, "max_iter must be positive");
    SA_CHECK_HIP(hipSetDevice(c->device));
    if (X_bsm) {
      SA_REQUIRE(X_bsm->nrow == n && X_bsm->ncol == n, "inv_sqrt: X_bsm must be n x n");
      SA_REQUIRE(!X_bsm->base, "inv_sqrt: X_bsm must not be a view");
    }
    const size_t nn = (size_t)n * (size_t)n;
    DevBuf& bM = c->buf("ns_m");
    DevBuf& bX = c->buf("ns_x");
    bM.reserve(sizeof(double) * nn);
    bX.reserve(sizeof(double) * nn);
    SA_CHECK_HIP(hipMemcpyAsync(bM.p, M, sizeof(double) * nn, hipMemcpyHostToDevice, c->stream));
    int it = 0;
    double r = 0.0;
    inv_sqrt_device(c, n, bM.as<double>(), bX.as<double>(), max_iter, tol, &it, &r);
    if (X) SA_CHECK_HIP(hipMemcpyAsync(X, bX.p, sizeof(double) * nn, hipMemcpyDeviceToHost, c->stream));
    if (X_bsm) bsm_from_dense_device(X_bsm, bX.as<double>(), n);
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    // the iteration's n x n work arrays are set-up memory: give them back
    for (const char* nm : {"ns_m", "ns_x", "ns_y", "ns_z", "ns_t", "ns_y2", "ns_z2"}) c->buf(nm).release();
    if (iters) *iters = it;
    if (residual) *residual = r;
  });
}

int sa_ri_k_step(sa_bsm_t X, sa_bsm_t B, sa_bsm_t Bp, sa_bsm_t K, int32_t n_mu, const int32_t* mu_sizes, int32_t n_ic,
                 const int32_t* i_chunks, double thresh, const sa_opts_t* opts_transform, const sa_opts_t* opts_k) {
  return guard([&] {
    SA_REQUIRE(X && B && Bp && K && mu_sizes && i_chunks, "null argument");
    SA_REQUIRE(B->ncb == n_mu * n_ic, "ri_k: B has %d column blocks, expected n_mu*n_ic = %d", B->ncb, n_mu * n_ic);
    for (int32_t mu = 0; mu < n_mu; ++mu)
      for (int32_t ic = 0; ic < n_ic; ++ic)
        SA_REQUIRE(B->csz[ic + (int64_t)n_ic * mu] == mu_sizes[mu] * i_chunks[ic],
                   "ri_k: B column block (mu %d, i-chunk %d) has width %d, expected %d", mu, ic,
                   B->csz[ic + (int64_t)n_ic * mu], mu_sizes[mu] * i_chunks[ic]);
    SA_REQUIRE(K->nrb == n_mu && K->ncb == n_mu, "ri_k: K must have n_mu x n_mu blocks");
    sa_opts_t ot, ok;
    sa_opts_default(&ot);
    sa_opts_default(&ok);
    if (opts_transform) ot = *opts_transform;
    if (opts_k) ok = *opts_k;
    SA_CHECK_HIP(hipSetDevice(B->ctx->device));
    // the split K contraction takes B''s block norms in the pass over B' that
    // also finds its slicing exponents (view blocks of at most 128 columns)
    // (env SA_RIK_FUSED_NORMS=0: norms and exponents in their own passes)
    const char* fz = std::getenv("SA_RIK_FUSED_NORMS");
    const bool fused_norms = ok.mode == SA_MODE_OZAKI && *std::max_element(mu_sizes, mu_sizes + n_mu) <= 128 &&
                             !(fz && std::atoi(fz) == 0);
    // transform B' = X B
    check_mm_geometry(Bp, X, false, B, false);
    {
      ProfScope ps(B->ctx, "rik_transform");
      beta_prepass(Bp, 0.0);
      if (ot.mode == SA_MODE_OZAKI)
        ozaki_multiply(Bp, X, false, B, false, thresh, 1.0, ot);
      else
        screened_multiply(Bp, X, false, B, false, thresh, 1.0, ot.tile, nullptr, nullptr, 0, false);
      // B' block norms (the K view's: the same blocks): here for the FP64 K
      // contraction, in the K stage's one pass over B' for the split one
      if (!fused_norms) launch_block_norms(Bp);
    }
    // K-view of B': rows (P, i-chunk), columns mu
    const int32_t npb = Bp->nrb;
    std::vector<int32_t> vr((size_t)npb * n_ic), vrs((size_t)npb * n_ic);
    for (int32_t ic = 0; ic < n_ic; ++ic)
      for (int32_t pb = 0; pb < npb; ++pb) {
        vr[pb + (size_t)npb * ic] = Bp->rsz[pb] * i_chunks[ic];
        vrs[pb + (size_t)npb * ic] = Bp->rscr[pb] * i_chunks[ic];
      }
    std::vector<int32_t> vc(mu_sizes, mu_sizes + n_mu);
    sa_bsm_t view = nullptr;
    int rc = sa_bsm_view_reshape(Bp, npb * n_ic, vr.data(), vrs.data(), n_mu, vc.data(), K->cscr.data(), &view);
    if (rc != SA_OK) throw Error{rc};
    try {
      check_mm_geometry(K, view, true, view, false);
      ProfScope ps(B->ctx, "rik_k");
      beta_prepass(K, 0.0);
      // K = B'^T B' is symmetric: the blocks above the diagonal are computed
      // and the ones below written as their transposes
      if (fused_norms) {
        // B' block norms and the exponents of the K operand's vectors (the
        // view's mu columns) in one pass over B'; the slicer then cuts digits only
        const OzPrepHandle prep = ozaki_prepare_view(Bp, view, ok);
        ozaki_multiply(K, view, true, view, false, thresh, 1.0, ok, true, prep.get());
      } else if (ok.mode == SA_MODE_OZAKI) {
        ozaki_multiply(K, view, true, view, false, thresh, 1.0, ok, true);
      } else
        screened_multiply(K, view, true, view, false, thresh, 1.0, ok.tile, nullptr, nullptr, 0, false, true);
    } catch (...) {
      sa_bsm_destroy(view);
      throw;
    }
    sa_bsm_destroy(view);
  });
}

int sa_occ_ri_k(sa_ctx_t c, int64_t n_bf, int64_t n_aux, int64_t n_occ, int64_t n_vec2, const double* ints3c,
                const double* mo, const uint64_t* keys, int32_t bs, double thresh, int32_t teams,
                const sa_opts_t* opts, double* K) {
  std::vector<sa_bsm_t> mats;
  std::vector<sa_ctx_t> ctxs;
  const int rc = guard([&] {
    SA_REQUIRE(c && mo && K && (n_vec2 == 0 || (ints3c && keys)), "null argument");
    SA_REQUIRE(n_bf > 0 && n_occ > 0 && n_aux >= 0 && n_vec2 >= 0 && bs > 0, "occ_ri_k: bad sizes");
    const uint64_t nn = (uint64_t)n_bf * (uint64_t)n_bf;
    for (int64_t n = 0; n < n_vec2; ++n)  // occ_ri_k_step.cpp:139
      SA_REQUIRE(keys[n] < nn, "occ_ri_k: key %lld = %llu is beyond the %lld x %lld matrix", (long long)n,
                 (unsigned long long)keys[n], (long long)n_bf, (long long)n_bf);
    SA_CHECK_HIP(hipSetDevice(c->device));
    auto ok = [](int e) {
      if (e != SA_OK) throw Error{e};
    };
    auto make = [&](sa_ctx_t x, int64_t r, int64_t q, double t) {
      sa_bsm_t m = nullptr;
      ok(sa_bsm_create(x, r, q, bs, bs, t, &m));
      mats.push_back(m);
      return m;
    };

    // slot layout of U: the blocks any key touches, column-major inside a block
    const int64_t nrb = ceil_div(n_bf, (int64_t)bs), nb = nrb * nrb;
    auto bsize = [&](int64_t i) { return (int64_t)std::min<int64_t>(bs, n_bf - i * bs); };
    std::vector<int64_t> slot((size_t)nb, -1), dest((size_t)n_vec2);
    for (int64_t n = 0; n < n_vec2; ++n) {
      const int64_t r = (int64_t)(keys[n] % (uint64_t)n_bf), q = (int64_t)(keys[n] / (uint64_t)n_bf);
      slot[(size_t)(r / bs + (q / bs) * nrb)] = 0;
    }
    int64_t total = 0;
    for (int64_t b = 0; b < nb; ++b)
      if (slot[(size_t)b] >= 0) {
        slot[(size_t)b] = total;
        total += bsize(b % nrb) * bsize(b / nrb);
      }
    for (int64_t n = 0; n < n_vec2; ++n) {
      const int64_t r = (int64_t)(keys[n] % (uint64_t)n_bf), q = (int64_t)(keys[n] / (uint64_t)n_bf);
      const int64_t rb = r / bs, cb = q / bs;
      dest[(size_t)n] = slot[(size_t)(rb + cb * nrb)] + (r - rb * bs) + (q - cb * bs) * bsize(rb);
    }
    // a key named twice: the reference's later write wins (one writer per position)
    {
      std::vector<int64_t> idx((size_t)n_vec2);
      for (int64_t n = 0; n < n_vec2; ++n) idx[(size_t)n] = n;
      std::stable_sort(idx.begin(), idx.end(), [&](int64_t a, int64_t b) { return keys[a] < keys[b]; });
      for (size_t i = 0; i + 1 < idx.size(); ++i)
        if (keys[idx[i]] == keys[idx[i + 1]]) dest[(size_t)idx[i]] = -1;
    }

    // The team of occ_ri_k_step.cpp:90-126: P split as OpenMP's static
    // schedule (the first n_aux % T members one iteration more), each member
    // with its own context (stream), decompression buffer, mui, ij, ijs and
    // K_sub; the members' K_sub are added in member order.
    const int64_t T = std::max<int64_t>(1, std::min<int64_t>(teams > 0 ? teams : 8, std::max<int64_t>(n_aux, 1)));
    ctxs.assign((size_t)T, nullptr);
    ctxs[0] = c;
    for (int64_t t = 1; t < T; ++t) ok(sa_ctx_create(c->device, &ctxs[(size_t)t]));
    std::vector<sa_bsm_t> Ks((size_t)T);
    for (int64_t t = 0; t < T; ++t) {
      Ks[(size_t)t] = make(ctxs[(size_t)t], n_bf, n_occ, 0.0);
      ok(sa_bsm_fill(Ks[(size_t)t], 0.0));
    }
    sa_bsm_t Kb = make(c, n_bf, n_occ, 0.0);
    ok(sa_bsm_fill(Kb, 0.0));
    ok(sa_ctx_synchronize(c));

    std::vector<int> trc((size_t)T, SA_OK);
    std::vector<std::string> terr((size_t)T);
    auto member = [&](int64_t t) {
      sa_ctx_t x = ctxs[(size_t)t];
      std::vector<sa_bsm_t> own;
      const int64_t q = n_aux / T, r = n_aux % T;
      const int64_t P0 = t * q + std::min<int64_t>(t, r), np = q + (t < r ? 1 : 0);
      trc[(size_t)t] = guard([&] {
        SA_CHECK_HIP(hipSetDevice(x->device));
        auto mk = [&](int64_t rr, int64_t qq, double th) {
          sa_bsm_t m = nullptr;
          ok(sa_bsm_create(x, rr, qq, bs, bs, th, &m));
          own.push_back(m);
          return m;
        };
        // occ_ri_k_step.cpp:84-98 (U's copy-in threshold is the reference's 1e-20)
        sa_bsm_t Cb = mk(n_bf, n_occ, 0.0);
        ok(sa_bsm_from_dense(Cb, mo, n_bf));
        DevBuf& bints = x->buf("occ_ints");
        DevBuf& bdest = x->buf("occ_dest");
        DevBuf& bslot = x->buf("occ_slot");
        bints.reserve(sizeof(double) * (size_t)std::max<int64_t>(1, n_vec2 * np));
        bdest.reserve(sizeof(int64_t) * (size_t)std::max<int64_t>(1, n_vec2));
        bslot.reserve(sizeof(int64_t) * (size_t)std::max<int64_t>(1, nb));
        if (n_vec2 * np)
          upload_staged(x, bints.p, ints3c + P0 * n_vec2, sizeof(double) * (size_t)(n_vec2 * np));
        if (n_vec2)
          SA_CHECK_HIP(hipMemcpyAsync(bdest.p, dest.data(), sizeof(int64_t) * (size_t)n_vec2, hipMemcpyHostToDevice,
                                      x->stream));
        SA_CHECK_HIP(
            hipMemcpyAsync(bslot.p, slot.data(), sizeof(int64_t) * (size_t)nb, hipMemcpyHostToDevice, x->stream));
        // occ_ri_k_step.cpp:100-118 over batches of nseg aux functions.  The
        // reference empties mui and ij (beta = 0) for every P, so the block
        // products of a batch are exactly those of its P one by one, screened
        // by the same norms; only the FP64 order of K_sub's sum over P changes
        // (one accumulation per batch).
        const int32_t nocb = (int32_t)ceil_div(n_occ, (int64_t)bs);
        const int64_t per_seg = std::max<int64_t>(1, total) * (int64_t)sizeof(double);
        const int64_t cap = occ_batch_cap();
        // block-diagonal C's table holds nseg^2 nrb nocb entries: at most 2^24
        const int64_t diag_cap = (int64_t)std::sqrt((double)((int64_t)1 << 24) / (double)(nrb * (int64_t)nocb));
        int64_t nseg = cap > 0 ? cap : std::min<int64_t>((256ll << 20) / per_seg, diag_cap);
        nseg = std::max<int64_t>(1, std::min<int64_t>(nseg, std::max<int64_t>(np, 1)));
        std::vector<int32_t> sbf((size_t)nrb), socc((size_t)nocb);
        for (int64_t i = 0; i < nrb; ++i) sbf[(size_t)i] = (int32_t)bsize(i);
        for (int32_t i = 0; i < nocb; ++i) socc[(size_t)i] = (int32_t)std::min<int64_t>(bs, n_occ - (int64_t)i * bs);
        auto rep = [&](const std::vector<int32_t>& v) {
          std::vector<int32_t> r;
          for (int64_t P = 0; P < nseg; ++P) r.insert(r.end(), v.begin(), v.end());
          return r;
        };
        const std::vector<int32_t> sbf_n = rep(sbf), socc_n = rep(socc);
        auto full = [&](size_t n) { return std::vector<int32_t>(n, bs); };  // screening sizes: the target bs
        auto mkb = [&](const std::vector<int32_t>& r, const std::vector<int32_t>& q, double th) {
          sa_bsm_t m = nullptr;
          const std::vector<int32_t> rs = full(r.size()), qs = full(q.size());
          ok(sa_bsm_create_blocked(x, (int32_t)r.size(), r.data(), rs.data(), (int32_t)q.size(), q.data(),
                                   qs.data(), th, &m));
          own.push_back(m);
          return m;
        };
        sa_bsm_t Ub = mkb(sbf, sbf_n, 1e-20);         // [U_1 ... U_nseg]; U's copy-in threshold
        sa_bsm_t Cd = mkb(sbf_n, socc_n, 0.0);        // diag(C, ..., C)
        sa_bsm_t mub = mkb(sbf, socc_n, 0.0);         // [mui_1 ... mui_nseg]
        sa_bsm_t ijb = mkb(socc, socc_n, 0.0);        // [ij_1 ... ij_nseg]
        sa_bsm_t ijsb = mkb(socc_n, socc, 0.0);       // [ijs_1; ...; ijs_nseg]
        bsm_ensure_capacity(Ub, nseg * total);
        Ub->used = nseg * total;
        fill_f64(x, Ub->data(), nseg * total, 0.0);
        bsm_ensure_capacity(Cd, std::max<int64_t>(Cb->used, 1));
        Cd->used = Cb->used;
        if (Cb->used)
          SA_CHECK_HIP(hipMemcpyAsync(Cd->data(), Cb->data(), sizeof(double) * (size_t)Cb->used,
                                      hipMemcpyDeviceToDevice, x->stream));
        hipLaunchKernelGGL(k_occ_blockdiag, dim3((unsigned)ceil_div(Cd->nb(), 256)), dim3(256), 0, x->stream,
                           (const int64_t*)Cb->d_off, (const double*)Cb->d_norm, (int32_t)nrb, nocb, (int32_t)nseg,
                           Cd->d_off, Cd->d_norm);
        bsm_ensure_capacity(ijsb, nseg * n_occ * n_occ);
        ijsb->used = nseg * n_occ * n_occ;
        DevBuf& btouch = x->buf("occ_touch");
        btouch.reserve(sizeof(int64_t) * (size_t)(ijsb->nb() + 1));
        SA_CHECK_HIP(hipGetLastError());
        for (int64_t Pb = 0; Pb < np; Pb += nseg) {
          const int64_t nP = std::min<int64_t>(nseg, np - Pb);
          if (n_vec2)
            hipLaunchKernelGGL(k_occ_scatter_batch, dim3((unsigned)ceil_div(n_vec2 * nP, 256)), dim3(256), 0,
                               x->stream, (const double*)bints.as<double>() + Pb * n_vec2,
                               (const int64_t*)bdest.as<int64_t>(), n_vec2, nP, total, Ub->data());
          hipLaunchKernelGGL(k_occ_block_norms_batch, dim3((unsigned)ceil_div(nb * nseg, 4)), dim3(256), 0,
                             x->stream, (const double*)Ub->data(), (const int64_t*)bslot.as<int64_t>(),
                             (const int32_t*)Ub->d_rsz, (const int32_t*)Ub->d_csz, (const int32_t*)Ub->d_rscr,
                             (const int32_t*)Ub->d_cscr, Ub->nrb, nb, nseg, nP, total, Ub->thresh, Ub->d_off,
                             Ub->d_norm);
          SA_CHECK_HIP(hipGetLastError());
          ok(sa_matmult(mub, Ub, 0, Cd, 0, thresh, 1.0, 0.0, opts));  // mui_P = U_P C
          ok(sa_bsm_calc_norms(mub));
          ok(sa_matmult(ijb, Cb, 1, mub, 0, thresh, 1.0, 0.0, opts));  // ij_P = C^T mui_P
          hipLaunchKernelGGL(k_occ_sym, dim3((unsigned)(nseg * nocb * nocb)), dim3(256), 0, x->stream,
                             (const double*)ijb->data(), (const int64_t*)ijb->d_off, (const int32_t*)ijb->d_rsz,
                             nocb, n_occ, (int32_t)nseg, ijsb->data(), ijsb->d_off, ijsb->d_norm,
                             btouch.as<int64_t>());
          SA_CHECK_HIP(hipGetLastError());
          launch_block_norms(ijsb, btouch.as<int64_t>());
          ok(sa_matmult(Ks[(size_t)t], mub, 0, ijsb, 0, thresh, 1.0, 1.0, opts));  // K_sub += sum_P mui_P ijs_P
        }
        SA_CHECK_HIP(hipStreamSynchronize(x->stream));
      });
      if (trc[(size_t)t] != SA_OK) terr[(size_t)t] = sa_last_error();
      for (sa_bsm_t m : own) sa_bsm_destroy(m);
    };
    {
      std::vector<std::thread> th;
      for (int64_t t = 1; t < T; ++t) th.emplace_back(member, t);
      member(0);
      for (auto& h : th) h.join();
    }
    for (int64_t t = 0; t < T; ++t)
      if (trc[(size_t)t] != SA_OK) {
        set_error("%s", terr[(size_t)t].c_str());
        throw Error{trc[(size_t)t]};
      }
    for (int64_t t = 0; t < T; ++t) ok(sa_bsm_add(Kb, Ks[(size_t)t], 1.0));  // K_mu_i_bsm += K_mu_i_sub_bsm
    for (int64_t i = 0; i < n_bf * n_occ; ++i) K[i] = 0.0;
    ok(sa_bsm_to_dense(Kb, K, n_bf));
  });
  for (sa_bsm_t m : mats) sa_bsm_destroy(m);
  for (size_t t = 1; t < ctxs.size(); ++t)
    if (ctxs[t]) sa_ctx_destroy(ctxs[t]);
  return rc;
}

}  // extern "C"
