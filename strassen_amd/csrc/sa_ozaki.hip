// sa_ozaki.hip -- split-precision (Ozaki-scheme) screened block multiply.
//
// The task list is the FP64 one (screening on FP64 norms), so the set of
// computed block products is unchanged; only the block products themselves are
// evaluated with low-precision matrix cores:
//  * slicing (once per operand per call, HBM-bound): every op(A) row / op(B)
//    column v of a block gets a power-of-two scale 2^e_v > max|x_v| and the
//    integer n = rint(x 2^(beta*S - e_v)) (|n| <= 2^(beta*S)) is cut into S
//    balanced beta-bit digits d_1..d_S (d_1 most significant), stored as FP16
//    (beta = 9) or BF16 (beta = 7) values, which hold them exactly;
//  * product: x_A x_B = sum_{p,q} d_p d_q 2^(e_r + f_c - beta (p+q)); pairs
//    with p + q <= S + 1 are kept (the error-free-transformation split of
//    Ozaki, Ogita, Oishi & Rump, Numer. Algorithms 59 (2012), in the integer
//    digit form of Ootomo & Yokota 2022 / Uchino et al. 2024);
//  * every digit product sum over one MFMA k-step (32) is an integer below
//    2^24, so the FP32 accumulation inside v_mfma_f32_16x16x32_{f16,bf16} is
//    exact; after each k-step the per-weight FP32 sums are scaled by
//    2^(e_r + f_c - beta w) and added into FP64 accumulators.
// Truncation error per product is about 2^(-beta*S) relative to the row/column
// maxima of the two blocks (S = 4 FP16 slices: 2^-36).
#include <algorithm>
#include <cmath>

#include "sa_internal.hpp"

namespace sa {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef double dbl4 __attribute__((ext_vector_type(4)));

constexpr int kOzK = 32;  // MFMA k of the 16x16x32 low-precision instructions

struct OzProd {
  const uint16_t* a;  // S slices of op(A) block, row-major: [p][row][kp]
  const uint16_t* b;  // S slices of op(B) block, column-major: [p][col][kp]
  const int32_t* ea;  // row exponents
  const int32_t* eb;  // column exponents
  int32_t kp;         // k padded to kOzK
  int32_t ma;         // rows of the op(A) block (slice stride)
  int32_t nb;         // columns of the op(B) block (slice stride)
  int32_t pad;
};

// --------------------------------------------------------------- slicing --
struct SliceArgs {
  const int64_t* off;     // block table of the operand
  const double* arena;
  const int32_t* rsz;
  const int32_t* csz;
  int32_t nrb;
  int64_t nb;
  int role_b;  // 0: operand A (vectors = rows of op(A)), 1: operand B (columns of op(B))
  int trans;
};

__device__ inline void slice_geom(const SliceArgs& s, int64_t b, int& nv, int& kd, int64_t& sv, int64_t& sk) {
  const int rb = (int)(b % s.nrb), cb = (int)(b / s.nrb);
  const int r = s.rsz[rb], c = s.csz[cb];
  if (!s.role_b) {           // op(A) rows
    if (!s.trans) { nv = r; kd = c; sv = 1; sk = r; }
    else          { nv = c; kd = r; sv = r; sk = 1; }
  } else {                   // op(B) columns
    if (!s.trans) { nv = c; kd = r; sv = r; sk = 1; }
    else          { nv = r; kd = c; sv = 1; sk = r; }
  }
}

__global__ void k_slice_need(SliceArgs s, int S, int64_t* __restrict__ need_sl, int64_t* __restrict__ need_ex) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= s.nb) return;
  int nv, kd;
  int64_t sv, sk;
  slice_geom(s, b, nv, kd, sv, sk);
  const bool pres = s.off[b] >= 0;
  const int64_t kp = (int64_t)ceil_div(kd, kOzK) * kOzK;
  need_sl[b] = pres ? (int64_t)S * nv * kp : 0;
  need_ex[b] = pres ? nv : 0;
}

template <bool BF>
__global__ void __launch_bounds__(256) k_slice(SliceArgs s, int S, int beta, const int64_t* __restrict__ sl_off,
                                               const int64_t* __restrict__ ex_off, uint16_t* __restrict__ slices,
                                               int32_t* __restrict__ exps) {
  const int64_t b = blockIdx.x;
  const int64_t o = s.off[b];
  if (o < 0) return;
  int nv, kd;
  int64_t sv, sk;
  slice_geom(s, b, nv, kd, sv, sk);
  const int kp = (int)ceil_div(kd, kOzK) * kOzK;
  const double* x = s.arena + o;
  uint16_t* out = slices + sl_off[b];
  int32_t* ex = exps + ex_off[b];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t mask = ((int64_t)1 << beta) - 1, half = (int64_t)1 << (beta - 1);
  for (int v = w; v < nv; v += 4) {
    double mx = 0.0;
    for (int kk = lane; kk < kd; kk += 64) mx = fmax(mx, fabs(x[v * sv + kk * sk]));
#pragma unroll
    for (int off = 32; off >= 1; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    int e = 0;
    if (mx > 0.0) (void)frexp(mx, &e);  // mx < 2^e
    if (lane == 0) ex[v] = e;
    for (int kk = lane; kk < kp; kk += 64) {
      const double a = kk < kd ? x[v * sv + kk * sk] : 0.0;
      int64_t n = mx > 0.0 ? (int64_t)rint(ldexp(a, beta * S - e)) : 0;
      // balanced digits, least significant first
      for (int p = S - 1; p >= 0; --p) {
        int64_t d;
        if (p > 0) {
          d = n & mask;
          if (d >= half) d -= (int64_t)1 << beta;
          n = (n - d) >> beta;
        } else {
          d = n;
        }
        uint16_t bits;
        if (BF) {
          const __bf16 h = (__bf16)(float)d;
          bits = __builtin_bit_cast(uint16_t, h);
        } else {
          const _Float16 h = (_Float16)(float)d;
          bits = __builtin_bit_cast(uint16_t, h);
        }
        out[(int64_t)p * nv * kp + (int64_t)v * kp + kk] = bits;
      }
    }
  }
}

// ------------------------------------------------------------- records --
struct OzTables {
  const int64_t* cnt;
  const int64_t* seg;
  const int64_t* need_scan;
  const int64_t* out_scan;
  const int32_t* kb;
  int64_t* coff;
  double* cnorm;
  double* carena;
  const int32_t* crsz;
  const int32_t* ccsz;
  int32_t nib, njb;
  int64_t cbase;
  const int64_t* aoff;
  int32_t a_nrb;
  const int64_t* boff;
  int32_t b_nrb;
  const int32_t* ksz;
  int32_t ta, tb;
  const uint16_t* asl;
  const int64_t* asl_off;
  const int32_t* aex;
  const int64_t* aex_off;
  const uint16_t* bsl;
  const int64_t* bsl_off;
  const int32_t* bex;
  const int64_t* bex_off;
};

__global__ void k_oz_records(OzTables t, GemmOut* __restrict__ outs, OzProd* __restrict__ prods) {
  const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = (int64_t)t.nib * t.njb;
  if (c >= nc) return;
  const int64_t cnt = t.cnt[c];
  if (cnt == 0) return;
  const int ib = (int)(c % t.nib), jb = (int)(c / t.nib);
  int64_t off = t.coff[c];
  const bool fresh = off < 0;
  if (fresh) {
    off = t.cbase + t.need_scan[c];
    t.coff[c] = off;
    t.cnorm[c] = 0.0;
  }
  const int64_t seg = t.seg[c];
  GemmOut o;
  o.c = t.carena + off;
  o.m = t.crsz[ib];
  o.n = t.ccsz[jb];
  o.p0 = (int32_t)seg;
  o.p1 = (int32_t)(seg + cnt);
  o.acc = fresh ? 0 : 1;
  o.pad = 0;
  outs[t.out_scan[c]] = o;
  for (int64_t p = seg; p < seg + cnt; ++p) {
    const int kb = t.kb[p];
    const int64_t ai = t.ta ? (int64_t)kb + (int64_t)ib * t.a_nrb : (int64_t)ib + (int64_t)kb * t.a_nrb;
    const int64_t bi = t.tb ? (int64_t)jb + (int64_t)kb * t.b_nrb : (int64_t)kb + (int64_t)jb * t.b_nrb;
    OzProd pr;
    pr.a = t.asl + t.asl_off[ai];
    pr.b = t.bsl + t.bsl_off[bi];
    pr.ea = t.aex + t.aex_off[ai];
    pr.eb = t.bex + t.bex_off[bi];
    pr.kp = (int32_t)ceil_div(t.ksz[kb], kOzK) * kOzK;
    pr.ma = o.m;
    pr.nb = o.n;
    pr.pad = 0;
    prods[p] = pr;
  }
}

// ----------------------------------------------------------------- GEMM --
// Wave-per-tile: each wave owns a (16*FM) x (16*FN) output tile.  Operand
// fragments (8 consecutive k per lane) are 16-byte loads from the slices.
template <int S, int BETA, bool BF, int FM, int FN, int NW>
__global__ void __launch_bounds__(64 * NW)
    k_gemm_ozaki(const GemmOut* __restrict__ outs, const OzProd* __restrict__ prods, const GemmItem* __restrict__ items,
                 int64_t nitems, double alpha) {
  constexpr int NWt = S;  // weights w = 2 .. S+1
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t it = (int64_t)blockIdx.x * NW + w;
  if (it >= nitems) return;
  const GemmItem item = items ? items[it] : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = outs[item.out];
  const int m = o.m, n = o.n;
  const int m0 = item.tm * 16 * FM, n0 = item.tn * 16 * FN;
  const int li = lane & 15, lg = lane >> 4;

  dbl4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  for (int p = o.p0; p < o.p1; ++p) {
    const OzProd pr = prods[p];
    // exponents of the rows (D layout: row 4*lg + v) and columns this lane folds
    int er[FM][4], ec[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int r = m0 + i * 16 + 4 * lg + v;
        er[i][v] = r < m ? pr.ea[r] : 0;
      }
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int cc = n0 + j * 16 + li;
      ec[j] = cc < n ? pr.eb[cc] : 0;
    }
    const int64_t sa_ = (int64_t)pr.ma * pr.kp, sb_ = (int64_t)pr.nb * pr.kp;
    for (int k0 = 0; k0 < pr.kp; k0 += kOzK) {
      uint4 af[S][FM], bfr[S][FN];
#pragma unroll
      for (int q = 0; q < S; ++q) {
#pragma unroll
        for (int i = 0; i < FM; ++i) {
          const int r = m0 + i * 16 + li;
          af[q][i] = r < m ? *reinterpret_cast<const uint4*>(pr.a + q * sa_ + (int64_t)r * pr.kp + k0 + 8 * lg)
                           : make_uint4(0, 0, 0, 0);
        }
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const int cc = n0 + j * 16 + li;
          bfr[q][j] = cc < n ? *reinterpret_cast<const uint4*>(pr.b + q * sb_ + (int64_t)cc * pr.kp + k0 + 8 * lg)
                             : make_uint4(0, 0, 0, 0);
        }
      }
      flt4 s32[NWt][FM][FN];
#pragma unroll
      for (int x = 0; x < NWt; ++x)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) s32[x][i][j] = flt4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int pa = 0; pa < S; ++pa)
#pragma unroll
        for (int pb = 0; pb < S; ++pb) {
          if (pa + pb > S - 1) continue;  // keep p + q <= S + 1 (1-based)
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j) {
              if (BF) {
                s32[pa + pb][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(
                    __builtin_bit_cast(bf16x8, af[pa][i]), __builtin_bit_cast(bf16x8, bfr[pb][j]), s32[pa + pb][i][j], 0,
                    0, 0);
              } else {
                s32[pa + pb][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(
                    __builtin_bit_cast(half8, af[pa][i]), __builtin_bit_cast(half8, bfr[pb][j]), s32[pa + pb][i][j], 0,
                    0, 0);
              }
            }
        }
      // fold: weight index x corresponds to p + q = x + 2 (1-based)
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double t = 0.0;
#pragma unroll
            for (int x = 0; x < NWt; ++x) t += ldexp((double)s32[x][i][j][v], -BETA * (x + 2));
            acc[i][j][v] += ldexp(t, er[i][v] + ec[j]);
          }
    }
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + j * 16 + li;
      if (col >= n) continue;
      double* cp = o.c + (int64_t)col * m;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = m0 + i * 16 + 4 * lg + v;
        if (row < m) {
          double x = alpha * acc[i][j][v];
          if (o.acc) x = x + cp[row];
          cp[row] = x;
        }
      }
    }
}

using OzFn = void (*)(const GemmOut*, const OzProd*, const GemmItem*, int64_t, double);

struct OzCfg {
  int S, beta;
  bool bf;
  int fm, fn, nw;
  OzFn fn_;
};

#define SA_OZ(S, BETA, BF, FM, FN, NW) OzCfg{S, BETA, BF, FM, FN, NW, k_gemm_ozaki<S, BETA, BF, FM, FN, NW>}

static const OzCfg kOzCfgs[] = {
    SA_OZ(3, 9, false, 2, 2, 4), SA_OZ(4, 9, false, 2, 2, 4), SA_OZ(5, 9, false, 2, 2, 4),
    SA_OZ(3, 9, false, 1, 1, 4), SA_OZ(4, 9, false, 1, 1, 4), SA_OZ(5, 9, false, 1, 1, 4),
    SA_OZ(4, 7, true, 2, 2, 4),  SA_OZ(5, 7, true, 2, 2, 4),  SA_OZ(6, 7, true, 2, 2, 4),
    SA_OZ(4, 7, true, 1, 1, 4),  SA_OZ(5, 7, true, 1, 1, 4),  SA_OZ(6, 7, true, 1, 1, 4),
};

__global__ void k_count_tiles_oz(const GemmOut* __restrict__ outs, int64_t nouts, int TM, int TN,
                                 int64_t* __restrict__ cnt) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  cnt[o] = ceil_div(outs[o].m, TM) * ceil_div(outs[o].n, TN);
}

__global__ void k_emit_items_oz(const GemmOut* __restrict__ outs, int64_t nouts, int TM, int TN,
                                const int64_t* __restrict__ start, GemmItem* __restrict__ items) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  const int tm = (int)ceil_div(outs[o].m, TM), tn = (int)ceil_div(outs[o].n, TN);
  int64_t s = start[o];
  for (int b = 0; b < tn; ++b)
    for (int a = 0; a < tm; ++a) items[s++] = GemmItem{(int32_t)o, (int16_t)a, (int16_t)b};
}

// slice one operand (all present blocks); returns device pointers
struct Sliced {
  const uint16_t* sl;
  const int64_t* sl_off;
  const int32_t* ex;
  const int64_t* ex_off;
};

static Sliced slice_operand(sa_context* c, sa_bsm* X, bool role_b, bool trans, int S, int beta, bool bf,
                            const char* tag) {
  const int64_t nb = X->nb();
  std::string t(tag);
  DevBuf& bso = c->buf((t + "_sloff").c_str());
  DevBuf& beo = c->buf((t + "_exoff").c_str());
  DevBuf& btot = c->buf((t + "_tot").c_str());
  bso.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  beo.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  btot.reserve(sizeof(int64_t) * 2);
  SliceArgs a{X->d_off, X->data(), X->d_rsz, X->d_csz, X->nrb, nb, role_b ? 1 : 0, trans ? 1 : 0};
  hipLaunchKernelGGL(k_slice_need, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, c->stream, a, S,
                     bso.as<int64_t>(), beo.as<int64_t>());
  scan_exclusive_i64_async(c, bso.as<int64_t>(), nb, btot.as<int64_t>());
  scan_exclusive_i64_async(c, beo.as<int64_t>(), nb, btot.as<int64_t>() + 1);
  int64_t* h = static_cast<int64_t*>(pinned(c, 2 * sizeof(int64_t)));
  SA_CHECK_HIP(hipMemcpyAsync(h, btot.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  const int64_t nsl = h[0], nex = h[1];
  DevBuf& bsl = c->buf((t + "_sl").c_str());
  DevBuf& bex = c->buf((t + "_ex").c_str());
  bsl.reserve(sizeof(uint16_t) * (size_t)(nsl + 8));
  bex.reserve(sizeof(int32_t) * (size_t)(nex + 1));
  {
    ProfScope ps(c, "ozaki_slice");
    if (nb) {
      if (bf)
        hipLaunchKernelGGL(k_slice<true>, dim3((unsigned)nb), dim3(256), 0, c->stream, a, S, beta,
                           (const int64_t*)bso.p, (const int64_t*)beo.p, bsl.as<uint16_t>(), bex.as<int32_t>());
      else
        hipLaunchKernelGGL(k_slice<false>, dim3((unsigned)nb), dim3(256), 0, c->stream, a, S, beta,
                           (const int64_t*)bso.p, (const int64_t*)beo.p, bsl.as<uint16_t>(), bex.as<int32_t>());
    }
  }
  SA_CHECK_HIP(hipGetLastError());
  return Sliced{bsl.as<uint16_t>(), bso.as<int64_t>(), bex.as<int32_t>(), beo.as<int64_t>()};
}

static int pick_cfg(int S, bool bf, int max_m, int max_n, int tile) {
  const int want_fm = (std::max(max_m, max_n) > 16) ? 2 : 1;
  for (int i = 0; i < (int)(sizeof(kOzCfgs) / sizeof(kOzCfgs[0])); ++i) {
    const OzCfg& k = kOzCfgs[i];
    if (k.S == S && k.bf == bf && (tile < 0 ? k.fm == want_fm : i == tile)) return i;
  }
  return -1;
}

int64_t ozaki_multiply(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh, double alpha,
                       const sa_opts_t& o) {
  sa_context* c = C->ctx;
  const bool bf = o.slice_type == SA_SLICE_BF16;
  const int S = o.slices;
  const int beta = bf ? 7 : 9;
  const int cfg = pick_cfg(S, bf, 64, 64, -1);
  SA_REQUIRE(cfg >= 0, "ozaki: %d %s slices not supported (FP16: 3..5, BF16: 4..6)", S, bf ? "BF16" : "FP16");
  const ScreenResult R = screen_and_allocate(C, A, ta, B, tb, thresh, nullptr, nullptr, 0, false);
  if (R.ntasks == 0) return 0;
  // slice operands; share when both roles read the same vectors of one matrix
  const sa_bsm* ra = A->base ? A->base : A;
  const sa_bsm* rb = B->base ? B->base : B;
  const Sliced sa_ = slice_operand(c, A, false, ta, S, beta, bf, "oz_a");
  const Sliced sb_ = (ra == rb && ta != tb && A == B) ? sa_ : slice_operand(c, B, true, tb, S, beta, bf, "oz_b");

  DevBuf& bouts = c->buf("mm_outs");
  DevBuf& bprods = c->buf("oz_prods");
  bouts.reserve(sizeof(GemmOut) * (size_t)(R.nouts + 1));
  bprods.reserve(sizeof(OzProd) * (size_t)(R.ntasks + 1));
  const std::vector<int32_t>& ksz_h = ta ? A->rsz : A->csz;
  (void)ksz_h;
  OzTables t;
  t.cnt = R.cnt; t.seg = R.seg; t.need_scan = R.need_scan; t.out_scan = R.out_scan; t.kb = R.kb;
  t.coff = C->d_off; t.cnorm = C->d_norm; t.carena = C->data(); t.crsz = C->d_rsz; t.ccsz = C->d_csz;
  t.nib = R.nib; t.njb = R.njb; t.cbase = R.cbase;
  t.aoff = A->d_off; t.a_nrb = A->nrb; t.boff = B->d_off; t.b_nrb = B->nrb;
  t.ksz = ta ? A->d_rsz : A->d_csz;
  t.ta = ta; t.tb = tb;
  t.asl = sa_.sl; t.asl_off = sa_.sl_off; t.aex = sa_.ex; t.aex_off = sa_.ex_off;
  t.bsl = sb_.sl; t.bsl_off = sb_.sl_off; t.bex = sb_.ex; t.bex_off = sb_.ex_off;
  hipLaunchKernelGGL(k_oz_records, dim3((unsigned)ceil_div(R.nc, 128)), dim3(128), 0, c->stream, t,
                     bouts.as<GemmOut>(), bprods.as<OzProd>());

  int32_t mm = 0, mn = 0;
  for (int32_t v : C->rsz) mm = std::max(mm, v);
  for (int32_t v : C->csz) mn = std::max(mn, v);
  const int cf = pick_cfg(S, bf, mm, mn, -1);
  const OzCfg& k = kOzCfgs[cf];
  const int TM = 16 * k.fm, TN = 16 * k.fn;
  const GemmItem* items = nullptr;
  int64_t nitems = R.nouts;
  if (mm > TM || mn > TN) {
    c->ws_gemm_items.reserve(sizeof(int64_t) * (size_t)(R.nouts + 1));
    int64_t* cnt = c->ws_gemm_items.as<int64_t>();
    hipLaunchKernelGGL(k_count_tiles_oz, dim3((unsigned)ceil_div(R.nouts, 256)), dim3(256), 0, c->stream,
                       bouts.as<GemmOut>(), R.nouts, TM, TN, cnt);
    nitems = scan_exclusive_i64(c, cnt, R.nouts);
    c->ws_gemm_items2.reserve(sizeof(GemmItem) * (size_t)(nitems + 1));
    hipLaunchKernelGGL(k_emit_items_oz, dim3((unsigned)ceil_div(R.nouts, 256)), dim3(256), 0, c->stream,
                       bouts.as<GemmOut>(), R.nouts, TM, TN, (const int64_t*)cnt, c->ws_gemm_items2.as<GemmItem>());
    items = c->ws_gemm_items2.as<GemmItem>();
  }
  SA_REQUIRE(nitems < ((int64_t)1 << 31), "ozaki: too many work items");
  {
    ProfScope ps(c, "gemm_ozaki");
    hipLaunchKernelGGL(k.fn_, dim3((unsigned)ceil_div(nitems, k.nw)), dim3(64 * k.nw), 0, c->stream,
                       (const GemmOut*)bouts.as<GemmOut>(), (const OzProd*)bprods.as<OzProd>(), items, nitems, alpha);
  }
  SA_CHECK_HIP(hipGetLastError());
  return R.ntasks;
}

}  // namespace sa
