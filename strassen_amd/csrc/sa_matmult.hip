// sa_matmult.hip -- norm-screened block GEMM C = alpha op(A) op(B) + beta C
// (blocksparsematrix.hpp:292-414) on the device.
//
// Pipeline (all on the context stream):
//  1. beta pre-pass (zero / scale), as the reference;
//  2. screening, pass 1: one wave per C block (ib, jb) walks kb in chunks of
//     64 lanes, tests  A(ib,kb), B(kb,jb) present  and
//     |A||B| >= thresh * s_i * s_j * s_k,  counts survivors with a wavefront
//     ballot;
//  3. device exclusive scans: task segments, new-C-block arena offsets, the
//     compacted list of non-empty outputs (one host readback of 3 totals);
//  4. screening, pass 2: the same walk writes each survivor kb at its
//     ballot-prefix position -- the canonical task list (C blocks column-major,
//     kb ascending), bit-identical to the oracle's;
//  5. record build: one GemmOut per non-empty C block, one GemmProd per task;
//  6. one grouped FP64 MFMA launch over all output tiles.
// A C block created here starts with a zero cached norm, like the reference's
// Mat(ni, nj, alloc) at blocksparsematrix.hpp:398-404.
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <vector>

#include "sa_internal.hpp"

namespace sa {

struct ScreenArgs {
  const int32_t* rowstart;  // CSR of op(A): present kb of row ib, ascending
  const int32_t* rowlist;
  const int64_t* aoff;
  const double* anorm;
  const int64_t* boff;
  const double* bnorm;
  const int32_t* iscr;  // screening sizes of op(A) rows
  const int32_t* jscr;  // screening sizes of op(B) columns
  const int32_t* kscr;  // screening sizes of the inner dimension
  int32_t a_nrb, b_nrb;
  int32_t nib, njb, nkb;
  int32_t ta, tb;
  double thresh;
  // optional Strassen coverage: tree bits of levels [0, nlev) at lvl_off[l]
  const uint8_t* tree;
  const int64_t* lvl_off;
  int32_t nlev;
  // optional: products into an empty C block are dropped (Strassen entry,
  // strassen.hpp:441-444 -- the leaf dgemm runs with m=n=0)
  const int64_t* coff;
  // symmetric C: blocks below the diagonal are not screened (their partners
  // above it carry the same triples)
  int32_t sym;
};

__device__ inline bool covered_by_strassen(const ScreenArgs& s, int ib, int jb, int kb) {
  for (int l = s.nlev - 1; l >= 0; --l) {
    const int sh = l + 1;
    const int64_t ns = ((int64_t)s.nib + (1 << sh) - 1) >> sh;
    const int64_t idx = (int64_t)(ib >> sh) + (int64_t)(jb >> sh) * ns + (int64_t)(kb >> sh) * ns * ns;
    if (s.tree[s.lvl_off[l] + idx]) return true;
  }
  return false;
}

// PASS 1: cnt[c] = number of surviving kb.  PASS 2: write kb at seg[c] + rank.
template <int PASS>
__global__ void __launch_bounds__(256) k_screen(ScreenArgs s, int64_t* __restrict__ cnt_or_seg,
                                                int32_t* __restrict__ kb_out) {
  const int lane = threadIdx.x & 63;
  const int64_t c = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int64_t nc = (int64_t)s.nib * s.njb;
  if (c >= nc) return;
  const int ib = (int)(c % s.nib), jb = (int)(c / s.nib);
  const int64_t sij = (int64_t)s.iscr[ib] * (int64_t)s.jscr[jb];
  int64_t pos = PASS == 2 ? cnt_or_seg[c] : 0;
  const int r0 = s.rowstart[ib];
  const int r1 = ((s.coff && s.coff[c] < 0) || (s.sym && ib > jb)) ? r0 : s.rowstart[ib + 1];
  for (int q0 = r0; q0 < r1; q0 += 64) {
    const int q = q0 + lane;
    bool ok = false;
    int kb = 0;
    if (q < r1) {
      kb = s.rowlist[q];
      const int64_t ai = s.ta ? (int64_t)kb + (int64_t)ib * s.a_nrb : (int64_t)ib + (int64_t)kb * s.a_nrb;
      const int64_t bi = s.tb ? (int64_t)jb + (int64_t)kb * s.b_nrb : (int64_t)kb + (int64_t)jb * s.b_nrb;
      if (s.boff[bi] >= 0) {
        const double est = s.anorm[ai] * s.bnorm[bi];
        const double thr = s.thresh * (double)(sij * (int64_t)s.kscr[kb]);
        ok = est >= thr;
        if (ok && s.nlev > 0) ok = !covered_by_strassen(s, ib, jb, kb);
      }
    }
    const uint64_t mask = __ballot(ok);
    if (PASS == 2 && ok) {
      const uint64_t below = mask & ((1ull << lane) - 1ull);
      kb_out[pos + __popcll(below)] = kb;
    }
    pos += __popcll(mask);
  }
  if (PASS == 1 && lane == 0) cnt_or_seg[c] = pos;
}

// CSR of the present blocks of op(A) by row: row ib, kb ascending.
// A block (ib, kb) of op(A) sits at table index ib*sr + kb*sk.  One wave per
// row (4 per workgroup), 64 kb per step.
__global__ void __launch_bounds__(256) k_row_count(const int64_t* __restrict__ aoff, int32_t nrows, int32_t nk,
                                                   int64_t sr, int64_t sk, int64_t* __restrict__ cnt) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r == 0 && lane == 0) cnt[nrows] = 0;  // the extra slot: its exclusive scan is the total
  if (r >= nrows) return;
  int64_t n = 0;
  for (int32_t k0 = 0; k0 < nk; k0 += 64) {
    const int32_t k = k0 + lane;
    n += __popcll(__ballot(k < nk && aoff[r * sr + (int64_t)k * sk] >= 0));
  }
  if (lane == 0) cnt[r] = n;
}

__global__ void __launch_bounds__(256) k_row_fill(const int64_t* __restrict__ aoff, int32_t nrows, int32_t nk,
                                                  int64_t sr, int64_t sk, const int64_t* __restrict__ start,
                                                  int32_t* __restrict__ rowstart, int32_t* __restrict__ list) {
  const int lane = threadIdx.x & 63;
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r > nrows) return;
  int64_t s = start[r];
  if (lane == 0) rowstart[r] = (int32_t)s;
  if (r == nrows) return;
  for (int32_t k0 = 0; k0 < nk; k0 += 64) {
    const int32_t k = k0 + lane;
    const bool present = k < nk && aoff[r * sr + (int64_t)k * sk] >= 0;
    const uint64_t mask = __ballot(present);
    if (present) list[s + __popcll(mask & ((1ull << lane) - 1ull))] = k;
    s += __popcll(mask);
  }
}

struct MMTables {
  const int64_t* cnt;
  const int64_t* seg;
  const int64_t* need_scan;  // exclusive scan of new-block sizes
  const int64_t* out_scan;   // exclusive scan of (cnt > 0)
  const int32_t* kb;
  int64_t* coff;
  double* cnorm;
  double* carena;
  const int32_t* crsz;
  const int32_t* ccsz;
  int32_t nib, njb;
  int64_t cbase;  // C.used before this multiply
  const int64_t* aoff;
  const double* aarena;
  int32_t a_nrb;
  const int64_t* boff;
  const double* barena;
  int32_t b_nrb;
  const int32_t* ksz;  // actual inner block sizes
  int32_t ta, tb;
};

// one thread per C block: assign new C blocks, emit the output record and the
// product records of its segment
__global__ void k_build_records(MMTables t, GemmOut* __restrict__ outs, GemmProd* __restrict__ prods) {
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
  o.group = (int32_t)(out_slot(ib, jb, t.nib) >> 2) + 1;  // group column (4 C blocks of one column), + 1
  o.hint = 0;
  o.pad_ = 0;
  outs[t.out_scan[out_slot(ib, jb, t.nib)]] = o;
  for (int64_t p = seg; p < seg + cnt; ++p) {
    const int kb = t.kb[p];
    const int64_t ai = t.ta ? (int64_t)kb + (int64_t)ib * t.a_nrb : (int64_t)ib + (int64_t)kb * t.a_nrb;
    const int64_t bi = t.tb ? (int64_t)jb + (int64_t)kb * t.b_nrb : (int64_t)kb + (int64_t)jb * t.b_nrb;
    GemmProd pr;
    pr.a = t.aarena + t.aoff[ai];
    pr.b = t.barena + t.boff[bi];
    pr.k = t.ksz[kb];
    pr.pad = 0;
    prods[p] = pr;
  }
}

// one thread per output slot (out_slot order, padded to whole groups): the
// slot's flag, 0 for padding, and for a real C block its arena need
__global__ void k_need_and_flags(const int64_t* __restrict__ cnt, const int64_t* __restrict__ coff,
                                 const int32_t* __restrict__ crsz, const int32_t* __restrict__ ccsz, int32_t nib,
                                 int32_t njb, int64_t nslots, int sym, int64_t* __restrict__ need,
                                 int64_t* __restrict__ outflag) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nslots) return;
  const int64_t ngi = ceil_div(nib, kOutGroup), g = s / (kOutGroup * kOutGroup);
  const int w = (int)(s % (kOutGroup * kOutGroup));
  const int ib = (int)(g % ngi) * kOutGroup + w % kOutGroup, jb = (int)(g / ngi) * kOutGroup + w / kOutGroup;
  if (ib >= nib || jb >= njb) {
    outflag[s] = 0;
    return;
  }
  const int64_t c = (int64_t)ib + (int64_t)jb * nib;
  const bool any = cnt[c] > 0;
  // a mirrored block (sym, below the diagonal) is written, not computed
  const bool mirrored = sym && ib > jb && cnt[(int64_t)jb + (int64_t)ib * nib] > 0;
  need[c] = ((any || mirrored) && coff[c] < 0) ? (int64_t)crsz[ib] * ccsz[jb] : 0;
  outflag[s] = any ? 1 : 0;
}

// Symmetric C: block (ib, jb), ib > jb, becomes the transpose of (jb, ib)
// wherever that block received products; a block created here starts with a
// zero cached norm, as every block a multiply creates.  One workgroup per
// lower block; the mirrored block's arena offset was assigned with the rest.
__global__ void __launch_bounds__(256) k_mirror_lower(const int64_t* __restrict__ cnt, int64_t* __restrict__ coff,
                                                      double* __restrict__ cnorm, double* __restrict__ carena,
                                                      const int32_t* __restrict__ crsz, const int64_t* __restrict__ need,
                                                      int64_t cbase, int32_t nib) {
  const int64_t c = blockIdx.x;
  const int ib = (int)(c % nib), jb = (int)(c / nib);
  if (ib <= jb) return;
  const int64_t up = (int64_t)jb + (int64_t)ib * nib;
  if (cnt[up] == 0) return;
  __shared__ int64_t dst;
  if (threadIdx.x == 0) {
    int64_t o = coff[c];
    if (o < 0) {
      o = cbase + need[c];
      coff[c] = o;
      cnorm[c] = 0.0;
    }
    dst = o;
  }
  __syncthreads();
  const int m = crsz[ib], n = crsz[jb];  // block (ib, jb) is m x n; its partner n x m
  const double* src = carena + coff[up];
  double* d = carena + dst;
  for (int64_t e = threadIdx.x; e < (int64_t)m * n; e += blockDim.x) {
    const int r = (int)(e % m), q = (int)(e / m);
    d[e] = src[q + (int64_t)r * n];
  }
}

// geometry of op(X) for screening / checking
struct OpGeo {
  int32_t nrb, ncb;
  const std::vector<int32_t>* rsz;
  const std::vector<int32_t>* csz;
  const std::vector<int32_t>* rscr;
  const std::vector<int32_t>* cscr;
  const int32_t* d_rscr;
  const int32_t* d_cscr;
  const int32_t* d_rsz;
  const int32_t* d_csz;
};

static OpGeo op_geo(const sa_bsm* m, bool t) {
  if (!t) return OpGeo{m->nrb, m->ncb, &m->rsz, &m->csz, &m->rscr, &m->cscr, m->d_rscr, m->d_cscr, m->d_rsz, m->d_csz};
  return OpGeo{m->ncb, m->nrb, &m->csz, &m->rsz, &m->cscr, &m->rscr, m->d_cscr, m->d_rscr, m->d_csz, m->d_rsz};
}

void check_mm_geometry(const sa_bsm* C, const sa_bsm* A, bool ta, const sa_bsm* B, bool tb) {
  const OpGeo ga = op_geo(A, ta), gb = op_geo(B, tb);
  SA_REQUIRE(ga.ncb == gb.nrb && *ga.csz == *gb.rsz, "matmult: inner block dimensions of op(A) and op(B) differ");
  SA_REQUIRE(ga.nrb == C->nrb && *ga.rsz == C->rsz && *ga.rscr == C->rscr,
             "matmult: row blocks of C and op(A) differ");
  SA_REQUIRE(gb.ncb == C->ncb && *gb.csz == C->csz && *gb.cscr == C->cscr,
             "matmult: column blocks of C and op(B) differ");
  SA_REQUIRE(C != A && C != B, "matmult: C must not alias A or B");
  SA_REQUIRE(!C->base, "matmult: C must not be a reshape view");
  SA_REQUIRE((A->base ? A->base : A) != C && (B->base ? B->base : B) != C, "matmult: C must not alias A or B");
  SA_REQUIRE(A->ctx == C->ctx && B->ctx == C->ctx, "matmult: matrices belong to different contexts");
}

// Screening and C-block allocation shared by the FP64 and split-precision
// multiplies (after the beta pre-pass).  tree/lvl_off (device, optional)
// exclude triples covered by executed Strassen nodes.
ScreenResult screen_and_allocate(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh,
                                 const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist) {
  TableMem mem(C->ctx);
  return screen_and_allocate(C, A, ta, B, tb, thresh, d_tree, d_lvl_off, nlev, c_must_exist, mem);
}

ScreenResult screen_and_allocate(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh,
                                 const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist,
                                 TableMem& mem, bool sym) {
  sa_context* c = C->ctx;
  const OpGeo ga = op_geo(A, ta), gb = op_geo(B, tb);
  const int32_t nib = ga.nrb, njb = gb.ncb, nkb = ga.ncb;
  const int64_t nc = (int64_t)nib * njb;
  ScreenResult R;
  R.nib = nib; R.njb = njb; R.nkb = nkb; R.nc = nc;
  R.sym = sym;
  c->last_nc = nc;
  c->last_nib = nib;
  c->last_ntasks = 0;
  c->last_sym = sym;
  if (nc == 0) return R;
  SA_REQUIRE(!sym || (nib == njb && C->rsz == C->csz), "matmult: a symmetric product needs a square block grid");

  const int64_t nslots = out_slots(nib, njb);
  int64_t* cnt = mem.get<int64_t>("mm_cnt", nc + 1);
  int64_t* seg = mem.get<int64_t>("mm_seg", nc + 1);
  int64_t* need = mem.get<int64_t>("mm_need", nc + 1);
  int64_t* outf = mem.get<int64_t>("mm_out", nslots + 1);
  int64_t* tot = mem.get<int64_t>("mm_totals", 8);
  // scan partials, shared by the scans below (they run in stream order)
  int64_t* sws = mem.fixed() ? mem.get<int64_t>("mm_scan_ws", scan_ws_elems(std::max<int64_t>(nc, nslots) + 1))
                             : nullptr;

  // CSR of op(A) rows (one extra slot so start[nib] is the total)
  int64_t* rc = mem.get<int64_t>("mm_rowcnt", nib + 2);
  int32_t* rstart = mem.get<int32_t>("mm_rowstart", nib + 2);
  int32_t* rlist = mem.get<int32_t>("mm_rowlist", (int64_t)nib * nkb + 1);
  {
    const int64_t sr = ta ? A->nrb : 1, sk = ta ? 1 : A->nrb;
    hipLaunchKernelGGL(k_row_count, dim3((unsigned)ceil_div(nib, 4)), dim3(256), 0, c->stream,
                       (const int64_t*)A->d_off, nib, nkb, sr, sk, rc);
    scan_exclusive_i64_async(c, rc, rc, nib + 1, tot + 3, sws);
    hipLaunchKernelGGL(k_row_fill, dim3((unsigned)ceil_div(nib + 1, 4)), dim3(256), 0, c->stream,
                       (const int64_t*)A->d_off, nib, nkb, sr, sk, (const int64_t*)rc, rstart, rlist);
  }

  ScreenArgs s;
  s.rowstart = rstart;
  s.rowlist = rlist;
  s.aoff = A->d_off; s.anorm = A->d_norm; s.boff = B->d_off; s.bnorm = B->d_norm;
  s.iscr = ga.d_rscr; s.jscr = gb.d_cscr; s.kscr = ga.d_cscr;
  s.a_nrb = A->nrb; s.b_nrb = B->nrb;
  s.nib = nib; s.njb = njb; s.nkb = nkb;
  s.ta = ta; s.tb = tb;
  s.thresh = thresh;
  s.tree = d_tree; s.lvl_off = d_lvl_off; s.nlev = nlev;
  s.coff = c_must_exist ? C->d_off : nullptr;
  s.sym = sym ? 1 : 0;

  {
    ProfScope ps(c, "screen");
    hipLaunchKernelGGL(k_screen<1>, dim3((unsigned)ceil_div(nc, 4)), dim3(256), 0, c->stream, s, cnt,
                       (int32_t*)nullptr);
  }
  hipLaunchKernelGGL(k_need_and_flags, dim3((unsigned)ceil_div(nslots, 256)), dim3(256), 0, c->stream,
                     (const int64_t*)cnt, (const int64_t*)C->d_off, (const int32_t*)C->d_rsz,
                     (const int32_t*)C->d_csz, nib, njb, nslots, sym ? 1 : 0, need, outf);
  {  // seg = exclusive scan of cnt; need and outf in place
    const int64_t* src[3] = {cnt, need, outf};
    int64_t* const dst[3] = {seg, need, outf};
    const int64_t n[3] = {nc, nc, nslots};
    int64_t* const totals[3] = {tot + 0, tot + 1, tot + 2};
    scan3_exclusive_i64_async(c, src, dst, n, totals, sws);
  }
  int64_t h[3];
  mem.read(tot, h, 3);
  R.ntasks = h[0];
  R.nnew = h[1];
  R.nouts = h[2];
  c->last_ntasks = R.ntasks;
  c->last_cnt = cnt;
  c->last_seg = seg;
  c->last_kb = nullptr;
  SA_REQUIRE(R.ntasks < ((int64_t)1 << 31), "matmult: task list exceeds 2^31 entries");

  R.cbase = C->used;
  if (mem.fixed()) {
    SA_REQUIRE((int64_t)(C->arena.bytes / sizeof(double)) >= R.cbase + R.nnew,
               "allocation-free multiply: C needs %lld more elements of block storage than its arena holds "
               "(reserve it first with sa_bsm_reserve)",
               (long long)(R.cbase + R.nnew - (int64_t)(C->arena.bytes / sizeof(double))));
  } else {
    bsm_ensure_capacity(C, R.cbase + R.nnew);
  }
  C->used = R.cbase + R.nnew;
  R.cnt = cnt; R.seg = seg; R.need_scan = need; R.out_scan = outf;
  if (R.ntasks == 0) return R;

  int32_t* kb = mem.get<int32_t>("mm_kb", R.ntasks + 1);
  {
    ProfScope ps(c, "screen");
    hipLaunchKernelGGL(k_screen<2>, dim3((unsigned)ceil_div(nc, 4)), dim3(256), 0, c->stream, s, seg, kb);
  }
  R.kb = kb;
  c->last_kb = kb;
  return R;
}

void mirror_lower(sa_bsm* C, const ScreenResult& R) {
  if (!R.sym || R.nc == 0 || R.ntasks == 0) return;
  hipLaunchKernelGGL(k_mirror_lower, dim3((unsigned)R.nc), dim3(256), 0, C->ctx->stream, R.cnt, C->d_off, C->d_norm,
                     C->data(), (const int32_t*)C->d_rsz, R.need_scan, R.cbase, R.nib);
  SA_CHECK_HIP(hipGetLastError());
}

MMTables mm_tables(const ScreenResult& R, sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb) {
  const OpGeo ga = op_geo(A, ta);
  MMTables t;
  t.cnt = R.cnt; t.seg = R.seg; t.need_scan = R.need_scan; t.out_scan = R.out_scan; t.kb = R.kb;
  t.coff = C->d_off; t.cnorm = C->d_norm; t.carena = C->data(); t.crsz = C->d_rsz; t.ccsz = C->d_csz;
  t.nib = R.nib; t.njb = R.njb; t.cbase = R.cbase;
  t.aoff = A->d_off; t.aarena = A->data(); t.a_nrb = A->nrb;
  t.boff = B->d_off; t.barena = B->data(); t.b_nrb = B->nrb;
  t.ksz = ga.d_csz; t.ta = ta; t.tb = tb;
  return t;
}

static int64_t screened_multiply_mem(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh,
                                     double alpha, int tile_cfg, const uint8_t* d_tree, const int64_t* d_lvl_off,
                                     int nlev, bool c_must_exist, bool sym, TableMem& mem) {
  sa_context* c = C->ctx;
  const ScreenResult R =
      screen_and_allocate(C, A, ta, B, tb, thresh, d_tree, d_lvl_off, nlev, c_must_exist, mem, sym);
  if (R.ntasks == 0) return 0;
  GemmOut* outs = mem.get<GemmOut>("mm_outs", R.nouts + 1);
  GemmProd* prods = mem.get<GemmProd>("mm_prods", R.ntasks + 1);
  const MMTables t = mm_tables(R, C, A, ta, B, tb);
  hipLaunchKernelGGL(k_build_records, dim3((unsigned)ceil_div(R.nc, 128)), dim3(128), 0, c->stream, t, outs, prods);

  const OpGeo ga = op_geo(A, ta);
  GemmBatch gb2;
  gb2.outs = outs;
  gb2.nouts = R.nouts;
  gb2.prods = prods;
  gb2.nprods = R.ntasks;
  gb2.ta = ta; gb2.tb = tb;
  gb2.alpha = alpha;
  int32_t mm = 0, mn = 0, mk = 0;
  for (int32_t v : C->rsz) mm = std::max(mm, v);
  for (int32_t v : C->csz) mn = std::max(mn, v);
  for (int32_t v : *ga.csz) mk = std::max(mk, v);
  gb2.max_m = mm; gb2.max_n = mn; gb2.max_k = mk;
  if (mem.fixed()) {  // one launch, tile choice without timing runs, items in the region
    c->last_kwindows = 1;
    const GemmPlan g = gemm_fp64_plan_fixed(gb2, tile_cfg, mem);
    c->last_gemm_cfg = g.cfg;
    gemm_fp64_launch(c, gb2, g);
  } else {
    const KWindows W = plan_kwindows(c, A, ta, B, tb, R, outs);
    c->last_kwindows = W.nwin;
    if (W.nwin == 1) {
      gemm_fp64_run(c, gb2, tile_cfg);
    } else {
      // one tile config (the shape class is the whole multiply's) and the
      // work items of all windows planned together; the launches then run
      // back to back in ascending kb
      GemmBatch g0 = gb2;
      for (int w = 0; w < W.nwin; ++w)
        if (W.off[(size_t)w + 1] > W.off[(size_t)w]) {
          g0.outs = W.outs + W.off[(size_t)w];
          g0.nouts = W.off[(size_t)w + 1] - W.off[(size_t)w];
          break;
        }
      const int cfg = gemm_fp64_choose(c, g0, tile_cfg);
      c->last_gemm_cfg = cfg;
      int TM = 0, TN = 0;
      gemm_fp64_tile(cfg, &TM, &TN);
      const WindowItems wi = window_items(c, W.outs, W.off, TM, TN, mm, mn);
      for (int w = 0; w < W.nwin; ++w) {
        GemmBatch gw = gb2;
        GemmPlan g;
        g.cfg = cfg;
        g.nitems = wi.off[(size_t)w + 1] - wi.off[(size_t)w];
        if (g.nitems == 0) continue;
        if (wi.items) {
          gw.outs = W.outs;
          g.items = wi.items + wi.off[(size_t)w];
        } else {
          gw.outs = W.outs + W.off[(size_t)w];
        }
        gw.nouts = W.off[(size_t)w + 1] - W.off[(size_t)w];
        ProfScope ps(c, "gemm_fp64");
        gemm_fp64_launch(c, gw, g);
      }
    }
  }
  mirror_lower(C, R);
  return R.ntasks;
}

int64_t screened_multiply(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh, double alpha,
                          int tile_cfg, const uint8_t* d_tree, const int64_t* d_lvl_off, int nlev, bool c_must_exist,
                          bool sym) {
  TableMem mem(C->ctx);
  return screened_multiply_mem(C, A, ta, B, tb, thresh, alpha, tile_cfg, d_tree, d_lvl_off, nlev, c_must_exist, sym,
                               mem);
}

// ------------------------------------------------------------- k-windows --
__global__ void k_kb_bytes(const int64_t* __restrict__ off, const int32_t* __restrict__ rsz,
                           const int32_t* __restrict__ csz, int32_t nrb, int64_t nb, int kb_is_row,
                           unsigned long long* __restrict__ bytes) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= nb || off[b] < 0) return;
  const int r = (int)(b % nrb), cc = (int)(b / nrb);
  atomicAdd(bytes + (kb_is_row ? r : cc), (unsigned long long)rsz[r] * (unsigned long long)csz[cc] * 8ull);
}

__global__ void k_out_bytes(const GemmOut* __restrict__ outs, int64_t nouts, unsigned long long* __restrict__ total) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  atomicAdd(total, (unsigned long long)outs[o].m * (unsigned long long)outs[o].n * 8ull);
}

// flags[w * nouts + o] = 1 when output o has a product in window w
__global__ void k_win_flags(const GemmOut* __restrict__ outs, int64_t nouts, const int32_t* __restrict__ kb,
                            const int32_t* __restrict__ win_of_kb, int64_t* __restrict__ flags) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  int last = -1;
  for (int32_t p = outs[o].p0; p < outs[o].p1; ++p) {
    const int w = win_of_kb[kb[p]];
    if (w != last) flags[(int64_t)w * nouts + o] = 1;
    last = w;
  }
}

__global__ void k_win_emit(const GemmOut* __restrict__ outs, int64_t nouts, const int32_t* __restrict__ kb,
                           const int32_t* __restrict__ win_of_kb, const int64_t* __restrict__ pos,
                           GemmOut* __restrict__ wouts) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  const GemmOut g = outs[o];
  int32_t p = g.p0;
  bool first = true;
  while (p < g.p1) {
    const int w = win_of_kb[kb[p]];
    int32_t q = p + 1;
    while (q < g.p1 && win_of_kb[kb[q]] == w) ++q;
    GemmOut x = g;
    x.p0 = p;
    x.p1 = q;
    x.acc = first ? g.acc : 1;
    wouts[pos[(int64_t)w * nouts + o]] = x;
    first = false;
    p = q;
  }
}

__global__ void k_gather_stride(const int64_t* __restrict__ src, int64_t stride, int64_t n, int64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i * stride];
}

static int64_t env_mb(const char* name, int64_t dflt) {
  const char* v = std::getenv(name);
  return v ? std::atoll(v) : dflt;
}

KWindows plan_kwindows(sa_context* c, sa_bsm* A, bool ta, sa_bsm* B, bool tb, const ScreenResult& R,
                       const GemmOut* outs) {
  KWindows W;
  W.outs = outs;
  W.off = {0, R.nouts};
  // operand budget per window and the cache size the C blocks must fit in
  const int64_t budget = env_mb("SA_KWINDOW_MB", 160) << 20;
  const int64_t cache = (int64_t)256 << 20;
  const int32_t nkb = R.nkb;
  if (budget <= 0 || nkb < 2 || R.nouts == 0) return W;
  // operands whose arenas hold at most two budgets in all need no byte count
  auto arena_bytes = [](const sa_bsm* X) { return 8 * (X->base ? X->base->used : X->used); };
  if (arena_bytes(A) + arena_bytes(B) <= 2 * budget) return W;
  DevBuf& bb = c->buf("kw_bytes");
  bb.reserve(sizeof(unsigned long long) * (size_t)(nkb + 1));
  unsigned long long* kbytes = bb.as<unsigned long long>();
  fill_i64(c, reinterpret_cast<int64_t*>(kbytes), nkb + 1, 0);
  hipLaunchKernelGGL(k_kb_bytes, dim3((unsigned)ceil_div(std::max<int64_t>(A->nb(), 1), 256)), dim3(256), 0,
                     c->stream, (const int64_t*)A->d_off, (const int32_t*)A->d_rsz, (const int32_t*)A->d_csz,
                     A->nrb, A->nb(), ta ? 1 : 0, kbytes);
  hipLaunchKernelGGL(k_kb_bytes, dim3((unsigned)ceil_div(std::max<int64_t>(B->nb(), 1), 256)), dim3(256), 0,
                     c->stream, (const int64_t*)B->d_off, (const int32_t*)B->d_rsz, (const int32_t*)B->d_csz,
                     B->nrb, B->nb(), tb ? 0 : 1, kbytes);
  hipLaunchKernelGGL(k_out_bytes, dim3((unsigned)ceil_div(R.nouts, 256)), dim3(256), 0, c->stream, outs, R.nouts,
                     kbytes + nkb);
  std::vector<unsigned long long> hb((size_t)nkb + 1);
  SA_CHECK_HIP(hipMemcpyAsync(hb.data(), kbytes, sizeof(unsigned long long) * (size_t)(nkb + 1),
                              hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  unsigned long long opnd = 0;
  for (int32_t k = 0; k < nkb; ++k) opnd += hb[(size_t)k];
  if ((int64_t)hb[(size_t)nkb] > cache || (int64_t)opnd <= 2 * budget) return W;
  std::vector<int32_t> win((size_t)nkb);
  int nwin = 0;
  unsigned long long acc = 0;
  for (int32_t k = 0; k < nkb; ++k) {
    if (acc > 0 && acc + hb[(size_t)k] > (unsigned long long)budget) {
      ++nwin;
      acc = 0;
    }
    acc += hb[(size_t)k];
    win[(size_t)k] = nwin;
  }
  ++nwin;
  if (nwin < 2) return W;
  DevBuf& bw = c->buf("kw_win");
  DevBuf& bf = c->buf("kw_flags");
  DevBuf& bo = c->buf("kw_outs");
  DevBuf& bofs = c->buf("kw_off");
  const int64_t nf = (int64_t)nwin * R.nouts;
  bw.reserve(sizeof(int32_t) * (size_t)nkb);
  bf.reserve(sizeof(int64_t) * (size_t)(nf + 1));
  bofs.reserve(sizeof(int64_t) * (size_t)(nwin + 1));
  SA_CHECK_HIP(hipMemcpyAsync(bw.p, win.data(), sizeof(int32_t) * (size_t)nkb, hipMemcpyHostToDevice, c->stream));
  fill_i64(c, bf.as<int64_t>(), nf + 1, 0);
  hipLaunchKernelGGL(k_win_flags, dim3((unsigned)ceil_div(R.nouts, 256)), dim3(256), 0, c->stream, outs, R.nouts,
                     R.kb, (const int32_t*)bw.p, bf.as<int64_t>());
  const int64_t total = scan_exclusive_i64(c, bf.as<int64_t>(), nf + 1);
  bo.reserve(sizeof(GemmOut) * (size_t)(total + 1));
  hipLaunchKernelGGL(k_win_emit, dim3((unsigned)ceil_div(R.nouts, 256)), dim3(256), 0, c->stream, outs, R.nouts,
                     R.kb, (const int32_t*)bw.p, (const int64_t*)bf.p, bo.as<GemmOut>());
  hipLaunchKernelGGL(k_gather_stride, dim3((unsigned)ceil_div(nwin + 1, 256)), dim3(256), 0, c->stream,
                     (const int64_t*)bf.p, R.nouts, (int64_t)(nwin + 1), bofs.as<int64_t>());
  W.off.assign((size_t)nwin + 1, 0);
  SA_CHECK_HIP(hipMemcpyAsync(W.off.data(), bofs.p, sizeof(int64_t) * (size_t)(nwin + 1), hipMemcpyDeviceToHost,
                              c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  SA_CHECK_HIP(hipGetLastError());
  W.nwin = nwin;
  W.outs = bo.as<GemmOut>();
  return W;
}

// ------------------------------------------------ allocation-free multiply --
// The tables of sa_matmult_noalloc live in this fixed device region.  Host
// code never passes `scratch` itself (its host-side name is only a stand-in
// for the device symbol): noalloc_setup looks up each device's own copy once,
// with that device current, and the entry point uses that address.
constexpr size_t kScratchBytes = (size_t)64 << 20;
static __device__ __align__(256) uint8_t scratch[kScratchBytes];

static std::mutex g_scratch_mu;
static std::vector<void*> g_scratch_ptr;  // per device

void noalloc_setup(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= device) g_scratch_ptr.resize((size_t)device + 1, nullptr);
  if (g_scratch_ptr[(size_t)device]) return;
  void* p = nullptr;
  SA_CHECK_HIP(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
  SA_REQUIRE(p != nullptr, "no device address for the scratch region on device %d", device);
  g_scratch_ptr[(size_t)device] = p;
}

static void* scratch_region(int device) {
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  SA_REQUIRE((int)g_scratch_ptr.size() > device && g_scratch_ptr[(size_t)device],
             "allocation-free multiply: scratch region of device %d not set up", device);
  return g_scratch_ptr[(size_t)device];
}

// blocksparsematrix.hpp:299-305
void beta_prepass(sa_bsm* C, double beta) {
  if (beta == 1.0) return;
  const int rc = beta == 0.0 ? sa_bsm_zero(C) : sa_bsm_scale(C, beta);
  if (rc != SA_OK) throw Error{rc};
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_matmult(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh, double alpha, double beta,
               const sa_opts_t* opts) {
  return guard([&] {
    SA_REQUIRE(C && A && B, "null matrix");
    sa_opts_t o;
    sa_opts_default(&o);
    if (opts) o = *opts;
    check_mm_geometry(C, A, transA != 0, B, transB != 0);
    SA_CHECK_HIP(hipSetDevice(C->ctx->device));
    beta_prepass(C, beta);
    if (o.mode == SA_MODE_FP64) {
      screened_multiply(C, A, transA != 0, B, transB != 0, thresh, alpha, o.tile, nullptr, nullptr, 0, false);
    } else if (o.mode == SA_MODE_OZAKI) {
      ozaki_multiply(C, A, transA != 0, B, transB != 0, thresh, alpha, o);
    } else {
      SA_REQUIRE(false, "matmult: unknown mode %d", o.mode);
    }
  });
}

int sa_matmult_noalloc(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh, double alpha,
                       double beta) {
  return guard([&] {
    SA_REQUIRE(C && A && B, "null matrix");
    check_mm_geometry(C, A, transA != 0, B, transB != 0);
    sa_context* c = C->ctx;
    SA_CHECK_HIP(hipSetDevice(c->device));
    TableMem mem(c, scratch_region(c->device), kScratchBytes);
    c->quiet = true;  // no profiling events
    try {
      beta_prepass(C, beta);
      screened_multiply_mem(C, A, transA != 0, B, transB != 0, thresh, alpha, -1, nullptr, nullptr, 0, false, false,
                            mem);
    } catch (...) {
      c->quiet = false;
      throw;
    }
    c->quiet = false;
  });
}

int sa_matmult_scratch_bytes(int64_t* bytes) {
  return guard([&] {
    SA_REQUIRE(bytes, "null argument");
    *bytes = (int64_t)kScratchBytes;
  });
}

int sa_tasklist_get(sa_ctx_t c, int32_t* triples, int64_t cap, int64_t* n) {
  return guard([&] {
    SA_REQUIRE(c && n, "null argument");
    *n = c->last_ntasks;
    if (c->last_ntasks == 0) return;
    if (c->last_sym) {  // full count: the mirrored blocks' lists as well
      std::vector<int64_t> cnt0((size_t)c->last_nc);
      SA_CHECK_HIP(hipMemcpyAsync(cnt0.data(), c->last_cnt, sizeof(int64_t) * c->last_nc, hipMemcpyDeviceToHost,
                                  c->stream));
      SA_CHECK_HIP(hipStreamSynchronize(c->stream));
      int64_t full = 0;
      for (int64_t ci = 0; ci < c->last_nc; ++ci) {
        const int64_t ib = ci % c->last_nib, jb = ci / c->last_nib;
        full += cnt0[(size_t)(ib > jb ? jb + ib * c->last_nib : ci)];
      }
      *n = full;
    }
    if (!triples || cap <= 0) return;
    const int64_t nc = c->last_nc, nt = c->last_ntasks;
    std::vector<int64_t> cnt(nc), seg(nc);
    std::vector<int32_t> kb(nt);
    SA_CHECK_HIP(hipMemcpyAsync(cnt.data(), c->last_cnt, sizeof(int64_t) * nc, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipMemcpyAsync(seg.data(), c->last_seg, sizeof(int64_t) * nc, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipMemcpyAsync(kb.data(), c->last_kb, sizeof(int32_t) * nt, hipMemcpyDeviceToHost, c->stream));
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));
    const int64_t nib = c->last_nib;
    int64_t w = 0;
    for (int64_t ci = 0; ci < nc && w < cap; ++ci) {
      const int64_t ib = ci % nib, jb = ci / nib;
      // a mirrored block of a symmetric product carries its partner's triples
      const int64_t src = (c->last_sym && ib > jb) ? jb + ib * nib : ci;
      for (int64_t p = seg[src]; p < seg[src] + cnt[src] && w < cap; ++p, ++w) {
        triples[3 * w + 0] = (int32_t)ib;
        triples[3 * w + 1] = (int32_t)jb;
        triples[3 * w + 2] = kb[p];
      }
    }
  });
}

}  // extern "C"
