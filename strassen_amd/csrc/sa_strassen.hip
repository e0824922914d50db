// sa_strassen.hip -- hybrid Strassen-sparse multiply (strassen.hpp) on the device.
//
//  * dry run (strassen.hpp:510-679): one thread per super-block node and level,
//    integer flop counts summed over the node's <= 8 children, compared with
//    the Winograd cost 7^(l+1)*2bs^3 + floor(15*fac[l]*4^(l+1)*bs^2) (computed
//    on the host exactly as the reference does), edge nodes excluded.  Bits are
//    identical to the oracle's.
//  * execution: the executed Strassen roots are the set bits with no set
//    ancestor ("Option A", strassen.hpp:12-19).  Every triple outside them is a
//    screened leaf (strassen.hpp:435-446) and runs through the screened task
//    builder with a tree mask.  Each root starts a tree of Winograd nodes
//    (strassen.hpp:359-405); nodes are processed breadth-first and batched:
//      - one pre-add launch per depth (A side and B side, strassen.hpp:153-213),
//      - one grouped FP64 MFMA launch for all leaf products of the batch
//        (dense 8-product form of strassen.hpp:311-323; also 1x1x1),
//      - one post-add launch per depth, deepest first (strassen.hpp:215-237
//        plus the M5..M7 accumulation and accumulate_into_bsm, :249-256).
//    Each Winograd node writes its 7 products into separate scratch grids
//    M1..M7 and the post-add forms C00 = M1+M7, C01 = M1+M3+M4+M5,
//    C10 = M1+M2+M4+M6, C11 = M1+M2+M3+M4 (the reference's schedule, with the
//    same association of the t = M4+M1 term).
//  Divergences (documented): empty blocks read as zero on the Strassen path;
//  scratch is sized exactly from the tree and roots are processed in batches
//  that fit a scratch budget.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "sa_internal.hpp"

namespace sa {


static const double kStrassenFac[20] = {10, 10, 10, 25, 35, 35, 35, 35, 35, 35,
                                        35, 35, 35, 35, 35, 35, 35, 35, 35, 35};

// ---------------------------------------------------------------- dry run --
struct DryArgs {
  const int64_t* aoff;
  const double* anorm;
  const int64_t* boff;
  const double* bnorm;
  const int32_t* rsz;  // actual sizes (square, uniform geometry)
  int32_t nb;          // blocks per side
  int64_t n;           // matrix size
  int32_t bs;
  double thr;          // thresh_per_block
};

// level 0 node flops from the <= 8 block triples of a 2x2x2 super-block
__global__ void k_dry_level0(DryArgs a, int64_t ns, uint64_t cost_s, uint64_t* __restrict__ cur,
                             uint8_t* __restrict__ bits, unsigned long long* __restrict__ totals) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  uint64_t f = 0, f0 = 0;
  for (int64_t di = 0; di < 2; ++di)
    for (int64_t dj = 0; dj < 2; ++dj)
      for (int64_t dk = 0; dk < 2; ++dk) {
        const int64_t ib = 2 * si + di, jb = 2 * sj + dj, kb = 2 * sk + dk;
        if (ib >= a.nb || jb >= a.nb || kb >= a.nb) continue;
        const int64_t ai = ib + kb * a.nb, bi = kb + jb * a.nb;
        const double est = a.anorm[ai] * a.bnorm[bi];
        const uint64_t mm = a.aoff[ai] >= 0 ? (uint64_t)a.rsz[ib] : 0;
        const uint64_t kk = a.aoff[ai] >= 0 ? (uint64_t)a.rsz[kb] : 0;
        const uint64_t nn = a.boff[bi] >= 0 ? (uint64_t)a.rsz[jb] : 0;
        const uint64_t t = (est >= a.thr ? 1u : 0u) * (2u * mm * nn * kk);
        f += t;
      }
  f0 = f;
  const int64_t sb = 2;
  const int64_t ib0 = si * sb, jb0 = sj * sb, kb0 = sk * sb;
  const int64_t ib1 = ib0 + sb < a.nb ? ib0 + sb : a.nb;
  const int64_t jb1 = jb0 + sb < a.nb ? jb0 + sb : a.nb;
  const int64_t kb1 = kb0 + sb < a.nb ? kb0 + sb : a.nb;
  const int64_t ie = ib1 * a.bs < a.n ? ib1 * a.bs : a.n;
  const int64_t je = jb1 * a.bs < a.n ? jb1 * a.bs : a.n;
  const int64_t ke = kb1 * a.bs < a.n ? kb1 * a.bs : a.n;
  const bool edge = (ie - ib0 * a.bs != a.bs * sb) || (je - jb0 * a.bs != a.bs * sb) || (ke - kb0 * a.bs != a.bs * sb);
  const bool dos = (f >= cost_s) && !edge;
  bits[s] = dos ? 1 : 0;
  cur[s] = dos ? cost_s : f;
  atomicAdd(&totals[0], (unsigned long long)f0);
  atomicAdd(&totals[1], (unsigned long long)cur[s]);
}

__global__ void k_dry_level(DryArgs a, int level, int64_t ns, int64_t pns, uint64_t cost_s,
                            const uint64_t* __restrict__ prev, uint64_t* __restrict__ cur, uint8_t* __restrict__ bits,
                            unsigned long long* __restrict__ total) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  const int64_t psb = (int64_t)1 << level, sb = psb * 2;
  const int64_t ib0 = si * sb, jb0 = sj * sb, kb0 = sk * sb;
  const int64_t ib1 = ib0 + sb < a.nb ? ib0 + sb : a.nb;
  const int64_t jb1 = jb0 + sb < a.nb ? jb0 + sb : a.nb;
  const int64_t kb1 = kb0 + sb < a.nb ? kb0 + sb : a.nb;
  const int64_t isub = ceil_div(ib1 - ib0, psb), jsub = ceil_div(jb1 - jb0, psb), ksub = ceil_div(kb1 - kb0, psb);
  uint64_t f = 0;
  for (int64_t x = 0; x < isub; ++x)
    for (int64_t y = 0; y < jsub; ++y)
      for (int64_t z = 0; z < ksub; ++z) f += prev[(2 * si + x) + (2 * sj + y) * pns + (2 * sk + z) * pns * pns];
  const int64_t ie = ib1 * a.bs < a.n ? ib1 * a.bs : a.n;
  const int64_t je = jb1 * a.bs < a.n ? jb1 * a.bs : a.n;
  const int64_t ke = kb1 * a.bs < a.n ? kb1 * a.bs : a.n;
  const bool edge = (ie - ib0 * a.bs != a.bs * sb) || (je - jb0 * a.bs != a.bs * sb) || (ke - kb0 * a.bs != a.bs * sb);
  const bool dos = (f >= cost_s) && !edge;
  bits[s] = dos ? 1 : 0;
  cur[s] = dos ? cost_s : f;
  atomicAdd(total, (unsigned long long)cur[s]);
}

struct Tree {
  int nlev = 0;
  std::vector<int64_t> lvl_off, lvl_n;  // node offset / count per level
  std::vector<uint64_t> lvl_flops;      // nlev+1 totals
  int64_t total = 0;
};

static bool uniform_square(const sa_bsm* A, const sa_bsm* B) {
  if (!A->uniform || !B->uniform) return false;
  if (A->nrb != A->ncb || B->nrb != B->ncb || A->nrb != B->nrb || A->nrb == 0) return false;
  if (A->nrow != A->ncol || B->nrow != B->ncol || A->nrow != B->nrow) return false;
  return A->bs_r == A->bs_c && B->bs_r == B->bs_c && A->bs_r == B->bs_r;
}

static Tree tree_shape(int32_t nb) {
  Tree t;
  t.nlev = (int)(size_t)std::log2((double)nb + 0.5);
  int64_t off = 0;
  for (int l = 0; l < t.nlev; ++l) {
    const int64_t sb = (int64_t)1 << (l + 1);
    const int64_t ns = ceil_div(nb, sb);
    t.lvl_off.push_back(off);
    t.lvl_n.push_back(ns * ns * ns);
    off += ns * ns * ns;
  }
  t.total = off;
  t.lvl_flops.assign(t.nlev + 1, 0);
  return t;
}

// device dry run; bits land in the ctx buffer "st_bits" (t.total bytes)
static void dryrun_device(sa_bsm* A, sa_bsm* B, double thr, Tree& t) {
  sa_context* c = A->ctx;
  DevBuf& bbits = c->buf("st_bits");
  DevBuf& bf0 = c->buf("st_f0");
  DevBuf& bf1 = c->buf("st_f1");
  DevBuf& btot = c->buf("st_tot");
  bbits.reserve((size_t)t.total + 16);
  btot.reserve(sizeof(uint64_t) * (size_t)(t.nlev + 2));
  fill_i64(c, btot.as<int64_t>(), t.nlev + 2, 0);
  if (t.nlev == 0) return;
  bf0.reserve(sizeof(uint64_t) * (size_t)(t.lvl_n[0] + 1));
  bf1.reserve(sizeof(uint64_t) * (size_t)(t.lvl_n[0] + 1));
  DryArgs a{A->d_off, A->d_norm, B->d_off, B->d_norm, A->d_rsz, A->nrb, A->nrow, A->bs_r, thr};
  const uint64_t bs = (uint64_t)A->bs_r;
  const uint64_t max_mult = 2u * bs * bs * bs, max_add = bs * bs;
  uint64_t* prev = bf0.as<uint64_t>();
  uint64_t* cur = bf1.as<uint64_t>();
  unsigned long long* tot = btot.as<unsigned long long>();
  ProfScope ps(c, "dryrun");
  for (int l = 0; l < t.nlev; ++l) {
    const uint64_t cost_mult = (uint64_t)std::round(std::pow(7.0, (double)(l + 1))) * max_mult;
    const uint64_t cost_add = (uint64_t)(15e0 * kStrassenFac[l] * std::pow(4.0, (double)(l + 1)) * (double)max_add);
    const uint64_t cost_s = cost_mult + cost_add;
    const int64_t ns = ceil_div(A->nrb, (int64_t)1 << (l + 1));
    const int64_t nn = ns * ns * ns;
    if (l == 0) {
      hipLaunchKernelGGL(k_dry_level0, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, a, ns, cost_s, cur,
                         bbits.as<uint8_t>() + t.lvl_off[0], tot);
    } else {
      const int64_t pns = ceil_div(A->nrb, (int64_t)1 << l);
      hipLaunchKernelGGL(k_dry_level, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, a, l, ns, pns, cost_s,
                         (const uint64_t*)prev, cur, bbits.as<uint8_t>() + t.lvl_off[l], tot + l + 1);
    }
    std::swap(prev, cur);
  }
  SA_CHECK_HIP(hipGetLastError());
}

static void download_tree(sa_context* c, Tree& t, std::vector<uint8_t>& bits) {
  bits.resize((size_t)t.total);
  if (t.total)
    SA_CHECK_HIP(hipMemcpyAsync(bits.data(), c->buf("st_bits").p, (size_t)t.total, hipMemcpyDeviceToHost, c->stream));
  std::vector<uint64_t> tot(t.nlev + 2);
  SA_CHECK_HIP(hipMemcpyAsync(tot.data(), c->buf("st_tot").p, sizeof(uint64_t) * (t.nlev + 2), hipMemcpyDeviceToHost,
                              c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  for (int l = 0; l <= t.nlev; ++l) t.lvl_flops[l] = tot[l];
}

// ------------------------------------------------------------ Winograd nodes --
struct GridRef {
  int32_t kind;  // 0 scratch, 1 A, 2 B, 3 C
  int32_t ld;    // scratch: leading dimension in blocks
  int64_t base;  // scratch: first block index
  int32_t r0, c0;
};

struct WNode {
  GridRef in[8];   // a, b, c, d (A quadrants); A(TL), B(BL), C(TR), D(BR) of the B operand
  GridRef out[4];  // 00, 01, 10, 11
  int64_t tmp_base;
  int64_t m_base;
  int32_t h;
  int32_t root;  // 1: outputs are C blocks, accumulate (accumulate_into_bsm)
};

struct LeafProd {
  GridRef x, y;  // h x h operands
  int64_t m_blk;  // first scratch block of the h x h result (ld = h)
  int64_t out0;   // first GemmOut index
  int64_t prod0;  // first GemmProd index
  int32_t h;
  int32_t pad;
};

struct PosItem {
  int32_t node;
  int16_t r, c;
  int32_t side;
};

struct Ptrs {
  double* scratch;
  int64_t bs2;
  const int64_t* aoff;
  const double* aarena;
  const int64_t* boff;
  const double* barena;
  const int64_t* coff;
  double* carena;
  int32_t nrb;
};

__device__ inline double* resolve(const GridRef& g, int r, int c, const Ptrs& P) {
  if (g.kind == 0) return P.scratch + (g.base + (int64_t)(g.r0 + r) + (int64_t)(g.c0 + c) * g.ld) * P.bs2;
  const int64_t i = (int64_t)(g.r0 + r) + (int64_t)(g.c0 + c) * P.nrb;
  if (g.kind == 1) {
    const int64_t o = P.aoff[i];
    return o < 0 ? nullptr : const_cast<double*>(P.aarena) + o;
  }
  if (g.kind == 2) {
    const int64_t o = P.boff[i];
    return o < 0 ? nullptr : const_cast<double*>(P.barena) + o;
  }
  const int64_t o = P.coff[i];
  return o < 0 ? nullptr : P.carena + o;
}

// Winograd pre-adds, A side (strassen.hpp:153-182) and B side (:184-213)
__global__ void __launch_bounds__(256) k_preadd(const WNode* __restrict__ nodes, const PosItem* __restrict__ items,
                                                Ptrs P) {
  const PosItem it = items[blockIdx.x];
  const WNode& nd = nodes[it.node];
  const int s = it.side;
  const double* x[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) x[q] = resolve(nd.in[s * 4 + q], it.r, it.c, P);
  double* o[4];
  const int64_t hh = (int64_t)nd.h * nd.h, pos = (int64_t)it.r + (int64_t)it.c * nd.h;
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = P.scratch + (nd.tmp_base + (int64_t)(s * 4 + q) * hh + pos) * P.bs2;
  for (int64_t e = threadIdx.x; e < P.bs2; e += blockDim.x) {
    const double x0 = x[0] ? x[0][e] : 0.0, x1 = x[1] ? x[1][e] : 0.0;
    const double x2 = x[2] ? x[2][e] : 0.0, x3 = x[3] ? x[3][e] : 0.0;
    if (s == 0) {  // a, b, c, d -> c-a, c+d, c+d-a, b-(c+d-a)
      o[0][e] = x2 - x0;
      const double v = x2 + x3;
      o[1][e] = v;
      const double w = v - x0;
      o[2][e] = w;
      o[3][e] = x1 - w;
    } else {       // A, B, C, D -> C-D, C-A, D-(C-A), B-(D-(C-A))
      o[0][e] = x2 - x3;
      const double v = x2 - x0;
      o[1][e] = v;
      const double w = x3 - v;
      o[2][e] = w;
      o[3][e] = x1 - w;
    }
  }
}

// post-adds (strassen.hpp:215-237 and the M5..M7 / accumulate_into_bsm steps)
__global__ void __launch_bounds__(256) k_postadd(const WNode* __restrict__ nodes, const PosItem* __restrict__ items,
                                                 Ptrs P) {
  const PosItem it = items[blockIdx.x];
  const WNode& nd = nodes[it.node];
  const int64_t hh = (int64_t)nd.h * nd.h, pos = (int64_t)it.r + (int64_t)it.c * nd.h;
  const double* M[7];
#pragma unroll
  for (int k = 0; k < 7; ++k) M[k] = P.scratch + (nd.m_base + k * hh + pos) * P.bs2;
  double* o[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) o[q] = resolve(nd.out[q], it.r, it.c, P);
  const bool acc = nd.root != 0;
  for (int64_t e = threadIdx.x; e < P.bs2; e += blockDim.x) {
    const double m1 = M[0][e], m2 = M[1][e], m3 = M[2][e], m4 = M[3][e];
    const double m5 = M[4][e], m6 = M[5][e], m7 = M[6][e];
    const double t = m4 + m1;
    const double c10 = m2 + t;
    const double v[4] = {m1 + m7, (m3 + t) + m5, c10 + m6, m3 + c10};
#pragma unroll
    for (int q = 0; q < 4; ++q)
      if (o[q]) o[q][e] = acc ? o[q][e] + v[q] : v[q];
  }
}

// Root post-adds accumulate into C blocks (accumulate_into_bsm,
// strassen.hpp:249-256).  Several roots can target the same C block (same i,j
// super-block, different k, or nested levels), so the work is C-block-centric:
// one workgroup per C block applies its contributions in a fixed order, each
// thread owning the same elements throughout -- no races, deterministic sums.
struct RootContrib {
  int32_t node;
  int16_t r, c;
  int32_t q;  // output quadrant 0..3
};

__global__ void __launch_bounds__(256) k_postadd_roots(const WNode* __restrict__ nodes,
                                                       const int64_t* __restrict__ cblk,
                                                       const int64_t* __restrict__ start,
                                                       const RootContrib* __restrict__ contrib, Ptrs P) {
  const int64_t g = blockIdx.x;
  const int64_t o = P.coff[cblk[g]];
  if (o < 0) return;  // += into an empty C block is a no-op in the reference
  double* out = P.carena + o;
  for (int64_t t = start[g]; t < start[g + 1]; ++t) {
    const RootContrib rc = contrib[t];
    const WNode& nd = nodes[rc.node];
    const int64_t hh = (int64_t)nd.h * nd.h, pos = (int64_t)rc.r + (int64_t)rc.c * nd.h;
    const double* M[7];
#pragma unroll
    for (int k = 0; k < 7; ++k) M[k] = P.scratch + (nd.m_base + k * hh + pos) * P.bs2;
    for (int64_t e = threadIdx.x; e < P.bs2; e += blockDim.x) {
      const double m1 = M[0][e], m2 = M[1][e], m3 = M[2][e], m4 = M[3][e];
      const double t4 = m4 + m1;
      const double c10 = m2 + t4;
      double v;
      if (rc.q == 0) v = m1 + M[6][e];
      else if (rc.q == 1) v = (m3 + t4) + M[4][e];
      else if (rc.q == 2) v = c10 + M[5][e];
      else v = m3 + c10;
      out[e] = out[e] + v;
    }
  }
}

// dense leaf products M = X Y (strassen.hpp:304-323) -> grouped GEMM records
__global__ void k_leaf_records(const LeafProd* __restrict__ lp, Ptrs P, int32_t bs, const double* dummy,
                               GemmOut* __restrict__ outs, GemmProd* __restrict__ prods) {
  const LeafProd L = lp[blockIdx.x];
  const int h = L.h;
  for (int64_t q = threadIdx.x; q < (int64_t)h * h; q += blockDim.x) {
    const int r = (int)(q % h), cc = (int)(q / h);
    const int64_t oi = L.out0 + q;
    const int64_t p0 = L.prod0 + q * h;
    GemmOut o;
    o.c = P.scratch + (L.m_blk + q) * P.bs2;
    o.m = bs;
    o.n = bs;
    o.p0 = (int32_t)p0;
    o.p1 = (int32_t)(p0 + h);
    o.acc = 0;
    o.group = 0;
    o.hint = 0;
    o.pad_ = 0;
    outs[oi] = o;
    for (int kb = 0; kb < h; ++kb) {
      const double* a = resolve(L.x, r, kb, P);
      const double* b = resolve(L.y, kb, cc, P);
      GemmProd pr;
      const bool live = a && b;
      pr.a = live ? a : dummy;
      pr.b = live ? b : dummy;
      pr.k = live ? bs : 0;
      pr.pad = 0;
      prods[p0 + kb] = pr;
    }
  }
}

// ------------------------------------------------------------ host planning --
struct Root {
  int level;
  int64_t ib0, jb0, kb0;
  int kind;  // 0 Strassen-Winograd, 1 dense
};

// topmost nodes whose Strassen bit (else dense bit, when dbits is given) is set
static void find_roots(const Tree& t, const std::vector<uint8_t>& bits, const std::vector<uint8_t>& dbits,
                       int32_t nb, std::vector<Root>& roots) {
  if (t.nlev == 0) return;
  struct F {
    const Tree& t;
    const std::vector<uint8_t>& bits;
    const std::vector<uint8_t>& dbits;
    int32_t nb;
    std::vector<Root>& roots;
    void visit(int l, int64_t si, int64_t sj, int64_t sk) {
      const int64_t ns = ceil_div(nb, (int64_t)1 << (l + 1));
      const int64_t node = t.lvl_off[l] + si + sj * ns + sk * ns * ns;
      const bool st = bits[node] != 0, dn = !dbits.empty() && dbits[node] != 0;
      if (st || dn) {
        const int64_t sb = (int64_t)1 << (l + 1);
        roots.push_back(Root{l, si * sb, sj * sb, sk * sb, st ? 0 : 1});
        return;
      }
      if (l == 0) return;
      const int64_t csb = (int64_t)1 << l;
      for (int64_t dj = 0; dj < 2; ++dj) {
        const int64_t cj = 2 * sj + dj;
        if (cj * csb >= nb) continue;
        for (int64_t di = 0; di < 2; ++di) {
          const int64_t ci = 2 * si + di;
          if (ci * csb >= nb) continue;
          for (int64_t dk = 0; dk < 2; ++dk) {
            const int64_t ck = 2 * sk + dk;
            if (ck * csb >= nb) continue;
            visit(l - 1, ci, cj, ck);
          }
        }
      }
    }
  } f{t, bits, dbits, nb, roots};
  const int top = t.nlev - 1;
  const int64_t ntop = ceil_div(nb, (int64_t)1 << t.nlev);
  for (int64_t sk = 0; sk < ntop; ++sk)
    for (int64_t sj = 0; sj < ntop; ++sj)
      for (int64_t si = 0; si < ntop; ++si) f.visit(top, si, sj, sk);
}

static GridRef quadrant(GridRef g, int qr, int qc, int32_t h) {
  g.r0 += qr * h;
  g.c0 += qc * h;
  return g;
}

struct Plan {
  std::vector<std::vector<WNode>> depth;  // nodes per depth
  std::vector<LeafProd> leaves;
  int64_t scratch_blocks = 0;
  int64_t nouts = 0, nprods = 0;
};

// children recurse iff the product grid is even and larger than the cutoff
static bool recurses(int32_t h, int64_t bs, int64_t cutoff) { return h % 2 == 0 && (int64_t)h * bs > cutoff; }

static void plan_node(Plan& p, int d, WNode nd, int64_t bs, int64_t cutoff) {
  const int32_t h = nd.h;
  const int64_t hh = (int64_t)h * h;
  nd.tmp_base = p.scratch_blocks;
  p.scratch_blocks += 8 * hh;
  nd.m_base = p.scratch_blocks;
  p.scratch_blocks += 7 * hh;
  if ((int)p.depth.size() <= d) p.depth.resize(d + 1);
  p.depth[d].push_back(nd);
  auto tmp = [&](int q) { return GridRef{0, h, nd.tmp_base + q * hh, 0, 0}; };
  // (X_k, Y_k) of M1..M7 (strassen.hpp:392-402)
  const GridRef X[7] = {nd.in[0], tmp(0), tmp(1), tmp(2), tmp(3), nd.in[3], nd.in[1]};
  const GridRef Y[7] = {nd.in[4], tmp(4), tmp(5), tmp(6), nd.in[7], tmp(7), nd.in[5]};
  for (int k = 0; k < 7; ++k) {
    const int64_t mb = nd.m_base + k * hh;
    if (recurses(h, bs, cutoff)) {
      const int32_t h2 = h / 2;
      WNode ch;
      ch.h = h2;
      ch.root = 0;
      // A operand quadrants a, b, c, d; B operand A(TL), B(BL), C(TR), D(BR)
      ch.in[0] = quadrant(X[k], 0, 0, h2);
      ch.in[1] = quadrant(X[k], 0, 1, h2);
      ch.in[2] = quadrant(X[k], 1, 0, h2);
      ch.in[3] = quadrant(X[k], 1, 1, h2);
      ch.in[4] = quadrant(Y[k], 0, 0, h2);
      ch.in[5] = quadrant(Y[k], 1, 0, h2);
      ch.in[6] = quadrant(Y[k], 0, 1, h2);
      ch.in[7] = quadrant(Y[k], 1, 1, h2);
      const GridRef mg{0, h, mb, 0, 0};
      ch.out[0] = quadrant(mg, 0, 0, h2);
      ch.out[1] = quadrant(mg, 0, 1, h2);
      ch.out[2] = quadrant(mg, 1, 0, h2);
      ch.out[3] = quadrant(mg, 1, 1, h2);
      plan_node(p, d + 1, ch, bs, cutoff);
    } else {
      LeafProd L;
      L.x = X[k];
      L.y = Y[k];
      L.m_blk = mb;
      L.out0 = p.nouts;
      L.prod0 = p.nprods;
      L.h = h;
      L.pad = 0;
      p.nouts += hh;
      p.nprods += hh * h;
      p.leaves.push_back(L);
    }
  }
}

static WNode root_node(const Root& r) {
  const int32_t h = 1 << r.level;
  WNode nd;
  nd.h = h;
  nd.root = 1;
  nd.in[0] = GridRef{1, 0, 0, (int32_t)r.ib0, (int32_t)r.kb0};
  nd.in[1] = GridRef{1, 0, 0, (int32_t)r.ib0, (int32_t)(r.kb0 + h)};
  nd.in[2] = GridRef{1, 0, 0, (int32_t)(r.ib0 + h), (int32_t)r.kb0};
  nd.in[3] = GridRef{1, 0, 0, (int32_t)(r.ib0 + h), (int32_t)(r.kb0 + h)};
  nd.in[4] = GridRef{2, 0, 0, (int32_t)r.kb0, (int32_t)r.jb0};             // A (TL)
  nd.in[5] = GridRef{2, 0, 0, (int32_t)(r.kb0 + h), (int32_t)r.jb0};       // B (BL)
  nd.in[6] = GridRef{2, 0, 0, (int32_t)r.kb0, (int32_t)(r.jb0 + h)};       // C (TR)
  nd.in[7] = GridRef{2, 0, 0, (int32_t)(r.kb0 + h), (int32_t)(r.jb0 + h)}; // D (BR)
  nd.out[0] = GridRef{3, 0, 0, (int32_t)r.ib0, (int32_t)r.jb0};
  nd.out[1] = GridRef{3, 0, 0, (int32_t)r.ib0, (int32_t)(r.jb0 + h)};
  nd.out[2] = GridRef{3, 0, 0, (int32_t)(r.ib0 + h), (int32_t)r.jb0};
  nd.out[3] = GridRef{3, 0, 0, (int32_t)(r.ib0 + h), (int32_t)(r.jb0 + h)};
  return nd;
}

template <class T>
static T* upload(sa_context* c, const char* name, const std::vector<T>& v) {
  DevBuf& b = c->buf(name);
  b.reserve(sizeof(T) * (v.size() + 1));
  if (!v.empty()) SA_CHECK_HIP(hipMemcpyAsync(b.p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice, c->stream));
  return b.as<T>();
}

static void run_plan(sa_bsm* C, sa_bsm* A, sa_bsm* B, Plan& p, int tile_cfg) {
  sa_context* c = C->ctx;
  const int64_t bs = A->bs_r, bs2 = bs * bs;
  DevBuf& scr = c->buf("st_scratch");
  scr.reserve(sizeof(double) * (size_t)(p.scratch_blocks * bs2 + 1));
  DevBuf& dum = c->buf("st_dummy");
  dum.reserve(64);
  Ptrs P{scr.as<double>(), bs2, A->d_off, A->data(), B->d_off, B->data(), C->d_off, C->data(), A->nrb};
  // flatten nodes, position items per depth
  std::vector<WNode> all;
  std::vector<int64_t> depth_first(p.depth.size() + 1, 0);
  for (size_t d = 0; d < p.depth.size(); ++d) {
    depth_first[d] = (int64_t)all.size();
    all.insert(all.end(), p.depth[d].begin(), p.depth[d].end());
  }
  depth_first[p.depth.size()] = (int64_t)all.size();
  std::vector<std::vector<PosItem>> pre(p.depth.size()), post(p.depth.size());
  for (size_t d = 0; d < p.depth.size(); ++d)
    for (int64_t n = depth_first[d]; n < depth_first[d + 1]; ++n) {
      const int32_t h = all[n].h;
      for (int32_t cc = 0; cc < h; ++cc)
        for (int32_t r = 0; r < h; ++r) {
          post[d].push_back(PosItem{(int32_t)n, (int16_t)r, (int16_t)cc, 0});
          pre[d].push_back(PosItem{(int32_t)n, (int16_t)r, (int16_t)cc, 0});
          pre[d].push_back(PosItem{(int32_t)n, (int16_t)r, (int16_t)cc, 1});
        }
    }
  const WNode* d_nodes = upload(c, "st_nodes", all);
  std::vector<PosItem*> d_pre(p.depth.size()), d_post(p.depth.size());
  for (size_t d = 0; d < p.depth.size(); ++d) {
    std::string a = "st_pre" + std::to_string(d), b = "st_post" + std::to_string(d);
    d_pre[d] = upload(c, a.c_str(), pre[d]);
    d_post[d] = upload(c, b.c_str(), post[d]);
  }
  const LeafProd* d_leaves = upload(c, "st_leaves", p.leaves);
  DevBuf& bouts = c->buf("st_outs");
  DevBuf& bprods = c->buf("st_prods");
  bouts.reserve(sizeof(GemmOut) * (size_t)(p.nouts + 1));
  bprods.reserve(sizeof(GemmProd) * (size_t)(p.nprods + 1));
  // leaf records (they only point into A, B, C and the scratch), then the
  // leaf GEMM's tile config and work items, which may synchronise
  if (!p.leaves.empty())
    hipLaunchKernelGGL(k_leaf_records, dim3((unsigned)p.leaves.size()), dim3(256), 0, c->stream, d_leaves, P,
                       (int32_t)bs, (const double*)dum.p, bouts.as<GemmOut>(), bprods.as<GemmProd>());
  GemmBatch gb;
  gb.outs = bouts.as<GemmOut>();
  gb.nouts = p.nouts;
  gb.prods = bprods.as<GemmProd>();
  gb.nprods = p.nprods;
  gb.alpha = 1.0;
  gb.max_m = gb.max_n = gb.max_k = (int32_t)bs;
  const GemmPlan gp = gemm_fp64_plan(c, gb, tile_cfg);
  // contributions of the roots (depth 0) grouped by target C block, in root order
  std::vector<std::pair<int64_t, RootContrib>> rc;
  for (int64_t n = depth_first[0]; n < depth_first[1]; ++n) {
    const WNode& nd = all[n];
    for (int q = 0; q < 4; ++q)
      for (int32_t cc = 0; cc < nd.h; ++cc)
        for (int32_t r = 0; r < nd.h; ++r) {
          const int64_t cb = (int64_t)(nd.out[q].r0 + r) + (int64_t)(nd.out[q].c0 + cc) * A->nrb;
          rc.push_back({cb, RootContrib{(int32_t)n, (int16_t)r, (int16_t)cc, q}});
        }
  }
  std::stable_sort(rc.begin(), rc.end(), [](const auto& x, const auto& y) { return x.first < y.first; });
  std::vector<int64_t> cblk, start;
  std::vector<RootContrib> contrib;
  contrib.reserve(rc.size());
  for (size_t i = 0; i < rc.size(); ++i) {
    if (i == 0 || rc[i].first != rc[i - 1].first) {
      cblk.push_back(rc[i].first);
      start.push_back((int64_t)i);
    }
    contrib.push_back(rc[i].second);
  }
  start.push_back((int64_t)rc.size());
  const int64_t* d_cblk = upload(c, "st_cblk", cblk);
  const int64_t* d_start = upload(c, "st_cstart", start);
  const RootContrib* d_contrib = upload(c, "st_contrib", contrib);
  // every level of the recursion as one HIP graph: the pre-adds per depth, the
  // grouped GEMM of all leaves, the post-adds per depth and the roots' sums
  {
    ProfScope ps(c, "winograd");
    run_as_graph(c, [&] {
      for (size_t d = 0; d < p.depth.size(); ++d)
        hipLaunchKernelGGL(k_preadd, dim3((unsigned)pre[d].size()), dim3(256), 0, c->stream, d_nodes,
                           (const PosItem*)d_pre[d], P);
      gemm_fp64_launch(c, gb, gp);
      for (size_t d = p.depth.size(); d-- > 1;)
        hipLaunchKernelGGL(k_postadd, dim3((unsigned)post[d].size()), dim3(256), 0, c->stream, d_nodes,
                           (const PosItem*)d_post[d], P);
      if (!cblk.empty())
        hipLaunchKernelGGL(k_postadd_roots, dim3((unsigned)cblk.size()), dim3(256), 0, c->stream, d_nodes, d_cblk,
                           d_start, d_contrib, P);
      SA_CHECK_HIP(hipGetLastError());
    });
  }
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));  // host item vectors go out of scope
}

// ------------------------------------------------------------- dense roots --
// A dense root at level l covers w = 2^(l+1) blocks per side: its A region
// (rows ib0.., columns kb0..) and B region (rows kb0.., columns jb0..) are
// packed column-major into W x W arrays (W = w bs, absent blocks as zeros),
// multiplied by the library's grouped FP64 MFMA GEMM (one output record and one
// product per root, tile autotuned for the W x W x W shape) per batch of
// equal-level roots, and the W x W products are added into the present C
// blocks of the region.
// Roots sharing a C region are added in ascending kb0 order, one launch per
// rank, so every C block sums its contributions in a fixed order.
__global__ void k_pack_regions(const int64_t* __restrict__ off, const double* __restrict__ arena, int32_t nrb,
                               const int32_t* __restrict__ r0, const int32_t* __restrict__ c0, int32_t w, int32_t bs,
                               double* __restrict__ dst) {
  const int64_t g = blockIdx.x;
  const int64_t ww = (int64_t)w * w;
  const int64_t root = g / ww;
  const int rem = (int)(g % ww);
  const int bi = rem % w, bj = rem / w;
  const int64_t W = (int64_t)w * bs;
  double* d = dst + root * W * W + (int64_t)bj * bs * W + (int64_t)bi * bs;
  const int64_t o = off[(int64_t)(r0[root] + bi) + (int64_t)(c0[root] + bj) * nrb];
  for (int e = threadIdx.x; e < bs * bs; e += blockDim.x) {
    const int r = e % bs, cc = e / bs;
    d[(int64_t)cc * W + r] = o >= 0 ? arena[o + e] : 0.0;
  }
}

__global__ void k_scatter_regions(const int64_t* __restrict__ coff, double* __restrict__ carena, int32_t nrb,
                                  const int32_t* __restrict__ r0, const int32_t* __restrict__ c0,
                                  const int32_t* __restrict__ sel, int32_t w, int32_t bs,
                                  const double* __restrict__ src) {
  const int64_t g = blockIdx.x;
  const int64_t ww = (int64_t)w * w;
  const int32_t root = sel[g / ww];
  const int rem = (int)(g % ww);
  const int bi = rem % w, bj = rem / w;
  const int64_t o = coff[(int64_t)(r0[root] + bi) + (int64_t)(c0[root] + bj) * nrb];
  if (o < 0) return;  // += into an empty block is dropped, as for Strassen roots
  const int64_t W = (int64_t)w * bs;
  const double* sp = src + (int64_t)root * W * W + (int64_t)bj * bs * W + (int64_t)bi * bs;
  for (int e = threadIdx.x; e < bs * bs; e += blockDim.x) {
    const int r = e % bs, cc = e / bs;
    carena[o + e] += sp[(int64_t)cc * W + r];
  }
}

// Scratch budget of one batch: free HBM / frac, capped at SA_SCRATCH_MB
// megabytes when that environment variable is set (smaller batches for
// callers that share the device, and for tests of the batching itself).
static int64_t scratch_budget(int frac) {
  size_t free_b = 0, total_b = 0;
  SA_CHECK_HIP(hipMemGetInfo(&free_b, &total_b));
  int64_t b = (int64_t)(free_b / (size_t)frac);
  if (const char* v = std::getenv("SA_SCRATCH_MB")) {
    const int64_t cap = std::atoll(v) << 20;
    if (cap > 0) b = std::min(b, cap);
  }
  return b;
}

// dense roots of one level; returns the number executed
static int64_t run_dense_level(sa_bsm* C, sa_bsm* A, sa_bsm* B, std::vector<Root> rs) {
  if (rs.empty()) return 0;
  sa_context* c = C->ctx;
  const int32_t bs = A->bs_r;
  const int32_t w = 1 << (rs[0].level + 1);
  const int64_t W = (int64_t)w * bs, W2 = W * W;
  std::stable_sort(rs.begin(), rs.end(), [](const Root& x, const Root& y) {
    if (x.jb0 != y.jb0) return x.jb0 < y.jb0;
    if (x.ib0 != y.ib0) return x.ib0 < y.ib0;
    return x.kb0 < y.kb0;
  });
  const int64_t per = 3 * W2 * (int64_t)sizeof(double);
  const int64_t chunk = std::max<int64_t>(1, std::min<int64_t>(scratch_budget(4) / per, 4096));
  DevBuf& bp = c->buf("dn_pack");
  bp.reserve((size_t)(std::min<int64_t>(chunk, (int64_t)rs.size()) * per) + 64);
  ProfScope ps(c, "dense_roots");
  for (size_t i0 = 0; i0 < rs.size(); i0 += (size_t)chunk) {
    const size_t n = std::min(rs.size() - i0, (size_t)chunk);
    std::vector<int32_t> ai(n), ak(n), bk(n), bj(n);
    std::vector<int32_t> ci(n), cj(n);
    for (size_t q = 0; q < n; ++q) {
      const Root& r = rs[i0 + q];
      ai[q] = (int32_t)r.ib0; ak[q] = (int32_t)r.kb0;
      bk[q] = (int32_t)r.kb0; bj[q] = (int32_t)r.jb0;
      ci[q] = (int32_t)r.ib0; cj[q] = (int32_t)r.jb0;
    }
    // ranks among roots of one C region (sorted by kb0 within it)
    std::vector<std::vector<int32_t>> ranks;
    for (size_t q = 0; q < n; ++q) {
      size_t rk = 0;
      if (q > 0) {
        size_t p = q;
        while (p > 0 && ci[p - 1] == ci[q] && cj[p - 1] == cj[q]) --p;
        rk = q - p;
      }
      if (ranks.size() <= rk) ranks.resize(rk + 1);
      ranks[rk].push_back((int32_t)q);
    }
    double* PA = bp.as<double>();
    double* PB = PA + (int64_t)n * W2;
    double* PC = PB + (int64_t)n * W2;
    const int32_t* dai = upload(c, "dn_ai", ai);
    const int32_t* dak = upload(c, "dn_ak", ak);
    const int32_t* dbk = upload(c, "dn_bk", bk);
    const int32_t* dbj = upload(c, "dn_bj", bj);
    const unsigned nblk = (unsigned)(n * (size_t)w * (size_t)w);
    hipLaunchKernelGGL(k_pack_regions, dim3(nblk), dim3(256), 0, c->stream, (const int64_t*)A->d_off,
                       (const double*)A->data(), A->nrb, dai, dak, w, bs, PA);
    hipLaunchKernelGGL(k_pack_regions, dim3(nblk), dim3(256), 0, c->stream, (const int64_t*)B->d_off,
                       (const double*)B->data(), B->nrb, dbk, dbj, w, bs, PB);
    std::vector<GemmOut> ho(n);
    std::vector<GemmProd> hp(n);
    for (size_t q = 0; q < n; ++q) {
      ho[q] = GemmOut{PC + (int64_t)q * W2, (int32_t)W, (int32_t)W, (int32_t)q, (int32_t)q + 1, 0, 0, 0, 0};
      hp[q] = GemmProd{PA + (int64_t)q * W2, PB + (int64_t)q * W2, (int32_t)W, 0};
    }
    GemmBatch gb;
    gb.outs = upload(c, "dn_outs", ho);
    gb.nouts = (int64_t)n;
    gb.prods = upload(c, "dn_prods", hp);
    gb.nprods = (int64_t)n;
    gb.max_m = gb.max_n = gb.max_k = (int32_t)W;
    gb.flops = 2.0 * (double)W * (double)W * (double)W * (double)n;
    gemm_fp64_run(c, gb, -1);
    for (size_t rk = 0; rk < ranks.size(); ++rk) {
      std::string nm = "dn_sel" + std::to_string(rk);
      const int32_t* dsel = upload(c, nm.c_str(), ranks[rk]);
      hipLaunchKernelGGL(k_scatter_regions, dim3((unsigned)(ranks[rk].size() * (size_t)w * (size_t)w)), dim3(256), 0,
                         c->stream, (const int64_t*)C->d_off, C->data(), C->nrb, dai, dbj, dsel, w, bs,
                         (const double*)PC);
    }
    SA_CHECK_HIP(hipGetLastError());
    SA_CHECK_HIP(hipStreamSynchronize(c->stream));  // host index vectors go out of scope
  }
  return (int64_t)rs.size();
}

// ------------------------------------------------------- learned planning --
// Strassen node cost of the reference model at level l (strassen.hpp:590-600)
static uint64_t strassen_model_cost(int l, uint64_t bs) {
  const uint64_t cost_mult = (uint64_t)std::round(std::pow(7.0, (double)(l + 1))) * (2u * bs * bs * bs);
  const uint64_t cost_add = (uint64_t)(15e0 * kStrassenFac[l] * std::pow(4.0, (double)(l + 1)) * (double)(bs * bs));
  return cost_mult + cost_add;
}

// summed-area table of block presence: sat[r + c (nb+1)] = #present blocks with row < r, col < c
__device__ inline int32_t sat_count(const int32_t* sat, int32_t nb, int64_t r0, int64_t r1, int64_t c0, int64_t c1) {
  const int64_t ld = nb + 1;
  return sat[r1 + c1 * ld] - sat[r0 + c1 * ld] - sat[r1 + c0 * ld] + sat[r0 + c0 * ld];
}

// screened flops of every level-0 node (the dry run's count before any decision)
__global__ void k_raw_level0(DryArgs a, int64_t ns, double* __restrict__ raw) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  double f = 0.0;
  for (int64_t di = 0; di < 2; ++di)
    for (int64_t dj = 0; dj < 2; ++dj)
      for (int64_t dk = 0; dk < 2; ++dk) {
        const int64_t ib = 2 * si + di, jb = 2 * sj + dj, kb = 2 * sk + dk;
        if (ib >= a.nb || jb >= a.nb || kb >= a.nb) continue;
        const int64_t ai = ib + kb * a.nb, bi = kb + jb * a.nb;
        if (a.aoff[ai] < 0 || a.boff[bi] < 0) continue;
        if (a.anorm[ai] * a.bnorm[bi] >= a.thr) f += 2.0 * a.rsz[ib] * a.rsz[jb] * (double)a.rsz[kb];
      }
  raw[s] = f;
}

__global__ void k_raw_level(int32_t nb, int level, int64_t ns, int64_t pns, const double* __restrict__ prev,
                            double* __restrict__ cur) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  const int64_t psb = (int64_t)1 << level, sb = psb * 2;
  const int64_t ib0 = si * sb, jb0 = sj * sb, kb0 = sk * sb;
  const int64_t isub = ceil_div(std::min<int64_t>(ib0 + sb, nb) - ib0, psb);
  const int64_t jsub = ceil_div(std::min<int64_t>(jb0 + sb, nb) - jb0, psb);
  const int64_t ksub = ceil_div(std::min<int64_t>(kb0 + sb, nb) - kb0, psb);
  double f = 0.0;
  for (int64_t x = 0; x < isub; ++x)
    for (int64_t y = 0; y < jsub; ++y)
      for (int64_t z = 0; z < ksub; ++z) f += prev[(2 * si + x) + (2 * sj + y) * pns + (2 * sk + z) * pns * pns];
  cur[s] = f;
}

struct FeatArgs {
  const int32_t *satA, *satB, *satC;
  int32_t nb;
  int64_t n;
  int32_t bs;
};

__device__ inline bool node_edge(int32_t nb, int64_t n, int32_t bs, int64_t sb, int64_t ib0, int64_t jb0,
                                 int64_t kb0) {
  const int64_t ib1 = std::min<int64_t>(ib0 + sb, nb), jb1 = std::min<int64_t>(jb0 + sb, nb),
                kb1 = std::min<int64_t>(kb0 + sb, nb);
  const int64_t ie = std::min<int64_t>(ib1 * bs, n), je = std::min<int64_t>(jb1 * bs, n),
                ke = std::min<int64_t>(kb1 * bs, n);
  return (ie - ib0 * bs != bs * sb) || (je - jb0 * bs != bs * sb) || (ke - kb0 * bs != bs * sb);
}

// The selector's node features (all scaled to about [0, 1]):
//  0 level/8, 1 log2(bs)/8, 2 log2(node side)/16, 3-4 block fill of the node's
//  A and B regions, 5 screened flops / dense flops of the region, 6 block fill
//  of its C region, 7 the cost-model (heuristic) decision, 8 edge node,
//  9 log2(1 + screened flops)/64
__global__ void k_node_features(FeatArgs fa, int level, int64_t ns, const double* __restrict__ raw,
                                const uint8_t* __restrict__ hbits, double* __restrict__ X) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  const int64_t w = (int64_t)1 << (level + 1);
  const int64_t ib0 = si * w, jb0 = sj * w, kb0 = sk * w;
  const int64_t ib1 = std::min<int64_t>(ib0 + w, fa.nb), jb1 = std::min<int64_t>(jb0 + w, fa.nb),
                kb1 = std::min<int64_t>(kb0 + w, fa.nb);
  const double ww = (double)(w * w);
  const double W = (double)w * fa.bs;
  double* x = X + s * kMlpFeatures;
  x[0] = level / 8.0;
  x[1] = log2((double)fa.bs) / 8.0;
  x[2] = log2(W) / 16.0;
  x[3] = sat_count(fa.satA, fa.nb, ib0, ib1, kb0, kb1) / ww;
  x[4] = sat_count(fa.satB, fa.nb, kb0, kb1, jb0, jb1) / ww;
  x[5] = raw[s] / (2.0 * W * W * W);
  x[6] = sat_count(fa.satC, fa.nb, ib0, ib1, jb0, jb1) / ww;
  x[7] = hbits[s] ? 1.0 : 0.0;
  x[8] = node_edge(fa.nb, fa.n, fa.bs, w, ib0, jb0, kb0) ? 1.0 : 0.0;
  x[9] = log2(1.0 + raw[s]) / 64.0;
}

// Bottom-up plan: best(node) = min(descend, Strassen, dense) in units of
// screened-path flops; descend = screened flops at level 0, else the sum of the
// children's best; Strassen / dense = screened flops x 2^(predicted log2
// ratio), interior nodes only.  Ties keep the descend choice.  The heuristic
// judges every learned choice: a path is admissible only where the reference
// cost model prices it (Strassen: its node cost; dense: the region's dense
// flops) at no more than kJudge x the node's screened flops.
constexpr double kJudge = 1.5;
__global__ void k_plan_level(int32_t nb, int64_t n, int32_t bs, int level, int64_t ns, int64_t pns,
                             double cost_s, const double* __restrict__ raw, const double* __restrict__ Y,
                             const double* __restrict__ prev, double* __restrict__ best, uint8_t* __restrict__ sbit,
                             uint8_t* __restrict__ dbit) {
  const int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= ns * ns * ns) return;
  const int64_t si = s % ns, sj = (s / ns) % ns, sk = s / (ns * ns);
  const int64_t w = (int64_t)1 << (level + 1), psb = w / 2;
  double desc = 0.0;
  if (level == 0) {
    desc = raw[s];
  } else {
    const int64_t ib0 = si * w, jb0 = sj * w, kb0 = sk * w;
    const int64_t isub = ceil_div(std::min<int64_t>(ib0 + w, nb) - ib0, psb);
    const int64_t jsub = ceil_div(std::min<int64_t>(jb0 + w, nb) - jb0, psb);
    const int64_t ksub = ceil_div(std::min<int64_t>(kb0 + w, nb) - kb0, psb);
    for (int64_t x = 0; x < isub; ++x)
      for (int64_t y = 0; y < jsub; ++y)
        for (int64_t z = 0; z < ksub; ++z) desc += prev[(2 * si + x) + (2 * sj + y) * pns + (2 * sk + z) * pns * pns];
  }
  double b = desc;
  uint8_t sv = 0, dv = 0;
  if (!node_edge(nb, n, bs, w, si * w, sj * w, sk * w)) {
    const double W = (double)w * bs;
    const double cs = raw[s] * exp2(Y[s * kMlpOutputs + 0]);
    const double cd = raw[s] * exp2(Y[s * kMlpOutputs + 1]);
    if (cs < b && cost_s <= kJudge * raw[s]) { b = cs; sv = 1; dv = 0; }
    if (cd < b && 2.0 * W * W * W <= kJudge * raw[s]) { b = cd; sv = 0; dv = 1; }
  }
  best[s] = b;
  sbit[s] = sv;
  dbit[s] = dv;
}

struct MlpPlan {
  bool used = false, overruled = false;
  double model_cost = 0.0, heur_cost = 0.0;  // reference flop model of both plans
};

static std::vector<int32_t> presence_sat(sa_context* c, const sa_bsm* X) {
  const int32_t nb = X->nrb;
  std::vector<int64_t> off((size_t)X->nb());
  if (!off.empty())
    SA_CHECK_HIP(hipMemcpyAsync(off.data(), X->d_off, sizeof(int64_t) * off.size(), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  const int64_t ld = nb + 1;
  std::vector<int32_t> sat((size_t)(ld * ld), 0);
  for (int64_t cc = 1; cc <= nb; ++cc)
    for (int64_t r = 1; r <= nb; ++r)
      sat[r + cc * ld] = (off[(r - 1) + (cc - 1) * nb] >= 0 ? 1 : 0) + sat[(r - 1) + cc * ld] + sat[r + (cc - 1) * ld] -
                         sat[(r - 1) + (cc - 1) * ld];
  return sat;
}

// Features of every dry-run node (rows in dry-run order) into buffer "pl_x";
// screened flops per node into "pl_raw".  Needs the dry run's bits in "st_bits".
static void node_features(sa_bsm* C, sa_bsm* A, sa_bsm* B, double thr, const Tree& t) {
  sa_context* c = A->ctx;
  const int32_t nb = A->nrb;
  std::vector<int32_t> sa_ = presence_sat(c, A), sb_ = presence_sat(c, B);
  std::vector<int32_t> sc_;
  if (C) {
    sc_ = presence_sat(c, C);
  } else {  // no C: every block counts as present
    const int64_t ld = nb + 1;
    sc_.assign((size_t)(ld * ld), 0);
    for (int64_t cc = 0; cc <= nb; ++cc)
      for (int64_t r = 0; r <= nb; ++r) sc_[r + cc * ld] = (int32_t)(r * cc);
  }
  const int32_t* dsa = upload(c, "pl_sat_a", sa_);
  const int32_t* dsb = upload(c, "pl_sat_b", sb_);
  const int32_t* dsc = upload(c, "pl_sat_c", sc_);
  DevBuf& braw = c->buf("pl_raw");
  DevBuf& bx = c->buf("pl_x");
  braw.reserve(sizeof(double) * (size_t)(t.total + 1));
  bx.reserve(sizeof(double) * (size_t)(t.total * kMlpFeatures + 1));
  double* raw = braw.as<double>();
  DryArgs a{A->d_off, A->d_norm, B->d_off, B->d_norm, A->d_rsz, A->nrb, A->nrow, A->bs_r, thr};
  FeatArgs fa{dsa, dsb, dsc, nb, A->nrow, A->bs_r};
  const uint8_t* hb = c->buf("st_bits").as<uint8_t>();
  for (int l = 0; l < t.nlev; ++l) {
    const int64_t ns = ceil_div(nb, (int64_t)1 << (l + 1));
    const int64_t nn = ns * ns * ns;
    if (l == 0) {
      hipLaunchKernelGGL(k_raw_level0, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, a, ns,
                         raw + t.lvl_off[0]);
    } else {
      const int64_t pns = ceil_div(nb, (int64_t)1 << l);
      hipLaunchKernelGGL(k_raw_level, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, nb, l, ns, pns,
                         (const double*)(raw + t.lvl_off[l - 1]), raw + t.lvl_off[l]);
    }
    hipLaunchKernelGGL(k_node_features, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, fa, l, ns,
                       (const double*)(raw + t.lvl_off[l]), hb + t.lvl_off[l],
                       bx.as<double>() + t.lvl_off[l] * kMlpFeatures);
  }
  SA_CHECK_HIP(hipGetLastError());
}

// The learned plan (each choice admitted by the reference cost model, see
// k_plan_level).  As a last guard the heuristic plan stands if the learned one
// costs more than 4x the heuristic plan's model flops.
static MlpPlan plan_mlp(sa_bsm* C, sa_bsm* A, sa_bsm* B, double thr, const Tree& t, std::vector<uint8_t>& sbits,
                        std::vector<uint8_t>& dbits) {
  sa_context* c = A->ctx;
  MlpPlan info;
  ProfScope ps(c, "mlp_plan");
  node_features(C, A, B, thr, t);
  DevBuf& by = c->buf("pl_y");
  DevBuf& bbest = c->buf("pl_best");
  DevBuf& bsb = c->buf("pl_sbits");
  DevBuf& bdb = c->buf("pl_dbits");
  by.reserve(sizeof(double) * (size_t)(t.total * kMlpOutputs + 1));
  bbest.reserve(sizeof(double) * (size_t)(t.total + 1));
  bsb.reserve((size_t)t.total + 16);
  bdb.reserve((size_t)t.total + 16);
  mlp_forward_device(c, c->buf("pl_x").as<double>(), t.total, by.as<double>());
  const double* raw = c->buf("pl_raw").as<double>();
  const int32_t nb = A->nrb;
  for (int l = 0; l < t.nlev; ++l) {
    const int64_t ns = ceil_div(nb, (int64_t)1 << (l + 1));
    const int64_t nn = ns * ns * ns;
    const int64_t pns = l > 0 ? ceil_div(nb, (int64_t)1 << l) : 0;
    hipLaunchKernelGGL(k_plan_level, dim3((unsigned)ceil_div(nn, 256)), dim3(256), 0, c->stream, nb, A->nrow, A->bs_r,
                       l, ns, pns, (double)strassen_model_cost(l, (uint64_t)A->bs_r), raw + t.lvl_off[l], (const double*)(by.as<double>() + t.lvl_off[l] * kMlpOutputs),
                       l > 0 ? (const double*)(bbest.as<double>() + t.lvl_off[l - 1]) : (const double*)nullptr,
                       bbest.as<double>() + t.lvl_off[l], bsb.as<uint8_t>() + t.lvl_off[l],
                       bdb.as<uint8_t>() + t.lvl_off[l]);
  }
  SA_CHECK_HIP(hipGetLastError());
  sbits.resize((size_t)t.total);
  dbits.resize((size_t)t.total);
  std::vector<double> hraw((size_t)t.total);
  SA_CHECK_HIP(hipMemcpyAsync(sbits.data(), bsb.p, (size_t)t.total, hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipMemcpyAsync(dbits.data(), bdb.p, (size_t)t.total, hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipMemcpyAsync(hraw.data(), raw, sizeof(double) * (size_t)t.total, hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  // judge: the reference cost model (strassen.hpp:579-673) on both plans
  std::vector<Root> roots;
  find_roots(t, sbits, dbits, nb, roots);
  const uint64_t bs = (uint64_t)A->bs_r;
  double covered = 0.0, model = 0.0;
  for (const Root& r : roots) {
    const int64_t w = (int64_t)1 << (r.level + 1);
    const int64_t ns = ceil_div(nb, w);
    const int64_t node = t.lvl_off[r.level] + r.ib0 / w + (r.jb0 / w) * ns + (r.kb0 / w) * ns * ns;
    covered += hraw[(size_t)node];
    const double W = (double)w * (double)bs;
    model += r.kind == 0 ? (double)strassen_model_cost(r.level, bs) : 2.0 * W * W * W;
  }
  model += (double)t.lvl_flops[0] - covered;
  info.used = true;
  info.model_cost = model;
  info.heur_cost = (double)t.lvl_flops[t.nlev];
  info.overruled = model > 4.0 * info.heur_cost;
  return info;
}

struct StrassenStats {
  int64_t nroots = 0, nnodes = 0, nleaf_products = 0, scratch_bytes = 0, batches = 0, screened = 0, dense = 0;
  int32_t scheduler = SA_SCHED_HEURISTIC;  // the schedule that ran (MLP only when used and not overruled)
  MlpPlan mlp;
};

static StrassenStats g_last_stats;

static void hybrid(sa_bsm* C, sa_bsm* A, sa_bsm* B, double thresh, const sa_opts_t& o) {
  sa_context* c = C->ctx;
  const int64_t bs = A->bs_r;
  const double thr = thresh * (double)(bs * bs * bs);
  Tree t = tree_shape(A->nrb);
  std::vector<uint8_t> bits, dbits;
  MlpPlan mlp_info;
  int32_t sched = o.scheduler == SA_SCHED_TREE ? SA_SCHED_TREE : SA_SCHED_HEURISTIC;
  if (o.scheduler == SA_SCHED_TREE) {
    SA_REQUIRE(o.tree || o.tree_dense || t.total == 0, "strassen: scheduler TREE needs opts.tree or opts.tree_dense");
    if (o.tree)
      bits.assign(o.tree, o.tree + t.total);
    else
      bits.assign((size_t)t.total, 0);
    if (o.tree_dense) dbits.assign(o.tree_dense, o.tree_dense + t.total);
  } else if (o.scheduler == SA_SCHED_HEURISTIC || o.scheduler == SA_SCHED_MLP) {
    dryrun_device(A, B, thr, t);
    download_tree(c, t, bits);
    // the learned selector, when a model is loaded; the heuristic otherwise,
    // and whenever the reference cost model overrules the learned plan
    if (o.scheduler == SA_SCHED_MLP && mlp_loaded(c) && t.nlev > 0) {
      std::vector<uint8_t> sb, db;
      mlp_info = plan_mlp(C, A, B, thr, t, sb, db);
      if (!mlp_info.overruled) {
        bits.swap(sb);
        dbits.swap(db);
        sched = SA_SCHED_MLP;
      }
    }
  } else {
    SA_UNSUPPORTED("strassen: scheduler %d not available", o.scheduler);
  }
  std::vector<Root> roots;
  find_roots(t, bits, dbits, A->nrb, roots);
  // screened leaves skip every triple under a Strassen or dense root
  {
    std::vector<uint8_t> cover(bits);
    for (size_t q = 0; q < dbits.size(); ++q) cover[q] |= dbits[q];
    upload(c, "st_bits", cover);
  }
  // an executed root must cover full blocks only (never an edge super-block)
  for (const Root& r : roots) {
    const int64_t sb = (int64_t)1 << (r.level + 1);
    SA_REQUIRE(r.ib0 + sb <= A->nrb && r.jb0 + sb <= A->nrb && r.kb0 + sb <= A->nrb &&
                   (r.ib0 + sb) * bs <= A->nrow && (r.jb0 + sb) * bs <= A->nrow && (r.kb0 + sb) * bs <= A->nrow,
               "strassen: decision tree selects edge super-block (level %d at %lld,%lld,%lld)", r.level,
               (long long)r.ib0, (long long)r.jb0, (long long)r.kb0);
  }
  std::vector<std::vector<Root>> dense_by_level(t.nlev > 0 ? t.nlev : 1);
  {
    std::vector<Root> sroots;
    for (const Root& r : roots) (r.kind == 0 ? sroots : dense_by_level[r.level]).push_back(r);
    roots.swap(sroots);
  }
  StrassenStats st;
  st.nroots = (int64_t)roots.size();
  st.scheduler = sched;
  st.mlp = mlp_info;
  // screened leaves outside every executed root; C blocks are never created
  if (t.nlev > 0) {
    std::vector<int64_t> lo(t.lvl_off.begin(), t.lvl_off.end());
    const int64_t* d_lo = upload(c, "st_lvl_off", lo);
    st.screened = screened_multiply(C, A, false, B, false, thresh, 1.0, o.tile, c->buf("st_bits").as<uint8_t>(), d_lo,
                                    t.nlev, true);
  } else {
    st.screened = screened_multiply(C, A, false, B, false, thresh, 1.0, o.tile, nullptr, nullptr, 0, true);
  }
  for (auto& lv : dense_by_level) st.dense += run_dense_level(C, A, B, lv);
  if (roots.empty()) {
    g_last_stats = st;
    return;
  }
  // batches of roots whose scratch fits the budget
  const int64_t budget = scratch_budget(2);
  const int64_t bs2b = bs * bs * (int64_t)sizeof(double);
  size_t i = 0;
  while (i < roots.size()) {
    Plan p;
    size_t j = i;
    while (j < roots.size()) {
      // tentative append, rolled back if the batch would exceed the budget
      std::vector<size_t> dsz;
      for (auto& d : p.depth) dsz.push_back(d.size());
      const size_t nleaves = p.leaves.size();
      const int64_t sb0 = p.scratch_blocks, no0 = p.nouts, np0 = p.nprods;
      plan_node(p, 0, root_node(roots[j]), bs, o.cutoff);
      const int64_t bytes = p.scratch_blocks * bs2b + p.nprods * (int64_t)sizeof(GemmProd) +
                            p.nouts * (int64_t)sizeof(GemmOut);
      if (j > i && (bytes > budget || p.nprods >= ((int64_t)1 << 31))) {
        p.depth.resize(dsz.size());
        for (size_t d = 0; d < dsz.size(); ++d) p.depth[d].resize(dsz[d]);
        p.leaves.resize(nleaves);
        p.scratch_blocks = sb0;
        p.nouts = no0;
        p.nprods = np0;
        break;
      }
      SA_REQUIRE(p.nprods < ((int64_t)1 << 31), "strassen: one root needs >= 2^31 leaf products");
      ++j;
    }
    for (auto& d : p.depth) st.nnodes += (int64_t)d.size();
    st.nleaf_products += p.nprods;
    st.scratch_bytes = std::max(st.scratch_bytes, p.scratch_blocks * bs2b);
    st.batches++;
    run_plan(C, A, B, p, o.tile);
    i = j;
  }
  g_last_stats = st;
}

}  // namespace sa

using namespace sa;

extern "C" {

int sa_strassen_dryrun(sa_bsm_t A, sa_bsm_t B, double thresh_per_block, uint8_t* bits, int64_t cap,
                       int64_t* level_size, uint64_t* level_flops, int32_t* nlevels) {
  return guard([&] {
    SA_REQUIRE(A && B && nlevels, "null argument");
    SA_REQUIRE(uniform_square(A, B), "dryrun: A and B must be square with equal uniform blocking");
    SA_CHECK_HIP(hipSetDevice(A->ctx->device));
    Tree t = tree_shape(A->nrb);
    dryrun_device(A, B, thresh_per_block, t);
    std::vector<uint8_t> hb;
    download_tree(A->ctx, t, hb);
    *nlevels = t.nlev;
    if (bits) {
      SA_REQUIRE(cap >= t.total, "dryrun: bits capacity %lld < %lld", (long long)cap, (long long)t.total);
      if (t.total) std::memcpy(bits, hb.data(), (size_t)t.total);
    }
    if (level_size)
      for (int l = 0; l < t.nlev; ++l) level_size[l] = t.lvl_n[l];
    if (level_flops)
      for (int l = 0; l <= t.nlev; ++l) level_flops[l] = t.lvl_flops[l];
  });
}

int sa_matmult_strassen_sparse(sa_bsm_t C, sa_bsm_t A, int transA, sa_bsm_t B, int transB, double thresh,
                               double alpha, double beta, const sa_opts_t* opts) {
  return guard([&] {
    (void)beta;  // unused, as strassen.hpp:694-697: C is accumulated into
    SA_REQUIRE(C && A && B, "null matrix");
    SA_REQUIRE(!transA && !transB, "strassen: transposes are not supported (strassen.hpp:693)");
    SA_REQUIRE(alpha == 1.0, "strassen: alpha must be 1 (strassen.hpp:693)");
    SA_REQUIRE(uniform_square(A, B), "strassen: A and B must be square with equal uniform blocking");
    SA_REQUIRE(C->rsz == A->rsz && C->csz == A->csz, "strassen: C geometry differs from A");
    SA_REQUIRE(C != A && C != B, "strassen: C must not alias A or B");
    SA_REQUIRE(!C->base, "strassen: C must not be a reshape view");
    sa_opts_t o;
    sa_opts_default(&o);
    if (opts) o = *opts;
    SA_REQUIRE(o.mode == SA_MODE_FP64, "strassen: only FP64 mode");
    SA_CHECK_HIP(hipSetDevice(C->ctx->device));
    hybrid(C, A, B, thresh, o);
  });
}

int sa_strassen_features(sa_bsm_t C, sa_bsm_t A, sa_bsm_t B, double thresh, double* feat, int64_t cap,
                         int64_t* nnodes) {
  return guard([&] {
    SA_REQUIRE(A && B && nnodes, "null argument");
    SA_REQUIRE(uniform_square(A, B), "features: A and B must be square with equal uniform blocking");
    SA_REQUIRE(!C || (C->rsz == A->rsz && C->csz == A->csz), "features: C geometry differs from A");
    SA_CHECK_HIP(hipSetDevice(A->ctx->device));
    const int64_t bs = A->bs_r;
    const double thr = thresh * (double)(bs * bs * bs);
    Tree t = tree_shape(A->nrb);
    *nnodes = t.total;
    if (t.total == 0) return;
    dryrun_device(A, B, thr, t);
    node_features(C, A, B, thr, t);
    if (feat) {
      SA_REQUIRE(cap >= t.total * kMlpFeatures, "features: capacity %lld < %lld", (long long)cap,
                 (long long)(t.total * kMlpFeatures));
      SA_CHECK_HIP(hipMemcpyAsync(feat, A->ctx->buf("pl_x").p, sizeof(double) * (size_t)(t.total * kMlpFeatures),
                                  hipMemcpyDeviceToHost, A->ctx->stream));
      SA_CHECK_HIP(hipStreamSynchronize(A->ctx->stream));
    }
  });
}

int sa_strassen_plan(sa_bsm_t C, sa_bsm_t A, sa_bsm_t B, double thresh, uint8_t* sbits, uint8_t* dbits, int64_t cap,
                     int32_t* overruled, double* model_cost, double* heuristic_cost) {
  return guard([&] {
    SA_REQUIRE(A && B && sbits && dbits, "null argument");
    SA_REQUIRE(uniform_square(A, B), "plan: A and B must be square with equal uniform blocking");
    SA_REQUIRE(mlp_loaded(A->ctx), "plan: no path model loaded");
    SA_CHECK_HIP(hipSetDevice(A->ctx->device));
    const int64_t bs = A->bs_r;
    const double thr = thresh * (double)(bs * bs * bs);
    Tree t = tree_shape(A->nrb);
    SA_REQUIRE(cap >= t.total, "plan: capacity %lld < %lld", (long long)cap, (long long)t.total);
    if (t.total == 0) return;
    std::vector<uint8_t> hb, sb, db;
    dryrun_device(A, B, thr, t);
    download_tree(A->ctx, t, hb);
    const MlpPlan m = plan_mlp(C, A, B, thr, t, sb, db);
    std::memcpy(sbits, sb.data(), (size_t)t.total);
    std::memcpy(dbits, db.data(), (size_t)t.total);
    if (overruled) *overruled = m.overruled ? 1 : 0;
    if (model_cost) *model_cost = m.model_cost;
    if (heuristic_cost) *heuristic_cost = m.heur_cost;
  });
}

int sa_strassen_last_plan(int32_t* scheduler, int32_t* overruled, double* model_cost, double* heuristic_cost) {
  return guard([&] {
    if (scheduler) *scheduler = g_last_stats.scheduler;
    if (overruled) *overruled = g_last_stats.mlp.overruled ? 1 : 0;
    if (model_cost) *model_cost = g_last_stats.mlp.model_cost;
    if (heuristic_cost) *heuristic_cost = g_last_stats.mlp.heur_cost;
  });
}

int sa_strassen_last_stats(int64_t* nroots, int64_t* nnodes, int64_t* nleaf_products, int64_t* scratch_bytes,
                           int64_t* batches, int64_t* screened, int64_t* dense_roots) {
  return guard([&] {
    if (dense_roots) *dense_roots = g_last_stats.dense;
    if (nroots) *nroots = g_last_stats.nroots;
    if (nnodes) *nnodes = g_last_stats.nnodes;
    if (nleaf_products) *nleaf_products = g_last_stats.nleaf_products;
    if (scratch_bytes) *scratch_bytes = g_last_stats.scratch_bytes;
    if (batches) *batches = g_last_stats.batches;
    if (screened) *screened = g_last_stats.screened;
  });
}

}  // extern "C"
