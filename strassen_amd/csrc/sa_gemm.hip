// sa_gemm.hip -- grouped FP64 block GEMM on CDNA4 matrix cores.
//
// One launch runs a whole batch of output blocks of uneven sizes.  Each work
// item is one output tile (TM x TN) of one output block; its workgroup keeps the
// tile's accumulators in registers across every product of the block's
// segment (C-stationary), stages op(A) / op(B) k-slices of KT through LDS with
// loads coalesced along each stored block's contiguous dimension, and issues
// v_mfma_f64_16x16x4_f64 per 16x16 sub-tile.  C is written once at the end:
//   C = alpha * sum_p op(A_p) op(B_p) + (acc ? C : 0).
//
// MFMA operand layout (wave64, 16x16x4 f64):
//   A: lane l holds A[l%16][l/16]     (row i, k index)
//   B: lane l holds B[l/16][l%16]     (k index, column j)
//   D: lane l, element v holds D[4*v + l/16][l%16]   (measured on gfx950)
// Work items are remapped so that consecutive items (which share A/B blocks)
// land on the same XCD (hardware dispatches workgroup i to XCD i % 8).
#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>

#include "sa_internal.hpp"

namespace sa {

typedef double dbl4 __attribute__((ext_vector_type(4)));


template <int WM, int WN, int FM, int FN, int KT, bool TA, bool TB>
__global__ void __launch_bounds__(64 * WM * WN)
    k_gemm_fp64(const GemmOut* __restrict__ outs, const GemmProd* __restrict__ prods,
                const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = 16 * FM * WM;
  constexpr int TN = 16 * FN * WN;
  constexpr int LDA = TM + 2;
  constexpr int LDB = TN + 2;
  static_assert((TM * KT) % NT == 0 && (TN * KT) % NT == 0, "tile/threads mismatch");
  static_assert(KT % 4 == 0, "KT must be a multiple of the MFMA k (4)");
  __shared__ double As[KT][LDA];
  __shared__ double Bs[KT][LDB];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid % WM, wn = wid / WM;
  const int64_t it = xcd_remap(blockIdx.x, nitems);
  const GemmItem item = items ? items[it] : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = outs[item.out];
  const int m = o.m, n = o.n;
  const int m0 = item.tm * TM, n0 = item.tn * TN;

  dbl4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  for (int p = o.p0; p < o.p1; ++p) {
    const GemmProd pr = prods[p];
    const int k = pr.k;
    for (int k0 = 0; k0 < k; k0 += KT) {
#pragma unroll
      for (int r = 0; r < (TM * KT) / NT; ++r) {
        const int e = tid + r * NT;
        const int i = TA ? e / KT : e % TM;
        const int kk = TA ? e % KT : e / TM;
        const int gi = m0 + i, gk = k0 + kk;
        double v = 0.0;
        if (gi < m && gk < k) v = TA ? pr.a[gk + (int64_t)gi * k] : pr.a[gi + (int64_t)gk * m];
        As[kk][i] = v;
      }
#pragma unroll
      for (int r = 0; r < (TN * KT) / NT; ++r) {
        const int e = tid + r * NT;
        const int j = TB ? e % TN : e / KT;
        const int kk = TB ? e / TN : e % KT;
        const int gj = n0 + j, gk = k0 + kk;
        double v = 0.0;
        if (gj < n && gk < k) v = TB ? pr.b[gj + (int64_t)gk * n] : pr.b[gk + (int64_t)gj * k];
        Bs[kk][j] = v;
      }
      __syncthreads();
#pragma unroll
      for (int ks = 0; ks < KT / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[kr][(wm * FM + i) * 16 + (lane & 15)];
#pragma unroll
        for (int j = 0; j < FN; ++j) bf[j] = Bs[kr][(wn * FN + j) * 16 + (lane & 15)];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
      }
      __syncthreads();
    }
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int col = n0 + (wn * FN + j) * 16 + (lane & 15);
      const int rowb = m0 + (wm * FM + i) * 16 + (lane >> 4);
      if (col >= n) continue;
      double* cp = o.c + (int64_t)col * m;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = rowb + 4 * v;
        if (row < m) {
          double x = alpha * acc[i][j][v];
          if (o.acc) x = x + cp[row];
          cp[row] = x;
        }
      }
    }
}

// Small (atom-shaped, 5-40 wide) blocks: each wave owns a (16 FM) x (16 FN)
// output tile on its own (no LDS, no barriers) and walks the stream of stages
// of its output -- KC 4-k steps of one product, the whole product when k <= 4
// KC -- with operands loaded straight from global memory into registers and
// the next stage's loads issued before the current stage's MFMAs.  A stage
// then costs one memory latency instead of one per 4-k step, and with few
// registers per wave (KC = 2 or 4: under 72) seven or eight waves per SIMD cover it.
// 4-k steps past the product's k are skipped (wave-uniform branch).
template <int FM, int FN, int KC, int NW, bool TA, bool TB>
__global__ void __launch_bounds__(64 * NW)
    k_gemm_fp64_stage(const GemmOut* __restrict__ outs, const GemmProd* __restrict__ prods,
                      const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t nblk = (nitems + NW - 1) / NW;
  const int64_t it = xcd_remap(blockIdx.x, nblk) * NW + w;
  if (it >= nitems) return;
  const GemmItem item = items ? items[it] : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = outs[item.out];
  const int m = o.m, n = o.n;
  const int m0 = item.tm * 16 * FM, n0 = item.tn * 16 * FN;
  const int li = lane & 15, lk = lane >> 4;

  dbl4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  // operands of 4-k steps s < KC of the stage at k0 of product pr: lane l
  // holds op(A)[row l % 16][k0 + 4 s + l / 16] and op(B)[k0 + 4 s + l / 16][col l % 16]
  auto load = [&](const GemmProd& pr, int k0, double (&a)[KC][FM], double (&b)[KC][FN]) {
    const int k = pr.k;
#pragma unroll
    for (int s = 0; s < KC; ++s) {
      const int kk = k0 + 4 * s + lk;
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = m0 + i * 16 + li;
        a[s][i] = (r < m && kk < k) ? (TA ? pr.a[kk + (int64_t)r * k] : pr.a[r + (int64_t)kk * m]) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int cc = n0 + j * 16 + li;
        b[s][j] = (cc < n && kk < k) ? (TB ? pr.b[cc + (int64_t)kk * n] : pr.b[kk + (int64_t)cc * k]) : 0.0;
      }
    }
  };

  int p = o.p0, k0 = 0;
  if (p < o.p1) {
    GemmProd pr = prods[p];
    double ac[KC][FM], bc[KC][FN], an[KC][FM], bn[KC][FN];
    load(pr, 0, ac, bc);
    while (true) {
      int np = p, nk0 = k0 + 4 * KC;
      GemmProd pn = pr;
      if (nk0 >= pr.k) {
        np = p + 1;
        nk0 = 0;
        if (np < o.p1) pn = prods[np];
      }
      if (np < o.p1) load(pn, nk0, an, bn);
      const int ns = (int)ceil_div(pr.k - k0, 4);  // live 4-k steps of this stage (>= 1)
#pragma unroll
      for (int s = 0; s < KC; ++s) {
        if (s < ns) {
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[s][i], bc[s][j], acc[i][j], 0, 0, 0);
        }
      }
      if (np >= o.p1) break;
#pragma unroll
      for (int s = 0; s < KC; ++s) {
#pragma unroll
        for (int i = 0; i < FM; ++i) ac[s][i] = an[s][i];
#pragma unroll
        for (int j = 0; j < FN; ++j) bc[s][j] = bn[s][j];
      }
      p = np;
      k0 = nk0;
      pr = pn;
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
        const int row = m0 + i * 16 + 4 * v + lk;
        if (row < m) {
          double x = alpha * acc[i][j][v];
          if (o.acc) x = x + cp[row];
          cp[row] = x;
        }
      }
    }
}

// Pipelined variant for large blocks: the workgroup walks the flattened stream
// of (product, k-slice) stages of its output tile; while the MFMAs of stage s
// read LDS buffer s%2, the global loads of stage s+1 are already in flight into
// registers and land in the other buffer afterwards -- one barrier per stage.
// Rows of the staging buffers carry one pad element, so the k-major staging
// writes of transposed operands hit distinct LDS banks.
// Q4: the same tile on v_mfma_f64_4x4x4f64 (four 4 x 4 x 4 blocks per
// instruction).  With the same tiles and staging it runs faster than the
// v_mfma_f64_16x16x4f64 form here (measured on gfx950: dense 2048-blocks
// 51-54 against 46-49 TFLOP/s with the 128 x 64 tile, the RI-K step's FP64
// GEMMs about 6 % faster).  Block g of an instruction takes A[i][k] from lane 16 k + 4 g + i,
// B[k][j] from lane 16 k + 4 g + j and returns D[i][j] in lane 16 i + 4 g + j.
// A 16 x 16 fragment at one 4-k step is four instructions q = 0..3: block g
// computes rows 4 g .. 4 g + 3 against columns 4 q .. 4 q + 3, so the A
// operand is the 16x16x4 one (lane l: row l % 16, k = l / 16) and the B
// operand of instruction q holds B[k = l / 16][4 q + l % 4] (the same in all
// four blocks); result q of lane l is C[4 ((l / 4) % 4) + l / 16][4 q + l % 4].
template <int WM, int WN, int FM, int FN, int KT, bool TA, bool TB, bool Q4 = false>
__global__ void __launch_bounds__(64 * WM * WN)
    k_gemm_fp64_pipe(const GemmOut* __restrict__ outs, const GemmProd* __restrict__ prods,
                     const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  constexpr int NT = 64 * WM * WN;
  constexpr int TM = 16 * FM * WM;
  constexpr int TN = 16 * FN * WN;
  constexpr int LDA = TM + 1;
  constexpr int LDB = TN + 1;
  constexpr int LA = (TM * KT) / NT;
  constexpr int LB = (TN * KT) / NT;
  static_assert((TM * KT) % NT == 0 && (TN * KT) % NT == 0, "tile/threads mismatch");
  __shared__ double As[2][KT][LDA];
  __shared__ double Bs[2][KT][LDB];

  const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
  const int wm = wid % WM, wn = wid / WM;
  const int64_t it = xcd_remap(blockIdx.x, nitems);
  const GemmItem item = items ? gload1(items + it) : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = gload1(outs + item.out);
  const int m = o.m, n = o.n;
  const int m0 = item.tm * TM, n0 = item.tn * TN;

  dbl4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  double ra[LA], rb[LB];
  auto gload = [&](const GemmProd& pr, int k0) {
    const int k = pr.k;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + r * NT;
      const int i = TA ? e / KT : e % TM;
      const int kk = TA ? e % KT : e / TM;
      const int gi = m0 + i, gk = k0 + kk;
      ra[r] = (gi < m && gk < k) ? gload1(pr.a + (TA ? gk + (int64_t)gi * k : gi + (int64_t)gk * m)) : 0.0;
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + r * NT;
      const int j = TB ? e % TN : e / KT;
      const int kk = TB ? e / TN : e % KT;
      const int gj = n0 + j, gk = k0 + kk;
      rb[r] = (gj < n && gk < k) ? gload1(pr.b + (TB ? gj + (int64_t)gk * n : gk + (int64_t)gj * k)) : 0.0;
    }
  };
  auto lstore = [&](int buf) {
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int e = tid + r * NT;
      const int i = TA ? e / KT : e % TM;
      const int kk = TA ? e % KT : e / TM;
      As[buf][kk][i] = ra[r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int e = tid + r * NT;
      const int j = TB ? e % TN : e / KT;
      const int kk = TB ? e / TN : e % KT;
      Bs[buf][kk][j] = rb[r];
    }
  };

  int p = o.p0, k0 = 0;
  if (p >= o.p1) goto epilogue;
  {
    GemmProd pr = gload1(prods + p);
    // record of the following product, loaded one product ahead
    GemmProd pf = p + 1 < o.p1 ? gload1(prods + p + 1) : pr;
    gload(pr, 0);
    lstore(0);
    __syncthreads();
    int buf = 0;
    while (true) {
      int np = p, nk0 = k0 + KT;
      if (nk0 >= pr.k) {
        np = p + 1;
        nk0 = 0;
      }
      const bool has_next = np < o.p1;
      GemmProd pn = pr;
      if (has_next) {
        if (np != p) {
          pn = pf;
          if (np + 1 < o.p1) pf = gload1(prods + np + 1);
        }
        gload(pn, nk0);
      }
#pragma unroll
      for (int ks = 0; ks < KT / 4; ++ks) {
        const int kr = ks * 4 + (lane >> 4);
        double af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i) af[i] = As[buf][kr][(wm * FM + i) * 16 + (lane & 15)];
        if constexpr (Q4) {
          double bq[FN][4];
#pragma unroll
          for (int j = 0; j < FN; ++j)
#pragma unroll
            for (int q = 0; q < 4; ++q) bq[j][q] = Bs[buf][kr][(wn * FN + j) * 16 + 4 * q + (lane & 3)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
#pragma unroll
              for (int q = 0; q < 4; ++q)
                acc[i][j][q] = __builtin_amdgcn_mfma_f64_4x4x4f64(af[i], bq[j][q], acc[i][j][q], 0, 0, 0);
        } else {
#pragma unroll
          for (int j = 0; j < FN; ++j) bf[j] = Bs[buf][kr][(wn * FN + j) * 16 + (lane & 15)];
#pragma unroll
          for (int i = 0; i < FM; ++i)
#pragma unroll
            for (int j = 0; j < FN; ++j)
              acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(af[i], bf[j], acc[i][j], 0, 0, 0);
        }
      }
      if (!has_next) break;
      lstore(buf ^ 1);
      __syncthreads();
      buf ^= 1;
      p = np;
      k0 = nk0;
      pr = pn;
    }
  }
epilogue:
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        // element v of lane l: 16x16x4: C[4 v + l / 16][l % 16];
        // 4x4x4 blocks: C[4 ((l / 4) % 4) + l / 16][4 v + l % 4]
        const int row = m0 + (wm * FM + i) * 16 + (Q4 ? 4 * ((lane >> 2) & 3) + (lane >> 4) : 4 * v + (lane >> 4));
        const int col = n0 + (wn * FN + j) * 16 + (Q4 ? 4 * v + (lane & 3) : (lane & 15));
        if (row < m && col < n) {
          double* cp = o.c + (int64_t)col * m + row;
          double x = alpha * acc[i][j][v];
          if (o.acc) x = x + *cp;
          *cp = x;
        }
      }
    }
}

// Wave-per-tile variant for small blocks: each wave owns one output tile of
// (16*FM) x (16*FN), loads its MFMA operand fragments straight from global
// memory (L1/L2 resident for atom-sized blocks) and prefetches the next k-step
// into registers while the current MFMAs issue.  No LDS and no barriers, so NW
// independent waves share a workgroup.
template <int FM, int FN, int NW, bool TA, bool TB>
__global__ void __launch_bounds__(64 * NW)
    k_gemm_fp64_wave(const GemmOut* __restrict__ outs, const GemmProd* __restrict__ prods,
                     const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int64_t nblk = (nitems + NW - 1) / NW;
  const int64_t it = xcd_remap(blockIdx.x, nblk) * NW + w;
  if (it >= nitems) return;
  const GemmItem item = items ? items[it] : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = outs[item.out];
  const int m = o.m, n = o.n;
  const int m0 = item.tm * 16 * FM, n0 = item.tn * 16 * FN;
  const int li = lane & 15, lk = lane >> 4;

  dbl4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};

  for (int p = o.p0; p < o.p1; ++p) {
    const GemmProd pr = prods[p];
    const int k = pr.k;
    double ac[FM], bc[FN], an[FM], bn[FN];
    auto load = [&](int k0, double* a, double* b) {
      const int kk = k0 + lk;
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int r = m0 + i * 16 + li;
        a[i] = (r < m && kk < k) ? (TA ? pr.a[kk + (int64_t)r * k] : pr.a[r + (int64_t)kk * m]) : 0.0;
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int cc = n0 + j * 16 + li;
        b[j] = (cc < n && kk < k) ? (TB ? pr.b[cc + (int64_t)kk * n] : pr.b[kk + (int64_t)cc * k]) : 0.0;
      }
    };
    load(0, ac, bc);
    for (int k0 = 0; k0 < k; k0 += 4) {
      if (k0 + 4 < k) load(k0 + 4, an, bn);
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ac[i], bc[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < FM; ++i) ac[i] = an[i];
#pragma unroll
      for (int j = 0; j < FN; ++j) bc[j] = bn[j];
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
        const int row = m0 + i * 16 + 4 * v + lk;
        if (row < m) {
          double x = alpha * acc[i][j][v];
          if (o.acc) x = x + cp[row];
          cp[row] = x;
        }
      }
    }
}

// ----------------------------------------------------------- work items --
__global__ void k_count_tiles(const GemmOut* __restrict__ outs, int64_t nouts, int TM, int TN,
                              int64_t* __restrict__ cnt) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  cnt[o] = ceil_div(outs[o].m, TM) * ceil_div(outs[o].n, TN);
}

__global__ void k_emit_items(const GemmOut* __restrict__ outs, int64_t nouts, int TM, int TN,
                             const int64_t* __restrict__ start, GemmItem* __restrict__ items) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  const int tm = (int)ceil_div(outs[o].m, TM), tn = (int)ceil_div(outs[o].n, TN);
  int64_t s = start[o];
  for (int b = 0; b < tn; ++b)
    for (int a = 0; a < tm; ++a) items[s++] = GemmItem{(int32_t)o, (int16_t)a, (int16_t)b};
}

// Work items in group-column order: the outputs of one group column (the up to
// 4 consecutive records with the same nonzero group: C blocks of 4 block rows
// in one block column) emit their tiles together, in chunks of CC column
// tiles: for each chunk, each output's tiles of those columns, row tile
// fastest.  The tiles running together on one XCD then span several block
// rows (sharing op(B) fragments, whose product lists mostly coincide) and a
// chunk of column tiles (sharing op(A) fragments), instead of the 3-4 row
// tiles of one block.  Outputs with group 0 keep their own order (row tile
// fastest).  One thread per run; start: exclusive scan of the tile counts.
__global__ void k_emit_items_grouped(const GemmOut* __restrict__ outs, int64_t nouts, int TM, int TN, int CC,
                                     const int64_t* __restrict__ start, GemmItem* __restrict__ items) {
  const int64_t o = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (o >= nouts) return;
  const int32_t g = outs[o].group;
  if (g != 0 && o > 0 && outs[o - 1].group == g) return;  // not the first of its run
  int64_t end = o + 1;
  if (g != 0)
    while (end < nouts && end - o < 4 && outs[end].group == g) ++end;
  int maxtn = 0;
  for (int64_t q = o; q < end; ++q) maxtn = max(maxtn, (int)ceil_div(outs[q].n, TN));
  int64_t s = start[o];
  for (int c0 = 0; c0 < maxtn; c0 += CC)
    for (int64_t q = o; q < end; ++q) {
      const int tm = (int)ceil_div(outs[q].m, TM), tn = (int)ceil_div(outs[q].n, TN);
      for (int b = c0; b < min(c0 + CC, tn); ++b)
        for (int a = 0; a < tm; ++a) items[s++] = GemmItem{(int32_t)q, (int16_t)a, (int16_t)b};
    }
}

// column-tile chunk of the group-column item order (env SA_ITEM_CHUNK; 0: each
// output's tiles on their own)
int item_chunk() {
  static const int cc = [] {
    const char* v = std::getenv("SA_ITEM_CHUNK");
    return v ? std::max(0, std::atoi(v)) : 8;
  }();
  return cc;
}

void emit_items(sa_context* c, const GemmOut* outs, int64_t nouts, int TM, int TN, const int64_t* start,
                GemmItem* items, bool grouped) {
  const int cc = item_chunk();
  if (grouped && cc > 0)
    hipLaunchKernelGGL(k_emit_items_grouped, dim3((unsigned)ceil_div(nouts, 256)), dim3(256), 0, c->stream, outs,
                       nouts, TM, TN, cc, start, items);
  else
    hipLaunchKernelGGL(k_emit_items, dim3((unsigned)ceil_div(nouts, 256)), dim3(256), 0, c->stream, outs, nouts, TM,
                       TN, start, items);
}

// ------------------------------------------------------------- configs --
using GemmFn = void (*)(const GemmOut*, const GemmProd*, const GemmItem*, int64_t, double);

struct Cfg {
  const char* name;
  int tm, tn, threads;
  GemmFn fn[4];  // [ta*2+tb]
  int items_per_wg = 1;
  bool pipelined = false;  // k_gemm_fp64_pipe: the only candidates for blocks of 256 and more
};

#define SA_CFG(NAME, WM, WN, FM, FN, KT)                                                           \
  Cfg {                                                                                           \
    NAME, 16 * FM * WM, 16 * FN * WN, 64 * WM * WN,                                               \
        {k_gemm_fp64<WM, WN, FM, FN, KT, false, false>, k_gemm_fp64<WM, WN, FM, FN, KT, false, true>, \
         k_gemm_fp64<WM, WN, FM, FN, KT, true, false>, k_gemm_fp64<WM, WN, FM, FN, KT, true, true>} \
  }

#define SA_WCFG(NAME, FM, FN, NW)                                                                   \
  Cfg {                                                                                             \
    NAME, 16 * FM, 16 * FN, 64 * NW,                                                                \
        {k_gemm_fp64_wave<FM, FN, NW, false, false>, k_gemm_fp64_wave<FM, FN, NW, false, true>,     \
         k_gemm_fp64_wave<FM, FN, NW, true, false>, k_gemm_fp64_wave<FM, FN, NW, true, true>},      \
        NW                                                                                          \
  }

#define SA_SCFG(NAME, FM, FN, KC, NW)                                                                       \
  Cfg {                                                                                                   \
    NAME, 16 * FM, 16 * FN, 64 * NW,                                                                      \
        {k_gemm_fp64_stage<FM, FN, KC, NW, false, false>, k_gemm_fp64_stage<FM, FN, KC, NW, false, true>, \
         k_gemm_fp64_stage<FM, FN, KC, NW, true, false>, k_gemm_fp64_stage<FM, FN, KC, NW, true, true>},  \
        NW                                                                                                \
  }
#define SA_QCFG(NAME, WM, WN, FM, FN, KT)                                                                 \
  Cfg {                                                                                                   \
    NAME, 16 * FM * WM, 16 * FN * WN, 64 * WM * WN,                                                       \
        {k_gemm_fp64_pipe<WM, WN, FM, FN, KT, false, false, true>,                                        \
         k_gemm_fp64_pipe<WM, WN, FM, FN, KT, false, true, true>,                                         \
         k_gemm_fp64_pipe<WM, WN, FM, FN, KT, true, false, true>, k_gemm_fp64_pipe<WM, WN, FM, FN, KT, true, true, true>}, \
        1, true                                                                                           \
  }
#define SA_PCFG(NAME, WM, WN, FM, FN, KT)                                                            \
  Cfg {                                                                                              \
    NAME, 16 * FM * WM, 16 * FN * WN, 64 * WM * WN,                                                  \
        {k_gemm_fp64_pipe<WM, WN, FM, FN, KT, false, false>, k_gemm_fp64_pipe<WM, WN, FM, FN, KT, false, true>, \
         k_gemm_fp64_pipe<WM, WN, FM, FN, KT, true, false>, k_gemm_fp64_pipe<WM, WN, FM, FN, KT, true, true>}, \
        1, true                                                                                      \
  }

static const Cfg kCfgs[] = {
    SA_CFG("64x64_w4_k16", 2, 2, 2, 2, 16),   // 0: uniform 64 blocks
    SA_CFG("128x128_w4_k16", 2, 2, 4, 4, 16), // 1: uniform 128 blocks
    SA_CFG("32x32_w1_k16", 1, 1, 2, 2, 16),   // 2: small blocks, one wave
    SA_CFG("48x48_w1_k8", 1, 1, 3, 3, 8),     // 3: atom blocks <= 48
    SA_CFG("16x16_w1_k8", 1, 1, 1, 1, 8),     // 4: tiny blocks
    SA_CFG("32x32_w4_k16", 2, 2, 1, 1, 16),   // 5: 32 tile, four waves
    SA_CFG("64x64_w4_k8", 2, 2, 2, 2, 8),     // 6
    SA_CFG("128x64_w4_k16", 2, 2, 4, 2, 16),  // 7
    SA_WCFG("wave16x16_n4", 1, 1, 4),         // 8: wave per tile, direct loads
    SA_WCFG("wave32x32_n4", 2, 2, 4),         // 9
    SA_WCFG("wave48x48_n4", 3, 3, 4),         // 10
    SA_WCFG("wave32x32_n2", 2, 2, 2),         // 11
    SA_WCFG("wave64x32_n2", 4, 2, 2),         // 12
    SA_WCFG("wave64x64_n1", 4, 4, 1),         // 13
    SA_PCFG("pipe128x128_k16", 2, 2, 4, 4, 16), // 14: pipelined, large blocks
    SA_PCFG("pipe128x64_k16", 2, 2, 4, 2, 16),  // 15
    SA_PCFG("pipe64x64_k16", 2, 2, 2, 2, 16),   // 16
    SA_PCFG("pipe64x128_k16", 2, 2, 2, 4, 16),  // 17
    SA_QCFG("q4pipe128x128_k16", 2, 2, 4, 4, 16),  // 18: pipelined, 4x4x4 FP64 MFMAs
    SA_QCFG("q4pipe128x64_k16", 2, 2, 4, 2, 16),   // 19
    SA_QCFG("q4pipe64x64_k16", 2, 2, 2, 2, 16),    // 20
    SA_QCFG("q4pipe64x128_k16", 2, 2, 2, 4, 16),   // 21
    SA_SCFG("stage16x16_k16_n4", 1, 1, 4, 4),      // 22: wave per tile, whole-stage register loads
    SA_SCFG("stage16x16_k8_n4", 1, 1, 2, 4),       // 23
    SA_SCFG("stage32x32_k16_n4", 2, 2, 4, 4),      // 24
    SA_SCFG("stage16x32_k16_n4", 1, 2, 4, 4),      // 25
    SA_SCFG("stage32x16_k16_n4", 2, 1, 4, 4),      // 26
    SA_SCFG("stage48x48_k8_n4", 3, 3, 2, 4),       // 27
};
static constexpr int kNumCfgs = (int)(sizeof(kCfgs) / sizeof(kCfgs[0]));

int gemm_fp64_num_configs() { return kNumCfgs; }
const char* gemm_fp64_config_name(int cfg) { return (cfg >= 0 && cfg < kNumCfgs) ? kCfgs[cfg].name : "?"; }

static int heuristic_cfg(const GemmBatch& b) {
  const int mx = std::max(b.max_m, b.max_n);
  if (mx > 96) return 19;  // q4pipe128x64_k16
  if (mx > 48) return 0;
  if (mx > 32) return 3;
  if (mx > 16) return 2;
  return 4;
}

// work items of one tile config (output tiles when a block exceeds the tile;
// sizing reads back one total)
static GemmPlan plan_items(sa_context* c, const GemmBatch& b, int cfg) {
  const Cfg& k = kCfgs[cfg];
  const GemmItem* items = nullptr;
  int64_t nitems = b.nouts;
  if (b.max_m > k.tm || b.max_n > k.tn) {
    c->ws_gemm_items.reserve(sizeof(int64_t) * (size_t)(b.nouts + 1));
    int64_t* cnt = c->ws_gemm_items.as<int64_t>();
    hipLaunchKernelGGL(k_count_tiles, dim3((unsigned)ceil_div(b.nouts, 256)), dim3(256), 0, c->stream, b.outs, b.nouts,
                       k.tm, k.tn, cnt);
    nitems = scan_exclusive_i64(c, cnt, b.nouts);
    c->ws_gemm_items2.reserve(sizeof(GemmItem) * (size_t)(nitems + 1));
    emit_items(c, b.outs, b.nouts, k.tm, k.tn, (const int64_t*)cnt, c->ws_gemm_items2.as<GemmItem>(), true);
    items = c->ws_gemm_items2.as<GemmItem>();
  }
  SA_REQUIRE(nitems < (int64_t)1 << 31, "too many GEMM work items");
  return GemmPlan{cfg, items, nitems};
}

// one kernel launch, no host interaction (capturable)
void gemm_fp64_launch(sa_context* c, const GemmBatch& b, const GemmPlan& g) {
  if (b.nouts == 0) return;
  const Cfg& k = kCfgs[g.cfg];
  const int64_t nwg = ceil_div(g.nitems, k.items_per_wg);
  hipLaunchKernelGGL(k.fn[b.ta * 2 + b.tb], dim3((unsigned)nwg), dim3(k.threads), 0, c->stream, b.outs, b.prods,
                     g.items, g.nitems, b.alpha);
  SA_CHECK_HIP(hipGetLastError());
}

static void launch_cfg(sa_context* c, const GemmBatch& b, int cfg) { gemm_fp64_launch(c, b, plan_items(c, b, cfg)); }

// ------------------------------------------------------- window items --
__global__ void k_gather_at(const int64_t* __restrict__ src, const int64_t* __restrict__ idx, int64_t n,
                            int64_t* __restrict__ dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[idx[i]];
}

WindowItems window_items(sa_context* c, const GemmOut* outs, const std::vector<int64_t>& off, int TM, int TN,
                         int32_t max_m, int32_t max_n) {
  WindowItems W;
  W.off = off;
  const int64_t n = off.empty() ? 0 : off.back();
  if (n == 0 || (max_m <= TM && max_n <= TN)) return W;
  const int64_t nw1 = (int64_t)off.size();
  DevBuf& bc = c->buf("win_cnt");
  DevBuf& bi = c->buf("win_idx");
  bc.reserve(sizeof(int64_t) * (size_t)(n + 1));
  bi.reserve(sizeof(int64_t) * (size_t)(2 * nw1));
  int64_t* cnt = bc.as<int64_t>();
  int64_t* idx = bi.as<int64_t>();
  SA_CHECK_HIP(hipMemcpyAsync(idx, off.data(), sizeof(int64_t) * (size_t)nw1, hipMemcpyHostToDevice, c->stream));
  hipLaunchKernelGGL(k_count_tiles, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, c->stream, outs, n, TM, TN, cnt);
  scan_exclusive_i64_async(c, cnt, n, cnt + n);  // cnt[n]: the total
  hipLaunchKernelGGL(k_gather_at, dim3((unsigned)ceil_div(nw1, 256)), dim3(256), 0, c->stream, (const int64_t*)cnt,
                     (const int64_t*)idx, nw1, idx + nw1);
  int64_t* h = static_cast<int64_t*>(pinned(c, sizeof(int64_t) * (size_t)nw1));
  SA_CHECK_HIP(hipMemcpyAsync(h, idx + nw1, sizeof(int64_t) * (size_t)nw1, hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  W.off.assign(h, h + nw1);
  const int64_t nitems = W.off.back();
  SA_REQUIRE(n < ((int64_t)1 << 31) && nitems < ((int64_t)1 << 31), "too many GEMM work items");
  DevBuf& bit = c->buf("win_items");
  bit.reserve(sizeof(GemmItem) * (size_t)(nitems + 1));
  // group-column order within one launch only: the windows of a windowed
  // multiply hold records of the same C blocks
  emit_items(c, outs, n, TM, TN, (const int64_t*)cnt, bit.as<GemmItem>(), nw1 == 2);
  SA_CHECK_HIP(hipGetLastError());
  W.items = bit.as<GemmItem>();
  return W;
}

// ---------------------------------------------------------------- autotuner --
// The tile configuration is chosen per shape class (transposes, largest block
// dimensions) by timing every plausible configuration on a sample of the real
// batch (at most 32768 outputs taken at a uniform stride) whose outputs are redirected to
// scratch with overwrite semantics, so the caller's C is untouched.  The
// choice is cached for the process.
__global__ void k_redirect_outs(const GemmOut* __restrict__ in, int64_t n, int64_t stride, double* __restrict__ scratch,
                                int64_t slot, GemmOut* __restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  GemmOut o = in[i * stride];
  o.c = scratch + i * slot;
  o.acc = 0;
  out[i] = o;
}

// shape class -> tile config; shared by all contexts (and the threads that
// drive them), guarded by tune_mutex
static std::map<std::tuple<int, int, int, int, int>, int>& tune_cache() {
  static std::map<std::tuple<int, int, int, int, int>, int> m;
  return m;
}
static std::mutex& tune_mutex() {
  static std::mutex mu;
  return mu;
}

static int round16(int x) { return (x + 15) / 16 * 16; }

static int autotune(sa_context* c, const GemmBatch& b) {
  const auto key = std::make_tuple(b.ta, b.tb, b.max_m, b.max_n, b.max_k);
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(key);
    if (it != tune_cache().end()) return it->second;
  }
  // sample size bounded by work: about 1.6e10 flops of products (64 to 32768
  // outputs) -- small-block batches need samples of a few hundred
  // microseconds before the timings rank the configs as the full batch does
  const double per_out = 2.0 * b.max_m * b.max_n * (double)b.max_k * ((double)b.nprods / (double)b.nouts);
  const int64_t S0 = std::max<int64_t>(1, std::min<int64_t>(
                         b.nouts, std::min<int64_t>(32768, std::max<int64_t>(64, (int64_t)(1.6e10 / std::max(per_out, 1.0))))));
  const int64_t slot = (int64_t)b.max_m * b.max_n;
  // at most 512 MiB of redirected outputs (large dense shapes: a few outputs)
  const int64_t S = std::max<int64_t>(1, std::min<int64_t>(S0, std::max<int64_t>(4, ((int64_t)1 << 26) / slot)));
  const int64_t stride = b.nouts / S;
  DevBuf& to = c->buf("tune_outs");
  DevBuf& tc = c->buf("tune_c");
  to.reserve(sizeof(GemmOut) * (size_t)(S + 1));
  tc.reserve(sizeof(double) * (size_t)(S * slot + 1));
  hipLaunchKernelGGL(k_redirect_outs, dim3((unsigned)ceil_div(S, 256)), dim3(256), 0, c->stream, b.outs, S, stride,
                     tc.as<double>(), slot, to.as<GemmOut>());
  GemmBatch bs = b;
  bs.outs = to.as<GemmOut>();
  bs.nouts = S;
  hipEvent_t e0, e1;
  SA_CHECK_HIP(hipEventCreate(&e0));
  SA_CHECK_HIP(hipEventCreate(&e1));
  int best = heuristic_cfg(b);
  float best_ms = 1e30f;
  const int pm = round16(b.max_m), pn = round16(b.max_n);
  for (int cfg = 0; cfg < kNumCfgs; ++cfg) {
    const Cfg& k = kCfgs[cfg];
    if (cfg != best && ((int64_t)k.tm * k.tn > 4LL * pm * pn || k.tm > 2 * pm + 16 || k.tn > 2 * pn + 16)) continue;
    if (cfg != best && (int64_t)k.tm * k.tn * 16 < std::min<int64_t>((int64_t)pm * pn, 128 * 128)) continue;  // tiny tiles
    if (cfg != best && pm >= 256 && pn >= 256 && !k.pipelined) continue;  // large blocks: pipelined tiles only
    launch_cfg(c, bs, cfg);  // warm-up (item lists, caches)
    float t = 1e30f;
    for (int rep = 0; rep < 2; ++rep) {
      SA_CHECK_HIP(hipEventRecord(e0, c->stream));
      launch_cfg(c, bs, cfg);
      SA_CHECK_HIP(hipEventRecord(e1, c->stream));
      SA_CHECK_HIP(hipEventSynchronize(e1));
      float ms = 0.f;
      SA_CHECK_HIP(hipEventElapsedTime(&ms, e0, e1));
      t = std::min(t, ms);
    }
    if (t < best_ms) {
      best_ms = t;
      best = cfg;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  {
    std::lock_guard<std::mutex> lk(tune_mutex());
    tune_cache()[key] = best;
  }
  return best;
}

GemmPlan gemm_fp64_plan_fixed(const GemmBatch& b, int cfg, TableMem& mem) {
  if (b.nouts == 0) return GemmPlan{};
  if (cfg < 0 || cfg >= kNumCfgs) {
    cfg = heuristic_cfg(b);
    std::lock_guard<std::mutex> lk(tune_mutex());
    auto it = tune_cache().find(std::make_tuple(b.ta, b.tb, b.max_m, b.max_n, b.max_k));
    if (it != tune_cache().end()) cfg = it->second;
  }
  const Cfg& k = kCfgs[cfg];
  const GemmItem* items = nullptr;
  int64_t nitems = b.nouts;
  if (b.max_m > k.tm || b.max_n > k.tn) {
    sa_context* c = mem.c;
    int64_t* cnt = mem.get<int64_t>("gi_cnt", b.nouts + 1);
    hipLaunchKernelGGL(k_count_tiles, dim3((unsigned)ceil_div(b.nouts, 256)), dim3(256), 0, c->stream, b.outs, b.nouts,
                       k.tm, k.tn, cnt);
    nitems = scan_exclusive_i64(mem, cnt, b.nouts);
    SA_REQUIRE(nitems < (int64_t)1 << 31, "too many GEMM work items");
    GemmItem* it = mem.get<GemmItem>("gi_items", nitems + 1);
    hipLaunchKernelGGL(k_emit_items, dim3((unsigned)ceil_div(b.nouts, 256)), dim3(256), 0, c->stream, b.outs,
                       b.nouts, k.tm, k.tn, (const int64_t*)cnt, it);
    items = it;
  }
  return GemmPlan{cfg, items, nitems};
}

int gemm_fp64_choose(sa_context* c, const GemmBatch& b, int cfg) {
  return (cfg >= 0 && cfg < kNumCfgs) ? cfg : autotune(c, b);
}

void gemm_fp64_tile(int cfg, int* tm, int* tn) {
  *tm = kCfgs[cfg].tm;
  *tn = kCfgs[cfg].tn;
}

GemmPlan gemm_fp64_plan(sa_context* c, const GemmBatch& b, int cfg) {
  if (b.nouts == 0) return GemmPlan{};
  if (cfg < 0 || cfg >= kNumCfgs) cfg = autotune(c, b);
  c->last_gemm_cfg = cfg;
  return plan_items(c, b, cfg);
}

void gemm_fp64_run(sa_context* c, const GemmBatch& b, int cfg) {
  if (b.nouts == 0) return;
  const GemmPlan g = gemm_fp64_plan(c, b, cfg);
  ProfScope ps(c, "gemm_fp64");
  gemm_fp64_launch(c, b, g);
}

}  // namespace sa
