// sa_ozaki.hip -- split-precision (Ozaki-scheme) screened block multiply.
//
// The task list is the FP64 one (screening on FP64 norms), so the set of
// computed block products is unchanged; only the block products themselves are
// evaluated with low-precision matrix cores:
//  * slicing (once per operand per call, HBM-bound): every op(A) row / op(B)
//    column v of a block gets a power-of-two scale 2^e_v > max|x_v| and the
//    integer n = rint(x 2^(beta*S - TS - e_v)) (|n| <= 2^(beta*S - TS)) is cut
//    into S balanced beta-bit digits d_0..d_{S-1} (d_0 most significant,
//    |d_0| <= 2^(beta - TS), |d_p| <= 2^(beta - 1)), stored as FP16 (beta = 10,
//    TS = 2) or BF16 (beta = 8, TS = 1) values, which hold them exactly;
//  * product: x_A x_B = sum_{p,q} d_p d_q 2^(e_r + f_c + 2 TS - beta (p+q+2));
//    pairs with p + q <= S - 1 are kept (the error-free-transformation split of
//    Ozaki, Ogita, Oishi & Rump, Numer. Algorithms 59 (2012), in the integer
//    digit form of Ootomo & Yokota 2022 / Uchino et al. 2024);
//  * digit products of one weight w = p + q accumulate in FP32 inside
//    v_mfma_f32_16x16x32_{f16,bf16}; the sums of the weights that carry the
//    leading bits stay exact (integers below 2^24) between two folds into the
//    FP64 accumulators: FP16 weights 0..1 for 2 k-steps of 32 (weight 1:
//    2*256*512*32*2 = 2^24), BF16 weights 0..2 for 8, so all weights are
//    folded every 2 (FP16) or 8 (BF16) k-steps.  Later weights may round in
//    FP32, which costs at most 2^-(2 beta + 21) of the leading term per fold.
// INT8 slices (an MI355X-first third form, v_mfma_i32_16x16x64_i8 at twice
// the FP16 rate): beta = 8, TS = 2 digits in int8, where a vector whose
// largest value leaves the leading digit of m = n + H at most 127 takes the
// exponent one lower (TS = 1 in effect: one more significant bit); digit
// products accumulate in int32, exact for 2^11 / S k-steps of 64
// (|d_p d_q| <= 2^14), so products are folded only at their end (or every
// 256 k-steps).
// Truncation error per product is about 2^-(beta*S - TS) relative to the
// row/column maxima of the two blocks (4 FP16 slices: 2^-38; 5 INT8: 2^-38).
//
// Slice layout: for each block, S slices of [vb][ks][lane][16 bytes]
// fragments, one fragment = 16 vectors x KS k = the operand of one 16x16xKS
// MFMA (1 KB; KS = 32 for FP16/BF16, 64 for INT8): lane l holds vector
// 16 vb + (l & 15) at k = KS ks + EL (l >> 4) + t, EL = 8 (16-bit digits) or
// 16 (8-bit).  Vectors are padded to a multiple of 16 and k to a multiple of
// KS with zeros, so every fragment load is one contiguous, coalesced 1 KB read.
#include <algorithm>
#include <cmath>
#include <cstdlib>

#include "sa_internal.hpp"

namespace sa {

typedef _Float16 half8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef double dbl4 __attribute__((ext_vector_type(4)));

typedef int int4v __attribute__((ext_vector_type(4)));

// slice kinds: SA_SLICE_FP16 (0), SA_SLICE_BF16 (1), SA_SLICE_INT8 (2)
__host__ __device__ constexpr int oz_beta(int kind) { return kind == 0 ? 10 : 8; }
__host__ __device__ constexpr int oz_tshift(int kind) { return kind == 1 ? 1 : 2; }
__host__ __device__ constexpr int oz_kstep(int kind) { return kind == 2 ? 64 : 32; }  // MFMA k
__host__ __device__ constexpr int oz_esize(int kind) { return kind == 2 ? 1 : 2; }    // bytes per digit
// k-steps between folds of all weights into FP64 (FP16 weight 1 is exact for 2
// k-steps of 32, BF16 weights 0..2 for 8, INT8 all weights for 256 of 64)
__host__ __device__ constexpr int oz_full_fold(int kind) { return kind == 2 ? 256 : kind == 1 ? 8 : 2; }

template <int KIND>
struct OzAcc {
  typedef flt4 type;
};
template <>
struct OzAcc<2> {
  typedef int4v type;
};

__device__ inline flt4 oz_mfma(std::integral_constant<int, 0>, uint4 a, uint4 b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(__builtin_bit_cast(half8, a), __builtin_bit_cast(half8, b), c, 0, 0,
                                                0);
}
__device__ inline flt4 oz_mfma(std::integral_constant<int, 1>, uint4 a, uint4 b, flt4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, a), __builtin_bit_cast(bf16x8, b), c, 0,
                                                 0, 0);
}
__device__ inline int4v oz_mfma(std::integral_constant<int, 2>, uint4 a, uint4 b, int4v c) {
  return __builtin_amdgcn_mfma_i32_16x16x64_i8(__builtin_bit_cast(int4v, a), __builtin_bit_cast(int4v, b), c, 0, 0, 0);
}
__host__ __device__ constexpr double pow2i(int e) {
  double r = 1.0;
  for (int i = 0; i < e; ++i) r *= 2.0;
  for (int i = 0; i > e; --i) r *= 0.5;
  return r;
}

struct OzProd {
  const uint8_t* a;   // S slices of the op(A) block (fragment layout)
  const uint8_t* b;   // S slices of the op(B) block
  const int32_t* ea;  // row exponents (padded to 16)
  const int32_t* eb;  // column exponents (padded to 16)
  int32_t nks;        // MFMA k-steps
  int32_t nvba;       // 16-row blocks of the op(A) block
  int32_t nvbb;       // 16-column blocks of the op(B) block
  // bit 0: same row/column exponents as the previous product of its output
  // (the products chain in one accumulation)
  int32_t flags;
  // leading all-zero slices of each 16-row group g of the op(A) block (S: all
  // zero), byte g & 3 of za[g >> 2]; group 15 holds the least over groups 15
  // and up.  Four words rather than a byte array: picked by selects, they
  // stay in registers
  uint32_t za0, za1, za2, za3;
};
__host__ __device__ inline bool oz_chain(int32_t flags) { return (flags & 1) != 0; }
// leading zero slices of 16-row group g of a product's op(A) block (shifts
// of combined words: a select between loaded fields would become a load at
// a computed offset and keep the record out of registers)
__device__ inline int oz_zero_g(const OzProd& pr, int g) {
  const int q = min(g, 15);
  const uint64_t lo = ((uint64_t)pr.za1 << 32) | pr.za0, hi = ((uint64_t)pr.za3 << 32) | pr.za2;
  return (int)(((q < 8 ? lo : hi) >> (8 * (q & 7))) & 0xffu);
}

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

// bytes of slices and number of exponents of every block
__global__ void k_slice_need(SliceArgs s, int S, int kstep, int esize, int64_t* __restrict__ need_sl,
                             int64_t* __restrict__ need_ex) {
  const int64_t b = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= s.nb) return;
  int nv, kd;
  int64_t sv, sk;
  slice_geom(s, b, nv, kd, sv, sk);
  const bool pres = s.off[b] >= 0;
  const int64_t kp = ceil_div(kd, kstep) * kstep;
  const int64_t nvp = ceil_div(nv, 16) * 16;
  need_sl[b] = pres ? (int64_t)S * nvp * kp * esize : 0;
  need_ex[b] = pres ? nvp : 0;
}

// One single-wave workgroup per 16-vector group of a block (groups numbered
// through the exclusive scan of the padded vector counts, ex_off).  Each
// k-step of the group, a 16 x KS tile of the operand, is read with coalesced
// loads (along k where k is contiguous, else along the vectors) into LDS; lane
// l then owns vector 16 vb + (l & 15) and the EL digits at k = KS ks +
// EL (l >> 4) + t, i.e. exactly lane l's 16 bytes of the fragment, so every
// slice of a fragment is one coalesced 1 KB store.  Pass 1 reduces max|x| per
// vector; pass 2 re-reads the tiles and cuts the values into digits.  Balanced
// digits in closed form: with m = n + H, H = sum_{j < S-1} 2^(beta-1) 2^(beta j),
// the digit at position j < S-1 is ((m >> beta j) & (2^beta - 1)) - 2^(beta-1)
// and the leading one m >> beta (S-1), the same digits as peeling n from its
// least significant end.
// mode 0: exponents and digits; 1: exponents only; 2: digits with the
// exponents already in exps (shared by all blocks of a vector's line)
constexpr int kNoExp = INT32_MIN / 2;  // exponent of an all-zero vector before sharing

// Scale exponent e of a vector whose largest magnitude is mx: mx < 2^e.  INT8:
// one guard bit of the leading digit is needed only when the vector's largest
// value would carry it to +128; otherwise the scale moves up one bit (every
// |x| <= mx then gives a leading digit in [-128, 127]), which halves the
// truncation of the vector's digits.  An all-zero vector gets `none`.
template <int KIND>
__device__ inline int oz_vector_exp(double mx, int S, int none) {
  constexpr int BETA = oz_beta(KIND);
  if (!(mx > 0.0)) return none;
  int e = 0;
  (void)frexp(mx, &e);  // mx < 2^e
  if (KIND == 2) {
    int64_t H = 0;
    for (int j = 0; j < S - 1; ++j) H += ((int64_t)1 << (BETA - 1)) << (BETA * j);
    if ((int64_t)rint(ldexp(mx, BETA * S - 1 - e)) + H < ((int64_t)1 << (BETA * S - 1))) e -= 1;
  }
  return e;
}

template <int KIND>
__global__ void __launch_bounds__(64) k_slice(SliceArgs s, int S, const int64_t* __restrict__ sl_off,
                                              const int64_t* __restrict__ ex_off, int64_t ngroups,
                                              uint8_t* __restrict__ slices, int32_t* __restrict__ exps, int mode,
                                              uint8_t* __restrict__ zg) {
  constexpr int BETA = oz_beta(KIND), TS = oz_tshift(KIND);
  constexpr int KS = oz_kstep(KIND), ES = oz_esize(KIND), EL = 16 / ES;
  __shared__ double T[16][KS + 1];  // one k-step tile; the pad keeps the 16 rows on distinct banks
  const int lane = threadIdx.x;
  const int64_t g = blockIdx.x;
  if (g >= ngroups) return;
  // block of group g: the last b with ex_off[b] <= 16 g (absent blocks have no
  // vectors, so equal offsets end at the present block)
  const int64_t key = 16 * g;
  int64_t lo = 0, hi = s.nb;
  while (hi - lo > 1) {
    const int64_t mid = (lo + hi) >> 1;
    if (ex_off[mid] <= key) lo = mid;
    else hi = mid;
  }
  const int64_t b = lo;
  int nv, kd;
  int64_t sv, sk;
  slice_geom(s, b, nv, kd, sv, sk);
  const int nks = (int)ceil_div(kd, KS), kp = nks * KS;
  const int nvp = (int)ceil_div(nv, 16) * 16;
  const int64_t pstride = (int64_t)nvp * kp * ES;  // bytes per slice
  const int vb = (int)((key - ex_off[b]) >> 4), v0 = vb * 16;
  const double* xb = s.arena + s.off[b];
  uint8_t* out = slices + sl_off[b] + (int64_t)vb * nks * 1024 + lane * 16;
  const int vl = lane & 15, kg = lane >> 4;

  // one k-step tile T[i][kl] = x(vector v0 + i, k = KS ks + kl), zero outside
  // the block: each lane loads NV values into registers (load) and stores them
  // to T (put), so the next tile's loads can be in flight while a tile is cut
  constexpr int NV = KS / 4;
  auto load = [&](int ks, double (&r)[NV]) {
    const int k0 = ks * KS;
    if (sk == 1) {  // k contiguous: each load covers 64 / KS vectors x KS k
      constexpr int VPI = 64 / KS;
      const int kl = lane % KS, vo = lane / KS;
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int i = j * VPI + vo, v = v0 + i, kk = k0 + kl;
        r[j] = (v < nv && kk < kd) ? xb[(int64_t)v * sv + kk] : 0.0;
      }
    } else {  // vectors contiguous (or neither): each load covers 16 vectors x 4 k
#pragma unroll
      for (int j = 0; j < NV; ++j) {
        const int kl = j * 4 + kg, v = v0 + vl, kk = k0 + kl;
        r[j] = (v < nv && kk < kd) ? xb[(int64_t)v * sv + (int64_t)kk * sk] : 0.0;
      }
    }
  };
  auto put = [&](const double (&r)[NV]) {
    if (sk == 1) {
      constexpr int VPI = 64 / KS;
      const int kl = lane % KS, vo = lane / KS;
#pragma unroll
      for (int j = 0; j < NV; ++j) T[j * VPI + vo][kl] = r[j];
    } else {
#pragma unroll
      for (int j = 0; j < NV; ++j) T[vl][j * 4 + kg] = r[j];
    }
  };
  double cur[NV], nxt[NV];

  const int64_t mask = ((int64_t)1 << BETA) - 1, half = (int64_t)1 << (BETA - 1);
  int64_t H = 0;
  for (int j = 0; j < S - 1; ++j) H += half << (BETA * j);
  int e = 0;
  if (mode == 2) {
    e = exps[ex_off[b] + v0 + vl];  // >= this block's own exponent: digits stay in range
  } else {
    double mx = 0.0;
    load(0, cur);
    for (int ks = 0; ks < nks; ++ks) {
      put(cur);
      __syncthreads();
      if (ks + 1 < nks) load(ks + 1, nxt);
#pragma unroll
      for (int t = 0; t < EL; ++t) mx = fmax(mx, fabs(T[vl][kg * EL + t]));
      __syncthreads();
#pragma unroll
      for (int j = 0; j < NV; ++j) cur[j] = nxt[j];
    }
    mx = fmax(mx, __shfl_xor(mx, 16, 64));
    mx = fmax(mx, __shfl_xor(mx, 32, 64));
    // a zero vector does not raise its line's exponent (mode 1)
    e = oz_vector_exp<KIND>(mx, S, mode == 1 ? kNoExp : 0);
    if (kg == 0) exps[ex_off[b] + v0 + vl] = e;
    if (mode == 1) return;
  }

  uint32_t nz = 0;  // slices in which this lane cut a nonzero digit
  load(0, cur);
  for (int ks = 0; ks < nks; ++ks) {
    put(cur);
    __syncthreads();
    if (ks + 1 < nks) load(ks + 1, nxt);
    int64_t mv[EL];
#pragma unroll
    for (int t = 0; t < EL; ++t) mv[t] = (int64_t)rint(ldexp(T[vl][kg * EL + t], BETA * S - TS - e)) + H;
    __syncthreads();
#pragma unroll
    for (int j = 0; j < NV; ++j) cur[j] = nxt[j];
    for (int p = 0; p < S; ++p) {
      const int sh = BETA * (S - 1 - p);
      uint32_t w[4] = {0u, 0u, 0u, 0u};
#pragma unroll
      for (int t = 0; t < EL; ++t) {
        const int64_t d = p == 0 ? (mv[t] >> sh) : ((mv[t] >> sh) & mask) - half;
        if (KIND == 2) {
          w[t >> 2] |= ((uint32_t)d & 0xffu) << (8 * (t & 3));
        } else {
          const uint16_t bits = KIND == 1 ? __builtin_bit_cast(uint16_t, (__bf16)(float)d)
                                          : __builtin_bit_cast(uint16_t, (_Float16)(float)d);
          w[t >> 1] |= (uint32_t)bits << (16 * (t & 1));
        }
      }
      *reinterpret_cast<uint4*>(out + (int64_t)p * pstride + (int64_t)ks * 1024) = make_uint4(w[0], w[1], w[2], w[3]);
      if (w[0] | w[1] | w[2] | w[3]) nz |= 1u << p;
    }
  }
  // the group's leading all-zero slices (S: all zero)
#pragma unroll
  for (int o = 1; o < 64; o <<= 1) nz |= __shfl_xor(nz, o, 64);
  if (lane == 0) zg[g] = (uint8_t)(nz ? min((int)__builtin_ctz(nz), S) : S);
}

// One pass over every block of a matrix M whose slicing view V (same blocks,
// same memory, another block geometry: view block b holds M's block b) takes
// V's columns as op(A) rows of a transposed operand, i.e. vectors of
// k-contiguous elements: element e of a block (flat, column-major) is k = e %
// L of vector e / L, L = V's block rows.  One wave per block computes
//  * M's canonical block norm (the order of wave_block_norm: lane l sums
//    fma(x, x) over e = l, l + 64, ..., then the halving tree), and
//  * each vector's largest magnitude, and from it the slicer's exponent
//    (oz_vector_exp), written where k_slice mode 2 reads it,
// so the slicer's own max pass over the operand is not needed.  Padding
// vectors (up to a multiple of 16) get the zero-vector exponent.
template <int KIND>
__global__ void __launch_bounds__(64) k_norms_view_exps(const double* __restrict__ arena,
                                                        const int64_t* __restrict__ off, const int32_t* __restrict__ mrsz,
                                                        const int32_t* __restrict__ mcsz, int32_t mnrb,
                                                        const int32_t* __restrict__ vrsz,
                                                        const int32_t* __restrict__ vcsz, int32_t vnrb, int64_t nb,
                                                        int S, int none, const int64_t* __restrict__ ex_off,
                                                        double* __restrict__ norm, int32_t* __restrict__ exps) {
  __shared__ unsigned long long colmax[128];  // |x| bits: non-negative doubles order as integers
  const int lane = threadIdx.x;
  const int64_t b = blockIdx.x;
  if (b >= nb) return;
  const int64_t o = off[b];
  if (o < 0) {
    if (lane == 0) norm[b] = 0.0;
    return;
  }
  const int64_t n = (int64_t)mrsz[b % mnrb] * mcsz[b / mnrb];
  const int L = vrsz[b % vnrb], nv = vcsz[b / vnrb];
  const int nvp = (int)ceil_div(nv, 16) * 16;
  for (int v = lane; v < 128; v += 64) colmax[v] = 0ull;
  __syncthreads();
  const double* x = arena + o;
  double a = 0.0, mx = 0.0;
  int v = lane / L;                    // vector of element lane
  int64_t next = (int64_t)(v + 1) * L;  // first element of the following vector
  auto take = [&](int64_t e, double y) {
    a = fma(y, y, a);
    if (e >= next) {  // crossed into another vector: flush the running max
      atomicMax(&colmax[v], (unsigned long long)__double_as_longlong(mx));
      v = (int)(e / L);
      next = (int64_t)(v + 1) * L;
      mx = 0.0;
    }
    mx = fmax(mx, fabs(y));
  };
  // loads 8 ahead of the fma chain, whose order stays e = l, l + 64, ...
  int64_t e = lane;
  for (; e + 7 * 64 < n; e += 8 * 64) {
    double y[8];
#pragma unroll
    for (int t = 0; t < 8; ++t) y[t] = x[e + 64 * t];
#pragma unroll
    for (int t = 0; t < 8; ++t) take(e + 64 * t, y[t]);
  }
  for (; e < n; e += 64) take(e, x[e]);
  if (lane < n) atomicMax(&colmax[v], (unsigned long long)__double_as_longlong(mx));
#pragma unroll
  for (int s = 32; s >= 1; s >>= 1) a = a + __shfl_down(a, s, 64);
  if (lane == 0) norm[b] = sqrt(a);
  __syncthreads();
  for (int c = lane; c < nvp; c += 64) {
    const double m = c < nv ? __longlong_as_double((long long)colmax[c]) : 0.0;
    exps[ex_off[b] + c] = oz_vector_exp<KIND>(m, S, none);
  }
}

// Shared exponents: every vector of a line (an op(A) row across all its
// blocks of one block row, an op(B) column across all blocks of one block
// column) gets the largest of its blocks' exponents, so all products of an
// output block scale each row and column alike and their digit products can
// share the int32 accumulators.  One workgroup per line.
__global__ void __launch_bounds__(256) k_exp_share(SliceArgs s, const int64_t* __restrict__ ex_off,
                                                   int32_t* __restrict__ exps) {
  const int line = blockIdx.x;
  const int32_t nrb = s.nrb, ncb = (int32_t)(s.nb / s.nrb);
  // lines run along the op-vector index: op(A) rows = A rows (block row) unless
  // transposed; op(B) columns = B columns (block column) unless transposed
  const bool by_row = s.role_b ? (s.trans != 0) : (s.trans == 0);
  const int32_t nblk = by_row ? ncb : nrb;
  const int nv = by_row ? s.rsz[line] : s.csz[line];
  const int nvp = (int)ceil_div(nv, 16) * 16;
  auto blk = [&](int32_t i) { return by_row ? (int64_t)line + (int64_t)i * nrb : (int64_t)i + (int64_t)line * nrb; };
  for (int v = threadIdx.x; v < nvp; v += blockDim.x) {
    int e = kNoExp;
    for (int32_t i = 0; i < nblk; ++i) {
      const int64_t b = blk(i);
      if (s.off[b] >= 0) e = max(e, exps[ex_off[b] + v]);
    }
    if (e == kNoExp) e = 0;
    for (int32_t i = 0; i < nblk; ++i) {
      const int64_t b = blk(i);
      if (s.off[b] >= 0) exps[ex_off[b] + v] = e;
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
  const uint8_t* asl;
  const int64_t* asl_off;
  const int32_t* aex;
  const int64_t* aex_off;
  const uint8_t* bsl;
  const int64_t* bsl_off;
  const int32_t* bex;
  const int64_t* bex_off;
  const uint8_t* azg;    // leading all-zero slices of each 16-vector group of op(A)'s slicing
  int32_t zskip;         // 0: record no zero slice (no pair is skipped)
  int32_t zmin;          // zero-slice path from this many sixteenths of zero slice-pair products
  int32_t slices;
  int32_t tile_groups;   // 16-row groups per tile row of the GEMM kernel
  int32_t tile_n;        // columns per tile of the GEMM kernel
  unsigned long long* pairs;  // += slice-pair k-steps x tiles: [0] multiplied, [1] skipped as zero
  int32_t kstep;
  int32_t shared;  // operands sliced with shared exponents: the products of an output chain
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
  o.group = (int32_t)(out_slot(ib, jb, t.nib) >> 2) + 1;  // group column (4 C blocks of one column), + 1
  o.pad_ = 0;
  const int S = t.slices;
  const int G = t.tile_groups;                  // 16-row groups per tile row
  const int ntr = (int)ceil_div(ceil_div(o.m, 16), G);
  int64_t w_all = 0;                            // slice-pair k-steps of a tile row
  for (int64_t p = seg; p < seg + cnt; ++p) {
    const int kb = t.kb[p];
    const int64_t ai = t.ta ? (int64_t)kb + (int64_t)ib * t.a_nrb : (int64_t)ib + (int64_t)kb * t.a_nrb;
    const int64_t bi = t.tb ? (int64_t)jb + (int64_t)kb * t.b_nrb : (int64_t)kb + (int64_t)jb * t.b_nrb;
    OzProd pr;
    pr.a = t.asl + t.asl_off[ai];
    pr.b = t.bsl + t.bsl_off[bi];
    pr.ea = t.aex + t.aex_off[ai];
    pr.eb = t.bex + t.bex_off[bi];
    pr.nks = (int32_t)ceil_div(t.ksz[kb], t.kstep);
    pr.nvba = (int32_t)ceil_div(o.m, 16);
    pr.nvbb = (int32_t)ceil_div(o.n, 16);
    pr.flags = (t.shared && p > seg) ? 1 : 0;
    // the op(A) block's groups: pairs (pa, pb) with pa < z of a group are zero
    // (the sum over pa < z of S - pa)
    const uint8_t* zg = t.azg + t.aex_off[ai] / 16;
    auto zgrp = [&](int g) { return g < pr.nvba ? (t.zskip ? (int)zg[g] : 0) : S; };
    uint64_t lo = 0, hi = 0;
    int z15 = S;
    for (int g = 0; g < 16 || g < pr.nvba; ++g) {
      const uint64_t z = (uint64_t)zgrp(g);
      if (g < 8) lo |= z << (8 * g);
      else if (g < 15) hi |= z << (8 * (g - 8));
      else z15 = min(z15, (int)z);
    }
    hi |= (uint64_t)z15 << 56;
    pr.za0 = (uint32_t)lo;
    pr.za1 = (uint32_t)(lo >> 32);
    pr.za2 = (uint32_t)hi;
    pr.za3 = (uint32_t)(hi >> 32);
    prods[p] = pr;
    w_all += (int64_t)pr.nks * (S * (S + 1) / 2);
  }
  // The kernel's zero-slice path (one branch per A slice) pays for a tile row
  // when enough of its slice-pair products are zero: the slices below the
  // least count of the row's groups (pairs pa < z: the sum over pa < z of
  // S - pa).  Bit 31 covers the rows from 31 on.
  uint32_t hint = 0;
  int64_t skipped = 0;
  for (int tb = 0; tb < min(ntr, 32); ++tb) {
    const int g0 = tb * G, g1 = tb < 31 ? g0 + G : ntr * G;
    int64_t w_zero = 0;
    for (int64_t p = seg; p < seg + cnt; ++p) {
      const int kb = t.kb[p];
      const int64_t ai = t.ta ? (int64_t)kb + (int64_t)ib * t.a_nrb : (int64_t)ib + (int64_t)kb * t.a_nrb;
      const uint8_t* zg = t.azg + t.aex_off[ai] / 16;
      const int nks = (int)ceil_div(t.ksz[kb], t.kstep), ng = (int)ceil_div(o.m, 16);
      for (int g = g0; g < min(g1, ng); g += G) {
        int z = S;
        for (int q = g; q < min(g + G, ng); ++q) z = min(z, t.zskip ? (int)zg[q] : 0);
        w_zero += (int64_t)nks * (z * S - z * (z - 1) / 2);
      }
    }
    const int64_t rows = tb < 31 ? 1 : ntr - 31;
    if (w_zero > 0 && 16 * w_zero >= (int64_t)t.zmin * w_all * rows) {
      hint |= 1u << tb;
      skipped += w_zero;
    }
  }
  o.hint = (int32_t)hint;
  const int64_t tcols = ceil_div(o.n, t.tile_n);
  atomicAdd(t.pairs, (unsigned long long)((w_all * ntr - skipped) * tcols));
  atomicAdd(t.pairs + 1, (unsigned long long)(skipped * tcols));
  outs[t.out_scan[out_slot(ib, jb, t.nib)]] = o;
}

// ----------------------------------------------------------------- GEMM --
// Wave-wide 16-byte load from global memory straight into LDS: lane l's 16
// bytes land at lds + 16 l (lds wave-uniform); completion is counted on vmcnt.
__device__ inline void lds_dma16(const uint4* src, uint4* lds) {
  __builtin_amdgcn_global_load_lds((const void*)src, (__attribute__((address_space(3))) void*)lds, 16, 0, 0);
}

// Calls f(NA, NB) with the wave's live fragment counts as compile-time
// constants: (FM, FN), the partial shapes, or (0, 0) for a wave with no live
// fragment (it still takes its share of the loads and every barrier).
// FM, FN <= 2.
template <int FM, int FN, class F>
__device__ inline void oz_dispatch(int na, int nb, F&& f) {
  static_assert(FM >= 1 && FM <= 2 && FN >= 1 && FN <= 2, "fragment counts per wave");
  using C0 = std::integral_constant<int, 0>;
  using C1 = std::integral_constant<int, 1>;
  using CM = std::integral_constant<int, FM>;
  using CN = std::integral_constant<int, FN>;
  if (na >= FM && nb >= FN) f(CM{}, CN{});
  else if (FN > 1 && na >= FM && nb == 1) f(CM{}, C1{});
  else if (FM > 1 && na == 1 && nb >= FN) f(C1{}, CN{});
  else if (FM > 1 && FN > 1 && na == 1 && nb == 1) f(C1{}, C1{});
  else f(C0{}, C0{});
}

template <int S, int NA, int NB, int PA>
__device__ inline void oz_sched_slice() {
  constexpr int n = (S - PA) * NA * NB;
  if constexpr (PA + 1 < S) {
    __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);      // MFMA
    __builtin_amdgcn_sched_group_barrier(0x100, NA, 0);     // LDS read
    __builtin_amdgcn_sched_group_barrier(0x008, n - 1, 0);  // MFMA
  } else {
    __builtin_amdgcn_sched_group_barrier(0x008, n, 0);
  }
}
template <int S, int NA, int NB, int... PA>
__device__ inline void oz_sched(std::integer_sequence<int, PA...>) {
  __builtin_amdgcn_sched_group_barrier(0x100, S * NB + NA, 0);
  (oz_sched_slice<S, NA, NB, PA>(), ...);
}

// One workgroup of WM x WN waves (default 2 x 2) owns a (16 FM WM) x
// (16 FN WN) output tile; each wave owns (16 FM) x (16 FN).  A stage is one MFMA k-step for all S slices of
// the tile's op(A) rows and op(B) columns, i.e. S (2 FM + 2 FN) fragments of
// 1 KB, double-buffered in LDS and loaded straight from global memory into
// LDS (no register staging).  The stream of stages runs over the products of
// the output; the loads of stage s+1 are in flight while the MFMAs of stage s
// run, with one barrier per stage.  Fragments beyond the output block are
// neither loaded nor multiplied: the whole stage loop is specialised on the
// wave's live fragment counts, so each specialisation issues its MFMAs as one
// straight-line block (a branch per fragment would put every MFMA in its own
// basic block and serialise it behind the LDS reads).
// Zero slices: with exponents shared along a line, a block far below its
// line's largest has leading slices that are all zero.  The slicer records
// the leading zero slices of every 16-vector group; the product record
// carries those of the op(A) block's row groups, and tiles where enough of
// the slice-pair products are zero (3 sixteenths by default) take a path that
// neither loads nor multiplies the slices below the tile rows' least count.
template <int S, int KIND, int FM, int FN, int WM = 2, int WN = 2>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2)))
    k_gemm_ozaki(const GemmOut* __restrict__ outs, const OzProd* __restrict__ prods,
                 const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  constexpr int BETA = oz_beta(KIND), TS = oz_tshift(KIND);
  typedef typename OzAcc<KIND>::type acc4;
  constexpr int NW = WM * WN;                  // waves
  constexpr int RA = WM * FM, RB = WN * FN;    // 16-blocks of rows / columns per tile
  constexpr int NCA = S * RA, NCB = S * RB;    // fragments per stage
  constexpr int LA = (NCA + NW - 1) / NW, LB = (NCB + NW - 1) / NW;
  constexpr int FULL = oz_full_fold(KIND);
  // the two stage buffers are separate objects: the compiler then sees that
  // the LDS reads of one stage do not wait for the loads into the other
  __shared__ uint4 As0[NCA][64], As1[NCA][64];
  __shared__ uint4 Bs0[NCB][64], Bs1[NCB][64];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int wm = wid % WM, wn = wid / WM;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t it = xcd_remap(blockIdx.x, nitems);
  const GemmItem item = items ? gload1(items + it) : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = gload1(outs + item.out);
  const int m = o.m, n = o.n;
  const int nrb = (int)ceil_div(m, 16), ncb = (int)ceil_div(n, 16);
  const int rb0 = item.tm * RA, cb0 = item.tn * RB;
  const int nai = __builtin_amdgcn_readfirstlane(nrb - rb0 - wm * FM);  // live fragment rows (may be <= 0)
  const int naj = __builtin_amdgcn_readfirstlane(ncb - cb0 - wn * FN);

  // leading zero slices of a product's op(A) rows in the tile: the least over
  // its live 16-row groups.  Every wave skips the same slices (a stage lasts
  // as long as its slowest wave)
  auto zero_tile = [&](const OzProd& pr) {
    int z = S;
#pragma unroll
    for (int i = 0; i < RA; ++i)
      if (rb0 + i < nrb) z = min(z, oz_zero_g(pr, rb0 + i));
    return z;
  };

  // stage (product pr, k-step ks) into LDS buffer buf: one 1 KB fragment per
  // wave-wide 16-byte load; fragments beyond the output block are skipped,
  // and on the zero-slice path (ZS) the A slices below zt, the first slice
  // with a nonzero digit in the tile's rows, and the B slices that meet only
  // those
  auto gload = [&](const OzProd& pr, int ks, auto par, auto zs_, int zt) {
    constexpr bool ZS = decltype(zs_)::value;
    uint4(*As)[64] = decltype(par)::value ? As1 : As0;
    uint4(*Bs)[64] = decltype(par)::value ? Bs1 : Bs0;
    const uint4* a4 = reinterpret_cast<const uint4*>(pr.a);
    const uint4* b4 = reinterpret_cast<const uint4*>(pr.b);
    // offsets within one block's slices, in 16-byte units: below 2^31 (the
    // host rejects larger blocks), so the per-stage address arithmetic is 32-bit
    const int32_t psa = pr.nvba * pr.nks * 64, psb = pr.nvbb * pr.nks * 64;
    const int za = ZS ? zt : 0;
    if constexpr (ZS && NW % RA == 0 && NW % RB == 0 && NCA % NW == 0 && NCB % NW == 0) {
      // wave wid's fragments: one 16-row (16-column) group, wid % RA (% RB),
      // and the slices wid / RA + (NW / RA) r: one address per stage plus a
      // fixed slice stride, and one bounds branch per operand (the scalar
      // work per load stays a 64-bit add).  Only on the zero-slice path: with
      // the plain path's loads issued this way (in either operand order, or
      // one per MFMA) the K contraction of the RI-K step ran 2-10 % slower
      const int vba = rb0 + wid % RA, vbb = cb0 + wid % RB;
      if (vba < nrb) {
        const uint4* q = a4 + ((wid / RA) * psa + (vba * pr.nks + ks) * 64);  // wave-uniform
#pragma unroll
        for (int r = 0; r < LA; ++r) {
          const int c = wid + NW * r;
          if (wid / RA + (NW / RA) * r >= za) lds_dma16(q + lane, &As[c][0]);
          q += (NW / RA) * psa;
        }
      }
      if (vbb < ncb) {
        const uint4* q = b4 + ((wid / RB) * psb + (vbb * pr.nks + ks) * 64);
#pragma unroll
        for (int r = 0; r < LB; ++r) {
          const int c = wid + NW * r;
          if (wid / RB + (NW / RB) * r + za < S) lds_dma16(q + lane, &Bs[c][0]);
          q += (NW / RB) * psb;
        }
      }
    } else {
#pragma unroll
      for (int r = 0; r < LA; ++r) {
        const int c = wid + NW * r;
        const int p = c / RA, vb = rb0 + c % RA;
        if (c < NCA && vb < nrb && (!ZS || p >= za))
          lds_dma16(a4 + (p * psa + (vb * pr.nks + ks) * 64) + lane, &As[c][0]);
      }
#pragma unroll
      for (int r = 0; r < LB; ++r) {
        const int c = wid + NW * r;
        const int p = c / RB, vb = cb0 + c % RB;
        if (c < NCB && vb < ncb && (!ZS || p + za < S))
          lds_dma16(b4 + (p * psb + (vb * pr.nks + ks) * 64) + lane, &Bs[c][0]);
      }
    }
  };

  // the stage loop and write-back for a wave with NA x NB live fragments
  auto run = [&](auto na_, auto nb_, auto zs_) {
    constexpr int NA = decltype(na_)::value, NB = decltype(nb_)::value;
    constexpr bool ZS = decltype(zs_)::value;
    constexpr int EA = NA > 0 ? NA : 1, EB = NB > 0 ? NB : 1;  // array extents
    dbl4 acc[EA][EB];
    acc4 s32[S][EA][EB];
#pragma unroll
    for (int i = 0; i < EA; ++i)
#pragma unroll
      for (int j = 0; j < EB; ++j) {
        acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
        for (int w = 0; w < S; ++w) s32[w][i][j] = acc4{0, 0, 0, 0};
      }
    // exponents of the current product (er, ec) and of the next one (ern,
    // ecn: loaded with the next product's first stage, so the load is in
    // flight during a stage of MFMAs instead of in front of a barrier)
    int er[EA][4], ec[EB], ern[EA][4], ecn[EB];
    auto load_exps = [&](const OzProd& pr, int(&r)[EA][4], int(&cc)[EB]) {
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int v = 0; v < 4; ++v) r[i][v] = gload1(pr.ea + (rb0 + wm * FM + i) * 16 + 4 * lg + v);
#pragma unroll
      for (int j = 0; j < NB; ++j) cc[j] = gload1(pr.eb + (cb0 + wn * FN + j) * 16 + li);
    };
    // the MFMAs of one stage.  Plain path: the B slices and the first A slice
    // read up front, then each A slice's MFMAs with the next A slice read
    // right after the first of them, so its LDS latency hides behind the rest.
    // Zero-slice path: every fragment read first, then the MFMAs of each A
    // slice from zt, the first slice with a nonzero digit in the tile's rows
    // of the product's op(A) block, on (one uniform branch per slice).  A skipped pair holds an all-zero slice and adds nothing to the
    // exact integer sums, so both paths give the same bits.
    auto compute = [&](auto par, int zt) {
      const uint4(*As)[64] = decltype(par)::value ? As1 : As0;
      const uint4(*Bs)[64] = decltype(par)::value ? Bs1 : Bs0;
      if constexpr (ZS) {
        const int za = zt;
        uint4 aq[S][EA], bq[S][EB];
#pragma unroll
        for (int q = 0; q < S; ++q) {
#pragma unroll
          for (int j = 0; j < NB; ++j) bq[q][j] = Bs[q * RB + wn * FN + j][lane];
#pragma unroll
          for (int i = 0; i < NA; ++i) aq[q][i] = As[q * RA + wm * FM + i][lane];
        }
#pragma unroll
        for (int pa = 0; pa < S; ++pa) {
          if (pa < za) continue;
#pragma unroll
          for (int pb = 0; pb < S - pa; ++pb)
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
              for (int j = 0; j < NB; ++j)
                s32[pa + pb][i][j] =
                    oz_mfma(std::integral_constant<int, KIND>{}, aq[pa][i], bq[pb][j], s32[pa + pb][i][j]);
        }
      } else {
        uint4 bq[S][EB];
#pragma unroll
        for (int q = 0; q < S; ++q)
#pragma unroll
          for (int j = 0; j < NB; ++j) bq[q][j] = Bs[q * RB + wn * FN + j][lane];
        uint4 aq[S][EA];
#pragma unroll
        for (int i = 0; i < NA; ++i) aq[0][i] = As[wm * FM + i][lane];
#pragma unroll
        for (int pa = 0; pa < S; ++pa) {
          if (pa + 1 < S)
#pragma unroll
            for (int i = 0; i < NA; ++i) aq[pa + 1][i] = As[(pa + 1) * RA + wm * FM + i][lane];
#pragma unroll
          for (int pb = 0; pb < S - pa; ++pb)
#pragma unroll
            for (int i = 0; i < NA; ++i)
#pragma unroll
              for (int j = 0; j < NB; ++j)
                s32[pa + pb][i][j] =
                    oz_mfma(std::integral_constant<int, KIND>{}, aq[pa][i], bq[pb][j], s32[pa + pb][i][j]);
        }
        if constexpr (NA > 0 && NB > 0) oz_sched<S, NA, NB>(std::make_integer_sequence<int, S>{});
      }
    };
    // acc += sum_w s_w 2^(e_r + f_c + 2 TS - beta (w + 2)), s = 0: t = sum_w
    // s_w 2^(-beta w) by Horner steps, one scaling per element
    auto fold = [&]() {
#pragma unroll
      for (int i = 0; i < NA; ++i)
#pragma unroll
        for (int j = 0; j < NB; ++j) {
#pragma unroll
          for (int v = 0; v < 4; ++v) {
            double t = (double)s32[S - 1][i][j][v];
#pragma unroll
            for (int w = S - 2; w >= 0; --w) t = fma(t, pow2i(-BETA), (double)s32[w][i][j][v]);
            acc[i][j][v] += ldexp(t, er[i][v] + ec[j] + 2 * TS - 2 * BETA);
          }
#pragma unroll
          for (int w = 0; w < S; ++w) s32[w][i][j] = acc4{0, 0, 0, 0};
        }
    };
    using I0 = std::integral_constant<int, 0>;
    using I1 = std::integral_constant<int, 1>;

    int p = o.p0;
    if (p < o.p1) {
      OzProd pr = gload1(prods + p);
      // record of the following product, loaded one product ahead
      OzProd pf = p + 1 < o.p1 ? gload1(prods + p + 1) : pr;
      int ks = 0;
      load_exps(pr, er, ec);
      // zero-slice path: the leading zero slices of the tile's rows, for the
      // current and the next product
      int zt = 0, ztn = 0;
      if constexpr (ZS) zt = zero_tile(pr);
      gload(pr, 0, I0{}, zs_, zt);
      __syncthreads();
      int run = 0;  // k-steps in the int32 / FP32 accumulators since the last fold
      // one stage computing from buffer P while the next loads into P ^ 1;
      // true at the end of the stream
      auto step = [&](auto par) -> bool {
        constexpr int P = decltype(par)::value;
        int np = p, nks = ks + 1;
        if (nks >= pr.nks) {
          np = p + 1;
          nks = 0;
        }
        const bool has_next = np < o.p1;
        OzProd pn = pr;
        if (has_next) {
          if (np != p) {
            pn = pf;
            if (np + 1 < o.p1) pf = gload1(prods + np + 1);
            load_exps(pn, ern, ecn);
            if constexpr (ZS) ztn = zero_tile(pn);
          }
          gload(pn, nks, std::integral_constant<int, P ^ 1>{}, zs_, np != p ? ztn : zt);
        }
        compute(par, zt);
        // fold at the end of a product unless the next one chains (same
        // exponents), and after FULL k-steps of one accumulation in any case
        ++run;
        const bool chained = np != p && has_next && oz_chain(pn.flags);
        if ((np != p && !chained) || run == FULL) {
          fold();
          run = 0;
        }
        if (!has_next) return true;
        if (np != p) {
#pragma unroll
          for (int i = 0; i < NA; ++i)
#pragma unroll
            for (int v = 0; v < 4; ++v) er[i][v] = ern[i][v];
#pragma unroll
          for (int j = 0; j < NB; ++j) ec[j] = ecn[j];
          zt = ztn;
        }
        __syncthreads();
        p = np;
        ks = nks;
        pr = pn;
        return false;
      };
      while (true) {
        if (step(I0{})) break;
        if (step(I1{})) break;
      }
    }

#pragma unroll
    for (int i = 0; i < NA; ++i)
#pragma unroll
      for (int j = 0; j < NB; ++j) {
        const int col = (cb0 + wn * FN + j) * 16 + li;
        if (col >= n) continue;
        double* cp = o.c + (int64_t)col * m;
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int row = (rb0 + wm * FM + i) * 16 + 4 * lg + v;
          if (row < m) {
            double x = alpha * acc[i][j][v];
            if (o.acc) x = x + cp[row];
            cp[row] = x;
          }
        }
      }
  };
  // each tile takes one path for all its stages (the record builder's hint:
  // bit t for tile row t, bit 31 for the rows from 31 on)
  if ((o.hint >> min((int)item.tm, 31)) & 1)
    oz_dispatch<FM, FN>(nai, naj, [&](auto a, auto b) { run(a, b, std::true_type{}); });
  else
    oz_dispatch<FM, FN>(nai, naj, [&](auto a, auto b) { run(a, b, std::false_type{}); });
}

// Deep-pipelined variant for large blocks: WM x WN waves (8: two per SIMD) of
// (16 FM) x (16 FN) tiles, and the loads of the next TWO stages in flight while
// a stage computes: stage s computes from LDS buffer s%2, register set (s+1)%2
// holds stage s+1 (in flight since the previous stage) and the loads of stage
// s+2 go to register set s%2.  The larger workgroup tile also cuts the operand
// bytes per flop (128x64: 48 KB per 32-k stage for 2x the work of 64x64).
struct OzCursor {
  const uint4* a;
  const uint4* b;
  int32_t p, ks, nks;
};

template <int S, int KIND, int WM, int WN, int FM, int FN>
__global__ void __launch_bounds__(64 * WM * WN) __attribute__((amdgpu_waves_per_eu(2, 2)))
    k_gemm_ozaki_deep(const GemmOut* __restrict__ outs, const OzProd* __restrict__ prods,
                      const GemmItem* __restrict__ items, int64_t nitems, double alpha) {
  constexpr int BETA = oz_beta(KIND), TS = oz_tshift(KIND);
  typedef typename OzAcc<KIND>::type acc4;
  constexpr int NW = WM * WN;
  constexpr int RA = WM * FM, RB = WN * FN;
  constexpr int NCA = S * RA, NCB = S * RB;
  constexpr int LA = (NCA + NW - 1) / NW, LB = (NCB + NW - 1) / NW;
  // FP16: weight 1 is exact for 2 k-steps of 32, BF16 weights 0..2 for 8, INT8 all for 256 of 64
  constexpr int FOLD_EVERY = oz_full_fold(KIND);
  __shared__ uint4 As[2][NCA][64];
  __shared__ uint4 Bs[2][NCB][64];

  const int tid = threadIdx.x, lane = tid & 63;
  const int wid = __builtin_amdgcn_readfirstlane(tid >> 6);
  const int wm = wid % WM, wn = wid / WM;
  const int li = lane & 15, lg = lane >> 4;
  const int64_t it = xcd_remap(blockIdx.x, nitems);
  const GemmItem item = items ? gload1(items + it) : GemmItem{(int32_t)it, 0, 0};
  const GemmOut o = gload1(outs + item.out);
  const int m = o.m, n = o.n;
  const int nrb = (int)ceil_div(m, 16), ncb = (int)ceil_div(n, 16);
  const int rb0 = item.tm * RA, cb0 = item.tn * RB;
  bool act_i[FM], act_j[FN];
#pragma unroll
  for (int i = 0; i < FM; ++i) act_i[i] = rb0 + wm * FM + i < nrb;
#pragma unroll
  for (int j = 0; j < FN; ++j) act_j[j] = cb0 + wn * FN + j < ncb;

  dbl4 acc[FM][FN];
  acc4 s32[S][FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      acc[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
#pragma unroll
      for (int w = 0; w < S; ++w) s32[w][i][j] = acc4{0, 0, 0, 0};
    }
  int er[FM][4], ec[FN];
  uint4 ra[2][LA], rbf[2][LB];
  using I0 = std::integral_constant<int, 0>;
  using I1 = std::integral_constant<int, 1>;

  auto gload = [&](auto par, const OzCursor& c) {
    constexpr int P = decltype(par)::value;
    const int64_t psa = (int64_t)nrb * c.nks * 64, psb = (int64_t)ncb * c.nks * 64;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int ci = wid + NW * r;
      const int p = ci / RA, vb = rb0 + ci % RA;
      ra[P][r] = (ci < NCA && vb < nrb) ? gload1(c.a + p * psa + ((int64_t)vb * c.nks + c.ks) * 64 + lane)
                                        : make_uint4(0, 0, 0, 0);
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int ci = wid + NW * r;
      const int p = ci / RB, vb = cb0 + ci % RB;
      rbf[P][r] = (ci < NCB && vb < ncb) ? gload1(c.b + p * psb + ((int64_t)vb * c.nks + c.ks) * 64 + lane)
                                         : make_uint4(0, 0, 0, 0);
    }
  };
  auto lstore = [&](auto par) {
    constexpr int P = decltype(par)::value;
#pragma unroll
    for (int r = 0; r < LA; ++r) {
      const int ci = wid + NW * r;
      if (ci < NCA) As[P][ci][lane] = ra[P][r];
    }
#pragma unroll
    for (int r = 0; r < LB; ++r) {
      const int ci = wid + NW * r;
      if (ci < NCB) Bs[P][ci][lane] = rbf[P][r];
    }
  };
  // exponents of product p (record re-read: cursors carry only the slice pointers)
  auto load_exps = [&](int p) {
    const int32_t* ea = gload1(&prods[p].ea);
    const int32_t* eb = gload1(&prods[p].eb);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int v = 0; v < 4; ++v) er[i][v] = act_i[i] ? gload1(ea + (rb0 + wm * FM + i) * 16 + 4 * lg + v) : 0;
#pragma unroll
    for (int j = 0; j < FN; ++j) ec[j] = act_j[j] ? gload1(eb + (cb0 + wn * FN + j) * 16 + li) : 0;
  };
  auto compute = [&](int buf) {
    uint4 bq[S][FN];
#pragma unroll
    for (int q = 0; q < S; ++q)
#pragma unroll
      for (int j = 0; j < FN; ++j) bq[q][j] = Bs[buf][q * RB + wn * FN + j][lane];
#pragma unroll
    for (int pa = 0; pa < S; ++pa) {
      uint4 aq[FM];
#pragma unroll
      for (int i = 0; i < FM; ++i) aq[i] = As[buf][pa * RA + wm * FM + i][lane];
#pragma unroll
      for (int pb = 0; pb < S - pa; ++pb)
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) {
            s32[pa + pb][i][j] = oz_mfma(std::integral_constant<int, KIND>{}, aq[i], bq[pb][j], s32[pa + pb][i][j]);
          }
    }
  };
  // acc += sum_w s_w 2^(e_r + f_c + 2 TS - beta (w + 2)); s = 0.  All weights
  // at every fold point (a compile-time range: the two parity copies of the
  // stage body would otherwise index the weights at run time).
  auto fold = [&]() {
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) {
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          double t = 0.0;
#pragma unroll
          for (int w = 0; w < S; ++w) t = fma((double)s32[w][i][j][v], pow2i(-BETA * w), t);
          acc[i][j][v] += ldexp(t, er[i][v] + ec[j] + 2 * TS - 2 * BETA);
        }
#pragma unroll
        for (int w = 0; w < S; ++w) s32[w][i][j] = acc4{0, 0, 0, 0};
      }
  };

  const int p0 = o.p0, p1 = o.p1;
  if (p0 < p1) {
    auto mk = [&](int p) {
      return OzCursor{reinterpret_cast<const uint4*>(gload1(&prods[p].a)),
                      reinterpret_cast<const uint4*>(gload1(&prods[p].b)), p, 0, gload1(&prods[p].nks)};
    };
    OzCursor pf = mk(p0 + 1 < p1 ? p0 + 1 : p0);  // product after the furthest cursor's
    // next stage after c (the furthest cursor); false at the end of the stream
    auto adv = [&](const OzCursor& c, OzCursor& nx) -> bool {
      if (c.ks + 1 < c.nks) {
        nx = c;
        nx.ks = c.ks + 1;
        return true;
      }
      if (c.p + 1 >= p1) return false;
      nx = pf;
      if (c.p + 2 < p1) pf = mk(c.p + 2);
      return true;
    };
    OzCursor cur = mk(p0), c1 = cur;
    bool v1 = adv(cur, c1);
    load_exps(p0);
    gload(I0{}, cur);
    lstore(I0{});
    if (v1) gload(I1{}, c1);
    __syncthreads();
    // stage with parity P: LDS buffer P holds cur, register set P^1 holds c1
    auto step = [&](auto par) -> bool {
      constexpr int P = decltype(par)::value;
      OzCursor c2 = c1;
      bool v2 = false;
      if (v1) {
        v2 = adv(c1, c2);
        if (v2) gload(par, c2);
      }
      compute(P);
      const int done = cur.ks + 1;
      if (done == cur.nks || (done & (FOLD_EVERY - 1)) == 0) fold();
      if (!v1) return true;
      lstore(std::integral_constant<int, P ^ 1>{});
      __syncthreads();
      if (c1.p != cur.p) load_exps(c1.p);
      cur = c1;
      c1 = c2;
      v1 = v2;
      return false;
    };
    while (true) {
      if (step(I0{})) break;
      if (step(I1{})) break;
    }
  }

#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      if (!(act_i[i] && act_j[j])) continue;
      const int col = (cb0 + wn * FN + j) * 16 + li;
      if (col >= n) continue;
      double* cp = o.c + (int64_t)col * m;
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int row = (rb0 + wm * FM + i) * 16 + 4 * lg + v;
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
  int S;
  int kind;
  int fm, fn;    // workgroup tile (32 fm) x (32 fn)
  bool large;    // only for C blocks of at least 128 on both sides (one workgroup per CU)
  double w;      // relative cost per useful flop, for the choice below
  int threads;
  OzFn fn_;
};

#define SA_OZ(S, K, FM, FN) OzCfg{S, K, FM, FN, false, FM == 2 ? 1.0 : 1.6, 256, k_gemm_ozaki<S, K, FM, FN>}
// wide: (RM x RN) waves of 32 x 32 with LDS-direct staging, tile (32 RM) x (32 RN)
#define SA_OZW(S, K, RM, RN, W) OzCfg{S, K, RM, RN, true, W, 64 * RM * RN, k_gemm_ozaki<S, K, 2, 2, RM, RN>}
// deep: (WM x WN) = (RM x RN) waves of 32 x 32, workgroup tile (32 RM) x (32 RN)
#define SA_OZD(S, K, RM, RN) OzCfg{S, K, RM, RN, true, 1.15, 64 * RM * RN, k_gemm_ozaki_deep<S, K, RM, RN, 2, 2>}

static const OzCfg kOzCfgs[] = {
    SA_OZ(3, 0, 2, 2),  SA_OZ(4, 0, 2, 2),  SA_OZ(5, 0, 2, 2),   // 0-2: FP16 64x64
    SA_OZ(3, 0, 1, 1),  SA_OZ(4, 0, 1, 1),  SA_OZ(5, 0, 1, 1),   // 3-5: FP16 32x32
    SA_OZ(4, 1, 2, 2),  SA_OZ(5, 1, 2, 2),  SA_OZ(6, 1, 2, 2),   // 6-8: BF16 64x64
    SA_OZ(4, 1, 1, 1),  SA_OZ(5, 1, 1, 1),  SA_OZ(6, 1, 1, 1),   // 9-11: BF16 32x32
    SA_OZD(3, 0, 4, 2), SA_OZD(4, 0, 4, 2), SA_OZD(5, 0, 4, 2),  // 12-14: FP16 128x64, deep
    SA_OZD(3, 0, 2, 4), SA_OZD(4, 0, 2, 4), SA_OZD(5, 0, 2, 4),  // 15-17: FP16 64x128, deep
    SA_OZD(4, 1, 4, 2), SA_OZD(5, 1, 4, 2),                      // 18-19: BF16 128x64, deep
    SA_OZD(4, 1, 2, 4), SA_OZD(5, 1, 2, 4),                      // 20-21: BF16 64x128, deep
    SA_OZ(4, 2, 2, 2),  SA_OZ(5, 2, 2, 2),  SA_OZ(6, 2, 2, 2),   // 22-24: INT8 64x64
    SA_OZ(4, 2, 1, 1),  SA_OZ(5, 2, 1, 1),  SA_OZ(6, 2, 1, 1),   // 25-27: INT8 32x32
    SA_OZD(4, 2, 4, 2), SA_OZD(5, 2, 4, 2),                      // 28-29: INT8 128x64, deep
    SA_OZD(4, 2, 2, 4), SA_OZD(5, 2, 2, 4),                      // 30-31: INT8 64x128, deep
    SA_OZW(5, 2, 2, 4, 1.2), SA_OZW(5, 2, 4, 2, 1.2),            // 32-33: INT8 64x128 / 128x64, wide
    SA_OZW(4, 2, 2, 4, 1.2), SA_OZW(4, 2, 4, 2, 1.2),            // 34-35: INT8 S4 64x128 / 128x64, wide
};

// slice one operand (all present blocks); returns device pointers
struct Sliced {
  const uint8_t* sl;
  const int64_t* sl_off;
  const int32_t* ex;
  const int64_t* ex_off;
  const uint8_t* zg;  // per 16-vector group (numbered as the slicer's groups): leading all-zero slices
};

// slice sizes and offsets of every block, buffers reserved (one read-back of
// the totals); the exponents are not written yet
static Sliced slice_plan(sa_context* c, const SliceArgs& a, int S, int kind, const char* tag, int64_t* nex_out) {
  const int64_t nb = a.nb;
  std::string t(tag);
  DevBuf& bso = c->buf((t + "_sloff").c_str());
  DevBuf& beo = c->buf((t + "_exoff").c_str());
  DevBuf& btot = c->buf((t + "_tot").c_str());
  bso.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  beo.reserve(sizeof(int64_t) * (size_t)(nb + 1));
  btot.reserve(sizeof(int64_t) * 2);
  hipLaunchKernelGGL(k_slice_need, dim3((unsigned)ceil_div(nb, 256)), dim3(256), 0, c->stream, a, S, oz_kstep(kind),
                     oz_esize(kind), bso.as<int64_t>(), beo.as<int64_t>());
  scan_exclusive_i64_async(c, bso.as<int64_t>(), nb, btot.as<int64_t>());
  scan_exclusive_i64_async(c, beo.as<int64_t>(), nb, btot.as<int64_t>() + 1);
  int64_t* h = static_cast<int64_t*>(pinned(c, 2 * sizeof(int64_t)));
  SA_CHECK_HIP(hipMemcpyAsync(h, btot.p, 2 * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  SA_CHECK_HIP(hipStreamSynchronize(c->stream));
  const int64_t nsl = h[0], nex = h[1];
  DevBuf& bsl = c->buf((t + "_sl").c_str());
  DevBuf& bex = c->buf((t + "_ex").c_str());
  bsl.reserve((size_t)(nsl + 64));
  bex.reserve(sizeof(int32_t) * (size_t)(nex + 1));
  DevBuf& bzg = c->buf((t + "_zg").c_str());
  bzg.reserve((size_t)(nex / 16 + 16));
  *nex_out = nex;
  return Sliced{bsl.as<uint8_t>(), bso.as<int64_t>(), bex.as<int32_t>(), beo.as<int64_t>(), bzg.as<uint8_t>()};
}

// cut the digits (and the exponents first, unless exps_ready: then they are
// in place per block, or per line before sharing when shared)
static void slice_run(sa_context* c, const SliceArgs& a, int32_t nrb, int32_t ncb, int S, int kind, bool shared,
                      const Sliced& sl, int64_t nex, bool exps_ready) {
  ProfScope ps(c, "ozaki_slice");
  if (a.nb == 0) return;
  const int64_t ng = nex / 16;  // 16-vector groups (nex: padded vectors of all blocks)
  const dim3 g((unsigned)std::max<int64_t>(1, ng)), t(64);
  uint8_t* slp = const_cast<uint8_t*>(sl.sl);
  int32_t* exp = const_cast<int32_t*>(sl.ex);
  auto* mk = const_cast<uint8_t*>(sl.zg);
  auto run = [&](int mode) {
    if (kind == 2)
      hipLaunchKernelGGL(k_slice<2>, g, t, 0, c->stream, a, S, sl.sl_off, sl.ex_off, ng, slp, exp, mode, mk);
    else if (kind == 1)
      hipLaunchKernelGGL(k_slice<1>, g, t, 0, c->stream, a, S, sl.sl_off, sl.ex_off, ng, slp, exp, mode, mk);
    else
      hipLaunchKernelGGL(k_slice<0>, g, t, 0, c->stream, a, S, sl.sl_off, sl.ex_off, ng, slp, exp, mode, mk);
  };
  if (shared) {
    if (!exps_ready) run(1);
    const bool by_row = a.role_b ? a.trans : !a.trans;
    hipLaunchKernelGGL(k_exp_share, dim3((unsigned)(by_row ? nrb : ncb)), dim3(256), 0, c->stream, a, sl.ex_off, exp);
    run(2);
  } else {
    run(exps_ready ? 2 : 0);
  }
  SA_CHECK_HIP(hipGetLastError());
}

static SliceArgs slice_args(sa_bsm* X, bool role_b, bool trans) {
  return SliceArgs{X->d_off, X->data(), X->d_rsz, X->d_csz, X->nrb, X->nb(), role_b ? 1 : 0, trans ? 1 : 0};
}

// shared: exponents per op-vector line across blocks (k_exp_share), else per block
static Sliced slice_operand(sa_context* c, sa_bsm* X, bool role_b, bool trans, int S, int kind, bool shared,
                            const char* tag) {
  const SliceArgs a = slice_args(X, role_b, trans);
  int64_t nex = 0;
  const Sliced sl = slice_plan(c, a, S, kind, tag, &nex);
  slice_run(c, a, X->nrb, X->ncb, S, kind, shared, sl, nex, false);
  return sl;
}

struct OzPrep {
  const sa_bsm* V = nullptr;  // the view the operand slices
  int S = 0, kind = 0;
  bool shared = false;
  Sliced sl{};
  int64_t nex = 0;
};

OzPrepHandle ozaki_prepare_view(sa_bsm* M, sa_bsm* V, const sa_opts_t& o) {
  sa_context* c = M->ctx;
  SA_REQUIRE(V->nb() == M->nb(), "ozaki: the view must hold the matrix's blocks");
  int32_t mx = 0;
  for (int32_t v : V->csz) mx = std::max(mx, v);
  SA_REQUIRE(mx <= 128, "ozaki: view blocks of at most 128 columns (largest %d)", mx);
  auto* p = new OzPrep();
  p->V = V;
  p->S = o.slices;
  p->kind = o.slice_type;
  p->shared = o.shared_exp != 0;
  const SliceArgs a = slice_args(V, false, true);
  p->sl = slice_plan(c, a, p->S, p->kind, "oz_a", &p->nex);
  const int none = p->shared ? kNoExp : 0;
  const int64_t nb = M->nb();
  ProfScope ps(c, "norms");
  if (nb > 0) {
    double* norm = M->d_norm;
    int32_t* ex = const_cast<int32_t*>(p->sl.ex);
    const double* arena = M->data();
    if (p->kind == 2)
      hipLaunchKernelGGL(k_norms_view_exps<2>, dim3((unsigned)nb), dim3(64), 0, c->stream, arena,
                         (const int64_t*)M->d_off, (const int32_t*)M->d_rsz, (const int32_t*)M->d_csz, M->nrb,
                         (const int32_t*)V->d_rsz, (const int32_t*)V->d_csz, V->nrb, nb, p->S, none, p->sl.ex_off,
                         norm, ex);
    else if (p->kind == 1)
      hipLaunchKernelGGL(k_norms_view_exps<1>, dim3((unsigned)nb), dim3(64), 0, c->stream, arena,
                         (const int64_t*)M->d_off, (const int32_t*)M->d_rsz, (const int32_t*)M->d_csz, M->nrb,
                         (const int32_t*)V->d_rsz, (const int32_t*)V->d_csz, V->nrb, nb, p->S, none, p->sl.ex_off,
                         norm, ex);
    else
      hipLaunchKernelGGL(k_norms_view_exps<0>, dim3((unsigned)nb), dim3(64), 0, c->stream, arena,
                         (const int64_t*)M->d_off, (const int32_t*)M->d_rsz, (const int32_t*)M->d_csz, M->nrb,
                         (const int32_t*)V->d_rsz, (const int32_t*)V->d_csz, V->nrb, nb, p->S, none, p->sl.ex_off,
                         norm, ex);
    SA_CHECK_HIP(hipGetLastError());
  }
  return OzPrepHandle(p);
}

void OzPrepDelete::operator()(OzPrep* p) const { delete p; }

static constexpr int kNumOzCfgs = (int)(sizeof(kOzCfgs) / sizeof(kOzCfgs[0]));

// Padded-tile work of one tile size over the C block grid (separable in rows
// and columns), relative to the true block area.
static double padded_ratio(const std::vector<int32_t>& rsz, const std::vector<int32_t>& csz, int TM, int TN) {
  double r = 0, rp = 0, c = 0, cp = 0;
  for (int32_t v : rsz) { r += v; rp += (double)ceil_div(v, TM) * TM; }
  for (int32_t v : csz) { c += v; cp += (double)ceil_div(v, TN) * TN; }
  return (r > 0 && c > 0) ? (rp / r) * (cp / c) : 1.0;
}

// Kernel choice: a forced table id that runs this slice count/type, else the
// lowest padded work weighted by the relative cost per useful flop of each
// kernel shape (64x64 < deep 128-wide tiles < 32x32).  The 64x64 kernel
// (LDS-direct staging, dead fragments skipped, two independent workgroups per
// CU) beats the register-staged deep kernels at equal padding (RI-K: 15 % on
// the transform's 176 x 5263 blocks, 10 % on the K contraction's ~105-wide
// ones); the deep kernels remain for C blocks of at least 128 on both sides
// where their wider tile pads less.
static int pick_cfg(int S, int kind, const sa_bsm* C, int tile) {
  if (tile >= 0 && tile < kNumOzCfgs && kOzCfgs[tile].S == S && kOzCfgs[tile].kind == kind) return tile;
  auto mean = [](const std::vector<int32_t>& v) {
    double t = 0;
    for (int32_t x : v) t += x;
    return v.empty() ? 0.0 : t / (double)v.size();
  };
  const bool large = std::min(mean(C->rsz), mean(C->csz)) >= 128.0;
  int best = -1;
  double best_cost = 1e300;
  for (int i = 0; i < kNumOzCfgs; ++i) {
    const OzCfg& k = kOzCfgs[i];
    if (k.S != S || k.kind != kind) continue;
    if (k.large && !large) continue;
    const double cost = k.w * padded_ratio(C->rsz, C->csz, 32 * k.fm, 32 * k.fn);
    if (cost < best_cost) {
      best_cost = cost;
      best = i;
    }
  }
  return best;
}

// Skip the slice pairs of leading all-zero slices (env SA_OZ_ZERO_SKIP=0:
// multiply every pair; the results are the same bits either way)
static bool zero_skip() {
  const char* v = std::getenv("SA_OZ_ZERO_SKIP");
  return !(v && std::atoi(v) == 0);
}

// The zero-slice path from this many sixteenths of zero slice-pair products
// of an output on (env SA_OZ_ZERO_MIN; 0: whenever any pair is zero).  Its
// branches and up-front LDS reads cost about what 2-3 sixteenths of the MFMAs
// do (RI-K step: 3 sixteenths 720 ms, 1 727 ms, 6 763 ms, never 771 ms).
static int zero_min() {
  const char* v = std::getenv("SA_OZ_ZERO_MIN");
  return v ? std::max(0, std::atoi(v)) : 3;
}

// The split-precision GEMM addresses a block's slices with 32-bit offsets in
// 16-byte units: S slices of ceil(rows / 16) x ceil(k / kstep) fragments of 64
// units, in either orientation, must stay below 2^31 (blocks up to ~80,000 on
// a side at 5 INT8 slices).
static void require_slice_offsets_32bit(const sa_bsm* X, int S, int kind) {
  int64_t mr = 0, mc = 0;
  for (int32_t r : X->rsz) mr = std::max<int64_t>(mr, r);
  for (int32_t cc : X->csz) mc = std::max<int64_t>(mc, cc);
  const int64_t ks = oz_kstep(kind);
  const int64_t frags = std::max(ceil_div(mr, 16) * ceil_div(mc, ks), ceil_div(mc, 16) * ceil_div(mr, ks));
  if ((int64_t)S * frags * 64 >= ((int64_t)1 << 31))
    SA_UNSUPPORTED("ozaki: blocks of up to %lld x %lld are too large for split precision", (long long)mr,
                   (long long)mc);
}

int64_t ozaki_multiply(sa_bsm* C, sa_bsm* A, bool ta, sa_bsm* B, bool tb, double thresh, double alpha,
                       const sa_opts_t& o, bool sym, const OzPrep* prep) {
  sa_context* c = C->ctx;
  const int kind = o.slice_type;
  SA_REQUIRE(kind == SA_SLICE_FP16 || kind == SA_SLICE_BF16 || kind == SA_SLICE_INT8, "ozaki: unknown slice type %d",
             kind);
  const int S = o.slices;
  require_slice_offsets_32bit(A, S, kind);
  require_slice_offsets_32bit(B, S, kind);
  const int cfg = pick_cfg(S, kind, C, -1);
  SA_REQUIRE(cfg >= 0, "ozaki: %d %s slices not supported (FP16: 3..5, BF16: 4..6, INT8: 4..6)", S,
             kind == 2 ? "INT8" : kind == 1 ? "BF16" : "FP16");
  TableMem mem(c);
  const ScreenResult R = screen_and_allocate(C, A, ta, B, tb, thresh, nullptr, nullptr, 0, false, mem, sym);
  if (R.ntasks == 0) return 0;
  // slice operands; share when both roles read the same vectors of one matrix
  const sa_bsm* ra = A->base ? A->base : A;
  const sa_bsm* rb = B->base ? B->base : B;
  const bool shared = o.shared_exp != 0;
  // an operand prepared with ozaki_prepare_view (its exponents computed in the
  // pass over the matrix that also took its block norms) only gets its digits
  Sliced sa_;
  if (prep && prep->V == A && ta && prep->S == S && prep->kind == kind && prep->shared == shared) {
    sa_ = prep->sl;
    slice_run(c, slice_args(A, false, true), A->nrb, A->ncb, S, kind, shared, sa_, prep->nex, true);
  } else {
    sa_ = slice_operand(c, A, false, ta, S, kind, shared, "oz_a");
  }
  const Sliced sb_ =
      (ra == rb && ta != tb && A == B) ? sa_ : slice_operand(c, B, true, tb, S, kind, shared, "oz_b");

  DevBuf& bouts = c->buf("mm_outs");
  DevBuf& bprods = c->buf("oz_prods");
  bouts.reserve(sizeof(GemmOut) * (size_t)(R.nouts + 1));
  bprods.reserve(sizeof(OzProd) * (size_t)(R.ntasks + 1));
  OzTables t;
  t.cnt = R.cnt; t.seg = R.seg; t.need_scan = R.need_scan; t.out_scan = R.out_scan; t.kb = R.kb;
  t.coff = C->d_off; t.cnorm = C->d_norm; t.carena = C->data(); t.crsz = C->d_rsz; t.ccsz = C->d_csz;
  t.nib = R.nib; t.njb = R.njb; t.cbase = R.cbase;
  t.aoff = A->d_off; t.a_nrb = A->nrb; t.boff = B->d_off; t.b_nrb = B->nrb;
  t.ksz = ta ? A->d_rsz : A->d_csz;
  t.ta = ta; t.tb = tb;
  t.asl = sa_.sl; t.asl_off = sa_.sl_off; t.aex = sa_.ex; t.aex_off = sa_.ex_off;
  t.bsl = sb_.sl; t.bsl_off = sb_.sl_off; t.bex = sb_.ex; t.bex_off = sb_.ex_off;
  t.azg = sa_.zg;
  t.zskip = zero_skip() ? 1 : 0;
  t.zmin = zero_min();
  t.slices = S;
  t.kstep = oz_kstep(kind);
  t.shared = shared ? 1 : 0;
  const int cf = pick_cfg(S, kind, C, o.tile);  // a forced id that runs other slices is ignored
  const OzCfg& k = kOzCfgs[cf];
  const int TM = 32 * k.fm, TN = 32 * k.fn;
  t.tile_groups = TM / 16;
  t.tile_n = TN;
  DevBuf& bpairs = c->buf("oz_pairs");
  if (!c->oz_pairs_ready) {
    bpairs.reserve(2 * sizeof(int64_t));
    fill_i64(c, bpairs.as<int64_t>(), 2, 0);
    c->oz_pairs_ready = true;
  }
  t.pairs = bpairs.as<unsigned long long>();
  hipLaunchKernelGGL(k_oz_records, dim3((unsigned)ceil_div(R.nc, 128)), dim3(128), 0, c->stream, t,
                     bouts.as<GemmOut>(), bprods.as<OzProd>());

  int32_t mm = 0, mn = 0;
  for (int32_t v : C->rsz) mm = std::max(mm, v);
  for (int32_t v : C->csz) mn = std::max(mn, v);
  const KWindows W = plan_kwindows(c, A, ta, B, tb, R, bouts.as<GemmOut>());
  c->last_kwindows = W.nwin;
  // work items of all windows planned together (one read-back); the window
  // launches then run back to back in ascending kb
  const WindowItems wi = window_items(c, W.outs, W.off, TM, TN, mm, mn);
  for (int w = 0; w < W.nwin; ++w) {
    const int64_t nitems = wi.off[(size_t)w + 1] - wi.off[(size_t)w];
    if (nitems == 0) continue;
    const GemmOut* wo = wi.items ? W.outs : W.outs + W.off[(size_t)w];
    const GemmItem* items = wi.items ? wi.items + wi.off[(size_t)w] : nullptr;
    ProfScope ps(c, "gemm_ozaki");
    hipLaunchKernelGGL(k.fn_, dim3((unsigned)nitems), dim3((unsigned)k.threads), 0, c->stream, wo,
                       (const OzProd*)bprods.as<OzProd>(), items, nitems, alpha);
  }
  SA_CHECK_HIP(hipGetLastError());
  mirror_lower(C, R);
  return R.ntasks;
}

}  // namespace sa
