// FP64 supernodal LDL^T, triangular solves and selected inversion for gfx950.
//
// Storage (per matrix of a batch, offsets in doubles):
//   L block of supernode s: m x w column-major (ld = m) at l_off[s]; rows are the
//     w supernode columns followed by R_s. D sits on the diagonal, unit-L below.
//   U workspace of s: r x r column-major; holds the Schur update during the
//     factorisation (at u_off[s]) and Sigma_{R_s R_s} during selected inversion
//     (at u_off_sel[s]); both offset maps share one pool sized by lifetimes.
//   S (selected inverse) uses the L layout.
// Batches are laid out matrix after matrix; blockIdx.y selects the matrix.
//
// Every level of the supernodal tree is one launch per front class: fronts with
// m <= 64 run one wavefront each with the front in LDS; larger fronts run a
// 256-thread workgroup that walks 32-wide panels and applies the trailing
// update with FP64 MFMA (v_mfma_f64_16x16x4f64) tiles.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "plan.hpp"

namespace gmrf {

typedef double d4 __attribute__((ext_vector_type(4)));

#define GMRF_ST_SINGULAR 1
#define GMRF_ST_ZERO_PIVOT 2

// Compiler-only fence after a round of independent loads: the loads stay issued
// above it, so the compiler cannot sink them into the conditional code that
// uses them (which would wait for each load separately). No instruction is
// emitted; the hardware waits only where a value is first used.
#define ISSUE_FENCE() asm volatile("" ::: "memory")

__device__ __forceinline__ d4 mfma16(double a, double b, d4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Element (i, j), i >= j, of a front whose first w columns live in the L block
// and whose trailing r x r block lives in the U workspace.
__device__ __forceinline__ double* front_at(double* Lb, double* Ub, int m, int w, int r, int i,
                                            int j) {
  return j < w ? Lb + (int64_t)j * m + i : Ub + (int64_t)(j - w) * r + (i - w);
}

// ---------------------------------------------------------------------------
// 64x64 tile product on FP64 MFMA.  T[i][j] += sum_k A(i, k) * B(j, k) for the
// tile's 64 rows i and 64 columns j. Four waves each own a 32x32 quadrant held
// as 2x2 MFMA accumulators. MFMA output layout (16x16x4 f64): lane l, register
// v holds D[4v + l/16][l%16]; the MFMA A operand is fed from B (so D's row is
// the tile column j) and the MFMA B operand from A (D's column is the tile row
// i), which makes the 16 lanes of a register row consecutive tile rows i —
// contiguous addresses in column-major storage.
// ---------------------------------------------------------------------------
// An operand element as returned by a tile-engine accessor: the value loaded from
// a valid (clamped) address, whether the element lies inside the operand and, for
// LVs, a factor (from LDS or registers) applied with the mask. Engines that prefetch
// resolve these only when the chunk goes to LDS, so the next chunk's loads stay in
// flight while the current one is on MFMA.
struct LV {
  double v;
  bool ok;
  __device__ double get() const { return ok ? v : 0.0; }
};
struct LVs {
  double v, s;
  bool ok;
  __device__ double get() const { return ok ? v * s : 0.0; }
};
__device__ __forceinline__ double lv_get(double x) { return x; }
__device__ __forceinline__ double lv_get(const LV& x) { return x.get(); }
__device__ __forceinline__ double lv_get(const LVs& x) { return x.get(); }

struct TileAcc {
  d4 c[2][2];  // [bj][bi]
};

__device__ __forceinline__ void tile_zero(TileAcc& acc) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc.c[a][b] = d4{0.0, 0.0, 0.0, 0.0};
}

__device__ __forceinline__ void tile_negate(TileAcc& acc) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc.c[a][b] = -acc.c[a][b];
}

// AK / BK: the operand is contiguous along k in memory, so consecutive lanes
// load consecutive k (otherwise consecutive rows). The accessors are called with
// k < K only and must load unconditionally from a valid (clamped) address and
// report validity beside the value (LV): a branch around a load serialises the
// memory latency. The next K chunk is fetched into registers while the current
// one is on MFMA; masks and factors are applied when a chunk goes to LDS, so no
// wait for the fetched values precedes the MFMA steps.
//
// NT = 512: eight waves; waves w and w + 4 own the same quadrant and take
// alternate k-steps of each chunk, so every SIMD interleaves two independent MFMA
// streams (one wave alone leaves the FP64 MFMA pipe about half idle); the two
// partial accumulators are summed through LDS at the end and the result is held
// by waves 0-3 (threads < 256) only.
template <bool AK = false, bool BK = false, int NT = 256, typename FA, typename FB>
__device__ __forceinline__ void mma_tile(FA fa, FB fb, int K, double (*As)[kPanel + 1],
                                         double (*Bs)[kPanel + 1], TileAcc& acc) {
  static_assert(NT == 256 || NT == 512, "tile engine runs 4 or 8 waves");
  constexpr int NL = 64 * kPanel / NT, NH = NT / 256;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int half = NH == 2 ? (wv >> 2) : 0, qi = wv & 1, qj = (wv >> 1) & 1;
  const int lr = lane & 15, lk = lane >> 4;
  decltype(fa(0, 0)) ra[NL];
  decltype(fb(0, 0)) rb[NL];
  auto fetch = [&](int k0) {
    const int kc = min(kPanel, K - k0);
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + NT * it;
      {
        const int ii = AK ? idx / kPanel : (idx & 63), kk = AK ? idx % kPanel : (idx >> 6);
        ra[it] = fa(ii, k0 + min(kk, kc - 1));
      }
      {
        const int ii = BK ? idx / kPanel : (idx & 63), kk = BK ? idx % kPanel : (idx >> 6);
        rb[it] = fb(ii, k0 + min(kk, kc - 1));
      }
    }
    ISSUE_FENCE();
  };
  if (K > 0) fetch(0);
  for (int k0 = 0; k0 < K; k0 += kPanel) {
    const int kc = min(kPanel, K - k0);
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + NT * it;
      {
        const int ii = AK ? idx / kPanel : (idx & 63), kk = AK ? idx % kPanel : (idx >> 6);
        As[ii][kk] = kk < kc ? lv_get(ra[it]) : 0.0;
      }
      {
        const int ii = BK ? idx / kPanel : (idx & 63), kk = BK ? idx % kPanel : (idx >> 6);
        Bs[ii][kk] = kk < kc ? lv_get(rb[it]) : 0.0;
      }
    }
    __syncthreads();
    if (k0 + kPanel < K) fetch(k0 + kPanel);
#pragma unroll
    for (int st = 0; st < kPanel / (4 * NH); ++st) {
      const int kk = 4 * (st * NH + half);
      if (4 * st * NH >= kc) break;  // uniform: short last chunk (zero padded)
      const double a0 = Bs[qj * 32 + lr][kk + lk];
      const double a1 = Bs[qj * 32 + 16 + lr][kk + lk];
      const double b0 = As[qi * 32 + lr][kk + lk];
      const double b1 = As[qi * 32 + 16 + lr][kk + lk];
      acc.c[0][0] = mfma16(a0, b0, acc.c[0][0]);
      acc.c[0][1] = mfma16(a0, b1, acc.c[0][1]);
      acc.c[1][0] = mfma16(a1, b0, acc.c[1][0]);
      acc.c[1][1] = mfma16(a1, b1, acc.c[1][1]);
    }
    __syncthreads();
  }
  if (NH == 2) {
    // quadrant q's partials: waves 4, 5 through the A operand area, 6, 7 through B's
    static_assert(2 * 16 * 64 <= 64 * (kPanel + 1), "reduction exceeds the operand LDS");
    const int q = wv & 3;
    double* red = (q < 2 ? &As[0][0] : &Bs[0][0]) + (q & 1) * 16 * 64;
    if (half == 1) {
#pragma unroll
      for (int e = 0; e < 16; ++e) red[e * 64 + lane] = acc.c[e >> 3][(e >> 2) & 1][e & 3];
      tile_zero(acc);  // so that a following product on the same accumulator adds once
    }
    __syncthreads();
    if (half == 0) {
#pragma unroll
      for (int e = 0; e < 16; ++e) acc.c[e >> 3][(e >> 2) & 1][e & 3] += red[e * 64 + lane];
    }
    __syncthreads();
  }
}

// Visit every accumulator element with its tile coordinates (i, j).
template <typename F>
__device__ __forceinline__ void tile_foreach(const TileAcc& acc, F f) {
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int qi = wv & 1, qj = wv >> 1;
#pragma unroll
  for (int bj = 0; bj < 2; ++bj)
#pragma unroll
    for (int bi = 0; bi < 2; ++bi)
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = qi * 32 + bi * 16 + (lane & 15);
        const int j = qj * 32 + bj * 16 + 4 * v + (lane >> 4);
        f(i, j, acc.c[bj][bi][v]);
      }
}

// 32 x 32 tile on 256 threads: T[i][j] += sum_k A(i, k) B(j, k) for 32 rows i and
// 32 columns j. Wave w owns the 16 x 16 MFMA block of row block w & 1 and column block
// w >> 1 over the whole K, so a workgroup does a quarter of a 64 x 64 tile's MFMA
// work per chunk: for small tiles with a long K in launches of few workgroups. AK / BK
// as in mma_tile (the operand is contiguous along k: lanes along k, else along rows);
// each chunk is staged in LDS (sh: 2 x 32 x (kPanel + 1) doubles) with masks and
// factors applied at the store, as in mma_tile. acc[v] of lane l holds
// T[16 (w & 1) + l % 16][16 (w >> 1) + 4 v + l / 16].
template <bool AK = true, bool BK = true, typename FA, typename FB>
__device__ __forceinline__ void mma_tile32(FA fa, FB fb, int K, double* sh, d4& acc) {
  static_assert(kPanel == 32, "the 32 x 32 engine stages 32 x 32 operand chunks");
  constexpr int LD = kPanel + 1, NL = 32 * kPanel / 256;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int rb = wv & 1, cb = wv >> 1, lr = lane & 15, lk = lane >> 4;
  double* As = sh;
  double* Bs = sh + 32 * LD;
  decltype(fa(0, 0)) ra[NL];
  decltype(fb(0, 0)) rb_[NL];
  auto fetch = [&](int k0) {
    const int kc = min(kPanel, K - k0);
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + 256 * it;
      {
        const int ii = AK ? (idx >> 5) : (idx & 31), kk = AK ? (idx & 31) : (idx >> 5);
        ra[it] = fa(ii, k0 + min(kk, kc - 1));
      }
      {
        const int ii = BK ? (idx >> 5) : (idx & 31), kk = BK ? (idx & 31) : (idx >> 5);
        rb_[it] = fb(ii, k0 + min(kk, kc - 1));
      }
    }
    ISSUE_FENCE();
  };
  if (K > 0) fetch(0);
  for (int k0 = 0; k0 < K; k0 += kPanel) {
    const int kc = min(kPanel, K - k0);
#pragma unroll
    for (int it = 0; it < NL; ++it) {
      const int idx = tid + 256 * it;
      {
        const int ii = AK ? (idx >> 5) : (idx & 31), kk = AK ? (idx & 31) : (idx >> 5);
        As[ii * LD + kk] = kk < kc ? lv_get(ra[it]) : 0.0;
      }
      {
        const int ii = BK ? (idx >> 5) : (idx & 31), kk = BK ? (idx & 31) : (idx >> 5);
        Bs[ii * LD + kk] = kk < kc ? lv_get(rb_[it]) : 0.0;
      }
    }
    __syncthreads();
    if (k0 + kPanel < K) fetch(k0 + kPanel);
#pragma unroll
    for (int st = 0; st < kPanel / 4; ++st) {
      const int kk = 4 * st;
      if (kk >= kc) break;  // uniform: short last chunk (zero padded)
      acc = mfma16(Bs[(cb * 16 + lr) * LD + kk + lk], As[(rb * 16 + lr) * LD + kk + lk], acc);
    }
    __syncthreads();
  }
}

// Visit the accumulator elements of mma_tile32 with their tile coordinates (i, j).
template <typename F>
__device__ __forceinline__ void tile_foreach32(const d4& acc, F f) {
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int i = (wv & 1) * 16 + (lane & 15);
#pragma unroll
  for (int v = 0; v < 4; ++v) f(i, (wv >> 1) * 16 + 4 * v + (lane >> 4), acc[v]);
}

// The engine on one LDS block sh of 2 x 64 x (kPanel + 1) doubles (A, then B).
template <bool AK, bool BK, typename FA, typename FB>
__device__ __forceinline__ void tile_product(FA fa, FB fb, int K, double* sh, TileAcc& acc) {
  mma_tile<AK, BK>(fa, fb, K, (double (*)[kPanel + 1])sh, (double (*)[kPanel + 1])(sh + 64 * (kPanel + 1)), acc);
}

// ---------------------------------------------------------------------------
// Probe: T = A * B^T for one 64x64 tile (A, B 64 x K column-major, K <= 64).
// Used by the tests to pin the MFMA lane layout against a host product.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_mma_probe(const double* A, const double* B, int K,
                                                   double* T) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  TileAcc acc;
  tile_zero(acc);
  mma_tile([&](int i, int k) { return LV{A[(int64_t)k * 64 + i], true}; },
           [&](int j, int k) { return LV{B[(int64_t)k * 64 + j], true}; }, K, As, Bs, acc);  // k < K
  tile_foreach(acc, [&](int i, int j, double v) { T[(int64_t)j * 64 + i] = v; });
}

// ---------------------------------------------------------------------------
// Blocked dense LDL^T of a front of at most 64 rows held in LDS.
//
// The front sits row-major in LDS (F[i * LD + j], lower triangle). Columns are
// eliminated in blocks of 16: one wave factors the 16 x 16 diagonal block in
// registers (lane exchanges and, in the MFMA form, 16x16x4 products; no barrier),
// then all waves form the panel
// W = A21 Z11^T (Z11 = L11^-1), L21 = W D^-1 and the trailing update
// A22 -= W L21^T as 16 x 16 x 16 FP64 MFMA products. Optionally the inverse
// Z = L^-1 of the whole block is built alongside (block Gauss transforms):
// Z[k][j] <- Z11 Z[k][j] for j < k and Z[i][j] -= L[i][k] Z[k][j] for i > k.
// ---------------------------------------------------------------------------

// acc(16 x 16) += A(16 x 16) B(16 x 16) on one wave; a(i, q) and b(q, j) must
// load unconditionally (clamped addresses) and mask afterwards. acc[v] holds
// row 4 v + lane / 16, column lane % 16.
template <typename FA, typename FB>
__device__ __forceinline__ d4 wgemm16(FA a, FB b, d4 acc) {
  const int lane = threadIdx.x & 63, r = lane & 15, q = lane >> 4;
  double av[4], bv[4];
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) {
    av[s4] = lv_get(a(r, 4 * s4 + q));
    bv[s4] = lv_get(b(4 * s4 + q, r));
  }
  ISSUE_FENCE();
#pragma unroll
  for (int s4 = 0; s4 < 4; ++s4) acc = mfma16(av[s4], bv[s4], acc);
  return acc;
}

// 1/d from the hardware reciprocal estimate refined by two Newton-Raphson steps
// x <- x + x (1 - d x) (each step squares the relative error); within about one
// ulp of the correctly rounded quotient, with a shorter dependent chain than the
// IEEE division sequence.
__device__ __forceinline__ double recip_nr(double d) {
  double x = __builtin_amdgcn_rcp(d);
  double e = fma(-d, x, 1.0);
  x = fma(x, e, x);
  e = fma(-d, x, 1.0);
  return fma(x, e, x);
}

// One wave: LDL^T of the kn x kn (kn <= 16) diagonal block at (k0, k0) of F in
// place (unit L below the diagonal, D on it) and Zk = L_kk^-1 (16 x 17, unit
// lower, identity beyond kn). Lane (i = lane % 16, c = lane / 16) holds
// columns c + 4t; column k and row k of Z travel by shuffles.
template <int LD>
__device__ __forceinline__ void ldlt_base16(double* F, int k0, int kn, double* Zk) {
  const int lane = threadIdx.x & 63, i = lane & 15, c = lane >> 4;
  double d[4], z[4];
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = c + 4 * t;
    const double fv = F[(k0 + min(i, kn - 1)) * LD + k0 + min(j, kn - 1)];
    d[t] = (i < kn && j < kn) ? (j <= i ? fv : 0.0) : (i == j ? 1.0 : 0.0);
    z[t] = (i == j) ? 1.0 : 0.0;
  }
  for (int k = 0; k < kn; ++k) {
    const int tk = k >> 2, src = 16 * (k & 3);
    double v = d[0];
#pragma unroll
    for (int t = 1; t < 4; ++t) v = (t == tk) ? d[t] : v;
    const double ci = __shfl(v, src + i, 64);  // D[i][k]
    const double dk = __shfl(v, src + k, 64);
    double cj[4], zk[4];
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      cj[t] = __shfl(v, src + c + 4 * t, 64);  // D[j][k]
      zk[t] = __shfl(z[t], 16 * c + k, 64);    // Z[k][j]
    }
    const bool act = i > k && i < kn;
    const double lik = ci * recip_nr(dk);
#pragma unroll
    for (int t = 0; t < 4; ++t) {
      const int j = c + 4 * t;
      const double nd = d[t] - lik * cj[t];  // D_ij -= l_ik d_k l_jk
      const double nz = z[t] - lik * zk[t];
      d[t] = act ? ((j > k && j <= i) ? nd : (j == k ? lik : d[t])) : d[t];
      z[t] = (act && j <= k) ? nz : z[t];
    }
  }
#pragma unroll
  for (int t = 0; t < 4; ++t) {
    const int j = c + 4 * t;
    if (i < kn && j <= i) F[(k0 + i) * LD + k0 + j] = d[t];
    Zk[i * 17 + j] = z[t];
  }
}

// Lane l's double of a uniform source lane (two 32-bit reads into scalar registers).
__device__ __forceinline__ double readlane_d(double x, int l) {
  const long long v = __double_as_longlong(x);
  const unsigned lo = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)v, l);
  const unsigned hi = (unsigned)__builtin_amdgcn_readlane((int)(unsigned)(v >> 32), l);
  return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// The same contract on the MFMA pipe. The symmetric block M is held in the MFMA
// accumulator layout (lane l = 16 g + c, register v: M[4 v + g][c]), Z = L^-1 as
// well. Per 4-column step p (pivot rows q0 = 4 p ..): the 4 x 4 pivot block is read
// from lanes into scalar registers and factored redundantly (Lp Dp Lp^T, Y =
// Lp^-1); lane (c, g) gathers row c's entries x_a = M[c][q0 + a] (the symmetric
// copy in register p of lanes 16 a + c) and forms L[c][q0 + g] = (x Y^T)_g / d_g;
// the trailing update M -= L Dp L^T is one 16x16x4 MFMA whose operands are the
// lane's own values (A[c][g] = L[c][q0 + g], B[g][c] = d_g L[c][q0 + g]); Z's pivot
// rows become Y Z_p and the rows below take -L Z_p, another MFMA. Rows and columns
// beyond kn are the identity throughout.
template <int LD>
__device__ __forceinline__ void ldlt_base16_mfma(double* F, int k0, int kn, double* Zk) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  d4 M, Zv;
#pragma unroll
  for (int v = 0; v < 4; ++v) {
    const int i = 4 * v + g;
    const int hi = max(i, c), lo = min(i, c);
    const double fv = F[(k0 + min(hi, kn - 1)) * LD + k0 + min(lo, kn - 1)];
    M[v] = (i < kn && c < kn) ? fv : (i == c ? 1.0 : 0.0);
    Zv[v] = i == c ? 1.0 : 0.0;
  }
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int q0 = 4 * p;
    if (q0 >= kn) break;
    // pivot block (lower), P[a][b] = M[q0 + a][q0 + b] at lane 16 a + q0 + b
    const double p00 = readlane_d(M[p], q0), p10 = readlane_d(M[p], 16 + q0),
                 p11 = readlane_d(M[p], 16 + q0 + 1), p20 = readlane_d(M[p], 32 + q0),
                 p21 = readlane_d(M[p], 32 + q0 + 1), p22 = readlane_d(M[p], 32 + q0 + 2),
                 p30 = readlane_d(M[p], 48 + q0), p31 = readlane_d(M[p], 48 + q0 + 1),
                 p32 = readlane_d(M[p], 48 + q0 + 2), p33 = readlane_d(M[p], 48 + q0 + 3);
    // row c's entries in the pivot columns (symmetric copy: M[q0 + a][c])
    double x[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) x[a] = __shfl(M[p], 16 * a + c, 64);
    double z[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) z[a] = __shfl(Zv[p], 16 * a + c, 64);
    // Lp Dp Lp^T of the pivot block, column by column (as the register form)
    const double d0 = p00, r0 = recip_nr(d0);
    const double l10 = p10 * r0, l20 = p20 * r0, l30 = p30 * r0;
    const double s11 = p11 - l10 * p10, s21 = p21 - l20 * p10, s31 = p31 - l30 * p10;
    const double s22 = p22 - l20 * p20, s32 = p32 - l30 * p20, s33 = p33 - l30 * p30;
    const double d1 = s11, r1 = recip_nr(d1);
    const double l21 = s21 * r1, l31 = s31 * r1;
    const double t22 = s22 - l21 * s21, t32 = s32 - l31 * s21, t33 = s33 - l31 * s31;
    const double d2 = t22, r2 = recip_nr(d2);
    const double l32 = t32 * r2;
    const double d3 = t33 - l32 * t32, r3 = recip_nr(d3);
    // Y = Lp^-1 (unit lower)
    const double y10 = -l10, y21 = -l21, y32 = -l32;
    const double y20 = -(l20 + l21 * y10), y31 = -(l31 + l32 * y21);
    const double y30 = -(l30 + l31 * y10 + l32 * y20);
    // this lane's row g of Y, d_g, 1 / d_g
    const double ya0 = g == 1 ? y10 : (g == 2 ? y20 : y30);
    const double ya1 = g == 2 ? y21 : y31;
    const double dg = g == 0 ? d0 : (g == 1 ? d1 : (g == 2 ? d2 : d3));
    const double rg = g == 0 ? r0 : (g == 1 ? r1 : (g == 2 ? r2 : r3));
    // L[c][q0 + g] = (sum_{a <= g} Y[g][a] x_a) / d_g; x[g] with the index as a select
    // chain (measured faster here than selecting among the four values)
    double u = x[g == 0 ? 0 : (g == 1 ? 1 : (g == 2 ? 2 : 3))];
    u = g >= 1 ? u + ya0 * x[0] : u;
    u = g >= 2 ? u + ya1 * x[1] : u;
    u = g >= 3 ? u + y32 * x[2] : u;
    double lcg = u * rg;
    // pivot rows: the block's own factor
    const int a = c - q0;
    if (a >= 0 && a < 4) {
      const double la0 = a == 1 ? l10 : (a == 2 ? l20 : l30), la1 = a == 2 ? l21 : l31;
      lcg = a == g ? 1.0 : (g == 0 ? la0 : (g == 1 ? la1 : (g == 2 ? l32 : 0.0)));
      lcg = a < g ? 0.0 : lcg;
    }
    const int col = q0 + g;
    if (c < kn && col < kn && c >= col) F[(k0 + c) * LD + k0 + col] = c == col ? dg : lcg;
    // Z: pivot rows Y Z_p, rows below minus L Z_p
    double zn = z[g == 0 ? 0 : (g == 1 ? 1 : (g == 2 ? 2 : 3))];
    zn = g >= 1 ? zn + ya0 * z[0] : zn;
    zn = g >= 2 ? zn + ya1 * z[1] : zn;
    zn = g >= 3 ? zn + y32 * z[2] : zn;
    Zv[p] = zn;
    const bool below = c >= q0 + 4;
    Zv = mfma16(below ? -lcg : 0.0, zn, Zv);
    // trailing update of the rows and columns below the pivot block
    M = mfma16(below ? lcg : 0.0, -dg * lcg, M);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) Zk[(4 * v + g) * 17 + c] = Zv[v];
}

// The 16 x 16 base case of a front class: the MFMA form for fronts of 48 rows or
// more and the 64-column diagonal blocks (fewer dependent steps), the register form
// for the 16- and 32-row classes (fewer registers: those launches run many
// workgroups per CU).
template <int MB, int LD>
__device__ __forceinline__ void base16(double* F, int k0, int kn, double* Zk) {
  if constexpr (MB >= 48)
    ldlt_base16_mfma<LD>(F, k0, kn, Zk);
  else
    ldlt_base16<LD>(F, k0, kn, Zk);
}

// Blocked LDL^T of the first ncols columns of an m x m front in F (m <= MB;
// the calling workgroup has MB / 16 waves; only F's lower triangle is read). With
// WITHZ (ncols == m, F's strict upper triangle zero on entry) F's strict upper
// triangle receives Z = L^-1, transposed: Z[i][j] (i > j) at F[j * LD + i]; Z's
// diagonal is 1 and its upper part 0. Zk (16 x 17) is scratch, and so is Wb
// (MB x 17, the panel rows times D) with WITHZ; without, W[i][p] of block k0 is kept
// in F's strict upper triangle at F[(k0 + p) * LD + i] and Wb is not used.
template <int MB, bool WITHZ>
__device__ __forceinline__ void front_ldlt(double* F, int m, int ncols, double* Zk, double* Wb) {
  constexpr int LD = MB + 1, NW = MB / 16;
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63;
  const int vr = lane >> 4, vc = lane & 15;  // accumulator element (4 v + vr, vc)
  if (wv == 0 && ncols > 0) base16<MB, LD>(F, 0, min(16, ncols), Zk);
  __syncthreads();
  for (int k0 = 0; k0 < ncols; k0 += 16) {
    const int kn = min(16, ncols - k0), t0 = k0 + kn, kb = k0 >> 4;
    // panel rows below the block; Z[kb][jb] <- Zk Z[kb][jb] (jb < kb); Z[kb][kb] = Zk
    const int nrb = (m - t0 + 15) >> 4;
    const int ntask2 = nrb + (WITHZ ? kb + 1 : 0);
    for (int task = wv; task < ntask2; task += NW) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      if (task < nrb) {
        const int r0 = t0 + 16 * task;
        acc = wgemm16(
            [&](int ii, int q) {
              return LV{F[min(r0 + ii, m - 1) * LD + k0 + min(q, kn - 1)], (r0 + ii < m && q < kn)};
            },
            [&](int q, int pp) {
              return LV{Zk[pp * 17 + q], q < kn};
            },
            acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ii = 4 * v + vr, pp = vc;
          if (r0 + ii < m && pp < kn) {
            if (WITHZ)
              Wb[(r0 - t0 + ii) * 17 + pp] = acc[v];
            else
              F[(k0 + pp) * LD + r0 + ii] = acc[v];  // W[r0 + ii][pp] in F's free upper triangle
            F[(r0 + ii) * LD + k0 + pp] = acc[v] / F[(k0 + pp) * LD + k0 + pp];
          }
        }
      } else if (task < nrb + kb) {
        const int jb = task - nrb;
        acc = wgemm16(
            [&](int ii, int q) {
              return LV{Zk[ii * 17 + q], q < kn};
            },
            [&](int q, int jj) {  // Z[k0 + q][16 jb + jj], strictly lower (jb < kb)
              return LV{F[(16 * jb + jj) * LD + k0 + min(q, kn - 1)], q < kn};
            },
            acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ii = 4 * v + vr;
          if (ii < kn) F[(16 * jb + vc) * LD + k0 + ii] = acc[v];
        }
      } else {
        // Z[kb][kb] = Zk: its strictly lower entries
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int ii = 4 * v + vr;
          if (ii < kn && vc < ii) F[(k0 + vc) * LD + k0 + ii] = Zk[ii * 17 + vc];
        }
      }
    }
    __syncthreads();
    // trailing update A22 -= W L21^T (lower blocks) and Z[ib][jb] -= L[ib][kb] Z[kb][jb].
    // Look-ahead: task 0 is the next diagonal block; wave 0 takes only that one
    // (when there are other waves) and then factors the next block right away,
    // while the other waves finish the remaining tasks.
    const int ntb = nrb * (nrb + 1) / 2;
    const int ntask3 = ntb + (WITHZ ? nrb * (kb + 1) : 0);
    const int k1 = k0 + 16;
    const int t_first = NW == 1 ? 0 : (wv == 0 ? 0 : wv);
    const int t_last = NW == 1 ? ntask3 : (wv == 0 ? min(1, ntask3) : ntask3);
    const int t_step = NW == 1 ? 1 : (wv == 0 ? 1 : NW - 1);
    for (int task = t_first; task < t_last; task += t_step) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      if (task < ntb) {
        int ib = 0, rem = task;
        while (rem > ib) {
          rem -= ib + 1;
          ++ib;
        }
        const int jb = rem;
        const int R = t0 + 16 * ib, C = t0 + 16 * jb;
        acc = wgemm16(
            [&](int ii, int pp) {
              const int wr = min(t0 + 16 * ib + ii, m - 1), wc = min(pp, kn - 1);
              return LV{WITHZ ? Wb[(wr - t0) * 17 + wc] : F[(k0 + wc) * LD + wr], pp < kn};
            },
            [&](int pp, int jj) {
              return LV{F[min(C + jj, m - 1) * LD + k0 + min(pp, kn - 1)], (pp < kn && C + jj < m)};
            },
            acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = R + 4 * v + vr, j = C + vc;
          if (i < m && j < m && i >= j) F[i * LD + j] -= acc[v];
        }
      } else {
        const int t2 = task - ntb, ib = t2 / (kb + 1), jb = t2 - ib * (kb + 1);
        const int R = t0 + 16 * ib;
        acc = wgemm16(
            [&](int ii, int pp) {
              return LV{F[min(R + ii, m - 1) * LD + k0 + min(pp, kn - 1)], pp < kn};
            },
            [&](int pp, int jj) {  // Z[k0 + pp][16 jb + jj], jb <= kb (the diagonal block: unit lower)
              const int zi = k0 + min(pp, kn - 1), zj = 16 * jb + jj;
              const double zv = zi == zj ? 1.0 : F[min(zi, zj) * LD + max(zi, zj)];
              return LV{zi >= zj ? zv : 0.0, pp < kn};
            },
            acc);
#pragma unroll
        for (int v = 0; v < 4; ++v) {
          const int i = R + 4 * v + vr;
          if (i < m) F[(16 * jb + vc) * LD + i] -= acc[v];
        }
      }
    }
    if (wv == 0 && k1 < ncols) base16<MB, LD>(F, k1, min(16, ncols - k1), Zk);
    __syncthreads();
  }
}

// ---------------------------------------------------------------------------
// Factorisation
// ---------------------------------------------------------------------------

// Small fronts (m <= MB, MB in {16, 32, 64}), 4 MB threads per front. The front
// is assembled in LDS (row-major, ld = MB + 1) from the input entries and the
// children's updates, every global load issued before its use; the w supernode
// columns are then eliminated with the blocked LDL^T above, and the L block and
// the lower r x r update are written out.
template <int MB>
__global__ __launch_bounds__(4 * MB) void k_factor_small(DevPlan P, const int32_t* __restrict__ list,
                                                         const double* __restrict__ vals, int64_t ldv,
                                                         double* __restrict__ L,
                                                         double* __restrict__ U) {
  constexpr int NT = 4 * MB, LD = MB + 1, NCM = 32;
  __shared__ double F[MB * LD];
  __shared__ double Zk[16 * 17];
  __shared__ int64_t ch_uoff[NCM], ch_rp[NCM];  // the first NCM children's U offset and row list
  __shared__ int ch_rc[NCM];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int cp0 = P.child_ptr[s], cp1 = P.child_ptr[s + 1];
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const double* Um = U + (int64_t)b * P.u_size;
  const double* vb = vals + (int64_t)b * ldv;
  const int64_t q0 = P.fa_ptr[s], q1 = P.fa_ptr[s + 1];
  // the children's metadata (for the child loop below) and the first round of input
  // entries are loaded together, before the front is cleared
  int cm_rc = 0;
  int64_t cm_uoff = 0, cm_rp = 0;
  if (tid < min(cp1 - cp0, NCM)) {
    const int c = P.child_idx[cp0 + tid];
    cm_rc = P.sn_r[c];
    cm_uoff = P.u_off[c];
    cm_rp = P.rows_ptr[c];
  }
  int slot[4];
  double v[4];
  auto load_inputs = [&](int64_t qb) {
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const int64_t q = min(qb + tid + NT * u, q1 - 1);
      slot[u] = P.fa_slot[q];
      v[u] = vb[P.fa_csc[q]];
    }
    ISSUE_FENCE();
  };
  if (q1 > q0) load_inputs(q0);
  for (int idx = tid; idx < m * LD; idx += NT) F[idx] = 0.0;
  if (tid < min(cp1 - cp0, NCM)) {
    ch_rc[tid] = cm_rc;
    ch_uoff[tid] = cm_uoff;
    ch_rp[tid] = cm_rp;
  }
  __syncthreads();
  for (int64_t qb = q0; qb < q1; qb += 4 * NT) {
    if (qb > q0) load_inputs(qb);
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (qb + tid + NT * u < q1) {
        const int j = slot[u] / m, i = slot[u] - j * m;
        F[i * LD + j] = v[u];
      }
  }
  __syncthreads();
  for (int cq = cp0; cq < cp1; ++cq) {
    const int k = cq - cp0;
    int rc;
    int64_t uo, rp;
    if (k < NCM) {
      rc = ch_rc[k];
      uo = ch_uoff[k];
      rp = ch_rp[k];
    } else {
      const int c = P.child_idx[cq];
      rc = P.sn_r[c];
      uo = P.u_off[c];
      rp = P.rows_ptr[c];
    }
    const double* Uc = Um + uo;
    const int32_t* relc = P.rel + rp;
    const int ne = rc * rc;
    for (int e0 = 0; e0 < ne; e0 += 4 * NT) {
      double v[4];
      int ri[4], rj[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = min(e0 + tid + NT * u, ne - 1);
        const int j = e / rc, i = e - j * rc;
        v[u] = Uc[e];
        ri[u] = relc[i];
        rj[u] = relc[j];
      }
      ISSUE_FENCE();
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        const int e = e0 + tid + NT * u;
        const int j = e / rc, i = e - j * rc;
        if (e < ne && i >= j) F[ri[u] * LD + rj[u]] += v[u];
      }
    }
    __syncthreads();
  }
  front_ldlt<MB, false>(F, m, w, Zk, nullptr);
  for (int idx = tid; idx < m * m; idx += NT) {
    const int j = idx / m, i = idx - j * m;
    if (j < w)
      Lb[(int64_t)j * m + i] = (j <= i) ? F[i * LD + j] : 0.0;
    else if (i >= j)
      Ub[(int64_t)(j - w) * r + (i - w)] = F[i * LD + j];
  }
}

// ---------------------------------------------------------------------------
// Large fronts (m > 64), tile-parallel. Each launch covers every large front of
// one level for every matrix of the batch; a task is an int32 triple
// (supernode, a, b) whose meaning depends on the kernel.
// ---------------------------------------------------------------------------

// LDS of one diagonal-block factorisation: F (64 x 65: L and D in the lower
// triangle, Z = L^-1 transposed in the strict upper one), Zk (16 x 17), Wb (64 x 17).
constexpr int kDiagLds = kTile * (kTile + 1) + 16 * 17 + kTile * 17;

__device__ __forceinline__ void fl_diag_block(const DevPlan& P, int s, int step, int b,
                                              double* __restrict__ L, double* __restrict__ Z,
                                              double* lds) {
  constexpr int LD = kTile + 1;
  double* F = lds;
  double* Zk = F + kTile * LD;
  double* Wb = Zk + 16 * 17;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int c0 = step * kTile, nb = min(kTile, w - c0);
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)step * kTile * kTile;
  const int i = tid & 63, c = tid >> 6;
  {
    double v[16];
#pragma unroll
    for (int t = 0; t < 16; ++t) v[t] = Lb[(int64_t)(c0 + min(c + 4 * t, nb - 1)) * m + c0 + min(i, nb - 1)];
    ISSUE_FENCE();
#pragma unroll
    for (int t = 0; t < 16; ++t) {
      const int j = c + 4 * t;
      F[i * LD + j] = (i < nb && j <= i) ? v[t] : 0.0;
    }
  }
  __syncthreads();
  front_ldlt<kTile, true>(F, nb, nb, Zk, Wb);
#pragma unroll
  for (int t = 0; t < 16; ++t) {
    const int j = c + 4 * t;
    if (i < nb && j < nb) {
      Lb[(int64_t)(c0 + j) * m + c0 + i] = (j <= i) ? F[i * LD + j] : 0.0;
      Zb[j * kTile + i] = (j < i) ? F[j * LD + i] : (j == i ? 1.0 : 0.0);
    }
  }
}

// Assembly of one 64x64 tile (tile row I, tile column J, I >= J) of a front:
// the sum of every child's update through the child's inverse row map (children
// in fixed order: deterministic), then the input entries that fall into the tile
// (every L slot receives at most one). Thread (row i = tid % 64) owns tile
// columns tid / 64 + 4u, u < 16; each round of loads covers all 16 elements
// before any is used. The tile passes through LDS for the input entries.
//
// With DIAG (launches of few tiles), the workgroup of tile (0, 0) of a front goes
// on to factor the front's first diagonal block (k_fl_diag, step 0), which lies
// inside that tile: one launch less per level.
// Without DIAG the register budget is held to four waves per SIMD (the kernel is
// bound by its load latency, so occupancy pays).
template <bool DIAG>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(DIAG ? 1 : 4))) void k_fl_assemble(DevPlan P, const int32_t* __restrict__ tasks,
                                                     int ain_first, const double* __restrict__ vals,
                                                     int64_t ldv, double* __restrict__ L,
                                                     double* __restrict__ U, double* __restrict__ Z) {
  constexpr int NCM = 16;
  __shared__ double lds[DIAG ? kDiagLds : kTile * (kTile + 1)];
  __shared__ int64_t ch_inv[NCM], ch_uoff[NCM];  // the first NCM children's inverse map and U offset
  __shared__ int ch_rc[NCM];
  double* Ft = lds;  // Ft[65 jj + ii]
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1], J = tk[2];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int cp0 = P.child_ptr[s], cp1 = P.child_ptr[s + 1];
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const double* Um = U + (int64_t)b * P.u_size;
  const int ii = threadIdx.x & 63, jb = threadIdx.x >> 6;
  const int i = I * kTile + ii, j0 = J * kTile + jb;
  const int ic = min(i, m - 1);
  if ((int)threadIdx.x < min(cp1 - cp0, NCM)) {
    const int cq = cp0 + threadIdx.x, c = P.child_idx[cq];
    ch_inv[threadIdx.x] = P.cinv_off[cq];
    ch_uoff[threadIdx.x] = P.u_off[c];
    ch_rc[threadIdx.x] = P.sn_r[c];
  }
  __syncthreads();
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) v[u] = 0.0;
  for (int cq = cp0; cq < cp1; ++cq) {
    const int k = cq - cp0;
    int64_t io, uo;
    int rc;
    if (k < NCM) {
      io = ch_inv[k];
      uo = ch_uoff[k];
      rc = ch_rc[k];
    } else {
      const int c = P.child_idx[cq];
      io = P.cinv_off[cq];
      uo = P.u_off[c];
      rc = P.sn_r[c];
    }
    const int32_t* inv = P.cinv + io;
    const double* Uc = Um + uo;
    const int ci = inv[ic];
    int cj[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) cj[u] = inv[min(j0 + 4 * u, m - 1)];
    double x[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const bool ok = ci >= 0 && cj[u] >= 0;
      x[u] = Uc[ok ? (int64_t)cj[u] * rc + ci : 0];
    }
    ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] += (ci >= 0 && cj[u] >= 0) ? x[u] : 0.0;
  }
  const int e0 = P.ain_ptr[ain_first + blockIdx.x], e1 = P.ain_ptr[ain_first + blockIdx.x + 1];
  if (e1 > e0) {
    const double* vb = vals + (int64_t)b * ldv;
#pragma unroll
    for (int u = 0; u < 16; ++u) Ft[(jb + 4 * u) * (kTile + 1) + ii] = v[u];
    __syncthreads();
    for (int e = e0 + (int)threadIdx.x; e < e1; e += 256) Ft[P.ain_pos[e]] += vb[P.ain_csc[e]];
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) v[u] = Ft[(jb + 4 * u) * (kTile + 1) + ii];
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = j0 + 4 * u;
    if (i < m && j < m && i >= j) *front_at(Lb, Ub, m, w, r, i, j) = v[u];
  }
  if (DIAG && I == 0 && J == 0) {
    __syncthreads();  // the tile's stores, before the block is read back
    fl_diag_block(P, s, 0, b, L, Z, lds);
  }
}

// Diagonal block of panel step `step` (columns [64 step, 64 step + nb)), one
// task per front: blocked LDL^T of the nb x nb block in LDS and Z = L11^-1. The
// factorised block goes back to L; Z is kept in the per-panel Z workspace
// (Zb[q * 64 + i] = Z[i][q]) for the panel solves and selected inversion.
__global__ __launch_bounds__(256) void k_fl_diag(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, double* __restrict__ L,
                                                 double* __restrict__ Z) {
  __shared__ double lds[kDiagLds];
  fl_diag_block(P, tasks[3 * blockIdx.x], step, blockIdx.y, L, Z, lds);
}

// Panel solve as a product: task (s, I) computes 64 rows below the diagonal block,
// L21[i][p] = sum_q A21[i][q] Z[p][q] / d_p  (L21 = A21 L11^-T D^-1), in place.
// With split, each tile runs as two workgroups of 32 rows (part = blockIdx.x & 1), for
// launches of few tiles; each forms both 32-column halves of its rows (mma_tile32 twice)
// before it stores them. The solve is in place: a workgroup's K operand is every panel
// column of its rows, so the column halves of one row block cannot be split between
// workgroups (one would read columns the other has already overwritten).
__global__ __launch_bounds__(256) void k_fl_trsm(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, int split, double* __restrict__ L,
                                                 const double* __restrict__ Z) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  __shared__ double dinv[kTile];
  const int32_t* tk = tasks + 3 * (split ? (blockIdx.x >> 1) : blockIdx.x);
  const int s = tk[0], I = tk[1];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int c0 = step * kTile, nb = min(kTile, w - c0);
  const int R0 = c0 + nb + I * kTile;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)step * kTile * kTile;
  if (threadIdx.x < nb) dinv[threadIdx.x] = 1.0 / Lb[(int64_t)(c0 + threadIdx.x) * m + c0 + threadIdx.x];
  __syncthreads();
  if (split) {
    const int r0 = R0 + (blockIdx.x & 1) * 32;
    if (r0 >= m) return;  // uniform: the row block lies beyond the front
    d4 acc[2] = {{0.0, 0.0, 0.0, 0.0}, {0.0, 0.0, 0.0, 0.0}};
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int q0 = 32 * h;
      if (q0 >= nb) break;  // uniform
      mma_tile32<false, false>(
          [&](int ii, int q) {
            const int i = r0 + ii;
            return LV{Lb[(int64_t)(c0 + q) * m + min(i, m - 1)], i < m};
          },
          [&](int pp, int q) {
            const int p = q0 + pp;
            return LVs{Zb[q * kTile + p], dinv[p], p < nb};
          },
          nb, &As[0][0], acc[h]);
    }
    // every read of the rows' panel columns is done (the engine ends on a barrier)
#pragma unroll
    for (int h = 0; h < 2; ++h)
      tile_foreach32(acc[h], [&](int ii, int pp, double v) {
        const int i = r0 + ii, p = 32 * h + pp;
        if (i < m && p < nb) Lb[(int64_t)(c0 + p) * m + i] = v;
      });
    return;
  }
  TileAcc acc;
  tile_zero(acc);
  mma_tile(
      [&](int ii, int q) {
        const int i = R0 + ii;
        return LV{Lb[(int64_t)(c0 + q) * m + min(i, m - 1)], i < m};
      },
      [&](int pp, int q) {
        return LVs{Zb[q * kTile + pp], dinv[pp], pp < nb};
      },
      nb, As, Bs, acc);
  tile_foreach(acc, [&](int ii, int pp, double v) {
    const int i = R0 + ii;
    if (i < m && pp < nb) Lb[(int64_t)(c0 + pp) * m + i] = v;
  });
}

// Trailing update of panel step `step`: task (s, I, J), I >= J, is the 64x64
// tile at rows t0 + 64 I, columns t0 + 64 J of the trailing lower triangle,
// t0 = 64 step + nb:  C -= L_I D L_J^T on FP64 MFMA.
//
// With DIAG (launches of few tiles), the workgroup of tile (0, 0) goes on to
// factor the front's next diagonal block (k_fl_diag, step + 1), which is that
// tile, when the supernode has a next panel.
template <bool DIAG>
__global__ __launch_bounds__(256) void k_fl_update(DevPlan P, const int32_t* __restrict__ tasks,
                                                   int step, double* __restrict__ L,
                                                   double* __restrict__ U, double* __restrict__ Z) {
  static_assert(2 * 64 * (kPanel + 1) <= kDiagLds, "diagonal block LDS below the engine's");
  __shared__ double sh[DIAG ? kDiagLds : 2 * 64 * (kPanel + 1)];  // the tile engine's operands, then the tile
  double (*As)[kPanel + 1] = reinterpret_cast<double (*)[kPanel + 1]>(sh);
  double (*Bs)[kPanel + 1] = reinterpret_cast<double (*)[kPanel + 1]>(sh + 64 * (kPanel + 1));
  __shared__ double dpiv[kTile];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1], J = tk[2];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int c0 = step * kTile, nb = min(kTile, w - c0), t0 = c0 + nb;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const int R0 = t0 + I * kTile, C0 = t0 + J * kTile;
  if (threadIdx.x < nb) dpiv[threadIdx.x] = Lb[(int64_t)(c0 + threadIdx.x) * m + c0 + threadIdx.x];
  __syncthreads();
  TileAcc acc;
  tile_zero(acc);
  mma_tile(
      [&](int ii, int p) {
        const int i = R0 + ii;
        return LVs{Lb[(int64_t)(c0 + p) * m + min(i, m - 1)], dpiv[p], i < m};
      },
      [&](int jj, int p) {
        const int j = C0 + jj;
        return LV{Lb[(int64_t)(c0 + p) * m + min(j, m - 1)], j < m};
      },
      nb, As, Bs, acc);
  // C -= tile: the tile goes through LDS (the tile engine ends with a barrier) so
  // that thread (row ii = tid % 64) updates columns tid / 64 + 4u: each wave covers
  // 64 consecutive rows of a column (coalesced), and all 16 loads are issued (at
  // clamped valid offsets) before any is used.
  static_assert(kTile * (kTile + 1) <= 2 * 64 * (kPanel + 1), "tile staging exceeds the operand LDS");
  double (*Ct)[kTile + 1] = reinterpret_cast<double (*)[kTile + 1]>(sh);
  tile_foreach(acc, [&](int ii, int jj, double v) { Ct[jj][ii] = v; });
  __syncthreads();
  const int ii = threadIdx.x & 63, jb = threadIdx.x >> 6;
  const int i = R0 + ii, ic = min(i, m - 1);
  int off[16];
  double old[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = C0 + jb + 4 * u;
    const bool ok = i < m && j < m && i >= j;
    const int jc = ok ? j : min(C0, ic), i2 = ok ? i : ic;
    // within-front offsets fit in 32 bits (m w and r^2 below 2^31)
    off[u] = jc >= w ? (jc - w) * r + (i2 - w) : jc * m + i2;
    old[u] = (jc >= w ? Ub : Lb)[off[u]];
  }
  ISSUE_FENCE();
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = C0 + jb + 4 * u;
    const double nv = old[u] - Ct[jb + 4 * u][ii];
    if (i < m && j < m && i >= j) (j >= w ? Ub : Lb)[off[u]] = nv;
  }
  if (DIAG && I == 0 && J == 0 && t0 < w) {
    __syncthreads();  // the tile's stores, before the block is read back
    fl_diag_block(P, s, step + 1, b, L, Z, sh);
  }
}

// log|Q| = sum log|D_j| per matrix, plus the pivot status: an exact zero pivot
// (the numeric phase fails: NotPositiveDefinite) or |D_j| < 1e-14 / non-finite
// (Singular), as crates/inla_math/src/sparse_ldlt.rs:104-149,217-230.
// Grid (kLogdetChunks, B): each block reduces one chunk of the pivots; the last
// block of a matrix to finish (device-scope counter) sums the chunk partials in
// chunk order, writes the result and resets the counter for the next launch.
constexpr int kLogdetChunks = 16;

__global__ __launch_bounds__(256) void k_logdet(DevPlan P, const double* __restrict__ L,
                                                double* __restrict__ part, uint32_t* __restrict__ counter,
                                                double* __restrict__ logdet,
                                                int32_t* __restrict__ status) {
  __shared__ double red[4];
  __shared__ int flags[4];
  __shared__ int last;
  const int ch = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const double* Lm = L + (int64_t)b * P.l_size;
  const int len = (P.n + kLogdetChunks - 1) / kLogdetChunks;
  const int j0 = ch * len, j1 = min(P.n, j0 + len);
  double acc = 0.0;
  int fl = 0;
  for (int jb = j0; jb < j1; jb += 4 * 256) {
    int64_t sl[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) sl[u] = P.diag_slot[min(jb + tid + 256 * u, P.n - 1)];
    ISSUE_FENCE();
    double d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) d[u] = Lm[sl[u]];
    ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (jb + tid + 256 * u < j1) {
        if (d[u] == 0.0) fl |= GMRF_ST_ZERO_PIVOT;
        if (!(fabs(d[u]) >= 1e-14) || !isfinite(d[u])) fl |= GMRF_ST_SINGULAR;
        acc += log(fabs(d[u]));
      }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o, 64);
    fl |= __shfl_xor(fl, o, 64);
  }
  if ((tid & 63) == 0) {
    red[tid >> 6] = acc;
    flags[tid >> 6] = fl;
  }
  __syncthreads();
  if (tid == 0) {
    double* pb = part + ((int64_t)b * kLogdetChunks + ch) * 2;
    pb[0] = red[0] + red[1] + red[2] + red[3];
    pb[1] = (double)(flags[0] | flags[1] | flags[2] | flags[3]);
    __threadfence();
    last = atomicAdd(&counter[b], 1u) == kLogdetChunks - 1;
  }
  __syncthreads();
  if (last && tid == 0) {
    __threadfence();
    const volatile double* pv = part + (int64_t)b * kLogdetChunks * 2;
    double t = 0.0;
    int f = 0;
    for (int q = 0; q < kLogdetChunks; ++q) {
      t += pv[2 * q];
      f |= (int)pv[2 * q + 1];
    }
    logdet[b] = t;
    status[b] = f;
    counter[b] = 0u;
  }
}

// Per-block log|D| of a block-diagonal union (heterogeneous batch): block q of
// matrix b sums the pivots at its factor positions pos[ptr[q] .. ptr[q + 1]) in
// index order; grid (k, B), one workgroup per (block, matrix).
__global__ __launch_bounds__(256) void k_logdet_seg(DevPlan P, const double* __restrict__ L,
                                                    const int64_t* __restrict__ ptr,
                                                    const int32_t* __restrict__ pos, int k,
                                                    double* __restrict__ logdet,
                                                    int32_t* __restrict__ status) {
  __shared__ double red[4];
  __shared__ int flags[4];
  const int q = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const double* Lm = L + (int64_t)b * P.l_size;
  const int64_t p0 = ptr[q], p1 = ptr[q + 1];
  double acc = 0.0;
  int fl = 0;
  for (int64_t jb = p0; jb < p1; jb += 4 * 256) {
    int32_t j[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) j[u] = pos[min(jb + tid + 256 * u, p1 - 1)];
    ISSUE_FENCE();
    double d[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) d[u] = Lm[P.diag_slot[j[u]]];
    ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (jb + tid + 256 * u < p1) {
        if (d[u] == 0.0) fl |= GMRF_ST_ZERO_PIVOT;
        if (!(fabs(d[u]) >= 1e-14) || !isfinite(d[u])) fl |= GMRF_ST_SINGULAR;
        acc += log(fabs(d[u]));
      }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    acc += __shfl_xor(acc, o, 64);
    fl |= __shfl_xor(fl, o, 64);
  }
  if ((tid & 63) == 0) {
    red[tid >> 6] = acc;
    flags[tid >> 6] = fl;
  }
  __syncthreads();
  if (tid == 0) {
    logdet[(int64_t)b * k + q] = red[0] + red[1] + red[2] + red[3];
    status[(int64_t)b * k + q] = flags[0] | flags[1] | flags[2] | flags[3];
  }
}

// ---------------------------------------------------------------------------
// Solves. A launch covers B * nr right-hand sides (grid x = front or task x nr,
// the nr vectors of one matrix side by side so that they share the front's L in
// the caches; grid y = matrix): vector v uses the factor of
// matrix v / nr and its own slices of X, V and the staging area.
// ---------------------------------------------------------------------------
__global__ void k_gather_perm(DevPlan P, int B, const double* __restrict__ rhs, int64_t ldr,
                              int period, double* __restrict__ X) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    const int v = period > 0 ? b % period : b;
    X[e] = rhs[(int64_t)v * ldr + P.perm[k]];
  }
}

__global__ void k_scatter_perm(DevPlan P, int B, const double* __restrict__ X,
                               double* __restrict__ out, int64_t ldo) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    out[(int64_t)b * ldo + P.perm[k]] = X[e];
  }
}

__global__ void k_diag_scale(DevPlan P, int B, int nr, const double* __restrict__ L,
                             double* __restrict__ X) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    X[e] /= L[(int64_t)(b / nr) * P.l_size + P.diag_slot[k]];
  }
}

// Small fronts (m <= 64): a workgroup takes NV right-hand sides of one front and
// matrix, one wave each (lane = front row or column); the front's L block is staged
// once in LDS (dynamic, sized by the launch's largest block) for all of them, and
// the column sweeps run in registers with the pivot value broadcast from its lane
// (no barrier per column). Grid x = front x ceil(nr / NV), y = matrix.
//
// Forward: f = [x_J; 0] + sum_c extend(v_c); y_J = L11^-1 f_J; v_s = f_R - L21 y_J.
template <int NV>
__global__ __launch_bounds__(64 * NV) void k_solve_fwd(DevPlan P, const int32_t* __restrict__ list, int nr,
                                                       const double* __restrict__ L,
                                                       double* __restrict__ X,
                                                       double* __restrict__ V) {
  extern __shared__ double Ls[];  // Ls[k * m + i] = L[i][k], i < m, k < w
  __shared__ double f[NV][kSmallFront];
  const int nvb = (nr + NV - 1) / NV;
  const int fi = blockIdx.x / nvb, vb = blockIdx.x - fi * nvb;
  const int s = list[fi];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int jv = vb * NV + wv;  // this wave's vector of the matrix (jv >= nr: idle, writes nothing)
  const int bl = blockIdx.y, b = bl * nr + min(jv, nr - 1);
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  double* Xb = X + (int64_t)b * P.n;
  const int nl = m * w;
  for (int e = tid; e < nl; e += 64 * NV) Ls[e] = Lb[e];
  {
    const double xv = Xb[a0 + min(lane, w - 1)];
    f[wv][lane] = lane < w ? xv : 0.0;
  }
  __syncthreads();
  for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
    const int c = P.child_idx[cq];
    const int rc = P.sn_r[c];  // rc <= m: the child's rows map into distinct rows of this front
    const double* Vc = V + (int64_t)b * P.v_size + P.v_off[c];
    const int32_t* relc = P.rel + P.rows_ptr[c];
    if (lane < rc) f[wv][relc[lane]] += Vc[lane];
    __syncthreads();
  }
  double fv = f[wv][lane];
  for (int k = 0; k < w; ++k) {
    const double yk = __shfl(fv, k, 64);
    if (lane > k && lane < m) fv -= Ls[k * m + lane] * yk;
  }
  if (jv < nr && lane < m) {
    if (lane < w)
      Xb[a0 + lane] = fv;
    else
      V[(int64_t)b * P.v_size + P.v_off[s] + lane - w] = fv;
  }
}

// Backward: x_J = L11^-T (z_J - L21^T x_R) with x_R gathered from final values
// (lane = column; L staged with an odd leading dimension so that the column-wise
// reads of a wave do not collide in the LDS banks).
template <int NV>
__global__ __launch_bounds__(64 * NV) void k_solve_bwd(DevPlan P, const int32_t* __restrict__ list, int nr,
                                                       const double* __restrict__ L,
                                                       double* __restrict__ X) {
  extern __shared__ double Ls[];  // Ls[j * ld + i] = L[i][j], ld = m | 1
  __shared__ double g[NV][kSmallFront];
  const int nvb = (nr + NV - 1) / NV;
  const int fi = blockIdx.x / nvb, vb = blockIdx.x - fi * nvb;
  const int s = list[fi];
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int jv = vb * NV + wv;
  const int bl = blockIdx.y, b = bl * nr + min(jv, nr - 1);
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r, ld = m | 1;
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  double* Xb = X + (int64_t)b * P.n;
  const int nl = m * w;
  for (int e = tid; e < nl; e += 64 * NV) {
    const int j = e / m, i = e - j * m;
    Ls[j * ld + i] = Lb[e];
  }
  const int32_t* rows = P.rows + P.rows_ptr[s];
  g[wv][lane] = lane < r ? Xb[rows[lane]] : 0.0;
  __syncthreads();
  double t = 0.0;
  if (lane < w) {
    double acc = Xb[a0 + lane];
    for (int i = 0; i < r; ++i) acc -= Ls[lane * ld + w + i] * g[wv][i];
    t = acc;
  }
  for (int k = w - 1; k >= 0; --k) {
    const double xk = __shfl(t, k, 64);
    if (lane < k) t -= Ls[lane * ld + k] * xk;
  }
  if (jv < nr && lane < w) Xb[a0 + lane] = t;
}

// Large fronts, forward. The triangular part (one workgroup per front) solves the
// supernode's own w x w block, blocked by the factorisation's 64-column panels with
// the stored Z = L_PP^-1: y_P = Z f_P, then f_B -= L[B,P] y_P for the rows B below
// the panel. Both products split their K dimension over waves so that each panel
// costs one or two rounds of loads, reduced through LDS. With FULL the workgroup
// also carries the rows R_s below (f of length m, children's updates added in
// full, V_s written): levels of short fronts. Otherwise only the entries landing
// in the supernode's columns are added here, and the rows below are a separate
// row-parallel launch (k_solve_fwd_rows): tall fronts.
template <int NT, bool FULL>
__global__ __launch_bounds__(NT) void k_solve_fwd_tri(DevPlan P, const int32_t* __restrict__ list, int nr,
                                                      const double* __restrict__ L,
                                                      const double* __restrict__ Z,
                                                      double* __restrict__ X,
                                                      double* __restrict__ V) {
  constexpr int NQ = NT / 64, QW = 64 / NQ, NR = NT / 4;  // y: NQ parts of QW columns; rows: NR per round
  extern __shared__ double f[];
  __shared__ double part[NQ][kTile];
  __shared__ double rp[4][NR];
  __shared__ double y[kTile];
  const int fi = blockIdx.x / nr;  // the nr vectors of a front and matrix: consecutive workgroups
  const int s = list[fi];
  const int bl = blockIdx.y, b = bl * nr + (blockIdx.x - fi * nr);
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, m = w + P.sn_r[s];
  const int mr = FULL ? m : w;  // rows carried by this workgroup
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)bl * P.z_size + P.z_off[s];
  double* Xb = X + (int64_t)b * P.n;
  for (int i = tid; i < mr; i += NT) f[i] = i < w ? Xb[a0 + i] : 0.0;
  __syncthreads();
  for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
    const int c = P.child_idx[cq];
    const int rc = P.sn_r[c];
    const double* Vc = V + (int64_t)b * P.v_size + P.v_off[c];
    const int32_t* relc = P.rel + P.rows_ptr[c];
    for (int i = tid; i < rc; i += NT) {
      const int pi = relc[i];
      if (pi < mr) f[pi] += Vc[i];
    }
    __syncthreads();
  }
  const int i = tid & 63, pt = tid >> 6;   // y: row i, q in [QW pt, QW pt + QW)
  const int rr = tid % NR, qp = tid / NR;  // rows below: row c1 + rr (+NR k), q in [16 qp, 16 qp + 16)
  for (int c0 = 0; c0 < w; c0 += kTile) {
    const int nb = min(kTile, w - c0), c1 = c0 + nb;
    const double* Zp = Zb + (int64_t)c0 * kTile;
    {
      double zv[QW];
#pragma unroll
      for (int u = 0; u < QW; ++u) zv[u] = Zp[(QW * pt + u) * kTile + i];
      ISSUE_FENCE();
      double acc = 0.0;
#pragma unroll
      for (int u = 0; u < QW; ++u) {
        const int q = QW * pt + u;
        const double prod = zv[u] * f[c0 + min(q, nb - 1)];
        acc += (q < nb && i < nb) ? prod : 0.0;
      }
      part[pt][i] = acc;
    }
    __syncthreads();
    if (tid < nb) {
      double v = 0.0;
#pragma unroll
      for (int q = 0; q < NQ; ++q) v += part[q][tid];
      y[tid] = v;
      f[c0 + tid] = v;
    }
    __syncthreads();
    for (int row0 = c1; row0 < mr; row0 += NR) {
      const int row = min(row0 + rr, mr - 1);
      double lv[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) lv[u] = Lb[(int64_t)(c0 + min(16 * qp + u, nb - 1)) * m + row];
      ISSUE_FENCE();
      double a = 0.0;
#pragma unroll
      for (int u = 0; u < 16; ++u) {
        const int q = 16 * qp + u;
        a += q < nb ? lv[u] * y[min(q, nb - 1)] : 0.0;
      }
      rp[qp][rr] = a;
      __syncthreads();
      if (tid < NR && row0 + tid < mr) f[row0 + tid] -= rp[0][tid] + rp[1][tid] + rp[2][tid] + rp[3][tid];
      __syncthreads();
    }
  }
  double* Vs = V + (int64_t)b * P.v_size + P.v_off[s];
  for (int k = tid; k < mr; k += NT) {
    if (k < w)
      Xb[a0 + k] = f[k];
    else
      Vs[k - w] = f[k];
  }
}

// Forward, rows below a large supernode: task (s, chunk) forms the update entries
// of rows 64 chunk .. of R_s, V_s[i] = sum_c V_c[cinv_c(w + i)] - sum_j L[w + i, j] y_j,
// children in fixed order, the w columns split over the 4 waves (lane = row:
// coalesced column segments), reduced through LDS in fixed order.
__global__ __launch_bounds__(256) void k_solve_fwd_rows(DevPlan P, const int32_t* __restrict__ tasks, int nr,
                                                        const double* __restrict__ L,
                                                        const double* __restrict__ X,
                                                        double* __restrict__ V) {
  extern __shared__ double yv[];
  __shared__ double red[4][kTile];
  const int fi = blockIdx.x / nr;  // the nr vectors of a task and matrix: consecutive workgroups
  const int32_t* tk = tasks + 3 * fi;
  const int s = tk[0], ch = tk[1];
  const int bl = blockIdx.y, b = bl * nr + (blockIdx.x - fi * nr);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  const double* Xb = X + (int64_t)b * P.n;
  for (int j = tid; j < w; j += 256) yv[j] = Xb[a0 + j];
  __syncthreads();
  const int i = ch * kTile + lane;
  const int row = w + min(i, r - 1);
  double acc = 0.0;
  for (int j0 = wv; j0 < w; j0 += 64) {
    double lv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) lv[u] = Lb[(int64_t)min(j0 + 4 * u, w - 1) * m + row];
    ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int j = j0 + 4 * u;
      acc += j < w ? lv[u] * yv[min(j, w - 1)] : 0.0;
    }
  }
  red[wv][lane] = acc;
  __syncthreads();
  if (tid < kTile && i < r) {
    double c = 0.0;
    for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
      const int cc = P.child_idx[cq];
      const int ci = P.cinv[P.cinv_off[cq] + w + i];
      if (ci >= 0) c += V[(int64_t)b * P.v_size + P.v_off[cc] + ci];
    }
    V[(int64_t)b * P.v_size + P.v_off[s] + i] = c - (red[0][tid] + red[1][tid] + red[2][tid] + red[3][tid]);
  }
}

// Backward, rows below a large supernode: task (s, cc, part) forms the partial sums
// t[part][q] = sum_{i in part} L[w + i, q] x_{R_s[i]} for the 64 columns q of chunk
// cc, over rows 1024 part .. of R_s; wave wv covers columns 64 cc + 16 wv + u
// (lane = row), reduced across lanes by shuffles. Partials go to the front's staging
// area (parts x w), summed in fixed order by k_solve_bwd_tri.
__global__ __launch_bounds__(256) void k_solve_bwd_rows(DevPlan P, const int32_t* __restrict__ tasks, int nr,
                                                        const double* __restrict__ L,
                                                        const double* __restrict__ X,
                                                        double* __restrict__ T) {
  const int fi = blockIdx.x / nr;  // the nr vectors of a task and matrix: consecutive workgroups
  const int32_t* tk = tasks + 3 * fi;
  const int s = tk[0], cc = tk[1], part = tk[2];
  const int bl = blockIdx.y, b = bl * nr + (blockIdx.x - fi * nr);
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  const double* Xb = X + (int64_t)b * P.n;
  const int32_t* rows = P.rows + P.rows_ptr[s];
  const int q0 = cc * kTile + 16 * wv;
  const int i1 = min(r, (part + 1) * kSolvePart);
  double acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0;
  for (int i = part * kSolvePart + lane; i < i1 + 63; i += 64) {
    const int ic = min(i, r - 1);
    const double xi = Xb[rows[ic]];
    double lv[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) lv[u] = Lb[(int64_t)min(q0 + u, w - 1) * m + w + ic];
    ISSUE_FENCE();
    const double xs = i < i1 ? xi : 0.0;
#pragma unroll
    for (int u = 0; u < 16; ++u) acc[u] += lv[u] * xs;
  }
  double* Tb = T + (int64_t)b * P.ts_size + P.ts_off[s] + (int64_t)part * w;
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    double a = acc[u];
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
    if (lane == 0 && q0 + u < w) Tb[q0 + u] = a;
  }
}

// Backward, triangular part of a large front, panels right to left:
// t_P = x_P - sum_parts (L[R,P]^T x_R) - L[B,P]^T x_B (B = rows below the panel
// inside the supernode), x_P = Z^T t_P. With FULL, B extends over R_s (x_R
// gathered into LDS) and there are no partial sums. Wave wv covers panel columns
// q = (wv & 3) + 4u (u < 16) over rows c1 + 64 (wv >> 2) + lane (+NT / 4 k).
template <int NT, bool FULL>
__global__ __launch_bounds__(NT) void k_solve_bwd_tri(DevPlan P, const int32_t* __restrict__ list, int nr,
                                                      const double* __restrict__ L,
                                                      const double* __restrict__ Z,
                                                      double* __restrict__ X,
                                                      const double* __restrict__ T) {
  // wave wv owns the panel columns q = wv + NW u (u < CPW) over all rows below the
  // panel (lane = row, RB row blocks of 64 per round of loads), so a column's sum
  // is one wave's lane reduction
  constexpr int NW = NT / 64, CPW = 64 / NW, RB = 16 / CPW;
  extern __shared__ double xf[];
  __shared__ double red[kTile];
  __shared__ double tq[kTile];
  const int fi = blockIdx.x / nr;  // the nr vectors of a front and matrix: consecutive workgroups
  const int s = list[fi];
  const int bl = blockIdx.y, b = bl * nr + (blockIdx.x - fi * nr);
  const int tid = threadIdx.x, lane = tid & 63, wv = tid >> 6;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)bl * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)bl * P.z_size + P.z_off[s];
  double* Xb = X + (int64_t)b * P.n;
  const double* Tb = T + (int64_t)b * P.ts_size + P.ts_off[s];
  const int mr = FULL ? m : w;
  if (FULL) {
    const int32_t* rows = P.rows + P.rows_ptr[s];
    for (int k = tid; k < m; k += NT) xf[k] = k < w ? Xb[a0 + k] : Xb[rows[k - w]];
  } else {
    const int nparts = (r + kSolvePart - 1) / kSolvePart;
    for (int k = tid; k < w; k += NT) {
      double v = Xb[a0 + k];
      for (int q = 0; q < nparts; ++q) v -= Tb[(int64_t)q * w + k];
      xf[k] = v;
    }
  }
  __syncthreads();
  const int np = (w + kTile - 1) / kTile;
  for (int pk = np - 1; pk >= 0; --pk) {
    const int c0 = pk * kTile, nb = min(kTile, w - c0), c1 = c0 + nb;
    double acc[CPW];
#pragma unroll
    for (int u = 0; u < CPW; ++u) acc[u] = 0.0;
    for (int k0 = c1 + lane; k0 < mr + 63; k0 += 64 * RB) {
      double lv[RB][CPW], xk[RB];
#pragma unroll
      for (int j = 0; j < RB; ++j) {
        const int k = k0 + 64 * j, kc = min(k, mr - 1);
        xk[j] = k < mr ? xf[kc] : 0.0;
#pragma unroll
        for (int u = 0; u < CPW; ++u) lv[j][u] = Lb[(int64_t)(c0 + min(wv + NW * u, nb - 1)) * m + kc];
      }
      ISSUE_FENCE();
#pragma unroll
      for (int j = 0; j < RB; ++j)
#pragma unroll
        for (int u = 0; u < CPW; ++u) acc[u] += lv[j][u] * xk[j];
    }
#pragma unroll
    for (int u = 0; u < CPW; ++u) {
      double a = acc[u];
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
      if (lane == 0) red[wv + NW * u] = a;
    }
    __syncthreads();
    if (tid < nb) tq[tid] = xf[c0 + tid] - red[tid];
    __syncthreads();
    const double* Zp = Zb + (int64_t)pk * kTile * kTile;
    double zv[64 / NW];
#pragma unroll
    for (int u = 0; u < 64 / NW; ++u) zv[u] = Zp[(wv + NW * u) * kTile + lane];
    ISSUE_FENCE();
    const double tl = lane < nb ? tq[lane] : 0.0;
#pragma unroll
    for (int u = 0; u < 64 / NW; ++u) {
      double a = lane < nb ? zv[u] * tl : 0.0;
#pragma unroll
      for (int o = 32; o > 0; o >>= 1) a += __shfl_xor(a, o, 64);
      const int q = wv + NW * u;
      if (lane == 0 && q < nb) xf[c0 + q] = a;
    }
    __syncthreads();
  }
  for (int k = tid; k < w; k += NT) Xb[a0 + k] = xf[k];
}

// ---------------------------------------------------------------------------
// Selected inversion (roots to leaves). For supernode s with parent p, the
// entries Sigma_{R_s R_s} are gathered from p's front (p's S block and p's
// Sigma_RR in the U workspace) through rel, then the columns of s are
// processed right to left: Sigma[i, k] = -sum_{l > k} Sigma[i, l] L[l, k] and
// Sigma[k, k] = 1/D_k - sum_{l > k} L[l, k] Sigma[l, k]  (Takahashi et al.
// 1973; the same recurrence as crates/inla_math/src/sparse_ldlt.rs:376-419,
// evaluated densely inside the front).
// ---------------------------------------------------------------------------

// One wave: Zk = L_PP^-1 of the unit lower kn x kn block at (c0, c0) of the
// row-major Ls (stride ldl), identity beyond kn (16 x 17), on the MFMA pipe: Z in
// the accumulator layout (lane 16 g + c, register v: Z[4 v + g][c]); per 4-row step
// p the pivot rows become Y Z_p (Y = inverse of the 4 x 4 diagonal block, from
// uniform reads) and the rows below take -L[below, p] Z_p as one 16x16x4 MFMA.
__device__ __forceinline__ void unit_lower_inv16_mfma(const double* Ls, int ldl, int c0, int kn,
                                                      double* Zk) {
  const int lane = threadIdx.x & 63, g = lane >> 4, c = lane & 15;
  auto lat = [&](int i, int j) {  // L[i][j] of the block, 0 beyond kn
    return (i < kn && j < kn) ? Ls[(c0 + min(i, kn - 1)) * ldl + c0 + min(j, kn - 1)] : 0.0;
  };
  d4 Zv;
#pragma unroll
  for (int v = 0; v < 4; ++v) Zv[v] = 4 * v + g == c ? 1.0 : 0.0;
#pragma unroll
  for (int p = 0; p < 4; ++p) {
    const int q0 = 4 * p;
    if (q0 >= kn) break;
    const double l10 = lat(q0 + 1, q0), l20 = lat(q0 + 2, q0), l21 = lat(q0 + 2, q0 + 1);
    const double l30 = lat(q0 + 3, q0), l31 = lat(q0 + 3, q0 + 1), l32 = lat(q0 + 3, q0 + 2);
    const bool below = c >= q0 + 4;
    const double lcg = below ? lat(c, q0 + g) : 0.0;
    double z[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) z[a] = __shfl(Zv[p], 16 * a + c, 64);
    const double y10 = -l10, y21 = -l21, y32 = -l32;
    const double y20 = -(l20 + l21 * y10), y31 = -(l31 + l32 * y21);
    const double y30 = -(l30 + l31 * y10 + l32 * y20);
    const double ya0 = g == 1 ? y10 : (g == 2 ? y20 : y30);
    const double ya1 = g == 2 ? y21 : y31;
    double zn = z[g == 0 ? 0 : (g == 1 ? 1 : (g == 2 ? 2 : 3))];
    zn = g >= 1 ? zn + ya0 * z[0] : zn;
    zn = g >= 2 ? zn + ya1 * z[1] : zn;
    zn = g >= 3 ? zn + y32 * z[2] : zn;
    Zv[p] = zn;
    Zv = mfma16(-lcg, zn, Zv);
  }
#pragma unroll
  for (int v = 0; v < 4; ++v) Zk[(4 * v + g) * 17 + c] = Zv[v];
}

// Small fronts (m <= MB), 4 MB threads, blocked by 16 columns from the right.
// The whole Sigma front (both triangles) and the L block sit in LDS; Sigma_{R R}
// is gathered from the parent first. For block P = [c0, c1) and B = [c1, m):
//   Zk = L_PP^-1 (one wave), Lh = L[B,P] Zk, Sigma[B,P] = -Sigma[B,B] Lh,
//   Sigma[P,P] = Zk^T D_P^-1 Zk - Lh^T Sigma[B,P]
// (the block form of the Takahashi recurrence), the products on FP64 MFMA.
template <int MB>
__global__ __launch_bounds__(4 * MB) void k_selinv_small(DevPlan P, const int32_t* __restrict__ list,
                                                         const double* __restrict__ L,
                                                         double* __restrict__ S,
                                                         double* __restrict__ U,
                                                         const double* __restrict__ W,
                                                         int ldl) {
  constexpr int NT = 4 * MB, NJ = MB / 4, LD = MB + 1, NW = MB / 16;
  __shared__ double Sg[MB * LD];
  extern __shared__ double Ls[];  // MB x ldl: the L block, ldl > widest supernode of the launch
  __shared__ double Zk[16 * 17];
  __shared__ double Lh[MB * 17];
  __shared__ double dinv[16];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int vr = lane >> 4, vc = lane & 15;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Sm = S + (int64_t)b * P.l_size;
  double* Um = U + (int64_t)b * P.u_size;
  const double* Wm = W + (int64_t)b * P.w_size;
  // the parent's metadata and this front's row map are loaded first, so that their
  // latency overlaps the L block's loads
  const int i = tid % MB, c = tid / MB;
  const int p = max(P.sn_parent[s], 0);
  const int wp = P.sn_start[p + 1] - P.sn_start[p], rp = P.sn_r[p], mp = wp + rp;
  const bool plarge = P.is_large[p] != 0;
  const int64_t woff = P.w_off[p], loff = P.l_off[p], uoff = P.u_off_sel[p];
  const int32_t* rl = P.rel + P.rows_ptr[s];
  int ri = 0, rj[NJ];
  if (r > 0) {
    ri = rl[min(max(i - w, 0), r - 1)];
#pragma unroll
    for (int t = 0; t < NJ; ++t) rj[t] = rl[min(max(c + 4 * t - w, 0), r - 1)];
  }
  // L block into LDS, row-major
  const int nl = m * w;
  for (int e0 = 0; e0 < nl; e0 += 8 * NT) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) v[u] = Lb[min(e0 + tid + NT * u, nl - 1)];
    ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + tid + NT * u;
      const int k = e / m, i = e - k * m;
      if (e < nl) Ls[i * ldl + k] = v[u];
    }
  }
  // Sigma_{R R} from the parent front (its S block and Sigma_RR workspace, or
  // its dense Sigma front when the parent is large); the lower part is also
  // kept in this front's U workspace for the children.
  if (r > 0) {
    double v[NJ];
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      const int pi = max(ri, rj[t]), pj = min(ri, rj[t]);
      const double* base;
      int64_t off;
      if (pj < wp) {  // the parent's own columns: its S block
        base = Sm;
        off = loff + (int64_t)pj * mp + pi;
      } else if (plarge) {  // its R x R block: the lower triangle of its dense Sigma front
        base = Wm;
        off = woff + (int64_t)pj * mp + pi;
      } else {
        base = Um;
        off = uoff + (int64_t)(pj - wp) * rp + (pi - wp);
      }
      v[t] = base[off];
    }
    ISSUE_FENCE();
    double* Us = Um + P.u_off_sel[s];
#pragma unroll
    for (int t = 0; t < NJ; ++t) {
      const int j = c + 4 * t;
      if (i >= w && i < m && j >= w && j < m) {
        Sg[i * LD + j] = v[t];
        if (i >= j) Us[(int64_t)(j - w) * r + (i - w)] = v[t];
      }
    }
  }
  __syncthreads();
  for (int c0 = ((w - 1) >> 4) << 4; c0 >= 0; c0 -= 16) {
    const int c1 = min(w, c0 + 16), kn = c1 - c0, nB = m - c1;
    const int nrb = (nB + 15) >> 4;
    if (wv == 0 && MB >= 48) {
      unit_lower_inv16_mfma(Ls, ldl, c0, kn, Zk);
      if (lane < 16) {
        const int q = c0 + min(lane, kn - 1);
        dinv[lane] = 1.0 / Ls[q * ldl + q];
      }
    } else if (wv == 0) {
      // Zk = L_PP^-1 by Gauss transforms; lane (i, c) holds Zk[i][c + 4t]
      const int i = lane & 15, c = lane >> 4;
      double z[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) z[t] = (i == c + 4 * t) ? 1.0 : 0.0;
      for (int k = 0; k < kn; ++k) {
        const double lik = Ls[(c0 + min(i, kn - 1)) * ldl + c0 + k];
        double zk[4];
#pragma unroll
        for (int t = 0; t < 4; ++t) zk[t] = __shfl(z[t], 16 * c + k, 64);
        const bool act = i > k && i < kn;
#pragma unroll
        for (int t = 0; t < 4; ++t) {
          const double nz = z[t] - lik * zk[t];
          z[t] = (act && c + 4 * t <= k) ? nz : z[t];
        }
      }
#pragma unroll
      for (int t = 0; t < 4; ++t) Zk[i * 17 + c + 4 * t] = z[t];
      if (lane < 16) {
        const int q = c0 + min(lane, kn - 1);
        dinv[lane] = 1.0 / Ls[q * ldl + q];
      }
    }
    __syncthreads();
    // Lh = L[B,P] Zk
    for (int task = wv; task < nrb; task += NW) {
      const int r0 = 16 * task;
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      acc = wgemm16(
          [&](int ii, int q) {
            return LV{Ls[(c1 + min(r0 + ii, nB - 1)) * ldl + c0 + min(q, kn - 1)], (r0 + ii < nB && q < kn)};
          },
          [&](int q, int pp) {
            return LV{Zk[q * 17 + pp], q < kn};
          },
          acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) Lh[(r0 + 4 * v + vr) * 17 + vc] = acc[v];
    }
    __syncthreads();
    // Sigma[B,P] = -Sigma[B,B] Lh, both triangles
    for (int task = wv; task < nrb; task += NW) {
      const int r0 = 16 * task;
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      for (int k0 = 0; k0 < nB; k0 += 16)
        acc = wgemm16(
            [&](int ii, int kk) {
              return LV{Sg[(c1 + min(r0 + ii, nB - 1)) * LD + c1 + min(k0 + kk, nB - 1)], (r0 + ii < nB && k0 + kk < nB)};
            },
            [&](int kk, int pp) {
              return LV{Lh[min(k0 + kk, nB - 1) * 17 + pp], k0 + kk < nB};
            },
            acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int i = r0 + 4 * v + vr;
        if (i < nB && vc < kn) {
          Sg[(c1 + i) * LD + c0 + vc] = -acc[v];
          Sg[(c0 + vc) * LD + c1 + i] = -acc[v];
        }
      }
    }
    __syncthreads();
    // Sigma[P,P] = Zk^T D^-1 Zk - Lh^T Sigma[B,P]
    if (wv == 0) {
      d4 acc = {0.0, 0.0, 0.0, 0.0};
      acc = wgemm16(
          [&](int pp, int k) {
            return LVs{Zk[k * 17 + pp], dinv[k], k < kn};
          },
          [&](int k, int qq) {
            return LV{Zk[k * 17 + qq], k < kn};
          },
          acc);
      for (int k0 = 0; k0 < nB; k0 += 16)
        acc = wgemm16(
            [&](int pp, int ii) {
              const double x = Lh[min(k0 + ii, nB - 1) * 17 + pp];
              return k0 + ii < nB ? -x : 0.0;
            },
            [&](int ii, int qq) {
              return LV{Sg[(c1 + min(k0 + ii, nB - 1)) * LD + c0 + qq], k0 + ii < nB};
            },
            acc);
#pragma unroll
      for (int v = 0; v < 4; ++v) {
        const int pp = 4 * v + vr;
        if (pp < kn && vc < kn) Sg[(c0 + pp) * LD + c0 + vc] = acc[v];
      }
    }
    __syncthreads();
  }
  double* Sb = Sm + P.l_off[s];
  for (int idx = tid; idx < m * w; idx += NT) {
    const int j = idx / m, i = idx - j * m;
    Sb[(int64_t)j * m + i] = (i >= j) ? Sg[i * LD + j] : 0.0;
  }
}

// ---------------------------------------------------------------------------
// Selected inversion of large fronts. Each large front keeps its whole Sigma
// front (m x m, both triangles) in the W workspace: Sigma_{R R} is gathered
// from the parent first, then panels of 64 columns are processed right to
// left. For panel P = [c0, c1) and B = [c1, m):
//   Lh = L[B,P] L_PP^-1,  Sigma[B,P] = -Sigma[B,B] Lh,
//   Sigma[P,P] = L_PP^-T D_P^-1 L_PP^-1 - Lh^T Sigma[B,P]
// (the block form of the Takahashi recurrence).
// ---------------------------------------------------------------------------

// Gather tile (I, J), I >= J, of Sigma_{R_s R_s} into the lower triangle of W.
// Sigma of a large front lives in two places, both lower triangles: its own w
// columns in S, the R x R block in W (W's first w columns are never written).
__device__ __forceinline__ void sl_gather_task(const DevPlan& P, const int32_t* tk, int b,
                                               const double* __restrict__ S,
                                               const double* __restrict__ U, double* __restrict__ W) {
  const int s = tk[0], I = tk[1], J = tk[2];
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int p = P.sn_parent[s];
  const int wp = P.sn_start[p + 1] - P.sn_start[p], rp = P.sn_r[p], mp = wp + rp;
  const bool plarge = P.is_large[p] != 0;
  const int64_t woff = P.w_off[p], loff = P.l_off[p], uoff = P.u_off_sel[p];
  const int32_t* rl = P.rel + P.rows_ptr[s];
  const double* Sm = S + (int64_t)b * P.l_size;
  const double* Um = U + (int64_t)b * P.u_size;
  double* Wm = W + (int64_t)b * P.w_size;
  double* Ws = Wm + P.w_off[s];
  const int ii = threadIdx.x & 63, jb = threadIdx.x >> 6;
  const int i = I * kTile + ii;
  const int ri = rl[min(i, r - 1)];
  int rj[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) rj[u] = rl[min(J * kTile + jb + 4 * u, r - 1)];
  double v[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int pi = max(ri, rj[u]), pj = min(ri, rj[u]);
    const double* base;
    int64_t off;
    if (pj < wp) {  // the parent's own columns: its S block
      base = Sm;
      off = loff + (int64_t)pj * mp + pi;
    } else if (plarge) {  // its R x R block: the lower triangle of its W
      base = Wm;
      off = woff + (int64_t)pj * mp + pi;
    } else {
      base = Um;
      off = uoff + (int64_t)(pj - wp) * rp + (pi - wp);
    }
    v[u] = base[off];
  }
  ISSUE_FENCE();
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int j = J * kTile + jb + 4 * u;
    if (i < r && j < r && i >= j) Ws[(int64_t)(w + j) * m + (w + i)] = v[u];
  }
}

// Selected-inversion panels reuse the factorisation's 64-column partition,
// walked right to left: selinv step k handles panel index np - 1 - k.
__device__ __forceinline__ void sel_panel(int w, int step, int& c0, int& c1, int& pk) {
  const int np = (w + kTile - 1) / kTile;
  pk = np - 1 - step;
  c0 = pk * kTile;
  c1 = min(w, c0 + kTile);
}

// Lh rows of panel P, as a product with the stored Z = L_PP^-1: task (s, I)
// computes rows c1 + 64 I .. of Lh[i][p] = sum_q L[i][q] Z[q][p] into the panel's
// L_hat buffer of the staging area.
__device__ __forceinline__ void sl_lhat_task(const DevPlan& P, const int32_t* tk, int step, int b,
                                             const double* __restrict__ L,
                                             const double* __restrict__ Z, double* __restrict__ T,
                                             double* sh) {
  const int s = tk[0], I = tk[1];
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int R0 = c1 + I * kTile;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  double* Lh = T + (int64_t)b * P.t_size + P.t_off[s] + stage_lhat_off(m, step);
  TileAcc acc;
  tile_zero(acc);
  tile_product<false, true>(
      [&](int ii, int q) {
        const int i = R0 + ii;
        return LV{Lb[(int64_t)(c0 + q) * m + min(i, m - 1)], i < m};
      },
      [&](int pp, int q) {
        return LV{Zb[pp * kTile + q], pp < nb};
      },
      nb, sh, acc);
  if (threadIdx.x < 256)
    tile_foreach(acc, [&](int ii, int pp, double v) {
      const int i = R0 + ii;
      if (i < m && pp < nb) Lh[(int64_t)pp * m + i] = v;
    });
}

__global__ __launch_bounds__(256) void k_sl_lhat(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, const double* __restrict__ L,
                                                 const double* __restrict__ Z,
                                                 double* __restrict__ T) {
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  sl_lhat_task(P, tasks + 3 * blockIdx.x, step, blockIdx.y, L, Z, T, sh);
}

// The gather of a level and the L_hat of its first panel step in one launch: the
// first ng workgroups gather Sigma_{R R} tiles (they write the R x R part of W),
// the rest form L_hat (from L and Z into the staging area); the two sets touch
// disjoint memory.
__global__ __launch_bounds__(256) void k_sl_gather_lhat(DevPlan P, const int32_t* __restrict__ g_tasks,
                                                        int ng, const int32_t* __restrict__ l_tasks,
                                                        const double* __restrict__ S,
                                                        const double* __restrict__ U,
                                                        double* __restrict__ W,
                                                        const double* __restrict__ L,
                                                        const double* __restrict__ Z,
                                                        double* __restrict__ T) {
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  if ((int)blockIdx.x < ng)
    sl_gather_task(P, g_tasks + 3 * blockIdx.x, blockIdx.y, S, U, W);
  else
    sl_lhat_task(P, l_tasks + 3 * (blockIdx.x - ng), 0, blockIdx.y, L, Z, T, sh);
}

// Workgroup -> (task, matrix) of a launch over `count` tasks and B matrices on a 1-D
// grid. Workgroups are dispatched to the 8 XCDs round-robin by linear id, and each
// XCD has its own L2. With xcd set, every task of matrix b runs on XCD b % 8 (grid
// 8 ceil(B / 8) count; the workgroups past the last matrix return at once), so that
// the tiles of one front share their operands through one L2: an off-diagonal
// entry of Sigma is read by two row tiles, and every tile reads the whole L_hat
// panel. Without it the grid is count x B in the usual order.
__device__ __forceinline__ bool wg_task(int count, int B, int xcd, int& task, int& b) {
  const int L = blockIdx.x;
  if (xcd) {
    const int x = L & 7, j = L >> 3, q = j / count;
    task = j - q * count;
    b = x + 8 * q;
    return b < B;
  }
  b = L / count;
  task = L - b * count;
  return true;
}

// acc += Sigma[R0 .. R0 + 63][k_lo .. k_hi) Lh[k_lo .. k_hi)[0 .. nb) from the lower
// triangle of Sigma only (the front's own columns in S, the R x R block in W, the
// same column stride m). Up to the end of the tile's diagonal block the entries are
// read down their columns (row max(i, k) of column min(i, k): lanes along rows);
// beyond it every entry lies above the tile's rows and is read as row k of column i
// (lanes along k). Both ranges go through the tile engine on one accumulator.
template <int NT>
__device__ __forceinline__ void sigma_strip(const double* __restrict__ Sb, const double* __restrict__ Ws,
                                            const double* __restrict__ Lh, int w, int m, int nb, int R0,
                                            int k_lo, int k_hi, double (*As)[kPanel + 1],
                                            double (*Bs)[kPanel + 1], TileAcc& acc) {
  const int kd = min(max(R0 + kTile, k_lo), k_hi);
  if (kd > k_lo)
    mma_tile<false, true, NT>(
        [&](int ii, int kk) {
          const int i = min(R0 + ii, m - 1), k = k_lo + kk;
          const int hi = max(i, k), lo = min(i, k);
          return LV{(lo < w ? Sb : Ws)[(int64_t)lo * m + hi], R0 + ii < m};
        },
        [&](int jj, int kk) {
          return LV{Lh[(int64_t)min(jj, nb - 1) * m + k_lo + kk], jj < nb};
        },
        kd - k_lo, As, Bs, acc);
  if (k_hi > kd)
    mma_tile<true, true, NT>(
        [&](int ii, int kk) {
          const int i = min(R0 + ii, m - 1);
          return LV{(i < w ? Sb : Ws)[(int64_t)i * m + kd + kk], R0 + ii < m};
        },
        [&](int jj, int kk) {
          return LV{Lh[(int64_t)min(jj, nb - 1) * m + kd + kk], jj < nb};
        },
        k_hi - kd, As, Bs, acc);
}

// Sigma[B,P]: task (s, I) computes rows c1 + 64 I .. of B as
// -sum_{k in B} Sigma[i][k] Lh[k][p] (K = m - c1, lower triangle only) and stores it
// into S (column c0 + p, the layout of L).
__global__ __launch_bounds__(512) void k_sl_gemm(DevPlan P, const int32_t* __restrict__ tasks, int count,
                                                 int B, int xcd, int step, double* __restrict__ S,
                                                 const double* __restrict__ W,
                                                 const double* __restrict__ T) {
  constexpr int NT = 512;
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  double (*As)[kPanel + 1] = (double (*)[kPanel + 1])sh;
  double (*Bs)[kPanel + 1] = (double (*)[kPanel + 1])(sh + 64 * (kPanel + 1));
  int task, b;
  if (!wg_task(count, B, xcd, task, b)) return;
  const int32_t* tk = tasks + 3 * task;
  const int s = tk[0], I = tk[1];
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int R0 = c1 + I * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Ws = W + (int64_t)b * P.w_size + P.w_off[s];
  const double* Lh = T + (int64_t)b * P.t_size + P.t_off[s] + stage_lhat_off(m, step);
  TileAcc acc;
  tile_zero(acc);
  sigma_strip<NT>(Sb, Ws, Lh, w, m, nb, R0, c1, m, As, Bs, acc);
  // waves 0-3 hold the result
  if (threadIdx.x < 256)
    tile_foreach(acc, [&](int ii, int jj, double v) {
      const int i = R0 + ii;
      if (i < m && jj < nb) Sb[(int64_t)(c0 + jj) * m + i] = -v;
    });
}

// sigma_strip on a 32 x 32 block: rows R0 .. R0 + 31 and the panel columns q0 .. q0 + 31
// of Lh (mma_tile32), the diagonal block boundary at R0 + 32.
__device__ __forceinline__ void sigma_strip32(const double* __restrict__ Sb, const double* __restrict__ Ws,
                                              const double* __restrict__ Lh, int w, int m, int nb, int R0,
                                              int q0, int k_lo, int k_hi, double* sh, d4& acc) {
  const int kd = min(max(R0 + 32, k_lo), k_hi), nq = nb - q0;
  if (kd > k_lo)
    mma_tile32<false, true>(
        [&](int ii, int kk) {
          const int i = min(R0 + ii, m - 1), k = k_lo + kk;
          const int hi = max(i, k), lo = min(i, k);
          return LV{(lo < w ? Sb : Ws)[(int64_t)lo * m + hi], R0 + ii < m};
        },
        [&](int jj, int kk) {
          return LV{Lh[(int64_t)(q0 + min(jj, nq - 1)) * m + k_lo + kk], jj < nq};
        },
        kd - k_lo, sh, acc);
  if (k_hi > kd)
    mma_tile32<true, true>(
        [&](int ii, int kk) {
          const int i = min(R0 + ii, m - 1);
          return LV{(i < w ? Sb : Ws)[(int64_t)i * m + kd + kk], R0 + ii < m};
        },
        [&](int jj, int kk) {
          return LV{Lh[(int64_t)(q0 + min(jj, nq - 1)) * m + kd + kk], jj < nq};
        },
        k_hi - kd, sh, acc);
}

// k_sl_gemm with each task's 64 x 64 tile as four 32 x 32 workgroups (task 4 t + part:
// rows 32 (part & 1) .., columns 32 (part >> 1) ..), for launches of few tiles; the
// same XCD grouping over the 4 count block tasks.
__global__ __launch_bounds__(256) void k_sl_gemm32(DevPlan P, const int32_t* __restrict__ tasks, int count,
                                                   int B, int xcd, int step, double* __restrict__ S,
                                                   const double* __restrict__ W,
                                                   const double* __restrict__ T) {
  __shared__ double sh[2 * 32 * (kPanel + 1)];
  int t4, b;
  if (!wg_task(4 * count, B, xcd, t4, b)) return;
  const int32_t* tk = tasks + 3 * (t4 >> 2);
  const int part = t4 & 3;
  const int s = tk[0], I = tk[1];
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int R0 = c1 + I * kTile + (part & 1) * 32, q0 = (part >> 1) * 32;
  if (R0 >= m || q0 >= nb) return;  // uniform: the block lies beyond the front or the panel
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Ws = W + (int64_t)b * P.w_size + P.w_off[s];
  const double* Lh = T + (int64_t)b * P.t_size + P.t_off[s] + stage_lhat_off(m, step);
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  sigma_strip32(Sb, Ws, Lh, w, m, nb, R0, q0, c1, m, sh, acc);
  tile_foreach32(acc, [&](int ii, int jj, double v) {
    const int i = R0 + ii, j = q0 + jj;
    if (i < m && j < nb) Sb[(int64_t)(c0 + j) * m + i] = -v;
  });
}

// Sigma[P,P]: task (s) = Z^T D^-1 Z - Lh^T Sigma[B,P], both products on MFMA
// (K = nb, then K = m - c1; Sigma[B,P] read back from S); lower part into S. The
// strict upper triangle of S's diagonal blocks is not part of the layout's
// contract and is left as is.
__device__ __forceinline__ void sl_pp_task(const DevPlan& P, const int32_t* tk, int step, int b,
                                           const double* __restrict__ L, const double* __restrict__ Z,
                                           double* __restrict__ S, const double* __restrict__ T, double* sh,
                                           double* dinv) {
  const int s = tk[0];
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Lh = T + (int64_t)b * P.t_size + P.t_off[s] + stage_lhat_off(m, step);
  if (tid < kTile) {
    const int q = min(tid, nb - 1);
    dinv[tid] = 1.0 / Lb[(int64_t)(c0 + q) * m + c0 + q];
  }
  __syncthreads();
  TileAcc acc;
  tile_zero(acc);
  // (Z^T D^-1 Z)[p][q] = sum_k Z[k][p] Z[k][q] / d_k, Zb[q * 64 + k] = Z[k][q]
  tile_product<true, true>(
      [&](int pp, int k) {
        return LVs{Zb[pp * kTile + k], dinv[k], pp < nb};
      },
      [&](int qq, int k) {
        return LV{Zb[qq * kTile + k], qq < nb};
      },
      nb, sh, acc);
  if (m > c1) {
    // acc - Lh^T Sigma[B,P] as -(-acc + Lh^T Sigma[B,P]) (negation is exact)
    tile_negate(acc);
    tile_product<true, true>(
        [&](int ii, int kk) {
          return LV{Lh[(int64_t)min(ii, nb - 1) * m + c1 + kk], ii < nb};
        },
        [&](int jj, int kk) {
          return LV{Sb[(int64_t)(c0 + min(jj, nb - 1)) * m + c1 + kk], jj < nb};
        },
        m - c1, sh, acc);
    tile_negate(acc);
  }
  if (threadIdx.x < 256)
    tile_foreach(acc, [&](int ii, int jj, double v) {
      // tile row ii = p, tile column jj = q; lower part p >= q
      if (ii < nb && jj < nb && ii >= jj) Sb[(int64_t)(c0 + jj) * m + c0 + ii] = v;
    });
}

// sl_pp_task as three 32 x 32 blocks of the lower triangle of the 64 x 64 Sigma[P,P]
// (part 0: rows and columns 0-31; part 1: rows 32-63, columns 0-31; part 2: rows and
// columns 32-63), one workgroup each: a quarter of the tile's MFMA work per workgroup
// and the strictly upper block skipped. Same products and stores as sl_pp_task.
__device__ __forceinline__ void sl_pp_task32(const DevPlan& P, const int32_t* tk, int part, int step, int b,
                                             const double* __restrict__ L, const double* __restrict__ Z,
                                             double* __restrict__ S, const double* __restrict__ T, double* sh,
                                             double* dinv) {
  const int s = tk[0];
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int p0 = part == 0 ? 0 : 32, q0 = part == 2 ? 32 : 0;
  if (p0 >= nb) return;  // uniform: the block lies beyond the panel
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Lh = T + (int64_t)b * P.t_size + P.t_off[s] + stage_lhat_off(m, step);
  if (tid < kTile) {
    const int q = min(tid, nb - 1);
    dinv[tid] = 1.0 / Lb[(int64_t)(c0 + q) * m + c0 + q];
  }
  __syncthreads();
  d4 acc = {0.0, 0.0, 0.0, 0.0};
  // (Z^T D^-1 Z)[p][q] = sum_k Z[k][p] Z[k][q] / d_k, Zb[q * 64 + k] = Z[k][q]
  mma_tile32(
      [&](int ii, int k) {
        const int pp = min(p0 + ii, nb - 1);
        return LVs{Zb[pp * kTile + k], dinv[k], p0 + ii < nb};
      },
      [&](int jj, int k) {
        const int qq = min(q0 + jj, nb - 1);
        return LV{Zb[qq * kTile + k], q0 + jj < nb};
      },
      nb, sh, acc);
  if (m > c1) {
    // acc - Lh^T Sigma[B,P] as -(-acc + Lh^T Sigma[B,P]) (negation is exact)
    acc = -acc;
    mma_tile32(
        [&](int ii, int kk) {
          return LV{Lh[(int64_t)min(p0 + ii, nb - 1) * m + c1 + kk], p0 + ii < nb};
        },
        [&](int jj, int kk) {
          return LV{Sb[(int64_t)(c0 + min(q0 + jj, nb - 1)) * m + c1 + kk], q0 + jj < nb};
        },
        m - c1, sh, acc);
    acc = -acc;
  }
  tile_foreach32(acc, [&](int ii, int jj, double v) {
    const int p = p0 + ii, q = q0 + jj;
    if (p < nb && q < nb && p >= q) Sb[(int64_t)(c0 + q) * m + c0 + p] = v;
  });
}

// Sigma[P,P] of panel step `step` and L_hat of step + 1 in one launch: the first
// npp workgroups take the Sigma[P,P] tasks, the rest the L_hat tasks (they read
// only L and Z and write the other L_hat buffer, so the two sets are independent).
// With split, each Sigma[P,P] task runs as three workgroups (sl_pp_task32).
__global__ __launch_bounds__(256) void k_sl_pp_lhat(DevPlan P, const int32_t* __restrict__ pp_tasks,
                                                    int npp, const int32_t* __restrict__ lhat_tasks,
                                                    int step, int split, const double* __restrict__ L,
                                                    const double* __restrict__ Z,
                                                    double* __restrict__ S, double* __restrict__ T) {
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  __shared__ double dinv[kTile];
  const int npp_wg = split ? 3 * npp : npp;
  if ((int)blockIdx.x < npp_wg) {
    if (split)
      sl_pp_task32(P, pp_tasks + 3 * (blockIdx.x / 3), (int)blockIdx.x % 3, step, blockIdx.y, L, Z, S, T, sh,
                   dinv);
    else
      sl_pp_task(P, pp_tasks + 3 * blockIdx.x, step, blockIdx.y, L, Z, S, T, sh, dinv);
  } else {
    sl_lhat_task(P, lhat_tasks + 3 * (blockIdx.x - npp_wg), step + 1, blockIdx.y, L, Z, T, sh);
  }
}

// Sigma[B,P] split over K for launches with few row blocks (tall fronts of a small
// batch): task (s, I, part) forms the partial product of rows c1 + 64 I .. over
// the rows c1 + kGemmSplitK part .. of B into the front's staging area;
// k_sl_gemm_sum adds the parts in order and stores as k_sl_gemm does.
__global__ __launch_bounds__(512) void k_sl_gemm_part(DevPlan P, const int32_t* __restrict__ tasks,
                                                      int step, const double* __restrict__ S,
                                                      const double* __restrict__ W, double* __restrict__ T) {
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  double (*As)[kPanel + 1] = (double (*)[kPanel + 1])sh;
  double (*Bs)[kPanel + 1] = (double (*)[kPanel + 1])(sh + 64 * (kPanel + 1));
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1], part = tk[2];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0, nB = m - c1;
  const int parts = max(1, (nB + kGemmSplitK - 1) / kGemmSplitK);
  const int ks = part * kGemmSplitK, ke = min(nB, ks + kGemmSplitK);
  const int R0 = c1 + I * kTile;
  const double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Ws = W + (int64_t)b * P.w_size + P.w_off[s];
  double* Tb = T + (int64_t)b * P.t_size + P.t_off[s];
  const double* Lh = Tb + stage_lhat_off(m, step);
  TileAcc acc;
  tile_zero(acc);
  sigma_strip<512>(Sb, Ws, Lh, w, m, nb, R0, c1 + ks, c1 + ke, As, Bs, acc);
  double* Pg = Tb + stage_gemm_off(m) + ((int64_t)I * parts + part) * kTile * kTile;
  if (threadIdx.x < 256) tile_foreach(acc, [&](int ii, int jj, double v) { Pg[jj * kTile + ii] = v; });
}

__global__ __launch_bounds__(256) void k_sl_gemm_sum(DevPlan P, const int32_t* __restrict__ tasks,
                                                     int step, double* __restrict__ S,
                                                     const double* __restrict__ T) {
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0, nB = m - c1;
  const int parts = max(1, (nB + kGemmSplitK - 1) / kGemmSplitK);
  const int R0 = c1 + I * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Pg = T + (int64_t)b * P.t_size + P.t_off[s] + stage_gemm_off(m) + (int64_t)I * parts * kTile * kTile;
  const int ii = threadIdx.x & 63, jb = threadIdx.x >> 6;
  const int i = R0 + ii;
#pragma unroll 4
  for (int u = 0; u < 16; ++u) {
    const int jj = jb + 4 * u;
    double v = 0.0;
    for (int q = 0; q < parts; ++q) v += Pg[(int64_t)q * kTile * kTile + jj * kTile + ii];
    if (i < m && jj < nb) Sb[(int64_t)(c0 + jj) * m + i] = -v;
  }
}

// Sigma[P,P] of tall fronts in two launches: task (s, part) forms the partial
// product over rows c1 + kSelSplitK part .. of B (part 0 also adds Z^T D^-1 Z) into
// the front's staging area after L_hat (parts x 64 x 64); k_sl_pp_sum adds the
// parts in order and stores as the Sigma[P,P] tasks of k_sl_pp_lhat do. A single
// workgroup over K = m - c1
// would walk thousands of rows serially.
__global__ __launch_bounds__(256) void k_sl_pp_part(DevPlan P, const int32_t* __restrict__ tasks,
                                                    int step, const double* __restrict__ L,
                                                    const double* __restrict__ Z,
                                                    const double* __restrict__ S,
                                                    double* __restrict__ T) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  __shared__ double dinv[kTile];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], part = tk[1];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int ks = part * kSelSplitK, ke = min(m - c1, ks + kSelSplitK);
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  const double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  double* Tb = T + (int64_t)b * P.t_size + P.t_off[s];
  TileAcc acc;
  tile_zero(acc);
  if (part == 0) {
    if (tid < kTile) {
      const int q = min(tid, nb - 1);
      dinv[tid] = 1.0 / Lb[(int64_t)(c0 + q) * m + c0 + q];
    }
    __syncthreads();
    mma_tile<true, true>(
        [&](int pp, int k) {
          return LVs{Zb[pp * kTile + k], dinv[k], pp < nb};
        },
        [&](int qq, int k) {
          return LV{Zb[qq * kTile + k], qq < nb};
        },
        nb, As, Bs, acc);
  }
  if (ke > ks) {
    // acc - Lh^T Sigma[B,P] over this part, as -(-acc + Lh^T Sigma[B,P])
    tile_negate(acc);
    mma_tile<true, true>(
        [&](int ii, int kk) {
          return LV{Tb[stage_lhat_off(m, step) + (int64_t)min(ii, nb - 1) * m + c1 + ks + kk], ii < nb};
        },
        [&](int jj, int kk) {
          return LV{Sb[(int64_t)(c0 + min(jj, nb - 1)) * m + c1 + ks + kk], jj < nb};
        },
        ke - ks, As, Bs, acc);
    tile_negate(acc);
  }
  double* Pp = Tb + stage_pp_off(m) + (int64_t)part * kTile * kTile;
  tile_foreach(acc, [&](int ii, int jj, double v) { Pp[jj * kTile + ii] = v; });
}

__global__ __launch_bounds__(256) void k_sl_pp_sum(DevPlan P, const int32_t* __restrict__ tasks,
                                                   int step, double* __restrict__ S,
                                                   const double* __restrict__ T) {
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int parts = max(1, (m - c1 + kSelSplitK - 1) / kSelSplitK);
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  const double* Pp = T + (int64_t)b * P.t_size + P.t_off[s] + stage_pp_off(m);
  const int ii = threadIdx.x & 63, jb = threadIdx.x >> 6;
#pragma unroll 4
  for (int u = 0; u < 16; ++u) {
    const int jj = jb + 4 * u;
    double v = 0.0;
    for (int q = 0; q < parts; ++q) v += Pp[(int64_t)q * kTile * kTile + jj * kTile + ii];
    if (ii < nb && jj < nb && ii >= jj) Sb[(int64_t)(c0 + jj) * m + c0 + ii] = v;
  }
}

// ---------------------------------------------------------------------------
// Extraction
// ---------------------------------------------------------------------------
__global__ void k_diag_out(DevPlan P, int B, const double* __restrict__ S, double* __restrict__ out,
                           int64_t ldo) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    out[(int64_t)b * ldo + P.perm[k]] = S[(int64_t)b * P.l_size + P.diag_slot[k]];
  }
}

__global__ void k_pattern_out(DevPlan P, int B, const double* __restrict__ S,
                              double* __restrict__ out, int64_t ldo) {
  const int64_t tot = (int64_t)B * P.nnz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.nnz);
    const int64_t q = e - (int64_t)b * P.nnz;
    out[(int64_t)b * ldo + q] = S[(int64_t)b * P.l_size + P.slot_full[q]];
  }
}

// ---------------------------------------------------------------------------
// Host launch wrappers
// ---------------------------------------------------------------------------
static inline int grid1(int64_t tot) {
  int64_t g = (tot + 255) / 256;
  if (g > 65535) g = 65535;
  return (int)(g < 1 ? 1 : g);
}

hipError_t launch_factor_small(const DevPlan& P, const int32_t* list, int count, int maxm, int B,
                               const double* vals, int64_t ldv, double* L, double* U,
                               hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (maxm <= 16)
    hipLaunchKernelGGL(k_factor_small<16>, dim3(count, B), dim3(64), 0, st, P, list, vals, ldv, L, U);
  else if (maxm <= 32)
    hipLaunchKernelGGL(k_factor_small<32>, dim3(count, B), dim3(128), 0, st, P, list, vals, ldv, L, U);
  else if (maxm <= 48)
    hipLaunchKernelGGL(k_factor_small<48>, dim3(count, B), dim3(192), 0, st, P, list, vals, ldv, L, U);
  else if (maxm <= 64)
    hipLaunchKernelGGL(k_factor_small<64>, dim3(count, B), dim3(256), 0, st, P, list, vals, ldv, L, U);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_fl_assemble(const DevPlan& P, const int32_t* tasks, int count, int ain_first, int B,
                              bool diag, const double* vals, int64_t ldv, double* L, double* U, double* Z,
                              hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (diag)
    hipLaunchKernelGGL(k_fl_assemble<true>, dim3(count, B), dim3(256), 0, st, P, tasks, ain_first, vals, ldv, L,
                       U, Z);
  else
    hipLaunchKernelGGL(k_fl_assemble<false>, dim3(count, B), dim3(256), 0, st, P, tasks, ain_first, vals, ldv,
                       L, U, Z);
  return hipGetLastError();
}

hipError_t launch_fl_diag(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          double* L, double* Z, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_diag, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z);
  return hipGetLastError();
}

hipError_t launch_fl_trsm(const DevPlan& P, const int32_t* tasks, int count, int step, int B, bool split,
                          double* L, const double* Z, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_trsm, dim3((split ? 2 : 1) * count, B), dim3(256), 0, st, P, tasks, step,
                     split ? 1 : 0, L, Z);
  return hipGetLastError();
}

hipError_t launch_fl_update(const DevPlan& P, const int32_t* tasks, int count, int step, int B, bool diag,
                            double* L, double* U, double* Z, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (diag)
    hipLaunchKernelGGL(k_fl_update<true>, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, U, Z);
  else
    hipLaunchKernelGGL(k_fl_update<false>, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, U, Z);
  return hipGetLastError();
}

hipError_t launch_logdet(const DevPlan& P, int B, const double* L, double* part, uint32_t* counter,
                         double* logdet, int32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_logdet, dim3(kLogdetChunks, B), dim3(256), 0, st, P, L, part, counter, logdet,
                     status);
  return hipGetLastError();
}

hipError_t launch_logdet_seg(const DevPlan& P, int B, const double* L, const int64_t* ptr,
                             const int32_t* pos, int k, double* logdet, int32_t* status,
                             hipStream_t st) {
  if (k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_logdet_seg, dim3(k, B), dim3(256), 0, st, P, L, ptr, pos, k, logdet, status);
  return hipGetLastError();
}

hipError_t launch_gather_perm(const DevPlan& P, int B, const double* rhs, int64_t ldr, double* X,
                              hipStream_t st, int period) {
  hipLaunchKernelGGL(k_gather_perm, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, rhs, ldr,
                     period, X);
  return hipGetLastError();
}

hipError_t launch_scatter_perm(const DevPlan& P, int B, const double* X, double* out, int64_t ldo,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_perm, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, X, out,
                     ldo);
  return hipGetLastError();
}

hipError_t launch_diag_scale(const DevPlan& P, int B, int nr, const double* L, double* X,
                             hipStream_t st) {
  hipLaunchKernelGGL(k_diag_scale, dim3(grid1((int64_t)B * nr * P.n)), dim3(256), 0, st, P, B * nr, nr,
                     L, X);
  return hipGetLastError();
}

// Small fronts: maxm / maxw bound the launch's fronts (m <= 64); four vectors per
// workgroup when a matrix has several right-hand sides.
hipError_t launch_solve_fwd(const DevPlan& P, const int32_t* list, int count, int maxm, int maxw, int B,
                            int nr, const double* L, double* X, double* V, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (maxm > kSmallFront || maxw > maxm) return hipErrorInvalidValue;
  const size_t sh = (size_t)maxm * maxw * sizeof(double);
  if (nr >= 4)
    hipLaunchKernelGGL(k_solve_fwd<4>, dim3(count * ((nr + 3) / 4), B), dim3(256), sh, st, P, list, nr, L, X, V);
  else
    hipLaunchKernelGGL(k_solve_fwd<1>, dim3(count * nr, B), dim3(64), sh, st, P, list, nr, L, X, V);
  return hipGetLastError();
}

hipError_t launch_solve_bwd(const DevPlan& P, const int32_t* list, int count, int maxm, int maxw, int B,
                            int nr, const double* L, double* X, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  if (maxm > kSmallFront || maxw > maxm) return hipErrorInvalidValue;
  const size_t sh = (size_t)(maxm | 1) * maxw * sizeof(double);
  if (nr >= 4)
    hipLaunchKernelGGL(k_solve_bwd<4>, dim3(count * ((nr + 3) / 4), B), dim3(256), sh, st, P, list, nr, L, X);
  else
    hipLaunchKernelGGL(k_solve_bwd<1>, dim3(count * nr, B), dim3(64), sh, st, P, list, nr, L, X);
  return hipGetLastError();
}

// Dynamic LDS beyond 64 KiB has to be requested per kernel (gfx950: 160 KiB per CU).
static hipError_t lds_limit(const void* fn, size_t bytes) {
  if (bytes > 160 * 1024) return hipErrorInvalidValue;
  if (bytes <= 64 * 1024) return hipSuccess;
  return hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes);
}

// The triangular parts hold the supernode's w entries (m with full) in LDS; few
// fronts get 1024 threads each for latency, many get 256 for occupancy.
template <int NT, bool FULL>
static hipError_t fwd_tri(const DevPlan& P, const int32_t* list, int count, size_t sh, int B, int nr,
                          const double* L, const double* Z, double* X, double* V, hipStream_t st) {
  hipError_t e = lds_limit((const void*)k_solve_fwd_tri<NT, FULL>, sh);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_solve_fwd_tri<NT, FULL>), dim3(count * nr, B), dim3(NT), sh, st, P, list, nr, L,
                     Z, X, V);
  return hipGetLastError();
}

hipError_t launch_solve_fwd_tri(const DevPlan& P, const int32_t* list, int count, int maxlen, int B,
                                int nr, bool full, const double* L, const double* Z, double* X,
                                double* V, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const size_t sh = (size_t)maxlen * sizeof(double);
  const bool few = (int64_t)count * B * nr < 256;
  if (full)
    return few ? fwd_tri<1024, true>(P, list, count, sh, B, nr, L, Z, X, V, st)
               : fwd_tri<256, true>(P, list, count, sh, B, nr, L, Z, X, V, st);
  return few ? fwd_tri<1024, false>(P, list, count, sh, B, nr, L, Z, X, V, st)
             : fwd_tri<256, false>(P, list, count, sh, B, nr, L, Z, X, V, st);
}

hipError_t launch_solve_fwd_rows(const DevPlan& P, const int32_t* tasks, int count, int maxw,
                                 int B, int nr, const double* L, const double* X, double* V, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const size_t sh = (size_t)maxw * sizeof(double);
  hipError_t e = lds_limit((const void*)k_solve_fwd_rows, sh);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_solve_fwd_rows, dim3(count * nr, B), dim3(256), sh, st, P, tasks, nr, L, X, V);
  return hipGetLastError();
}

hipError_t launch_solve_bwd_rows(const DevPlan& P, const int32_t* tasks, int count, int B, int nr,
                                 const double* L, const double* X, double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_solve_bwd_rows, dim3(count * nr, B), dim3(256), 0, st, P, tasks, nr, L, X, T);
  return hipGetLastError();
}

template <int NT, bool FULL>
static hipError_t bwd_tri(const DevPlan& P, const int32_t* list, int count, size_t sh, int B, int nr,
                          const double* L, const double* Z, double* X, const double* T, hipStream_t st) {
  hipError_t e = lds_limit((const void*)k_solve_bwd_tri<NT, FULL>, sh);
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL((k_solve_bwd_tri<NT, FULL>), dim3(count * nr, B), dim3(NT), sh, st, P, list, nr, L,
                     Z, X, T);
  return hipGetLastError();
}

hipError_t launch_solve_bwd_tri(const DevPlan& P, const int32_t* list, int count, int maxlen, int B,
                                int nr, bool full, const double* L, const double* Z, double* X,
                                const double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const size_t sh = (size_t)maxlen * sizeof(double);
  const bool few = (int64_t)count * B * nr < 256;
  if (full)
    return few ? bwd_tri<1024, true>(P, list, count, sh, B, nr, L, Z, X, T, st)
               : bwd_tri<256, true>(P, list, count, sh, B, nr, L, Z, X, T, st);
  return few ? bwd_tri<1024, false>(P, list, count, sh, B, nr, L, Z, X, T, st)
             : bwd_tri<256, false>(P, list, count, sh, B, nr, L, Z, X, T, st);
}

// The L block lives in dynamic LDS sized by the widest supernode of the launch
// (row stride odd for the banks): narrow supernodes leave room for more
// workgroups per CU.
hipError_t launch_selinv_small(const DevPlan& P, const int32_t* list, int count, int maxm, int maxw,
                               int B, const double* L, double* S, double* U, const double* W,
                               hipStream_t st) {
  if (count <= 0) return hipSuccess;
  const int ldl = std::max(maxw, 1) | 1;
  if (maxm <= 16)
    hipLaunchKernelGGL(k_selinv_small<16>, dim3(count, B), dim3(64), (size_t)16 * ldl * sizeof(double), st,
                       P, list, L, S, U, W, ldl);
  else if (maxm <= 32)
    hipLaunchKernelGGL(k_selinv_small<32>, dim3(count, B), dim3(128), (size_t)32 * ldl * sizeof(double),
                       st, P, list, L, S, U, W, ldl);
  else if (maxm <= 48)
    hipLaunchKernelGGL(k_selinv_small<48>, dim3(count, B), dim3(192), (size_t)48 * ldl * sizeof(double),
                       st, P, list, L, S, U, W, ldl);
  else if (maxm <= 64)
    hipLaunchKernelGGL(k_selinv_small<64>, dim3(count, B), dim3(256), (size_t)64 * ldl * sizeof(double),
                       st, P, list, L, S, U, W, ldl);
  else
    return hipErrorInvalidValue;
  return hipGetLastError();
}

hipError_t launch_sl_gather_lhat(const DevPlan& P, const int32_t* g_tasks, int ng, const int32_t* l_tasks,
                                int nl, int B, const double* S, const double* U, double* W, const double* L,
                                const double* Z, double* T, hipStream_t st) {
  if (ng + nl <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_gather_lhat, dim3(ng + nl, B), dim3(256), 0, st, P, g_tasks, ng, l_tasks, S, U, W,
                     L, Z, T);
  return hipGetLastError();
}

hipError_t launch_sl_lhat(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          const double* L, const double* Z, double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_lhat, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z, T);
  return hipGetLastError();
}

hipError_t launch_sl_gemm(const DevPlan& P, const int32_t* tasks, int count, int step, int B, int xcd_max,
                          bool split, double* S, const double* W, const double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  // XCD-grouped when every XCD gets a matrix and a matrix's tiles fit one XCD
  const int xcd = (B >= 8 && count <= xcd_max) ? 1 : 0;
  const int64_t n = (split ? 4 : 1) * (int64_t)count;
  const int64_t grid = xcd ? (int64_t)8 * ((B + 7) / 8) * n : n * B;
  if (split)
    hipLaunchKernelGGL(k_sl_gemm32, dim3((unsigned)grid), dim3(256), 0, st, P, tasks, count, B, xcd, step, S, W,
                       T);
  else
    hipLaunchKernelGGL(k_sl_gemm, dim3((unsigned)grid), dim3(512), 0, st, P, tasks, count, B, xcd, step, S, W,
                       T);
  return hipGetLastError();
}

hipError_t launch_sl_gemm_part(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                               const double* S, const double* W, double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_gemm_part, dim3(count, B), dim3(512), 0, st, P, tasks, step, S, W, T);
  return hipGetLastError();
}

hipError_t launch_sl_gemm_sum(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                              double* S, const double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_gemm_sum, dim3(count, B), dim3(256), 0, st, P, tasks, step, S, T);
  return hipGetLastError();
}

hipError_t launch_sl_pp_lhat(const DevPlan& P, const int32_t* pp_tasks, int npp,
                             const int32_t* lhat_tasks, int nlhat, int step, int B, bool split,
                             const double* L, const double* Z, double* S, double* T, hipStream_t st) {
  if (npp + nlhat <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_pp_lhat, dim3((split ? 3 : 1) * npp + nlhat, B), dim3(256), 0, st, P, pp_tasks, npp,
                     lhat_tasks, step, split ? 1 : 0, L, Z, S, T);
  return hipGetLastError();
}

hipError_t launch_sl_pp_part(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                             const double* L, const double* Z, const double* S, double* T,
                             hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_pp_part, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z, S, T);
  return hipGetLastError();
}

hipError_t launch_sl_pp_sum(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                            double* S, const double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_pp_sum, dim3(count, B), dim3(256), 0, st, P, tasks, step, S, T);
  return hipGetLastError();
}

hipError_t launch_diag_out(const DevPlan& P, int B, const double* S, double* out, int64_t ldo,
                           hipStream_t st) {
  hipLaunchKernelGGL(k_diag_out, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, S, out,
                     ldo);
  return hipGetLastError();
}

hipError_t launch_pattern_out(const DevPlan& P, int B, const double* S, double* out, int64_t ldo,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_pattern_out, dim3(grid1((int64_t)B * P.nnz)), dim3(256), 0, st, P, B, S, out,
                     ldo);
  return hipGetLastError();
}

hipError_t launch_mma_probe(const double* A, const double* B, int K, double* T, hipStream_t st) {
  hipLaunchKernelGGL(k_mma_probe, dim3(1), dim3(256), 0, st, A, B, K, T);
  return hipGetLastError();
}

}  // namespace gmrf
