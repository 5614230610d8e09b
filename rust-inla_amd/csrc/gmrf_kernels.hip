// FP64 supernodal LDL^T, triangular solves and selected inversion for gfx950.
//
// Storage (per matrix of a batch, offsets in doubles):
//   L block of supernode s: m x w column-major (ld = m) at l_off[s]; rows are the
//     w supernode columns followed by R_s. D sits on the diagonal, unit-L below.
//   U workspace of s: r x r column-major at u_off[s]; holds the Schur update
//     during the factorisation and Sigma_{R_s R_s} during selected inversion.
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
struct TileAcc {
  d4 c[2][2];  // [bj][bi]
};

template <typename FA, typename FB>
__device__ __forceinline__ void mma_tile(FA fa, FB fb, int K, double (*As)[kPanel + 1],
                                         double (*Bs)[kPanel + 1], TileAcc& acc) {
  const int tid = threadIdx.x, wv = tid >> 6, lane = tid & 63;
  const int qi = wv & 1, qj = wv >> 1;
  const int lr = lane & 15, lk = lane >> 4;
  for (int k0 = 0; k0 < K; k0 += kPanel) {
    const int kc = min(kPanel, K - k0);
    for (int idx = tid; idx < 64 * kPanel; idx += 256) {
      const int ii = idx & 63, kk = idx >> 6;
      As[ii][kk] = kk < kc ? fa(ii, k0 + kk) : 0.0;
      Bs[ii][kk] = kk < kc ? fb(ii, k0 + kk) : 0.0;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < kPanel; kk += 4) {
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

__device__ __forceinline__ void tile_zero(TileAcc& acc) {
#pragma unroll
  for (int a = 0; a < 2; ++a)
#pragma unroll
    for (int b = 0; b < 2; ++b) acc.c[a][b] = d4{0.0, 0.0, 0.0, 0.0};
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
  mma_tile([&](int i, int k) { return A[(int64_t)k * 64 + i]; },
           [&](int j, int k) { return B[(int64_t)k * 64 + j]; }, K, As, Bs, acc);
  tile_foreach(acc, [&](int i, int j, double v) { T[(int64_t)j * 64 + i] = v; });
}

// ---------------------------------------------------------------------------
// Factorisation
// ---------------------------------------------------------------------------

// Small fronts: one wavefront, front in LDS row-major (ld = m + 1); lane t owns
// row t. Assembles the input entries and the children's updates, eliminates the
// w supernode columns, writes the L block and the r x r update.
__global__ __launch_bounds__(64) void k_factor_small(DevPlan P, const int32_t* __restrict__ list,
                                                     const double* __restrict__ vals, int64_t ldv,
                                                     double* __restrict__ L,
                                                     double* __restrict__ U) {
  extern __shared__ double F[];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int ld = m + 1;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const double* vb = vals + (int64_t)b * ldv;
  for (int idx = t; idx < m * ld; idx += 64) F[idx] = 0.0;
  __syncthreads();
  for (int64_t q = P.fa_ptr[s] + t; q < P.fa_ptr[s + 1]; q += 64) {
    const int slot = P.fa_slot[q];
    const int j = slot / m, i = slot - j * m;
    F[i * ld + j] = vb[P.fa_csc[q]];
  }
  __syncthreads();
  for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
    const int c = P.child_idx[cq];
    const int rc = P.sn_r[c];
    const double* Uc = U + (int64_t)b * P.u_size + P.u_off[c];
    const int32_t* relc = P.rel + P.rows_ptr[c];
    for (int j = 0; j < rc; ++j) {
      const int pj = relc[j];
      for (int i = j + t; i < rc; i += 64) F[relc[i] * ld + pj] += Uc[(int64_t)j * rc + i];
    }
    __syncthreads();
  }
  for (int k = 0; k < w; ++k) {
    const double dk = F[k * ld + k];
    double lik = 0.0;
    if (t > k && t < m) {
      lik = F[t * ld + k] / dk;
      F[t * ld + k] = lik;
    }
    __syncthreads();
    if (t > k && t < m) {
      const double y = lik * dk;
      double* row = F + t * ld;
      for (int j = k + 1; j <= t; ++j) row[j] -= y * F[j * ld + k];
    }
    __syncthreads();
  }
  for (int j = 0; j < w; ++j)
    for (int i = t; i < m; i += 64) Lb[(int64_t)j * m + i] = (i >= j) ? F[i * ld + j] : 0.0;
  for (int j = 0; j < r; ++j)
    for (int i = j + t; i < r; i += 64) Ub[(int64_t)j * r + i] = F[(w + i) * ld + (w + j)];
}

// ---------------------------------------------------------------------------
// Large fronts (m > 64), tile-parallel. Each launch covers every large front of
// one level for every matrix of the batch; a task is an int32 triple
// (supernode, a, b) whose meaning depends on the kernel.
// ---------------------------------------------------------------------------

// Assembly of one 64x64 tile (tile row a, tile column b, a >= b) of a front:
// input entries through the dense L-slot map plus every child's update through
// the child's inverse row map (children in fixed order: deterministic).
__global__ __launch_bounds__(256) void k_fl_assemble(DevPlan P, const int32_t* __restrict__ tasks,
                                                     const double* __restrict__ vals, int64_t ldv,
                                                     double* __restrict__ L,
                                                     double* __restrict__ U) {
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1], J = tk[2];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const double* vb = vals + (int64_t)b * ldv;
  const int32_t* am = P.dense_amap + P.l_off[s];
  const double* Um = U + (int64_t)b * P.u_size;
  for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
    const int i = I * kTile + (idx & (kTile - 1)), j = J * kTile + (idx / kTile);
    if (i >= m || j >= m || i < j) continue;
    double v = 0.0;
    if (j < w) {
      const int e = am[(int64_t)j * m + i];
      if (e >= 0) v = vb[e];
    }
    for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
      const int32_t* inv = P.cinv + P.cinv_off[cq];
      const int ci = inv[i], cj = inv[j];
      if (ci >= 0 && cj >= 0) {
        const int c = P.child_idx[cq];
        const int rc = P.sn_r[c];
        v += Um[P.u_off[c] + (int64_t)cj * rc + ci];
      }
    }
    *front_at(Lb, Ub, m, w, r, i, j) = v;
  }
}

// Diagonal block of panel step `step` (columns [64 step, 64 step + nb)), one
// task per front: LDL^T of the nb x nb block in LDS with 256 threads, and, in the
// same sweep, Z = L11^-1 accumulated from the Gauss transforms
// Z <- (I - l_k e_k^T) Z. The factorised block goes back to L; Z is kept in the
// per-panel Z workspace for the panel solve and for selected inversion.
__global__ __launch_bounds__(256) void k_fl_diag(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, double* __restrict__ L,
                                                 double* __restrict__ Z) {
  __shared__ double D[kTile][kTile + 1];   // D[i][j], lower
  __shared__ double Zs[kTile][kTile + 1];  // Zs[i][j], lower, unit diagonal
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int c0 = step * kTile, nb = min(kTile, w - c0);
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)step * kTile * kTile;
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int j = idx / nb, i = idx - j * nb;
    D[i][j] = (i >= j) ? Lb[(int64_t)(c0 + j) * m + c0 + i] : 0.0;
    Zs[i][j] = (i == j) ? 1.0 : 0.0;
  }
  __syncthreads();
  for (int k = 0; k < nb; ++k) {
    const double dk = D[k][k];
    for (int i = k + 1 + tid; i < nb; i += 256) D[i][k] /= dk;
    __syncthreads();
    // trailing update D[i][j] -= l_ik d_k l_jk (k < j <= i) and Z[i][q] -= l_ik Z[k][q] (q <= k < i)
    const int rest = nb - k - 1;
    for (int idx = tid; idx < rest * rest; idx += 256) {
      const int ii = idx % rest, jj = idx / rest;
      if (ii >= jj) {
        const int i = k + 1 + ii, j = k + 1 + jj;
        D[i][j] -= D[i][k] * dk * D[j][k];
      }
    }
    for (int idx = tid; idx < rest * (k + 1); idx += 256) {
      const int i = k + 1 + idx % rest, q = idx / rest;
      Zs[i][q] -= D[i][k] * Zs[k][q];
    }
    __syncthreads();
  }
  for (int idx = tid; idx < nb * nb; idx += 256) {
    const int j = idx / nb, i = idx - j * nb;
    Lb[(int64_t)(c0 + j) * m + c0 + i] = (i >= j) ? D[i][j] : 0.0;
    Zb[j * kTile + i] = (i >= j) ? Zs[i][j] : 0.0;
  }
}

// Panel solve as a product: task (s, I) computes 64 rows below the diagonal block,
// L21[i][p] = sum_q A21[i][q] Z[p][q] / d_p  (L21 = A21 L11^-T D^-1), in place.
__global__ __launch_bounds__(256) void k_fl_trsm(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, double* __restrict__ L,
                                                 const double* __restrict__ Z) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  __shared__ double dinv[kTile];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int c0 = step * kTile, nb = min(kTile, w - c0);
  const int R0 = c0 + nb + I * kTile;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)step * kTile * kTile;
  if (threadIdx.x < nb) dinv[threadIdx.x] = 1.0 / Lb[(int64_t)(c0 + threadIdx.x) * m + c0 + threadIdx.x];
  __syncthreads();
  TileAcc acc;
  tile_zero(acc);
  mma_tile(
      [&](int ii, int q) {
        const int i = R0 + ii;
        return (i < m && q < nb) ? Lb[(int64_t)(c0 + q) * m + i] : 0.0;
      },
      [&](int pp, int q) { return (pp < nb && q < nb) ? Zb[q * kTile + pp] * dinv[pp] : 0.0; }, nb, As,
      Bs, acc);
  tile_foreach(acc, [&](int ii, int pp, double v) {
    const int i = R0 + ii;
    if (i < m && pp < nb) Lb[(int64_t)(c0 + pp) * m + i] = v;
  });
}

// Trailing update of panel step `step`: task (s, I, J), I >= J, is the 64x64
// tile at rows t0 + 64 I, columns t0 + 64 J of the trailing lower triangle,
// t0 = 64 step + nb:  C -= L_I D L_J^T on FP64 MFMA.
__global__ __launch_bounds__(256) void k_fl_update(DevPlan P, const int32_t* __restrict__ tasks,
                                                   int step, double* __restrict__ L,
                                                   double* __restrict__ U) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
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
        return (i < m && p < nb) ? Lb[(int64_t)(c0 + p) * m + i] * dpiv[p] : 0.0;
      },
      [&](int jj, int p) {
        const int j = C0 + jj;
        return (j < m && p < nb) ? Lb[(int64_t)(c0 + p) * m + j] : 0.0;
      },
      nb, As, Bs, acc);
  tile_foreach(acc, [&](int ii, int jj, double v) {
    const int i = R0 + ii, j = C0 + jj;
    if (i < m && j < m && i >= j) *front_at(Lb, Ub, m, w, r, i, j) -= v;
  });
}

// log|Q| = sum log|D_j| per matrix, plus the pivot status: an exact zero pivot
// (the numeric phase fails: NotPositiveDefinite) or |D_j| < 1e-14 / non-finite
// (Singular), as crates/inla_math/src/sparse_ldlt.rs:104-149,217-230.
__global__ __launch_bounds__(256) void k_logdet(DevPlan P, const double* __restrict__ L,
                                                double* __restrict__ logdet,
                                                int32_t* __restrict__ status) {
  __shared__ double red[256];
  __shared__ int flags[256];
  const int b = blockIdx.x;
  const double* Lm = L + (int64_t)b * P.l_size;
  double acc = 0.0;
  int fl = 0;
  for (int j = threadIdx.x; j < P.n; j += 256) {
    const double d = Lm[P.diag_slot[j]];
    if (d == 0.0) fl |= GMRF_ST_ZERO_PIVOT;
    if (!(fabs(d) >= 1e-14) || !isfinite(d)) fl |= GMRF_ST_SINGULAR;
    acc += log(fabs(d));
  }
  red[threadIdx.x] = acc;
  flags[threadIdx.x] = fl;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) {
      red[threadIdx.x] += red[threadIdx.x + o];
      flags[threadIdx.x] |= flags[threadIdx.x + o];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    logdet[b] = red[0];
    status[b] = flags[0];
  }
}

// ---------------------------------------------------------------------------
// Solves (one right-hand side per matrix)
// ---------------------------------------------------------------------------
__global__ void k_gather_perm(DevPlan P, int B, const double* __restrict__ rhs, int64_t ldr,
                              double* __restrict__ X) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    X[e] = rhs[(int64_t)b * ldr + P.perm[k]];
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

__global__ void k_diag_scale(DevPlan P, int B, const double* __restrict__ L,
                             double* __restrict__ X) {
  const int64_t tot = (int64_t)B * P.n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / P.n), k = (int)(e - (int64_t)b * P.n);
    X[e] /= L[(int64_t)b * P.l_size + P.diag_slot[k]];
  }
}

// Forward: f = [x_J; 0] + sum_c extend(v_c); y_J = L11^-1 f_J; v_s = f_R - L21 y_J.
__global__ __launch_bounds__(128) void k_solve_fwd(DevPlan P, const int32_t* __restrict__ list,
                                                   const double* __restrict__ L,
                                                   double* __restrict__ X,
                                                   double* __restrict__ V) {
  extern __shared__ double f[];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Xb = X + (int64_t)b * P.n;
  for (int i = tid; i < m; i += blockDim.x) f[i] = i < w ? Xb[a0 + i] : 0.0;
  __syncthreads();
  for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
    const int c = P.child_idx[cq];
    const int rc = P.sn_r[c];
    const double* Vc = V + (int64_t)b * P.v_size + P.v_off[c];
    const int32_t* relc = P.rel + P.rows_ptr[c];
    for (int i = tid; i < rc; i += blockDim.x) f[relc[i]] += Vc[i];
    __syncthreads();
  }
  for (int k = 0; k < w; ++k) {
    const double yk = f[k];
    for (int i = k + 1 + tid; i < m; i += blockDim.x) f[i] -= Lb[(int64_t)k * m + i] * yk;
    __syncthreads();
  }
  double* Vs = V + (int64_t)b * P.v_size + P.v_off[s];
  for (int i = tid; i < m; i += blockDim.x) {
    if (i < w)
      Xb[a0 + i] = f[i];
    else
      Vs[i - w] = f[i];
  }
}

// Backward: x_J = L11^-T (z_J - L21^T x_R) with x_R gathered from final values.
__global__ __launch_bounds__(128) void k_solve_bwd(DevPlan P, const int32_t* __restrict__ list,
                                                   const double* __restrict__ L,
                                                   double* __restrict__ X) {
  extern __shared__ double sh[];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Xb = X + (int64_t)b * P.n;
  double* g = sh;      // r
  double* t = sh + r;  // w
  const int32_t* rows = P.rows + P.rows_ptr[s];
  for (int i = tid; i < r; i += blockDim.x) g[i] = Xb[rows[i]];
  __syncthreads();
  for (int j = tid; j < w; j += blockDim.x) {
    double acc = Xb[a0 + j];
    const double* col = Lb + (int64_t)j * m + w;
    for (int i = 0; i < r; ++i) acc -= col[i] * g[i];
    t[j] = acc;
  }
  __syncthreads();
  for (int k = w - 1; k >= 0; --k) {
    const double xk = t[k];
    for (int j = tid; j < k; j += blockDim.x) t[j] -= Lb[(int64_t)j * m + k] * xk;
    __syncthreads();
  }
  for (int j = tid; j < w; j += blockDim.x) Xb[a0 + j] = t[j];
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
__device__ __forceinline__ double sigma_parent(const DevPlan& P, const double* Sm, const double* Um,
                                               const double* Wm, int p, int pi, int pj) {
  const int wp = P.sn_start[p + 1] - P.sn_start[p];
  const int mp = wp + P.sn_r[p];
  if (P.is_large[p]) return Wm[P.w_off[p] + (int64_t)pj * mp + pi];
  if (pj < wp) return Sm[P.l_off[p] + (int64_t)pj * mp + pi];
  const int rp = P.sn_r[p];
  return Um[P.u_off[p] + (int64_t)(pj - wp) * rp + (pi - wp)];
}

__global__ __launch_bounds__(64) void k_selinv_small(DevPlan P, const int32_t* __restrict__ list,
                                                     const double* __restrict__ L,
                                                     double* __restrict__ S,
                                                     double* __restrict__ U,
                                                     const double* __restrict__ W) {
  extern __shared__ double sh[];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int t = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int ld = m + 1;
  double* Sg = sh;             // m x (m+1)
  double* Lc = sh + m * ld;    // m
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Sm = S + (int64_t)b * P.l_size;
  double* Um = U + (int64_t)b * P.u_size;
  const double* Wm = W + (int64_t)b * P.w_size;
  if (r > 0) {
    const int p = P.sn_parent[s];
    const int32_t* rl = P.rel + P.rows_ptr[s];
    double* Us = Um + P.u_off[s];
    for (int j = 0; j < r; ++j) {
      const int pj = rl[j];
      for (int i = j + t; i < r; i += 64) {
        const double v = sigma_parent(P, Sm, Um, Wm, p, rl[i], pj);
        Sg[(w + i) * ld + w + j] = v;
        Sg[(w + j) * ld + w + i] = v;
        Us[(int64_t)j * r + i] = v;
      }
    }
  }
  __syncthreads();
  for (int k = w - 1; k >= 0; --k) {
    if (t > k && t < m) Lc[t] = Lb[(int64_t)k * m + t];
    __syncthreads();
    if (t > k && t < m) {
      double acc = 0.0;
      const double* row = Sg + t * ld;
      for (int l = k + 1; l < m; ++l) acc += row[l] * Lc[l];
      Sg[t * ld + k] = -acc;
      Sg[k * ld + t] = -acc;
    }
    __syncthreads();
    double prod = (t > k && t < m) ? Lc[t] * Sg[t * ld + k] : 0.0;
    for (int o = 32; o > 0; o >>= 1) prod += __shfl_xor(prod, o, 64);
    if (t == 0) Sg[k * ld + k] = 1.0 / Lb[(int64_t)k * m + k] - prod;
    __syncthreads();
  }
  double* Sb = Sm + P.l_off[s];
  for (int j = 0; j < w; ++j)
    for (int i = t; i < m; i += 64) Sb[(int64_t)j * m + i] = (i >= j) ? Sg[i * ld + j] : 0.0;
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

// Gather tile (I, J), I >= J, of Sigma_{R_s R_s} into W (both triangles).
__global__ __launch_bounds__(256) void k_sl_gather(DevPlan P, const int32_t* __restrict__ tasks,
                                                   const double* __restrict__ S,
                                                   const double* __restrict__ U,
                                                   double* __restrict__ W) {
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1], J = tk[2];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const int p = P.sn_parent[s];
  const int32_t* rl = P.rel + P.rows_ptr[s];
  const double* Sm = S + (int64_t)b * P.l_size;
  const double* Um = U + (int64_t)b * P.u_size;
  double* Wm = W + (int64_t)b * P.w_size;
  double* Ws = Wm + P.w_off[s];
  for (int idx = threadIdx.x; idx < kTile * kTile; idx += 256) {
    const int i = I * kTile + (idx & (kTile - 1)), j = J * kTile + (idx / kTile);
    if (i >= r || j >= r || i < j) continue;
    const double v = sigma_parent(P, Sm, Um, Wm, p, rl[i], rl[j]);
    Ws[(int64_t)(w + j) * m + (w + i)] = v;
    Ws[(int64_t)(w + i) * m + (w + j)] = v;
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
// computes rows c1 + 64 I .. of Lh[i][p] = sum_q L[i][q] Z[q][p] into T.
__global__ __launch_bounds__(256) void k_sl_lhat(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, const double* __restrict__ L,
                                                 const double* __restrict__ Z,
                                                 double* __restrict__ T) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int R0 = c1 + I * kTile;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  double* Tb = T + (int64_t)b * P.l_size + P.l_off[s];
  TileAcc acc;
  tile_zero(acc);
  mma_tile(
      [&](int ii, int q) {
        const int i = R0 + ii;
        return (i < m && q < nb) ? Lb[(int64_t)(c0 + q) * m + i] : 0.0;
      },
      [&](int pp, int q) { return (pp < nb && q < nb) ? Zb[pp * kTile + q] : 0.0; }, nb, As, Bs, acc);
  tile_foreach(acc, [&](int ii, int pp, double v) {
    const int i = R0 + ii;
    if (i < m && pp < nb) Tb[(int64_t)(c0 + pp) * m + i] = v;
  });
}

// Sigma[B,P]: task (s, I) computes rows c1 + 64 I .. of B as
// -sum_{k in B} W[i][k] Lh[k][p] (K = m - c1) and stores it into S (lower) and
// both triangles of W.
__global__ __launch_bounds__(256) void k_sl_gemm(DevPlan P, const int32_t* __restrict__ tasks,
                                                 int step, double* __restrict__ S,
                                                 double* __restrict__ W,
                                                 const double* __restrict__ T) {
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0], I = tk[1];
  const int b = blockIdx.y;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const int R0 = c1 + I * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  double* Ws = W + (int64_t)b * P.w_size + P.w_off[s];
  const double* Tb = T + (int64_t)b * P.l_size + P.l_off[s];
  TileAcc acc;
  tile_zero(acc);
  mma_tile(
      [&](int ii, int kk) {
        const int i = R0 + ii;
        return i < m ? Ws[(int64_t)(c1 + kk) * m + i] : 0.0;
      },
      [&](int jj, int kk) { return jj < nb ? Tb[(int64_t)(c0 + jj) * m + c1 + kk] : 0.0; }, m - c1,
      As, Bs, acc);
  tile_foreach(acc, [&](int ii, int jj, double v) {
    const int i = R0 + ii;
    if (i < m && jj < nb) {
      Sb[(int64_t)(c0 + jj) * m + i] = -v;
      Ws[(int64_t)(c0 + jj) * m + i] = -v;
      Ws[(int64_t)i * m + c0 + jj] = -v;
    }
  });
}

// Sigma[P,P]: task (s) = Z^T D^-1 Z - Lh^T Sigma[B,P] (MFMA, K = m - c1);
// lower part into S, both triangles into W.
__global__ __launch_bounds__(256) void k_sl_pp(DevPlan P, const int32_t* __restrict__ tasks,
                                               int step, const double* __restrict__ L,
                                               const double* __restrict__ Z,
                                               double* __restrict__ S, double* __restrict__ W,
                                               const double* __restrict__ T) {
  __shared__ double sh[2 * 64 * (kPanel + 1)];
  __shared__ double dinv[kTile];
  const int32_t* tk = tasks + 3 * blockIdx.x;
  const int s = tk[0];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  int c0, c1, pk;
  sel_panel(w, step, c0, c1, pk);
  const int nb = c1 - c0;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  const double* Zb = Z + (int64_t)b * P.z_size + P.z_off[s] + (int64_t)pk * kTile * kTile;
  double* Sb = S + (int64_t)b * P.l_size + P.l_off[s];
  double* Ws = W + (int64_t)b * P.w_size + P.w_off[s];
  const double* Tb = T + (int64_t)b * P.l_size + P.l_off[s];
  double (*As)[kPanel + 1] = (double (*)[kPanel + 1])sh;
  double (*Bs)[kPanel + 1] = (double (*)[kPanel + 1])(sh + 64 * (kPanel + 1));
  TileAcc acc;
  tile_zero(acc);
  if (m > c1)
    mma_tile([&](int ii, int kk) { return ii < nb ? Tb[(int64_t)(c0 + ii) * m + c1 + kk] : 0.0; },
             [&](int jj, int kk) { return jj < nb ? Ws[(int64_t)(c0 + jj) * m + c1 + kk] : 0.0; },
             m - c1, As, Bs, acc);
  __syncthreads();
  // Zt[q][k] = Z[k][q] in the LDS freed by the tile engine
  double (*Zt)[kTile + 1] = (double (*)[kTile + 1])sh;
  for (int idx = tid; idx < kTile * kTile; idx += 256) {
    const int q = idx / kTile, k = idx - q * kTile;
    Zt[q][k] = (q < nb && k < nb) ? Zb[q * kTile + k] : 0.0;
  }
  if (tid < nb) dinv[tid] = 1.0 / Lb[(int64_t)(c0 + tid) * m + c0 + tid];
  __syncthreads();
  tile_foreach(acc, [&](int ii, int jj, double v) {
    // tile row ii = p, tile column jj = q; lower part p >= q
    if (ii < nb && jj < nb && ii >= jj) {
      double g = 0.0;
      for (int k = ii; k < nb; ++k) g += Zt[ii][k] * Zt[jj][k] * dinv[k];
      const double val = g - v;
      Sb[(int64_t)(c0 + jj) * m + c0 + ii] = val;
      Ws[(int64_t)(c0 + jj) * m + c0 + ii] = val;
      Ws[(int64_t)(c0 + ii) * m + c0 + jj] = val;
    }
  });
  if (c0 == 0)  // strictly upper part of the diagonal block: unused, kept zero
    for (int idx = tid; idx < w * w; idx += 256) {
      const int j = idx / w, i = idx - j * w;
      if (i < j) Sb[(int64_t)j * m + i] = 0.0;
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
  size_t sh = (size_t)maxm * (maxm + 1) * sizeof(double);
  hipLaunchKernelGGL(k_factor_small, dim3(count, B), dim3(64), sh, st, P, list, vals, ldv, L, U);
  return hipGetLastError();
}

hipError_t launch_fl_assemble(const DevPlan& P, const int32_t* tasks, int count, int B,
                              const double* vals, int64_t ldv, double* L, double* U, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_assemble, dim3(count, B), dim3(256), 0, st, P, tasks, vals, ldv, L, U);
  return hipGetLastError();
}

hipError_t launch_fl_diag(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          double* L, double* Z, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_diag, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z);
  return hipGetLastError();
}

hipError_t launch_fl_trsm(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          double* L, const double* Z, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_trsm, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z);
  return hipGetLastError();
}

hipError_t launch_fl_update(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                            double* L, double* U, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fl_update, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, U);
  return hipGetLastError();
}

hipError_t launch_logdet(const DevPlan& P, int B, const double* L, double* logdet, int32_t* status,
                         hipStream_t st) {
  hipLaunchKernelGGL(k_logdet, dim3(B), dim3(256), 0, st, P, L, logdet, status);
  return hipGetLastError();
}

hipError_t launch_gather_perm(const DevPlan& P, int B, const double* rhs, int64_t ldr, double* X,
                              hipStream_t st) {
  hipLaunchKernelGGL(k_gather_perm, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, rhs, ldr,
                     X);
  return hipGetLastError();
}

hipError_t launch_scatter_perm(const DevPlan& P, int B, const double* X, double* out, int64_t ldo,
                               hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_perm, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, X, out,
                     ldo);
  return hipGetLastError();
}

hipError_t launch_diag_scale(const DevPlan& P, int B, const double* L, double* X, hipStream_t st) {
  hipLaunchKernelGGL(k_diag_scale, dim3(grid1((int64_t)B * P.n)), dim3(256), 0, st, P, B, L, X);
  return hipGetLastError();
}

hipError_t launch_solve_fwd(const DevPlan& P, const int32_t* list, int count, int maxm, int B,
                            const double* L, double* X, double* V, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_solve_fwd, dim3(count, B), dim3(128), (size_t)maxm * sizeof(double), st, P,
                     list, L, X, V);
  return hipGetLastError();
}

hipError_t launch_solve_bwd(const DevPlan& P, const int32_t* list, int count, int maxm, int B,
                            const double* L, double* X, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_solve_bwd, dim3(count, B), dim3(128), (size_t)maxm * sizeof(double), st, P,
                     list, L, X);
  return hipGetLastError();
}

hipError_t launch_selinv_small(const DevPlan& P, const int32_t* list, int count, int maxm, int B,
                               const double* L, double* S, double* U, const double* W,
                               hipStream_t st) {
  if (count <= 0) return hipSuccess;
  size_t sh = ((size_t)maxm * (maxm + 1) + maxm) * sizeof(double);
  hipLaunchKernelGGL(k_selinv_small, dim3(count, B), dim3(64), sh, st, P, list, L, S, U, W);
  return hipGetLastError();
}

hipError_t launch_sl_gather(const DevPlan& P, const int32_t* tasks, int count, int B,
                            const double* S, const double* U, double* W, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_gather, dim3(count, B), dim3(256), 0, st, P, tasks, S, U, W);
  return hipGetLastError();
}

hipError_t launch_sl_lhat(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          const double* L, const double* Z, double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_lhat, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z, T);
  return hipGetLastError();
}

hipError_t launch_sl_gemm(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                          double* S, double* W, const double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_gemm, dim3(count, B), dim3(256), 0, st, P, tasks, step, S, W, T);
  return hipGetLastError();
}

hipError_t launch_sl_pp(const DevPlan& P, const int32_t* tasks, int count, int step, int B,
                        const double* L, const double* Z, double* S, double* W, const double* T,
                        hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_sl_pp, dim3(count, B), dim3(256), 0, st, P, tasks, step, L, Z, S, W, T);
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
