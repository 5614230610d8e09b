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

// Large fronts: one 256-thread workgroup per (front, matrix), front in global
// memory (L block + U workspace). Right-looking over 32-wide panels: factor the
// panel's diagonal block in LDS, solve the rows below (TRSM), then subtract
// L_panel D L_panel^T from the trailing lower triangle with MFMA tiles.
__global__ __launch_bounds__(256) void k_factor_large(DevPlan P, const int32_t* __restrict__ list,
                                                      const double* __restrict__ vals, int64_t ldv,
                                                      double* __restrict__ L,
                                                      double* __restrict__ U) {
  __shared__ double Dg[kPanel][kPanel + 1];
  __shared__ double dpiv[kPanel];
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Ub = U + (int64_t)b * P.u_size + P.u_off[s];
  const double* vb = vals + (int64_t)b * ldv;
  // ---- assemble
  for (int64_t idx = tid; idx < (int64_t)m * w; idx += 256) Lb[idx] = 0.0;
  for (int64_t idx = tid; idx < (int64_t)r * r; idx += 256) Ub[idx] = 0.0;
  __syncthreads();
  for (int64_t q = P.fa_ptr[s] + tid; q < P.fa_ptr[s + 1]; q += 256) Lb[P.fa_slot[q]] = vb[P.fa_csc[q]];
  __syncthreads();
  for (int cq = P.child_ptr[s]; cq < P.child_ptr[s + 1]; ++cq) {
    const int c = P.child_idx[cq];
    const int rc = P.sn_r[c];
    const double* Uc = U + (int64_t)b * P.u_size + P.u_off[c];
    const int32_t* relc = P.rel + P.rows_ptr[c];
    const int64_t tot = (int64_t)rc * rc;
    for (int64_t e = tid; e < tot; e += 256) {
      const int j = (int)(e / rc), i = (int)(e - (int64_t)j * rc);
      if (i < j) continue;
      *front_at(Lb, Ub, m, w, r, relc[i], relc[j]) += Uc[e];
    }
    __syncthreads();
  }
  // ---- panels
  for (int c0 = 0; c0 < w; c0 += kPanel) {
    const int nb = min(kPanel, w - c0);
    for (int idx = tid; idx < nb * nb; idx += 256) {
      const int j = idx / nb, i = idx - j * nb;
      Dg[i][j] = (i >= j) ? Lb[(int64_t)(c0 + j) * m + c0 + i] : 0.0;
    }
    __syncthreads();
    for (int k = 0; k < nb; ++k) {
      const double dk = Dg[k][k];
      if (tid > k && tid < nb) Dg[tid][k] /= dk;
      __syncthreads();
      if (tid > k && tid < nb) {
        const double y = Dg[tid][k] * dk;
        for (int j = k + 1; j <= tid; ++j) Dg[tid][j] -= y * Dg[j][k];
      }
      __syncthreads();
    }
    if (tid < nb) dpiv[tid] = Dg[tid][tid];
    for (int idx = tid; idx < nb * nb; idx += 256) {
      const int j = idx / nb, i = idx - j * nb;
      Lb[(int64_t)(c0 + j) * m + c0 + i] = (i >= j) ? Dg[i][j] : 0.0;
    }
    __syncthreads();
    // rows below the diagonal block: y L11^T = a_i, l_i = y / d
    for (int i = c0 + nb + tid; i < m; i += 256) {
      double y[kPanel];
#pragma unroll
      for (int p = 0; p < kPanel; ++p) {
        if (p < nb) {
          double v = Lb[(int64_t)(c0 + p) * m + i];
          for (int q = 0; q < p; ++q) v -= y[q] * Dg[p][q];
          y[p] = v;
        }
      }
#pragma unroll
      for (int p = 0; p < kPanel; ++p)
        if (p < nb) Lb[(int64_t)(c0 + p) * m + i] = y[p] / dpiv[p];
    }
    __syncthreads();
    // trailing update on [t0, m)^2, lower triangle, 64x64 tiles
    const int t0 = c0 + nb;
    const int nt = (m - t0 + 63) / 64;
    for (int tj = 0; tj < nt; ++tj)
      for (int ti = tj; ti < nt; ++ti) {
        const int R0 = t0 + ti * 64, C0 = t0 + tj * 64;
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
        __syncthreads();
      }
  }
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
                                               int p, int pi, int pj) {
  const int wp = P.sn_start[p + 1] - P.sn_start[p];
  if (pj < wp) return Sm[P.l_off[p] + (int64_t)pj * (wp + P.sn_r[p]) + pi];
  const int rp = P.sn_r[p];
  return Um[P.u_off[p] + (int64_t)(pj - wp) * rp + (pi - wp)];
}

__global__ __launch_bounds__(64) void k_selinv_small(DevPlan P, const int32_t* __restrict__ list,
                                                     const double* __restrict__ L,
                                                     double* __restrict__ S,
                                                     double* __restrict__ U) {
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
  if (r > 0) {
    const int p = P.sn_parent[s];
    const int32_t* rl = P.rel + P.rows_ptr[s];
    double* Us = Um + P.u_off[s];
    for (int j = 0; j < r; ++j) {
      const int pj = rl[j];
      for (int i = j + t; i < r; i += 64) {
        const double v = sigma_parent(P, Sm, Um, p, rl[i], pj);
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

// Large fronts: 256 threads; columns processed in 32-wide panels from the right.
// For panel P=[c0,c1) with B=[c1,m): Lh = L[B,P] L_PP^-1 (into T),
// Sigma[B,P] = -Sigma[B,B] Lh (MFMA tiles), Sigma[P,P] = L_PP^-T D^-1 L_PP^-1 - Lh^T Sigma[B,P].
__global__ __launch_bounds__(256) void k_selinv_large(DevPlan P, const int32_t* __restrict__ list,
                                                      const double* __restrict__ L,
                                                      double* __restrict__ S,
                                                      double* __restrict__ U,
                                                      double* __restrict__ T) {
  __shared__ double Dg[kPanel][kPanel + 1];  // L_PP, then Z = L_PP^-1
  __shared__ double dinv[kPanel];
  __shared__ double As[64][kPanel + 1];
  __shared__ double Bs[64][kPanel + 1];
  const int s = list[blockIdx.x];
  const int b = blockIdx.y;
  const int tid = threadIdx.x;
  const int a0 = P.sn_start[s], w = P.sn_start[s + 1] - a0, r = P.sn_r[s], m = w + r;
  const double* Lb = L + (int64_t)b * P.l_size + P.l_off[s];
  double* Sm = S + (int64_t)b * P.l_size;
  double* Um = U + (int64_t)b * P.u_size;
  double* Sb = Sm + P.l_off[s];
  double* Us = Um + P.u_off[s];
  double* Tb = T + (int64_t)b * P.l_size + P.l_off[s];
  if (r > 0) {
    const int p = P.sn_parent[s];
    const int32_t* rl = P.rel + P.rows_ptr[s];
    const int64_t tot = (int64_t)r * r;
    for (int64_t e = tid; e < tot; e += 256) {
      const int j = (int)(e / r), i = (int)(e - (int64_t)j * r);
      if (i >= j) Us[e] = sigma_parent(P, Sm, Um, p, rl[i], rl[j]);
    }
  }
  __syncthreads();
  // symmetric lookup of Sigma inside this front for (i, k) both >= current c1
  auto sig = [&](int i, int k) -> double {
    const int a = i > k ? i : k, c = i > k ? k : i;
    return c < w ? Sb[(int64_t)c * m + a] : Us[(int64_t)(c - w) * r + (a - w)];
  };
  for (int c1 = w; c1 > 0;) {
    const int c0 = c1 > kPanel ? c1 - kPanel : 0;
    const int nb = c1 - c0;
    for (int idx = tid; idx < nb * nb; idx += 256) {
      const int j = idx / nb, i = idx - j * nb;
      Dg[i][j] = (i > j) ? Lb[(int64_t)(c0 + j) * m + c0 + i] : (i == j ? 1.0 : 0.0);
    }
    if (tid < nb) dinv[tid] = 1.0 / Lb[(int64_t)(c0 + tid) * m + c0 + tid];
    __syncthreads();
    // Lh rows: x L_PP = l_i  (x_p = l_ip - sum_{q>p} x_q L_PP[q][p])
    for (int i = c1 + tid; i < m; i += 256) {
      double x[kPanel];
#pragma unroll
      for (int pp = kPanel - 1; pp >= 0; --pp) {
        x[pp] = 0.0;
        if (pp < nb) {
          double v = Lb[(int64_t)(c0 + pp) * m + i];
#pragma unroll
          for (int q = pp + 1; q < kPanel; ++q)
            if (q < nb) v -= x[q] * Dg[q][pp];
          x[pp] = v;
        }
      }
#pragma unroll
      for (int pp = 0; pp < kPanel; ++pp)
        if (pp < nb) Tb[(int64_t)(c0 + pp) * m + i] = x[pp];
    }
    __syncthreads();
    // Sigma[B, P] = -Sigma[B, B] Lh
    const int nB = m - c1;
    const int nt = (nB + 63) / 64;
    for (int ti = 0; ti < nt; ++ti) {
      const int R0 = c1 + ti * 64;
      TileAcc acc;
      tile_zero(acc);
      mma_tile([&](int ii, int kk) {
                 const int i = R0 + ii;
                 return i < m ? sig(i, c1 + kk) : 0.0;
               },
               [&](int jj, int kk) {
                 return jj < nb ? Tb[(int64_t)(c0 + jj) * m + c1 + kk] : 0.0;
               },
               nB, As, Bs, acc);
      tile_foreach(acc, [&](int ii, int jj, double v) {
        const int i = R0 + ii;
        if (i < m && jj < nb) Sb[(int64_t)(c0 + jj) * m + i] = -v;
      });
      __syncthreads();
    }
    // Z = L_PP^-1 (unit lower), column q by thread q; Dg is overwritten only
    // after every thread has finished reading L_PP.
    {
      const int q = tid;
      double z[kPanel];
#pragma unroll
      for (int i = 0; i < kPanel; ++i) {
        z[i] = 0.0;
        if (q < nb && i < nb && i >= q) {
          double v = (i == q) ? 1.0 : 0.0;
#pragma unroll
          for (int k = 0; k < i; ++k) v -= Dg[i][k] * z[k];  // z[k] = 0 for k < q
          z[i] = v;
        }
      }
      __syncthreads();
      if (q < nb) {
#pragma unroll
        for (int i = 0; i < kPanel; ++i)
          if (i < nb) Dg[i][q] = z[i];
      }
    }
    __syncthreads();
    // Sigma[P,P] lower: sum_k Z[k][p] Z[k][q] / d_k - sum_{i in B} Lh[i][p] Sigma[i][q]
    for (int idx = tid; idx < nb * nb; idx += 256) {
      const int q = idx / nb, pp = idx - q * nb;
      if (pp < q) continue;
      double g = 0.0;
      for (int k = pp; k < nb; ++k) g += Dg[k][pp] * Dg[k][q] * dinv[k];
      double h = 0.0;
      const double* lp = Tb + (int64_t)(c0 + pp) * m;
      const double* sq = Sb + (int64_t)(c0 + q) * m;
      for (int i = c1; i < m; ++i) h += lp[i] * sq[i];
      Sb[(int64_t)(c0 + q) * m + c0 + pp] = g - h;
    }
    __syncthreads();
    c1 = c0;
  }
  // zero the strictly upper part of the diagonal block (unused, kept clean)
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

hipError_t launch_factor_large(const DevPlan& P, const int32_t* list, int count, int B,
                               const double* vals, int64_t ldv, double* L, double* U,
                               hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_factor_large, dim3(count, B), dim3(256), 0, st, P, list, vals, ldv, L, U);
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
                               const double* L, double* S, double* U, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  size_t sh = ((size_t)maxm * (maxm + 1) + maxm) * sizeof(double);
  hipLaunchKernelGGL(k_selinv_small, dim3(count, B), dim3(64), sh, st, P, list, L, S, U);
  return hipGetLastError();
}

hipError_t launch_selinv_large(const DevPlan& P, const int32_t* list, int count, int B,
                               const double* L, double* S, double* U, double* T, hipStream_t st) {
  if (count <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_selinv_large, dim3(count, B), dim3(256), 0, st, P, list, L, S, U, T);
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
