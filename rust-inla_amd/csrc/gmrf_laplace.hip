// Kernels of the batched Laplace step: Q(theta) assembly, likelihood
// derivatives, A^T W A scatter (gather form), sparse products, reductions and
// the adjoint trace contractions. Every kernel works on `B` theta-points laid
// out one after another (row b of a [B x len] array).
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../include/gmrf_c.h"
#include "plan.hpp"

namespace gmrf {

// Compiler-only fence after a round of independent loads: keeps them issued
// ahead of the conditional code that consumes them (no instruction emitted).
#define LAP_ISSUE_FENCE() asm volatile("" ::: "memory")

// Gather lists longer than kLongGather (a dense row such as an intercept's column
// of A, summing every observation) go to a workgroup each (k_*_long) instead of one
// thread.
// The same sum by a 256-thread workgroup: thread t takes t0 + t, t0 + t + 256, ...
// (coalesced index and value loads, 8 per round), then a fixed-order tree over the
// threads. Every thread of the block must call it; the block total is returned to
// all threads.
template <typename IT>
__device__ __forceinline__ double gather_dot_block(int64_t t0, int64_t t1, const IT* __restrict__ idx,
                                                   const double* __restrict__ c,
                                                   const double* __restrict__ x, double* red) {
  const int tid = threadIdx.x;
  double s = 0.0;
  const int64_t tl = t1 > t0 ? t1 - 1 : 0;
  for (int64_t tb = t0; tb < t1; tb += 8 * 256) {
    IT ix[8];
    double cv[8], xv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t t = min(tb + tid + 256 * u, tl);
      ix[u] = idx[t];
      cv[u] = c[t];
    }
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 8; ++u) xv[u] = x[ix[u]];
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 8; ++u) s += (tb + tid + 256 * u < t1) ? cv[u] * xv[u] : 0.0;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((tid & 63) == 0) red[tid >> 6] = s;
  __syncthreads();
  const double tot = (red[0] + red[1]) + (red[2] + red[3]);
  __syncthreads();
  return tot;
}

// sum_{t in [t0, t1)} c[t] * x[idx[t]] with the index and value loads issued in
// rounds of 8 (clamped addresses, masked terms). idx and c are never empty
// arrays (allocations hold at least one element).
template <typename IT, int R = 8>
__device__ __forceinline__ double gather_dot(int64_t t0, int64_t t1, const IT* __restrict__ idx,
                                             const double* __restrict__ c,
                                             const double* __restrict__ x) {
  double s = 0.0;
  const int64_t tl = t1 > t0 ? t1 - 1 : 0;
  for (int64_t tb = t0; tb < t1; tb += R) {
    IT ix[R];
    double cv[R], xv[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
      const int64_t t = min(tb + u, tl);
      ix[u] = idx[t];
      cv[u] = c[t];
    }
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < R; ++u) xv[u] = x[ix[u]];
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < R; ++u) s += (tb + u < t1) ? cv[u] * xv[u] : 0.0;
  }
  return s;
}

// Likelihood families / links (crates/inla_stats/src/inference.rs:16-21,170-311).
enum { FAM_GAUSSIAN = 0, FAM_POISSON = 1, FAM_BINOMIAL = 2, FAM_NBINOMIAL = 3 };
enum { LINK_IDENTITY = 0, LINK_LOG = 1, LINK_LOGIT = 2 };

// Lanczos log-gamma (g = 7, n = 9; the method of inference.rs:1896-1920).
__device__ double lanczos_lgamma(double z) {
  const double PI = 3.14159265358979323846;
  double refl = 0.0, sgn = 1.0;
  if (z < 0.5) {
    refl = log(PI) - log(sin(PI * z));
    z = 1.0 - z;
    sgn = -1.0;
  }
  const double c[8] = {676.5203681218851,     -1259.1392167224028,  771.32342877765313,
                       -176.61502916214059,   12.507343278686905,   -0.13857109526572012,
                       9.9843695780195716e-6, 1.5056327351493116e-7};
  const double z1 = z - 1.0;
  double x = 0.99999999999980993;
  for (int i = 0; i < 8; ++i) x += c[i] / (z1 + (double)i + 1.0);
  const double t = z1 + 7.5;
  const double v = 0.5 * log(2.0 * PI) + (z1 + 0.5) * log(t) - t + log(x);
  return sgn < 0 ? refl - v : v;
}

// link_forward (inference.rs:1641-1664): mu, dmu/deta, d2mu/deta2, plus d3mu.
__device__ __forceinline__ void link_fwd(int link, double eta, double& mu, double& d1, double& d2,
                                         double& d3) {
  if (link == LINK_IDENTITY) {
    mu = eta; d1 = 1.0; d2 = 0.0; d3 = 0.0;
  } else if (link == LINK_LOG) {
    mu = exp(eta); d1 = mu; d2 = mu; d3 = mu;
  } else {
    double p;
    if (eta >= 0.0) {
      const double z = exp(-eta);
      p = 1.0 / (1.0 + z);
    } else {
      const double z = exp(eta);
      p = z / (1.0 + z);
    }
    mu = p;
    d1 = p * (1.0 - p);
    d2 = d1 * (1.0 - 2.0 * p);
    d3 = d1 * (1.0 - 6.0 * p + 6.0 * p * p);
  }
}

// Per observation: logp, grad = dlogp/deta, w = -hess (the Newton weight the
// reference uses: Fisher for Poisson / binomial, observed otherwise), dw/deta,
// and for the Gaussian precision hyperparameter d(logp)/dlog(prec) (dl_par).
// ok = 0 when the reference would return an error or a non-finite logp; inv = 1
// only for the inputs on which the reference returns an error (Ok(-inf) for a zero
// Poisson rate with y > 0 is not one).
struct LikOut {
  double logp, grad, w, dw, dl_par;
  int ok, inv;
};

__device__ LikOut eval_obs(int fam, int link, double eta, double y, double a, double b, double lpar) {
  LikOut o{0, 0, 0, 0, 0, 1, 0};
  if (!isfinite(eta)) {
    o.ok = 0;
    o.inv = 1;
    return o;
  }
  double mu, d1, d2, d3;
  link_fwd(link, eta, mu, d1, d2, d3);
  if (fam == FAM_GAUSSIAN) {  // 170-186; precision = lpar * a
    const double prec = lpar * a;
    const double r = mu - y;
    o.logp = -0.9189385332046728 + 0.5 * log(prec) - 0.5 * prec * r * r;
    o.grad = -prec * r * d1;
    o.w = prec * (d1 * d1 + r * d2);
    o.dw = prec * (3.0 * d1 * d2 + r * d3);
    o.dl_par = 0.5 - 0.5 * prec * r * r;
    if (!(prec > 0.0) || !isfinite(prec) || !isfinite(y)) o.ok = 0, o.inv = 1;
  } else if (fam == FAM_POISSON) {  // 188-223; a = exposure
    const double m = a * mu, dm = a * d1;
    if (m > 0.0) {
      o.logp = y * log(m) - m - lanczos_lgamma(y + 1.0);
      o.grad = (y / m - 1.0) * dm;
      o.w = dm * dm / m;
      // d/deta (dm^2 / m) with dm = a d1, d(dm)/deta = a d2
      o.dw = (2.0 * dm * a * d2 * m - dm * dm * dm) / (m * m);
    } else {
      o.logp = (y == 0.0) ? 0.0 : -INFINITY;
      o.grad = -dm;
      o.w = fabs(a * d2);
      o.dw = 0.0;
    }
    if (a < 0.0 || !isfinite(a) || y < 0.0 || !isfinite(y) || mu < 0.0) o.ok = 0, o.inv = 1;
  } else if (fam == FAM_BINOMIAL) {  // 225-267; a = trials
    const double n = a, p = mu;
    const double lc = lanczos_lgamma(n + 1.0) - lanczos_lgamma(y + 1.0) - lanczos_lgamma(n - y + 1.0);
    if (p == 0.0)
      o.logp = (y == 0.0) ? lc : -INFINITY;
    else if (p == 1.0)
      o.logp = (y == n) ? lc : -INFINITY;
    else
      o.logp = lc + y * log(p) + (n - y) * log(1.0 - p);
    const double g = (p > 0.0 && p < 1.0) ? y / p - (n - y) / (1.0 - p) : 0.0;
    o.grad = g * d1;
    o.w = n * d1;
    o.dw = n * d2;
    if (n < 0.0 || y < 0.0 || y > n || !isfinite(n) || !isfinite(y) || p < 0.0 || p > 1.0) o.ok = 0, o.inv = 1;
  } else {  // 269-311; a = exposure, b = size
    const double m = a * mu, dm = a * d1, d2m = a * d2, d3m = a * d3, r = b, mr = m + r;
    if (!(m > 0.0) || !(r > 0.0) || !isfinite(r) || y < 0.0 || !isfinite(y) || a < 0.0 || !isfinite(a)) {
      o.ok = 0;
      o.inv = 1;
      return o;
    }
    o.logp = lanczos_lgamma(y + r) - lanczos_lgamma(r) - lanczos_lgamma(y + 1.0) + r * log(r / mr) +
             y * log(m / mr);
    const double dl = y / m - (y + r) / mr;
    const double d2l = -y / (m * m) + (y + r) / (mr * mr);
    const double d3l = 2.0 * y / (m * m * m) - 2.0 * (y + r) / (mr * mr * mr);
    o.grad = dl * dm;
    const double hess = d2l * dm * dm + dl * d2m;
    o.w = -hess;
    o.dw = -(d3l * dm * dm * dm + 3.0 * d2l * dm * d2m + dl * d3m);
  }
  if (!isfinite(o.logp)) o.ok = 0;
  return o;
}

__global__ void k_lik(int B, int64_t nobs, int fam, int link, const double* __restrict__ eta,
                      const double* __restrict__ y, const double* __restrict__ pa,
                      const double* __restrict__ pb, const double* __restrict__ lpar,
                      double* __restrict__ logp, double* __restrict__ grad, double* __restrict__ w,
                      double* __restrict__ dw, double* __restrict__ dlpar, int32_t* __restrict__ bad,
                      int32_t* __restrict__ invalid, int nb, const int32_t* __restrict__ oblk) {
  const int64_t tot = (int64_t)B * nobs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nobs);
    const int64_t i = e - (int64_t)b * nobs;
    const int vp = b * nb + (oblk ? oblk[i] : 0);  // (point, block)
    const LikOut o = eval_obs(fam, link, eta[e], y[i], pa[i], pb ? pb[i] : 0.0, lpar ? lpar[vp] : 1.0);
    if (logp) logp[e] = o.logp;
    if (grad) grad[e] = o.grad;
    if (w) w[e] = o.w;
    if (dw) dw[e] = o.dw;
    if (dlpar) dlpar[e] = o.dl_par;
    if (!o.ok && bad) bad[vp] = 1;  // benign race: every writer stores 1
    if (o.inv && invalid) invalid[vp] = 1;
  }
}

// vals[b][e] = sum_k coefs[b][k] comps[k][e]
// (with blocks: the coefficients of the entry's block, coefs[(b nb + eblk[e]) K + k])
__global__ void k_affine(int B, int K, int64_t nnz, const double* __restrict__ coefs,
                         const double* __restrict__ comps, double* __restrict__ out, int nb,
                         const int32_t* __restrict__ eblk) {
  const int64_t tot = (int64_t)B * nnz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nnz);
    const int64_t q = e - (int64_t)b * nnz;
    const double* cb = coefs + (int64_t)(b * nb + (eblk ? eblk[q] : 0)) * K;
    double s = 0.0;
    for (int k = 0; k < K; ++k) s += cb[k] * comps[(int64_t)k * nnz + q];
    out[e] = s;
  }
}

// out[b][diag_idx[i]] += add (prior jitter on the stored diagonal)
__global__ void k_add_diag(int B, int64_t n, int64_t nnz, const int64_t* __restrict__ diag_idx,
                           double add, double* __restrict__ out) {
  const int64_t tot = (int64_t)B * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / n);
    const int64_t i = e - (int64_t)b * n;
    out[(int64_t)b * nnz + diag_idx[i]] += add;
  }
}

// Q_post = Q + A^T diag(w) A on the fixed pattern: every CSC entry gathers its
// observation products (ptr/obs/coef built once per model), no atomics.
__global__ void k_add_atwa(int B, int64_t nnz, int64_t nobs, const int64_t* __restrict__ ptr,
                           const int32_t* __restrict__ obs, const double* __restrict__ coef,
                           const double* __restrict__ w, const double* __restrict__ base,
                           double* __restrict__ out) {
  const int64_t tot = (int64_t)B * nnz;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nnz);
    const int64_t q = e - (int64_t)b * nnz;
    const double* wb = w + (int64_t)b * nobs;
    const int64_t p0 = ptr[q], p1 = ptr[q + 1];
    if (p1 - p0 > kLongGather) continue;  // k_add_atwa_long
    // rounds of 4 terms: most lists hold 1-4 observation pairs
    out[e] = base[e] + gather_dot<int32_t, 4>(p0, p1, obs, coef, wb);
  }
}

// The entries with long observation lists: one workgroup per (entry, matrix).
__global__ __launch_bounds__(256) void k_add_atwa_long(int64_t nnz, int64_t nobs, const int64_t* __restrict__ longs,
                                                       const int64_t* __restrict__ ptr,
                                                       const int32_t* __restrict__ obs,
                                                       const double* __restrict__ coef,
                                                       const double* __restrict__ w,
                                                       const double* __restrict__ base,
                                                       double* __restrict__ out) {
  __shared__ double red[4];
  const int64_t q = longs[blockIdx.x];
  const int b = blockIdx.y;
  const double s = gather_dot_block(ptr[q], ptr[q + 1], obs, coef, w + (int64_t)b * nobs, red);
  const int64_t e = (int64_t)b * nnz + q;
  if (threadIdx.x == 0) out[e] = base[e] + s;
}

// y[b][o] = alpha * sum_{p in [ptr[o], ptr[o+1])} vals[b*vld + p] x[b][idx[p]] + beta * y[b][o]
__global__ void k_gather_spmv(int B, int64_t nout, const int64_t* __restrict__ ptr,
                              const int32_t* __restrict__ idx, const double* __restrict__ vals,
                              int64_t vld, const double* __restrict__ x, int64_t ldx, double alpha,
                              double beta, double* __restrict__ y, int64_t ldy) {
  const int64_t tot = (int64_t)B * nout;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nout);
    const int64_t o = e - (int64_t)b * nout;
    const double* vb = vals + (int64_t)b * vld;
    const double* xb = x + (int64_t)b * ldx;
    const int64_t p0 = ptr[o], p1 = ptr[o + 1];
    if (p1 - p0 > kLongGather) continue;  // k_gather_spmv_long
    const double s = gather_dot(p0, p1, idx, vb, xb);
    double* yo = y + (int64_t)b * ldy + o;
    *yo = alpha * s + (beta == 0.0 ? 0.0 : beta * *yo);
  }
}

// The outputs with long lists: one workgroup per (output, matrix).
__global__ __launch_bounds__(256) void k_gather_spmv_long(const int64_t* __restrict__ longs,
                                                          const int64_t* __restrict__ ptr,
                                                          const int32_t* __restrict__ idx,
                                                          const double* __restrict__ vals, int64_t vld,
                                                          const double* __restrict__ x, int64_t ldx,
                                                          double alpha, double beta, double* __restrict__ y,
                                                          int64_t ldy) {
  __shared__ double red[4];
  const int64_t o = longs[blockIdx.x];
  const int b = blockIdx.y;
  const double s = gather_dot_block(ptr[o], ptr[o + 1], idx, vals + (int64_t)b * vld, x + (int64_t)b * ldx, red);
  if (threadIdx.x == 0) {
    double* yo = y + (int64_t)b * ldy + o;
    *yo = alpha * s + (beta == 0.0 ? 0.0 : beta * *yo);
  }
}

// out[b][i] = a[b][i] + alpha[b] * c[b][i]
// (with blocks: alpha of the element's block, alpha[b nb + blk[i]])
__global__ void k_axpy_b(int B, int64_t len, const double* __restrict__ a,
                         const double* __restrict__ alpha, const double* __restrict__ c,
                         double* __restrict__ out, int nb, const int32_t* __restrict__ blk) {
  const int64_t tot = (int64_t)B * len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / len);
    const int64_t i = e - (int64_t)b * len;
    out[e] = a[e] + alpha[b * nb + (blk ? blk[i] : 0)] * c[e];
  }
}

// Per-row reductions over [B x len]: op 0 max|x|, 1 sum x, 2 sum x*y, 3 max|x - y|.
// One 1024-thread block per row, 8 loads per thread per round; fixed order.
__device__ __forceinline__ double red_combine(int op, double a, double v) {
  return (op == 0 || op == 3) ? fmax(a, v) : a + v;
}

// With blocks (seg non-null, nb segments per row) block b nb + q reduces the row's
// segment [seg[q], seg[q + 1]) into out[b nb + q].
__global__ __launch_bounds__(1024) void k_rowreduce(int64_t len, int op, const double* __restrict__ x,
                                                    const double* __restrict__ y,
                                                    double* __restrict__ out, int nb,
                                                    const int64_t* __restrict__ seg) {
  __shared__ double red[16];
  const int bq = blockIdx.x, tid = threadIdx.x;
  const int b = bq / nb, q = bq - b * nb;
  const int64_t s0 = seg ? seg[q] : 0, s1 = seg ? seg[q + 1] : len;
  const double* xb = x + (int64_t)b * len;
  const double* yb = y ? y + (int64_t)b * len : xb;
  double acc = 0.0;
  for (int64_t i0 = s0; i0 < s1; i0 += 8 * 1024) {
    double xv[8], yv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int64_t i = min(i0 + tid + 1024 * u, s1 - 1);
      xv[u] = xb[i];
      yv[u] = yb[i];
    }
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      if (i0 + tid + 1024 * u >= s1) continue;
      const double v = xv[u];
      if (op == 0)
        acc = isfinite(v) && isfinite(acc) ? fmax(acc, fabs(v)) : INFINITY;
      else if (op == 1)
        acc += v;
      else if (op == 2)
        acc += v * yv[u];
      else
        acc = fmax(acc, fabs(v - yv[u]));
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc = red_combine(op, acc, __shfl_xor(acc, o, 64));
  if ((tid & 63) == 0) red[tid >> 6] = acc;
  __syncthreads();
  if (tid == 0) {
    double t = red[0];
    for (int r = 1; r < 16; ++r) t = red_combine(op, t, red[r]);
    out[bq] = t;
  }
}

// Adjoint contractions over the CSC pattern, two passes (deterministic order):
//   tr[b][k]   = sum_e comps[k][e] * Sigma_b[slot_full[e]]     (= tr(Sigma_b Q_k))
//   quad[b][k] = sum_e comps[k][e] * x_b[row_e] * z_b[col_e]   (= x_b^T Q_k z_b)
// Pass 1: block (chunk, g, b) reduces one chunk of the entries into part[] for the
// components k = kContractGroup g .. (at most kContractGroup at once): the entry's
// index and Sigma / x / z gathers are shared by the group's components.
constexpr int kContractChunks = 64;
constexpr int kContractGroup = 8;

// With blocks (eseg non-null) the sums run per (b, block q) over the block's entry
// range [eseg[q], eseg[q + 1]), chunked the same way; grid z = B nb.
__global__ __launch_bounds__(256) void k_contract_part(int K, int64_t nnz,
                                                       const double* __restrict__ comps,
                                                       const double* __restrict__ S, int64_t l_size,
                                                       const int64_t* __restrict__ slot_full,
                                                       const int32_t* __restrict__ qrow,
                                                       const int32_t* __restrict__ qcol,
                                                       const double* __restrict__ x,
                                                       const double* __restrict__ z, int64_t n,
                                                       double* __restrict__ part, int nb,
                                                       const int64_t* __restrict__ eseg) {
  constexpr int G = kContractGroup;
  __shared__ double red[4][2 * G];
  const int ch = blockIdx.x, k0 = blockIdx.y * G, bq = blockIdx.z;
  const int b = bq / nb, qb = bq - b * nb;
  const int kn = min(G, K - k0);
  const int64_t g0 = eseg ? eseg[qb] : 0, g1 = eseg ? eseg[qb + 1] : nnz;
  const int64_t len = (g1 - g0 + kContractChunks - 1) / kContractChunks;
  const int64_t e0 = g0 + ch * len, e1 = min(g1, e0 + len);
  const double* Sb = S ? S + (int64_t)b * l_size : nullptr;
  const double* xb = x ? x + (int64_t)b * n : nullptr;
  const double* zb = z ? z + (int64_t)b * n : nullptr;
  double a1[G], a2[G];
#pragma unroll
  for (int g = 0; g < G; ++g) a1[g] = a2[g] = 0.0;
  for (int64_t eb = e0 + threadIdx.x; eb < e1; eb += 2 * 256) {
    double sv[2], xz[2];
    int64_t sl[2];
    int32_t qr[2], qc[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const int64_t e = min(eb + 256 * u, e1 - 1);
      sl[u] = Sb ? slot_full[e] : 0;
      qr[u] = xb ? qrow[e] : 0;
      qc[u] = xb ? qcol[e] : 0;
    }
    LAP_ISSUE_FENCE();
    double sg[2], xg[2], zg[2];
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      sg[u] = Sb ? Sb[max(sl[u], (int64_t)0)] : 0.0;
      xg[u] = xb ? xb[qr[u]] : 0.0;
      zg[u] = xb ? zb[qc[u]] : 0.0;
    }
    LAP_ISSUE_FENCE();
#pragma unroll
    for (int u = 0; u < 2; ++u) {
      const bool ok = eb + 256 * u < e1;
      sv[u] = (ok && sl[u] >= 0) ? sg[u] : 0.0;
      xz[u] = ok ? xg[u] * zg[u] : 0.0;
    }
#pragma unroll
    for (int g = 0; g < G; ++g) {
      if (g < kn) {
        const double* ck = comps + (int64_t)(k0 + g) * nnz;
#pragma unroll
        for (int u = 0; u < 2; ++u) {
          const double cv = ck[min(eb + 256 * u, e1 - 1)];
          a1[g] += cv * sv[u];
          a2[g] += cv * xz[u];
        }
      }
    }
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
  for (int g = 0; g < G; ++g) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      a1[g] += __shfl_xor(a1[g], o, 64);
      a2[g] += __shfl_xor(a2[g], o, 64);
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int g = 0; g < G; ++g) {
      red[wv][2 * g] = a1[g];
      red[wv][2 * g + 1] = a2[g];
    }
  }
  __syncthreads();
  if ((int)threadIdx.x < 2 * kn) {
    const int t = threadIdx.x;
    const double v = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
    const int64_t q = ((int64_t)bq * K + k0 + (t >> 1)) * kContractChunks + ch;
    part[2 * q + (t & 1)] = v;
  }
}

// Pass 2: one thread per (b, k) sums its chunks in order.
__global__ void k_contract_sum(int BK, const double* __restrict__ part, double* __restrict__ tr,
                               double* __restrict__ quad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BK) return;
  double a1 = 0.0, a2 = 0.0;
  for (int ch = 0; ch < kContractChunks; ++ch) {
    a1 += part[2 * ((int64_t)t * kContractChunks + ch)];
    a2 += part[2 * ((int64_t)t * kContractChunks + ch) + 1];
  }
  if (tr) tr[t] = a1;
  if (quad) quad[t] = a2;
}

// v[b][i] = (A Sigma_b A^T)_ii = sum over the row's entry pairs of A_ia A_ib Sigma_b[ja, jb]
__global__ void k_vdiag(int B, int64_t nobs, const int64_t* __restrict__ vptr,
                        const int64_t* __restrict__ vslot, const double* __restrict__ vcoef,
                        const double* __restrict__ S, int64_t l_size, double* __restrict__ v) {
  const int64_t tot = (int64_t)B * nobs;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / nobs);
    const int64_t i = e - (int64_t)b * nobs;
    const double* Sb = S + (int64_t)b * l_size;
    v[e] = gather_dot(vptr[i], vptr[i + 1], vslot, vcoef, Sb);
  }
}

// Constraint projection pieces: r[b][k] = sum_j ca[k][j] x[b][j]; x[b][j] -= sum_k ca[k][j] lam[b][k]
__global__ __launch_bounds__(256) void k_constr_dot(int kc, int64_t n, const double* __restrict__ ca,
                                                    const double* __restrict__ x,
                                                    double* __restrict__ r) {
  __shared__ double red[256];
  const int k = blockIdx.x, b = blockIdx.y;
  double acc = 0.0;
  for (int64_t j = threadIdx.x; j < n; j += 256) acc += ca[(int64_t)k * n + j] * x[(int64_t)b * n + j];
  red[threadIdx.x] = acc;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) r[b * kc + k] = red[0];
}

__global__ void k_constr_apply(int B, int kc, int64_t n, const double* __restrict__ ca,
                               const double* __restrict__ lam, double scale, double* __restrict__ x) {
  const int64_t tot = (int64_t)B * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / n);
    const int64_t j = e - (int64_t)b * n;
    double corr = 0.0;
    for (int k = 0; k < kc; ++k) corr += ca[(int64_t)k * n + j] * lam[b * kc + k];
    x[e] -= scale * corr;
  }
}

// ---------------------------------------------------------------------------
// Low-rank constraint augmentation. With k constraint rows C, the augmented
// precision Q + kappa C^T C (constraints.rs:275-313 forms it explicitly, which
// makes the pattern dense for long rows) is written as Q_g + U E U^T with the
// grounded Q_g = Q + sum_t alpha_t e_{p_t} e_{p_t}^T (same pattern; non-singular
// when the constraints span the null space of Q), U = [C^T, e_p] (n x 2k) and
// E = diag(kappa I, -alpha). Log-determinants, solves, diagonals and traces of
// the augmented matrix follow from Y = Q_g^-1 U and K = E^-1 + U^T Y.
// ---------------------------------------------------------------------------

// dst_b[idx_t] += alpha_bt with alpha_bt = src_b[idx_t] (1 where that is not
// positive); alpha is also written out (B x k).
__global__ void k_ground(int B, int k, int64_t nnz, const int64_t* __restrict__ idx,
                         const double* __restrict__ src, double* __restrict__ dst,
                         double* __restrict__ alpha) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= B * k) return;
  const int b = e / k, t = e - b * k;
  const int64_t q = (int64_t)b * nnz + idx[t];
  const double v = src[q];
  const double a = v > 0.0 ? v : 1.0;
  dst[q] += a;
  alpha[e] = a;
}

// Partial Gram products over one chunk of 2048 entries:
// part[(b nchunk + c) RU RY + i RY + j] = sum_{x in chunk c} U_b[i][x] Y_b[j][x],
// U_b[i] = U + b ldub + i ldu (ldub = 0: one U for all b), Y likewise. Operands
// are staged through LDS 32 entries at a time; each thread owns up to 16 (i, j).
constexpr int kGramChunk = 2048, kGramSub = 32;

__global__ __launch_bounds__(256) void k_gram_part(int RU, int RY, int64_t n,
                                                   const double* __restrict__ U, int64_t ldu,
                                                   int64_t ldub, const double* __restrict__ Y,
                                                   int64_t ldy, int64_t ldyb, int nchunk,
                                                   double* __restrict__ part) {
  __shared__ double us[64][kGramSub + 1];
  __shared__ double ys[64][kGramSub + 1];
  const int c = blockIdx.x, b = blockIdx.y, tid = threadIdx.x;
  const int np = RU * RY;
  const double* Ub = U + (int64_t)b * ldub;
  const double* Yb = Y + (int64_t)b * ldyb;
  double acc[16];
#pragma unroll
  for (int u = 0; u < 16; ++u) acc[u] = 0.0;
  const int64_t x0 = (int64_t)c * kGramChunk, x1 = min(n, x0 + kGramChunk);
  for (int64_t xs = x0; xs < x1; xs += kGramSub) {
    for (int e = tid; e < (RU + RY) * kGramSub; e += 256) {
      const int row = e / kGramSub, q = e - row * kGramSub;
      const int64_t x = min(xs + q, x1 - 1);
      const bool isu = row < RU;
      const double* src = isu ? Ub + (int64_t)row * ldu : Yb + (int64_t)(row - RU) * ldy;
      const double v = src[x];
      const double vv = xs + q < x1 ? v : 0.0;
      if (isu)
        us[row][q] = vv;
      else
        ys[row - RU][q] = vv;
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < 16; ++u) {
      const int p = tid + 256 * u;
      if (p < np) {
        const int i = p / RY, j = p - (p / RY) * RY;
        double a = acc[u];
#pragma unroll 8
        for (int q = 0; q < kGramSub; ++q) a += us[i][q] * ys[j][q];
        acc[u] = a;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int u = 0; u < 16; ++u) {
    const int p = tid + 256 * u;
    if (p < np) part[((int64_t)b * nchunk + c) * np + p] = acc[u];
  }
}

// out[b np + p] = sum_c part[(b nchunk + c) np + p], chunks in order.
__global__ void k_gram_sum(int np, int nchunk, const double* __restrict__ part,
                           double* __restrict__ out) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x, b = blockIdx.y;
  if (p >= np) return;
  double s = 0.0;
  for (int c = 0; c < nchunk; ++c) s += part[((int64_t)b * nchunk + c) * np + p];
  out[(int64_t)b * np + p] = s;
}

// x_b[i] -= sum_j Y_b[j][i] h_b[j]
__global__ void k_lowrank_axpy(int B, int R, int64_t n, const double* __restrict__ Y, int64_t ldy,
                               int64_t ldyb, const double* __restrict__ h, double* __restrict__ x,
                               int64_t ldx) {
  const int64_t tot = (int64_t)B * n;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / n);
    const int64_t i = e - (int64_t)b * n;
    const double* Yb = Y + (int64_t)b * ldyb + i;
    double a = 0.0;
    for (int j = 0; j < R; ++j) a += Yb[(int64_t)j * ldy] * h[b * R + j];
    x[(int64_t)b * ldx + i] -= a;
  }
}

// d_b[i] -= y^T Kinv_b y with y_j = Y_b[j][i]: the diagonal of Y Kinv Y^T.
// 128 entries per workgroup; y (R x 128) and Kinv_b (R x R) in LDS.
__global__ __launch_bounds__(128) void k_lowrank_diag(int R, int64_t n, const double* __restrict__ Y,
                                                      int64_t ldy, int64_t ldyb,
                                                      const double* __restrict__ Kinv,
                                                      double* __restrict__ d, int64_t ldd) {
  extern __shared__ double sm[];
  double* ks = sm;            // R x R
  double* yv = sm + R * R;    // R x 128
  const int b = blockIdx.y, tid = threadIdx.x;
  const int64_t i = (int64_t)blockIdx.x * 128 + tid;
  const int64_t ic = min(i, n - 1);
  for (int e = tid; e < R * R; e += 128) ks[e] = Kinv[(int64_t)b * R * R + e];
  const double* Yb = Y + (int64_t)b * ldyb + ic;
  for (int j = 0; j < R; ++j) yv[j * 128 + tid] = Yb[(int64_t)j * ldy];
  __syncthreads();
  double q = 0.0;
  for (int j = 0; j < R; ++j) {
    double s = 0.0;
    for (int l = 0; l < R; ++l) s += ks[j * R + l] * yv[l * 128 + tid];
    q += yv[j * 128 + tid] * s;
  }
  if (i < n) d[(int64_t)b * ldd + i] -= q;
}

// out = a .* b (flat)
// out[b][perm[k]] = in[b][k]: rows of B vectors from the internal observation order
// to the caller's (perm[k] = the caller's index of internal observation k)
__global__ void k_scatter_rows(int B, int64_t len, const int64_t* __restrict__ perm,
                               const double* __restrict__ in, double* __restrict__ out) {
  const int64_t tot = (int64_t)B * len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int64_t b = e / len, k = e - b * len;
    out[b * len + (perm ? perm[k] : k)] = in[e];
  }
}

__global__ void k_mul(int64_t tot, const double* __restrict__ a, const double* __restrict__ b,
                      double* __restrict__ out) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x)
    out[e] = a[e] * b[e];
}

// x[b] <- keep[b] ? xt[b] : x[b]
__global__ void k_blend(int B, int64_t len, const double* __restrict__ keep,
                        const double* __restrict__ xt, double* __restrict__ x, int nb,
                        const int32_t* __restrict__ blk) {
  const int64_t tot = (int64_t)B * len;
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < tot;
       e += (int64_t)gridDim.x * blockDim.x) {
    const int b = (int)(e / len);
    const int64_t i = e - (int64_t)b * len;
    if (keep[b * nb + (blk ? blk[i] : 0)] != 0.0) x[e] = xt[e];
  }
}

// ---------------------------------------------------------------------------
static inline int grid_for(int64_t tot) {
  int64_t g = (tot + 255) / 256;
  if (g > 65535) g = 65535;
  return (int)(g < 1 ? 1 : g);
}

// Fill kernels: clearing inside a captured graph goes through a kernel, not a memset node.
__global__ void k_fill_f64(int64_t len, double v, double* __restrict__ p) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}
__global__ void k_fill_i32(int64_t len, int32_t v, int32_t* __restrict__ p) {
  for (int64_t e = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; e < len; e += (int64_t)gridDim.x * blockDim.x)
    p[e] = v;
}

// End of the Gaussian-identity evaluation on the device, one thread per point b:
// status (the Newton step's validity, then the prior's and the posterior's factor
// flags, in the order of the host path), the Laplace marginal likelihood
// sum log p - x^T Q x / 2 + log|Q_prior| / 2 - log|Q_post| / 2 (inference.rs:972) and
// the gradient sum_k dcoef_k (tr(Sigma_prior Q_k) - tr(Sigma_post Q_k) - x^T Q_k x) / 2
// plus, for the Gaussian precision, dlog(prec) (sum dlogp/dlog(prec) - sum v w / 2).
// fstat[b] is the posterior's factor status, fstat[B + b] the prior's; ld likewise.
__global__ void k_gauss_finish(int B, int K, int mt, int fam, const int32_t* __restrict__ fstat,
                               const int32_t* __restrict__ bad, const double* __restrict__ maxstep,
                               const double* __restrict__ ld, const double* __restrict__ quad,
                               const double* __restrict__ loglik, const double* __restrict__ tr_post,
                               const double* __restrict__ tr_prior, const double* __restrict__ qx,
                               const double* __restrict__ dcoefs, const double* __restrict__ dlog_lpar,
                               const double* __restrict__ sum_dl, const double* __restrict__ sum_vw,
                               double* __restrict__ mlik, double* __restrict__ grad, int32_t* __restrict__ status) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= B) return;
  // factor flags: bit 1 a zero pivot (not positive definite), bit 0 a pivot below
  // 1e-14 or non-finite (singular)
  auto code = [](int32_t f) { return (f & 2) ? (int)GMRF_NOT_POSITIVE_DEFINITE : (f & 1) ? (int)GMRF_SINGULAR : 0; };
  int st = (bad[b] || !isfinite(maxstep[b])) ? (int)GMRF_SOLVE_FAILURE : 0;
  const int fp = code(fstat[B + b]), fq = code(fstat[b]);
  if (fp != 0)
    st = fp;
  else if (st == 0 && fq != 0)
    st = fq;
  status[b] = st;
  mlik[b] = st == 0 ? loglik[b] - 0.5 * quad[b] + 0.5 * ld[B + b] - 0.5 * ld[b] : -INFINITY;
  if (grad) {
    for (int l = 0; l < mt; ++l) {
      double gq = 0.0;
      for (int k = 0; k < K; ++k) {
        const double c = dcoefs[((int64_t)b * mt + l) * K + k];
        gq += c * (-0.5 * qx[b * K + k] + 0.5 * tr_prior[b * K + k] - 0.5 * tr_post[b * K + k]);
      }
      if (fam == FAM_GAUSSIAN && dlog_lpar) gq += dlog_lpar[(int64_t)b * mt + l] * (sum_dl[b] - 0.5 * sum_vw[b]);
      grad[(int64_t)b * mt + l] = st == 0 ? gq : NAN;
    }
  }
}

hipError_t lap_lik(int B, int64_t nobs, int fam, int link, const double* eta, const double* y,
                   const double* pa, const double* pb, const double* lpar, double* logp, double* grad,
                   double* w, double* dw, double* dlpar, int32_t* bad, int32_t* invalid, hipStream_t st,
                   const LapBlocks* bk) {
  hipLaunchKernelGGL(k_lik, dim3(grid_for((int64_t)B * nobs)), dim3(256), 0, st, B, nobs, fam, link,
                     eta, y, pa, pb, lpar, logp, grad, w, dw, dlpar, bad, invalid, bk ? bk->nb : 1,
                     bk ? bk->oblk : nullptr);
  return hipGetLastError();
}

hipError_t lap_affine(int B, int K, int64_t nnz, const double* coefs, const double* comps,
                      double* out, hipStream_t st, const LapBlocks* bk) {
  hipLaunchKernelGGL(k_affine, dim3(grid_for((int64_t)B * nnz)), dim3(256), 0, st, B, K, nnz, coefs,
                     comps, out, bk ? bk->nb : 1, bk ? bk->eblk : nullptr);
  return hipGetLastError();
}

hipError_t lap_add_diag(int B, int64_t n, int64_t nnz, const int64_t* diag_idx, double add,
                        double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_add_diag, dim3(grid_for((int64_t)B * n)), dim3(256), 0, st, B, n, nnz,
                     diag_idx, add, out);
  return hipGetLastError();
}

// `longs` lists the entries / outputs whose gather list exceeds kLongGather
// (nlong of them, built with the map); the per-thread kernel skips them.
hipError_t lap_add_atwa(int B, int64_t nnz, int64_t nobs, const int64_t* ptr, const int32_t* obs,
                        const double* coef, const double* w, const double* base, double* out,
                        const int64_t* longs, int64_t nlong, hipStream_t st) {
  hipLaunchKernelGGL(k_add_atwa, dim3(grid_for((int64_t)B * nnz)), dim3(256), 0, st, B, nnz, nobs,
                     ptr, obs, coef, w, base, out);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nlong <= 0) return e;
  hipLaunchKernelGGL(k_add_atwa_long, dim3((unsigned)nlong, B), dim3(256), 0, st, nnz, nobs, longs, ptr, obs,
                     coef, w, base, out);
  return hipGetLastError();
}

hipError_t lap_spmv(int B, int64_t nout, const int64_t* ptr, const int32_t* idx, const double* vals,
                    int64_t vld, const double* x, int64_t ldx, double alpha, double beta, double* y,
                    int64_t ldy, const int64_t* longs, int64_t nlong, hipStream_t st) {
  hipLaunchKernelGGL(k_gather_spmv, dim3(grid_for((int64_t)B * nout)), dim3(256), 0, st, B, nout,
                     ptr, idx, vals, vld, x, ldx, alpha, beta, y, ldy);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess || nlong <= 0) return e;
  hipLaunchKernelGGL(k_gather_spmv_long, dim3((unsigned)nlong, B), dim3(256), 0, st, longs, ptr, idx, vals,
                     vld, x, ldx, alpha, beta, y, ldy);
  return hipGetLastError();
}

hipError_t lap_axpy(int B, int64_t len, const double* a, const double* alpha, const double* c,
                    double* out, hipStream_t st, const LapBlocks* bk) {
  hipLaunchKernelGGL(k_axpy_b, dim3(grid_for((int64_t)B * len)), dim3(256), 0, st, B, len, a, alpha,
                     c, out, bk ? bk->nb : 1, bk ? bk->lblk : nullptr);
  return hipGetLastError();
}

hipError_t lap_rowreduce(int B, int64_t len, int op, const double* x, const double* y, double* out,
                         hipStream_t st, int nb, const int64_t* seg) {
  hipLaunchKernelGGL(k_rowreduce, dim3(B * nb), dim3(1024), 0, st, len, op, x, y, out, nb, seg);
  return hipGetLastError();
}

hipError_t lap_contract(int B, int K, int64_t nnz, const double* comps, const double* S,
                        int64_t l_size, const int64_t* slot_full, const int32_t* qrow,
                        const int32_t* qcol, const double* x, const double* z, int64_t n, double* tr,
                        double* quad, double* part, hipStream_t st, const LapBlocks* bk) {
  const int nb = bk ? bk->nb : 1;
  hipLaunchKernelGGL(k_contract_part, dim3(kContractChunks, (K + kContractGroup - 1) / kContractGroup, B * nb),
                     dim3(256), 0, st, K, nnz, comps, S,
                     l_size, slot_full, qrow, qcol, x, z, n, part, nb, bk ? bk->eseg : nullptr);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return e;
  hipLaunchKernelGGL(k_contract_sum, dim3((B * nb * K + 63) / 64), dim3(64), 0, st, B * nb * K, part, tr, quad);
  return hipGetLastError();
}

hipError_t lap_vdiag(int B, int64_t nobs, const int64_t* vptr, const int64_t* vslot,
                     const double* vcoef, const double* S, int64_t l_size, double* v,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_vdiag, dim3(grid_for((int64_t)B * nobs)), dim3(256), 0, st, B, nobs, vptr,
                     vslot, vcoef, S, l_size, v);
  return hipGetLastError();
}

hipError_t lap_constr_dot(int B, int kc, int64_t n, const double* ca, const double* x, double* r,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_constr_dot, dim3(kc, B), dim3(256), 0, st, kc, n, ca, x, r);
  return hipGetLastError();
}

hipError_t lap_constr_apply(int B, int kc, int64_t n, const double* ca, const double* lam, double* x,
                            hipStream_t st, double scale) {
  hipLaunchKernelGGL(k_constr_apply, dim3(grid_for((int64_t)B * n)), dim3(256), 0, st, B, kc, n, ca,
                     lam, scale, x);
  return hipGetLastError();
}

hipError_t lap_ground(int B, int k, int64_t nnz, const int64_t* idx, const double* src, double* dst,
                      double* alpha, hipStream_t st) {
  if (B * k <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_ground, dim3((B * k + 255) / 256), dim3(256), 0, st, B, k, nnz, idx, src, dst,
                     alpha);
  return hipGetLastError();
}

// out (B x RU x RY) = U_b Y_b^T over n entries; part holds B * ceil(n / 2048) * RU * RY doubles.
hipError_t lap_gram(int B, int RU, int RY, int64_t n, const double* U, int64_t ldu, int64_t ldub,
                    const double* Y, int64_t ldy, int64_t ldyb, double* part, double* out,
                    hipStream_t st) {
  if (RU < 1 || RY < 1 || RU > 64 || RY > 64) return hipErrorInvalidValue;
  const int nchunk = (int)((n + kGramChunk - 1) / kGramChunk), np = RU * RY;
  hipLaunchKernelGGL(k_gram_part, dim3(nchunk, B), dim3(256), 0, st, RU, RY, n, U, ldu, ldub, Y, ldy,
                     ldyb, nchunk, part);
  hipLaunchKernelGGL(k_gram_sum, dim3((np + 255) / 256, B), dim3(256), 0, st, np, nchunk, part, out);
  return hipGetLastError();
}

hipError_t lap_lowrank_axpy(int B, int R, int64_t n, const double* Y, int64_t ldy, int64_t ldyb,
                            const double* h, double* x, int64_t ldx, hipStream_t st) {
  hipLaunchKernelGGL(k_lowrank_axpy, dim3(grid_for((int64_t)B * n)), dim3(256), 0, st, B, R, n, Y, ldy,
                     ldyb, h, x, ldx);
  return hipGetLastError();
}

hipError_t lap_lowrank_diag(int B, int R, int64_t n, const double* Y, int64_t ldy, int64_t ldyb,
                            const double* Kinv, double* d, int64_t ldd, hipStream_t st) {
  if (R < 1 || R > 64) return hipErrorInvalidValue;
  const size_t sh = (size_t)(R * R + R * 128) * sizeof(double);
  if (sh > 64 * 1024) {
    hipError_t e = hipFuncSetAttribute((const void*)k_lowrank_diag,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(k_lowrank_diag, dim3((unsigned)((n + 127) / 128), B), dim3(128), sh, st, R, n, Y,
                     ldy, ldyb, Kinv, d, ldd);
  return hipGetLastError();
}

hipError_t lap_scatter_rows(int B, int64_t len, const int64_t* perm, const double* in, double* out,
                            hipStream_t st) {
  hipLaunchKernelGGL(k_scatter_rows, dim3(grid_for((int64_t)B * len)), dim3(256), 0, st, B, len, perm, in, out);
  return hipGetLastError();
}

hipError_t lap_mul(int64_t tot, const double* a, const double* b, double* out, hipStream_t st) {
  hipLaunchKernelGGL(k_mul, dim3(grid_for(tot)), dim3(256), 0, st, tot, a, b, out);
  return hipGetLastError();
}

hipError_t lap_blend(int B, int64_t len, const double* keep, const double* xt, double* x,
                     hipStream_t st, const LapBlocks* bk) {
  hipLaunchKernelGGL(k_blend, dim3(grid_for((int64_t)B * len)), dim3(256), 0, st, B, len, keep, xt, x,
                     bk ? bk->nb : 1, bk ? bk->lblk : nullptr);
  return hipGetLastError();
}

hipError_t lap_fill_f64(int64_t len, double v, double* p, hipStream_t st) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_f64, dim3(grid_for(len)), dim3(256), 0, st, len, v, p);
  return hipGetLastError();
}

hipError_t lap_fill_i32(int64_t len, int32_t v, int32_t* p, hipStream_t st) {
  if (len <= 0) return hipSuccess;
  hipLaunchKernelGGL(k_fill_i32, dim3(grid_for(len)), dim3(256), 0, st, len, v, p);
  return hipGetLastError();
}

hipError_t lap_gauss_finish(int B, int K, int mt, int fam, const int32_t* fstat, const int32_t* bad,
                            const double* maxstep, const double* ld, const double* quad, const double* loglik,
                            const double* tr_post, const double* tr_prior, const double* qx, const double* dcoefs,
                            const double* dlog_lpar, const double* sum_dl, const double* sum_vw, double* mlik,
                            double* grad, int32_t* status, hipStream_t st) {
  hipLaunchKernelGGL(k_gauss_finish, dim3((B + 63) / 64), dim3(64), 0, st, B, K, mt, fam, fstat, bad, maxstep, ld,
                     quad, loglik, tr_post, tr_prior, qx, dcoefs, dlog_lpar, sum_dl, sum_vw, mlik, grad, status);
  return hipGetLastError();
}

}  // namespace gmrf
