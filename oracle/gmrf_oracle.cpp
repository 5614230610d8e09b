// ORACLE — test infrastructure, not product code.
//
// A plain CPU restatement of rust-inla's numeric GMRF path, used only by
// tests/ (as the parity checker), __graft_entry__.smoke() and the cpu_baseline
// leg of bench.py. The product library never links, loads or calls anything
// here. Every function cites the reference file:line it follows.
//
// Parity status of each piece (what pins it; see tests/test_oracle_*.py):
//  * ordering chooser (orc_choose_order & helpers): integer-exact restatement of
//    crates/inla_math/src/ordering.rs:8-162; pinned by the reference's own test
//    ordering.rs:191-204 and by hand-computed bandwidths.
//  * fill-reducing order for the Amd branch: the reference calls faer 0.24's AMD,
//    which is not available; the oracle uses exact minimum degree instead. Its
//    permutation is NOT the reference's (parity unpinned); numeric results are
//    permutation-invariant up to rounding, which is what the tests compare.
//  * LDL^T, solve, log|Q|: up-looking simplicial LDL^T (no pivoting, no
//    regularisation), singular threshold 1e-14 as sparse_ldlt.rs:118-119;
//    pinned by dense brute force, residuals and closed forms (AR1 log-det). The
//    ordering of the last pattern factorised on a thread is reused for the same
//    pattern, as the reference's thread-local pattern cache does
//    (sparse_ldlt.rs:200-207,234; scratch.rs:90-134).
//  * Takahashi diag: restated loop-for-loop from sparse_ldlt.rs:357-432,
//    including the binary-search lookup; pinned against dense inverses and the
//    reference's own known-answer tests (sparse_ldlt.rs:613-665).
//  * Laplace Newton loop, mlik: restated from inference.rs:776-975 with the
//    likelihoods of inference.rs:170-311; pinned by the Gaussian closed form,
//    stationarity of every family's mode, and a dense brute force of the
//    constrained (kappa-augmented, projected) Gaussian case.
//  * compute_cpo_pit with check_cpo_failure_for_obs and the CDFs of observation_cdf
//    (crates/inla_stats/src/model_selection.rs:222-511); pinned by the reference's own
//    tests there (failure cases, CDF known answers), scipy's Poisson / binomial CDFs,
//    the erfc identity of the normal_cdf formula, and the harmonic mean of CPO.
//  * ccd_design / grid_design / invert_symmetric_matrix / jacobi_eigen: restated
//    from crates/inla_math/src/integration.rs; pinned by what the mathematics fixes
//    (the CCD mixture of a standard normal has identity covariance, the grid of a
//    quadratic log-density is a closed-form lattice set, A A^-1 = I, V L V^T = A).
//
// orc_plant(code) switches on ONE deliberate mistake (the Plant list below) so
// that tests/test_oracle_pins.py can show every pinning check fails when the
// function it pins is wrong. It is 0 (no mistake) everywhere else.
#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <deque>
#include <numeric>
#include <vector>

typedef int64_t i64;

namespace {

enum Plant {
  P_NONE = 0,
  P_LOWER_BW,        // off-by-one bandwidth
  P_BW_FWD,          // column index not relabelled
  P_TIME_MAJOR,      // identity relabelling instead of time-major
  P_RCM,             // Cuthill-McKee order not reversed
  P_CHOOSE,          // RCM accepted without the halving test
  P_MIN_DEGREE,      // natural order instead of minimum degree
  P_LDLT,            // pivot update term dropped
  P_LOGDET,          // log D instead of log |D|
  P_SOLVE,           // diagonal scaling skipped
  P_TAKAHASHI,       // sign of the off-diagonal recurrence flipped
  P_SELINV_LOOKUP,   // entries looked up with the wrong permutation
  P_DENSE_LDLT,      // D dropped from the off-diagonal update
  P_LGAMMA,          // one Lanczos coefficient perturbed
  P_LIK_GAUSS,       // Gaussian normalising term dropped
  P_LIK_POISSON,     // exposure missing from the Poisson weight
  P_LIK_BINOMIAL,    // (1 - p) missing from the binomial weight
  P_LIK_NBINOMIAL,   // second-derivative term of the link dropped
  P_NEWTON_RHS,      // - Q x missing from the Newton right-hand side
  P_PROJECT,         // projection correction added instead of subtracted
  P_MLIK,            // log|Q_prior| without the factor 1/2
  P_KAPPA,           // augmentation constant halved
  P_CCD5,            // half-fraction sign of the fifth factor flipped
  P_CCD_WEIGHT,      // f0 instead of f0^2 in the weight
  P_GRID,            // first point beyond the threshold kept
  P_JACOBI,          // rotation applied with the wrong sign to the eigenvectors
  P_INVERT,          // pivot row scaled only in its left half
  P_NNZ_L,           // diagonal not counted
  P_PERM,            // inverse permutation returned
  P_D,               // D returned in original instead of factor order
  P_CPO_FAIL,        // CPO_DIFF border test at 1 instead of 3
  P_CPO,             // max shift not restored in 1 / CPO
  P_NORMAL_CDF,      // exp(-x^2) instead of exp(-x^2 / 2)
  P_POISSON_CDF,     // sum stops one term early
  P_BINOMIAL_CDF,    // p and 1 - p swapped
};
int g_plant = P_NONE;

// ----------------------------------------------------------------- ordering.rs
i64 lower_bandwidth(i64 n, const i64* ip, const i64* ix) {  // ordering.rs:8-18
  i64 bw = 0;
  for (i64 c = 0; c < n; ++c)
    for (i64 p = ip[c]; p < ip[c + 1]; ++p)
      if (ix[p] >= c) bw = std::max(bw, ix[p] - c + (g_plant == P_LOWER_BW ? 1 : 0));
  return bw;
}

i64 bandwidth_with_fwd(i64 n, const i64* ip, const i64* ix, const std::vector<i64>& fwd) {
  i64 bw = 0;  // ordering.rs:21-30
  for (i64 c = 0; c < n; ++c)
    for (i64 p = ip[c]; p < ip[c + 1]; ++p) {
      i64 dd = fwd[ix[p]] - (g_plant == P_BW_FWD ? c : fwd[c]);
      bw = std::max(bw, dd < 0 ? -dd : dd);
    }
  return bw;
}

bool time_major(i64 n, const i64* ip, const i64* ix, std::vector<i64>& fwd) {  // ordering.rs:56-89
  if (n < 16) return false;
  i64 bw = lower_bandwidth(n, ip, ix);
  if (bw < n / 8) return false;
  bool have = false;
  i64 best = 0;
  std::vector<i64> bestf;
  for (i64 nc = 2; nc <= 8; ++nc) {
    if (n % nc) continue;
    i64 nt = n / nc;
    if (nt < 8) continue;
    std::vector<i64> f(n);
    for (i64 c = 0; c < nc; ++c)
      for (i64 t = 0; t < nt; ++t) f[c * nt + t] = g_plant == P_TIME_MAJOR ? c * nt + t : t * nc + c;  // ordering.rs:36-49
    i64 b = bandwidth_with_fwd(n, ip, ix, f);
    if (b < bw && !(have && best <= b)) {
      have = true;
      best = b;
      bestf = f;
    }
  }
  if (!have) return false;
  i64 avg = std::max<i64>(ip[n] / std::max<i64>(n, 1), 2);
  if (best * 4 < bw || best <= 4 * avg) {
    fwd = bestf;
    return true;
  }
  return false;
}

void rcm(i64 n, const i64* ip, const i64* ix, std::vector<i64>& fwd) {  // ordering.rs:92-140
  std::vector<std::vector<i64>> adj(n);
  for (i64 c = 0; c < n; ++c)
    for (i64 p = ip[c]; p < ip[c + 1]; ++p)
      if (ix[p] != c) {
        adj[c].push_back(ix[p]);
        adj[ix[p]].push_back(c);
      }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  std::vector<i64> deg(n);
  for (i64 i = 0; i < n; ++i) deg[i] = (i64)adj[i].size();
  auto key = [&](i64 a, i64 b) { return deg[a] != deg[b] ? deg[a] < deg[b] : a < b; };
  for (auto& a : adj) std::sort(a.begin(), a.end(), key);
  std::vector<char> vis(n, 0);
  std::vector<i64> order;
  while (true) {
    i64 start = -1;
    for (i64 i = 0; i < n; ++i)
      if (!vis[i] && (start < 0 || key(i, start))) start = i;
    if (start < 0) break;
    std::deque<i64> q{start};
    vis[start] = 1;
    while (!q.empty()) {
      i64 u = q.front();
      q.pop_front();
      order.push_back(u);
      for (i64 v : adj[u])
        if (!vis[v]) {
          vis[v] = 1;
          q.push_back(v);
        }
    }
  }
  if (g_plant != P_RCM) std::reverse(order.begin(), order.end());
  fwd.assign(n, 0);
  for (i64 w = 0; w < n; ++w) fwd[order[w]] = w;
}

int choose(i64 n, const i64* ip, const i64* ix, std::vector<i64>& fwd) {  // ordering.rs:149-162
  i64 bw = lower_bandwidth(n, ip, ix);
  if (time_major(n, ip, ix, fwd)) return 1;
  if (bw > 32) {
    rcm(n, ip, ix, fwd);
    if (g_plant == P_CHOOSE || bandwidth_with_fwd(n, ip, ix, fwd) * 2 < bw) return 2;
  }
  fwd.clear();
  return 0;
}

// Exact minimum degree on the explicit elimination graph (lowest index on ties).
void min_degree(i64 n, const i64* ip, const i64* ix, std::vector<i64>& order) {
  std::vector<std::vector<i64>> adj(n);
  for (i64 c = 0; c < n; ++c)
    for (i64 p = ip[c]; p < ip[c + 1]; ++p)
      if (ix[p] > c) {
        adj[c].push_back(ix[p]);
        adj[ix[p]].push_back(c);
      }
  for (auto& a : adj) {
    std::sort(a.begin(), a.end());
    a.erase(std::unique(a.begin(), a.end()), a.end());
  }
  if (g_plant == P_MIN_DEGREE) {
    order.resize(n);
    for (i64 i = 0; i < n; ++i) order[i] = i;
    return;
  }
  std::vector<char> done(n, 0);
  std::vector<std::vector<i64>> bucket(n + 1);
  std::vector<i64> deg(n);
  for (i64 i = n - 1; i >= 0; --i) {
    deg[i] = (i64)adj[i].size();
    bucket[deg[i]].push_back(i);
  }
  order.clear();
  i64 md = 0;
  std::vector<i64> tmp;
  while ((i64)order.size() < n) {
    i64 p = -1;
    while (p < 0) {
      while (md <= n && bucket[md].empty()) ++md;
      i64 c = bucket[md].back();
      bucket[md].pop_back();
      if (!done[c] && deg[c] == md) p = c;  // skip stale entries
    }
    done[p] = 1;
    order.push_back(p);
    const std::vector<i64> nb = adj[p];
    for (i64 u : nb) {
      // N(u) <- (N(u) U N(p)) \ {u, p}
      tmp.clear();
      std::set_union(adj[u].begin(), adj[u].end(), nb.begin(), nb.end(), std::back_inserter(tmp));
      auto& au = adj[u];
      au.clear();
      for (i64 v : tmp)
        if (v != u && v != p) au.push_back(v);
      deg[u] = (i64)au.size();
      bucket[deg[u]].push_back(u);
      if (deg[u] < md) md = deg[u];
    }
    adj[p].clear();
  }
}

// --------------------------------------------------------------- factor
struct Factor {
  i64 n = 0;
  std::vector<i64> perm, iperm;      // perm[k] = original index at factor position k
  std::vector<i64> lp, li;           // unit L strictly below the diagonal, CSC, sorted rows
  std::vector<double> lx, d;
  int status = 0;                    // 0 ok, 1 singular
};

// Up-looking simplicial LDL^T of P A P^T using the lower triangle of A in the
// original numbering (rows >= cols), as crates/inla_math/src/sparse_ldlt.rs:74-102
// feeds faer. No pivoting, no regularisation; |d| < 1e-14 or non-finite is
// Singular (sparse_ldlt.rs:118-119).
Factor factor(i64 n, const i64* ip, const i64* ix, const double* x, const std::vector<i64>& perm) {
  Factor F;
  F.n = n;
  F.perm = perm;
  F.iperm.assign(n, 0);
  for (i64 k = 0; k < n; ++k) F.iperm[perm[k]] = k;
  // permuted upper triangle by columns: entry (i, k) with i <= k in factor order
  std::vector<std::vector<std::pair<i64, double>>> up(n);
  for (i64 c = 0; c < n; ++c)
    for (i64 p = ip[c]; p < ip[c + 1]; ++p) {
      i64 r = ix[p];
      if (r < c) continue;
      i64 a = F.iperm[r], b = F.iperm[c];
      if (a > b) std::swap(a, b);
      up[b].push_back({a, x[p]});
    }
  // etree
  std::vector<i64> parent(n, -1), flag(n, -1), lnz(n, 0);
  for (i64 k = 0; k < n; ++k) {
    flag[k] = k;
    for (auto& e : up[k]) {
      i64 i = e.first;
      for (; flag[i] != k; i = parent[i]) {
        if (parent[i] == -1) parent[i] = k;
        lnz[i]++;
        flag[i] = k;
      }
    }
  }
  F.lp.assign(n + 1, 0);
  for (i64 k = 0; k < n; ++k) F.lp[k + 1] = F.lp[k] + lnz[k];
  F.li.assign(F.lp[n], 0);
  F.lx.assign(F.lp[n], 0.0);
  F.d.assign(n, 0.0);
  std::vector<double> y(n, 0.0);
  std::vector<i64> pattern(n), fill(n);
  for (i64 k = 0; k < n; ++k) fill[k] = F.lp[k];
  for (i64 k = 0; k < n; ++k) {
    i64 top = n;
    flag[k] = k;
    for (auto& e : up[k]) {
      i64 i = e.first;
      y[i] += e.second;
      i64 len = 0;
      for (; flag[i] != k; i = parent[i]) {
        pattern[len++] = i;
        flag[i] = k;
      }
      while (len > 0) pattern[--top] = pattern[--len];
    }
    F.d[k] = y[k];
    y[k] = 0.0;
    for (; top < n; ++top) {
      i64 i = pattern[top];
      double yi = y[i];
      y[i] = 0.0;
      for (i64 p = F.lp[i]; p < fill[i]; ++p) y[F.li[p]] -= F.lx[p] * yi;
      double l = yi / F.d[i];
      if (g_plant != P_LDLT) F.d[k] -= l * yi;
      F.li[fill[i]] = k;
      F.lx[fill[i]] = l;
      fill[i]++;
    }
    if (!std::isfinite(F.d[k]) || std::fabs(F.d[k]) < 1e-14) F.status = 1;
  }
  return F;
}

// Takahashi on the filled pattern; loop-for-loop sparse_ldlt.rs:376-432.
struct Filled {
  std::vector<i64> cp, ri;
  std::vector<double> v;  // D_j on the diagonal, unit-L below
};

Filled filled(const Factor& F) {  // sparse_ldlt.rs:279-310 (sorted per-column CSC)
  Filled L;
  i64 n = F.n;
  L.cp.assign(n + 1, 0);
  for (i64 j = 0; j < n; ++j) L.cp[j + 1] = L.cp[j] + 1 + (F.lp[j + 1] - F.lp[j]);
  L.ri.resize(L.cp[n]);
  L.v.resize(L.cp[n]);
  for (i64 j = 0; j < n; ++j) {
    i64 q = L.cp[j];
    L.ri[q] = j;
    L.v[q] = F.d[j];
    ++q;
    for (i64 p = F.lp[j]; p < F.lp[j + 1]; ++p, ++q) {
      L.ri[q] = F.li[p];
      L.v[q] = F.lx[p];
    }
  }
  return L;
}

// sparse_ldlt.rs:357-368
bool lookup_s(const Filled& l, const std::vector<double>& s, i64 i, i64 j, double& out) {
  i64 row = i >= j ? i : j, col = i >= j ? j : i;
  auto a = l.ri.begin() + l.cp[col], b = l.ri.begin() + l.cp[col + 1];
  auto it = std::lower_bound(a, b, row);
  if (it == b || *it != row) return false;
  out = s[it - l.ri.begin()];
  return true;
}

// sparse_ldlt.rs:376-419: returns 0 ok, 1 singular, 2 missing entry.
int takahashi(const Filled& l, i64 n, std::vector<double>& s) {
  s.assign(l.v.size(), 0.0);
  for (i64 j = n - 1; j >= 0; --j) {
    i64 start = l.cp[j], end = l.cp[j + 1];
    if (start >= end) return 1;
    i64 diag_p = -1;
    for (i64 p = start; p < end; ++p)
      if (l.ri[p] == j) {
        diag_p = p;
        break;
      }
    if (diag_p < 0) return 1;
    double dj = l.v[diag_p];
    if (!std::isfinite(dj) || std::fabs(dj) < 1e-14) return 1;
    for (i64 p = end - 1; p >= start; --p) {
      i64 i = l.ri[p];
      if (i <= j) continue;
      double acc = 0.0;
      for (i64 pk = start; pk < end; ++pk) {
        i64 k = l.ri[pk];
        if (k <= j) continue;
        double sv;
        if (!lookup_s(l, s, k, i, sv)) return 2;
        acc += l.v[pk] * sv;
      }
      s[p] = g_plant == P_TAKAHASHI ? acc : -acc;
    }
    double sjj = 1.0 / dj;
    for (i64 pk = start; pk < end; ++pk) {
      if (l.ri[pk] <= j) continue;
      sjj -= l.v[pk] * s[pk];
    }
    s[diag_p] = sjj;
  }
  return 0;
}

struct Handle {
  Factor F;
  Filled L;
  std::vector<double> S;
  bool have_s = false;
};

double lanczos_log_gamma(double z) {
  // Lanczos approximation, g = 7, n = 9 (published coefficients); the same
  // method as crates/inla_stats/src/inference.rs:1896-1920.
  static const double C[9] = {0.99999999999980993,  676.5203681218851,   -1259.1392167224028,
                              771.32342877765313,   -176.61502916214059, 12.507343278686905,
                              -0.13857109526572012, 9.9843695780195716e-6, 1.5056327351493116e-7};
  const double PI = 3.14159265358979323846;
  if (z < 0.5) return std::log(PI) - std::log(std::sin(PI * z)) - lanczos_log_gamma(1.0 - z);
  double z1 = z - 1.0;
  double x = C[0];
  for (int i = 0; i < 8; ++i) x += (C[i + 1] + (g_plant == P_LGAMMA && i == 0 ? 1e-6 : 0.0)) / (z1 + (double)i + 1.0);
  double t = z1 + 7.5;
  return 0.5 * std::log(2.0 * PI) + (z1 + 0.5) * std::log(t) - t + std::log(x);
}

// Likelihoods (inference.rs:170-311) with link_forward (1641-1664).
struct Eval {
  double logp, grad, hess;
  bool ok;
};
void link_fwd(double eta, int link, double& mu, double& d1, double& d2) {
  if (link == 0) {
    mu = eta; d1 = 1.0; d2 = 0.0;
  } else if (link == 1) {
    mu = std::exp(eta); d1 = mu; d2 = mu;
  } else {
    double p = eta >= 0 ? 1.0 / (1.0 + std::exp(-eta)) : std::exp(eta) / (1.0 + std::exp(eta));
    mu = p; d1 = p * (1 - p); d2 = d1 * (1 - 2 * p);
  }
}
// family: 0 gaussian (a = precision), 1 poisson (a = exposure), 2 binomial (a = trials),
// 3 negative binomial (a = exposure, b = size)
Eval eval_lik(int fam, int link, double eta, double y, double a, double b) {
  Eval e{0, 0, 0, true};
  double mu, d1, d2;
  if (!std::isfinite(eta)) { e.ok = false; return e; }
  link_fwd(eta, link, mu, d1, d2);
  if (fam == 0) {  // 170-186
    double r = mu - y;
    e.logp = -0.9189385332046728 + (g_plant == P_LIK_GAUSS ? 0.0 : 0.5 * std::log(a)) - 0.5 * a * r * r;
    e.grad = -a * r * d1;
    e.hess = -a * (d1 * d1 + r * d2);
  } else if (fam == 1) {  // 188-223
    double m = a * mu, dm = a * d1;
    if (m > 0) e.logp = y * std::log(m) - m - lanczos_log_gamma(y + 1.0);
    else e.logp = (y == 0) ? 0.0 : -INFINITY;
    e.grad = m > 0 ? (y / m - 1.0) * dm : -dm;
    e.hess = m > 0 ? (g_plant == P_LIK_POISSON ? -(d1 * d1) / m : -(dm * dm) / m) : -std::fabs(a * d2);
  } else if (fam == 2) {  // 225-267
    double n = a, p = mu;
    double lc = lanczos_log_gamma(n + 1) - lanczos_log_gamma(y + 1) - lanczos_log_gamma(n - y + 1);
    if (p == 0) e.logp = y == 0 ? lc : -INFINITY;
    else if (p == 1) e.logp = y == n ? lc : -INFINITY;
    else e.logp = lc + y * std::log(p) + (n - y) * std::log(1 - p);
    double g = (p > 0 && p < 1) ? y / p - (n - y) / (1 - p) : 0.0;
    e.grad = g * d1;
    e.hess = g_plant == P_LIK_BINOMIAL ? -n * p : -n * d1;
  } else {  // 269-311
    double m = a * mu, dm = a * d1, d2m = a * d2, r = b, mr = m + r;
    if (!(m > 0)) { e.ok = false; return e; }
    e.logp = lanczos_log_gamma(y + r) - lanczos_log_gamma(r) - lanczos_log_gamma(y + 1) +
             r * std::log(r / mr) + y * std::log(m / mr);
    double dl = y / m - (y + r) / mr;
    double d2l = -y / (m * m) + (y + r) / (mr * mr);
    e.grad = dl * dm;
    e.hess = d2l * dm * dm + (g_plant == P_LIK_NBINOMIAL ? 0.0 : dl * d2m);
  }
  return e;
}

}  // namespace

extern "C" {

i64 orc_lower_bandwidth(i64 n, const i64* ip, const i64* ix) { return lower_bandwidth(n, ip, ix); }

i64 orc_bandwidth_with_fwd(i64 n, const i64* ip, const i64* ix, const i64* fwd) {
  std::vector<i64> f(fwd, fwd + n);
  return bandwidth_with_fwd(n, ip, ix, f);
}

// Returns 0 (Amd), 1 (time-major), 2 (RCM); fwd_out[old] = new when nonzero.
int orc_choose_order(i64 n, const i64* ip, const i64* ix, i64* fwd_out) {
  std::vector<i64> f;
  int k = choose(n, ip, ix, f);
  if (k) std::copy(f.begin(), f.end(), fwd_out);
  return k;
}

int orc_time_major(i64 n, const i64* ip, const i64* ix, i64* fwd_out) {
  std::vector<i64> f;
  if (!time_major(n, ip, ix, f)) return 0;
  std::copy(f.begin(), f.end(), fwd_out);
  return 1;
}

void orc_rcm(i64 n, const i64* ip, const i64* ix, i64* fwd_out) {
  std::vector<i64> f;
  rcm(n, ip, ix, f);
  std::copy(f.begin(), f.end(), fwd_out);
}

void orc_min_degree(i64 n, const i64* ip, const i64* ix, i64* order_out) {
  std::vector<i64> o;
  min_degree(n, ip, ix, o);
  std::copy(o.begin(), o.end(), order_out);
}

// Factorise; perm (perm[k] = original index) may be NULL: then the chooser's
// branch is followed (time-major / RCM permutation, else exact minimum degree).
// The ordering of the last pattern factorised on this thread is kept and reused when
// the same pattern comes again, as factorize_sparse reuses its symbolic analysis
// through the thread-local scratch's single-entry pattern cache
// (crates/inla_math/src/sparse_ldlt.rs:200-207,234; scratch.rs:90-134). The cache
// is bypassed while a deliberate mistake is switched on (orc_plant), so that the
// ordering pins see the mistake. *status: 0 ok, 1 singular.
void* orc_factor(i64 n, const i64* ip, const i64* ix, const double* x, const i64* perm, int* status) {
  struct PatternCache {
    std::vector<i64> indptr, indices, perm;
  };
  thread_local PatternCache cache;
  std::vector<i64> p(n);
  if (perm) {
    std::copy(perm, perm + n, p.begin());
  } else if (g_plant == 0 && (i64)cache.indptr.size() == n + 1 && std::equal(ip, ip + n + 1, cache.indptr.begin()) &&
             std::equal(ix, ix + ip[n], cache.indices.begin())) {
    p = cache.perm;
  } else {
    std::vector<i64> fwd;
    if (choose(n, ip, ix, fwd)) {
      for (i64 o = 0; o < n; ++o) p[fwd[o]] = o;
    } else {
      min_degree(n, ip, ix, p);
    }
    if (g_plant == 0) {
      cache.indptr.assign(ip, ip + n + 1);
      cache.indices.assign(ix, ix + ip[n]);
      cache.perm = p;
    }
  }
  Handle* h = new Handle;
  h->F = factor(n, ip, ix, x, p);
  *status = h->F.status;
  return h;
}

void orc_free(void* h) { delete (Handle*)h; }

i64 orc_nnz_l(void* hp) {  // including the diagonal
  Handle* h = (Handle*)hp;
  return h->F.lp[h->F.n] + (g_plant == P_NNZ_L ? 0 : h->F.n);
}

void orc_perm(void* hp, i64* out) {
  Handle* h = (Handle*)hp;
  const std::vector<i64>& p = g_plant == P_PERM ? h->F.iperm : h->F.perm;
  std::copy(p.begin(), p.end(), out);
}

void orc_d(void* hp, double* out) {  // factor order
  Handle* h = (Handle*)hp;
  if (g_plant == P_D) {
    for (i64 k = 0; k < h->F.n; ++k) out[h->F.perm[k]] = h->F.d[k];
    return;
  }
  std::copy(h->F.d.begin(), h->F.d.end(), out);
}

// ldlt.rs:64-67: sum log|D_i|
double orc_log_abs_det(void* hp) {
  Handle* h = (Handle*)hp;
  double s = 0.0;
  for (double v : h->F.d) s += std::log(g_plant == P_LOGDET ? v : std::fabs(v));
  return s;
}

// L y = P b, D z = y, L^T w = z, x = P^T w (sparse_ldlt.rs:243-265 semantics).
void orc_solve(void* hp, const double* rhs, double* out) {
  Handle* h = (Handle*)hp;
  const Factor& F = h->F;
  i64 n = F.n;
  std::vector<double> y(n);
  for (i64 k = 0; k < n; ++k) y[k] = rhs[F.perm[k]];
  for (i64 j = 0; j < n; ++j)
    for (i64 p = F.lp[j]; p < F.lp[j + 1]; ++p) y[F.li[p]] -= F.lx[p] * y[j];
  if (g_plant != P_SOLVE)
    for (i64 j = 0; j < n; ++j) y[j] /= F.d[j];
  for (i64 j = n - 1; j >= 0; --j) {
    double s = y[j];
    for (i64 p = F.lp[j]; p < F.lp[j + 1]; ++p) s -= F.lx[p] * y[F.li[p]];
    y[j] = s;
  }
  for (i64 k = 0; k < n; ++k) out[F.perm[k]] = y[k];
}

// Takahashi diag(Q^-1) in original order; returns 0 ok, 1 singular, 2 missing.
int orc_takahashi_diag(void* hp, double* out) {
  Handle* h = (Handle*)hp;
  if (!h->have_s) {
    h->L = filled(h->F);
    int st = takahashi(h->L, h->F.n, h->S);
    if (st) return st;
    h->have_s = true;
  }
  for (i64 j = 0; j < h->F.n; ++j) out[h->F.perm[j]] = h->S[h->L.cp[j]];  // sparse_ldlt.rs:421-430
  return 0;
}

// Selected-inverse entries Sigma[rows[q], cols[q]] (original numbering) for
// positions inside the filled pattern; returns 2 if one is outside.
int orc_selinv_entries(void* hp, i64 nq, const i64* rows, const i64* cols, double* out) {
  Handle* h = (Handle*)hp;
  if (!h->have_s) {
    std::vector<double> tmp(h->F.n);
    int st = orc_takahashi_diag(hp, tmp.data());
    if (st) return st;
  }
  for (i64 q = 0; q < nq; ++q) {
    const std::vector<i64>& ip = g_plant == P_SELINV_LOOKUP ? h->F.perm : h->F.iperm;
    i64 a = ip[rows[q]], b = ip[cols[q]];
    if (!lookup_s(h->L, h->S, a, b, out[q])) return 2;
  }
  return 0;
}

// Dense brute force: scalar LDL^T of a row-major n x n matrix (ldlt.rs:126-169),
// log|A| and diag(A^-1) via Y = L^-1 (ldlt.rs:232-258). Returns 1 if singular.
int orc_dense_ldlt(i64 n, const double* a, double* logdet, double* diag_inv) {
  std::vector<double> l(n * n, 0.0), d(n, 0.0);
  for (i64 i = 0; i < n; ++i) l[i * n + i] = 1.0;
  for (i64 j = 0; j < n; ++j) {
    double dj = a[j * n + j];
    for (i64 k = 0; k < j; ++k) dj -= l[j * n + k] * l[j * n + k] * d[k];
    if (!std::isfinite(dj) || std::fabs(dj) < 1e-14) return 1;
    d[j] = dj;
    for (i64 i = j + 1; i < n; ++i) {
      double lij = a[i * n + j];
      for (i64 k = 0; k < j; ++k) lij -= l[i * n + k] * l[j * n + k] * (g_plant == P_DENSE_LDLT ? 1.0 : d[k]);
      l[i * n + j] = lij / d[j];
    }
  }
  double s = 0;
  for (i64 i = 0; i < n; ++i) s += std::log(std::fabs(d[i]));
  *logdet = s;
  if (diag_inv) {
    std::vector<double> y(n * n, 0.0);
    for (i64 i = 0; i < n; ++i) {
      y[i * n + i] = 1.0;
      for (i64 j = 0; j < i; ++j) {
        double t = 0;
        for (i64 k = j; k < i; ++k) t -= l[i * n + k] * y[k * n + j];
        y[i * n + j] = t;
      }
    }
    for (i64 i = 0; i < n; ++i) {
      double t = 0;
      for (i64 k = i; k < n; ++k) t += y[k * n + i] * y[k * n + i] / d[k];
      diag_inv[i] = t;
    }
  }
  return 0;
}

double orc_log_gamma(double z) { return lanczos_log_gamma(z); }

// ---- CPO / PIT (crates/inla_stats/src/model_selection.rs:222-511)
// The normal CDF formula of model_selection.rs:450-465 (the Abramowitz-Stegun 7.1.26
// rational form applied with exp(-x^2 / 2), as the reference writes it).
double orc_normal_cdf(double x) {
  const double A1 = 0.254829592, A2 = -0.284496736, A3 = 1.421413741, A4 = -1.453152027,
               A5 = 1.061405429, P = 0.3275911;
  const double sign = x < 0.0 ? -1.0 : 1.0, ax = std::fabs(x);
  const double t = 1.0 / (1.0 + P * ax);
  const double e = g_plant == P_NORMAL_CDF ? std::exp(-ax * ax) : std::exp(-ax * ax / 2.0);
  const double y = 1.0 - (((((A5 * t + A4) * t) + A3) * t + A2) * t + A1) * t * e;
  return 0.5 * (1.0 + sign * y);
}

// Pr(X <= k), X ~ Poisson(lambda), by direct summation (model_selection.rs:470-485)
double orc_poisson_cdf(double lambda, i64 k) {
  if (lambda <= 0.0) return 1.0;
  double cdf = 0.0, lp = -lambda;
  const i64 last = g_plant == P_POISSON_CDF ? k - 1 : k;
  for (i64 j = 0; j <= last; ++j) {
    cdf += std::exp(lp);
    if (j < k) lp += std::log(lambda) - std::log((double)(j + 1));
  }
  return std::min(cdf, 1.0);
}

// Pr(X <= k), X ~ Binomial(n, p) (model_selection.rs:488-503)
double orc_binomial_cdf(i64 n, double p, i64 k) {
  if (p <= 0.0) return 1.0;
  if (p >= 1.0) return k >= n ? 1.0 : 0.0;
  const double q = g_plant == P_BINOMIAL_CDF ? p : 1.0 - p, pp = g_plant == P_BINOMIAL_CDF ? 1.0 - p : p;
  double cdf = 0.0;
  for (i64 j = 0; j <= k; ++j) {
    const double lc = lanczos_log_gamma((double)n + 1.0) - lanczos_log_gamma((double)j + 1.0) -
                      lanczos_log_gamma((double)(n - j) + 1.0);
    cdf += std::exp(lc + (double)j * std::log(pp) + (double)(n - j) * std::log(q));
  }
  return std::min(cdf, 1.0);
}

// check_cpo_failure_for_obs (model_selection.rs:327-396): 1 when unreliable
int orc_cpo_failure(int fam, int link, double y, double a, double b, double mu, double var) {
  if (!(var > 0.0) || !std::isfinite(var)) return 1;
  const Eval ev = eval_lik(fam, link, mu, y, a, b);
  if (!ev.ok) return 1;
  const double g = ev.grad, c = -ev.hess;
  const double tau_loo = 1.0 / var - c;
  if (!(tau_loo > 0.0) || !std::isfinite(tau_loo)) return 1;
  const double var_loo = 1.0 / tau_loo, sd = std::sqrt(var_loo);
  const double mu_loo = var_loo * (mu / var - c * mu + g);
  const double steps[5] = {-3.0, -1.5, 0.0, 1.5, 3.0};
  double d[5];
  for (int j = 0; j < 5; ++j) {
    const double x = mu_loo + steps[j] * sd;
    const Eval e = eval_lik(fam, link, x, y, a, b);
    if (!e.ok) return 1;
    d[j] = -0.5 * (x - mu) * (x - mu) / var + e.logp;
  }
  bool inc = true, dec = true;
  for (int j = 1; j < 5; ++j) {
    if (d[j] < d[j - 1]) inc = false;
    if (d[j] > d[j - 1]) dec = false;
  }
  if (inc || dec) return 1;
  double mx = -INFINITY;
  for (int j = 0; j < 5; ++j) mx = std::max(mx, d[j]);
  const double diff = g_plant == P_CPO_FAIL ? 1.0 : 3.0;  // CPO_DIFF
  return (mx - d[0] < diff || mx - d[4] < diff) ? 1 : 0;
}

// observation_cdf (model_selection.rs:398-447) for the four families: NAN for None
// (negative binomial); the link is taken as identity / log / logit as the reference does
double obs_cdf(int fam, double eta, double variance, double y, double a) {
  if (fam == 0) {
    const double tv = 1.0 / a + variance;
    if (tv <= 0.0) return NAN;
    return orc_normal_cdf((y - eta) / std::sqrt(tv));
  }
  if (fam == 1) {
    const double lambda = a * std::exp(eta);
    if (lambda < 0.0 || !std::isfinite(lambda)) return NAN;
    const i64 yi = (i64)std::floor(y);
    if (yi < 0) return 0.0;
    return orc_poisson_cdf(lambda, yi);
  }
  if (fam == 2) {
    const double p = eta >= 0.0 ? 1.0 / (1.0 + std::exp(-eta)) : std::exp(eta) / (1.0 + std::exp(eta));
    const i64 n = (i64)a;
    const i64 yi = (i64)std::floor(y);
    if (yi < 0) return 0.0;
    return orc_binomial_cdf(n, p, std::min(yi, n));
  }
  return NAN;
}

// compute_cpo_pit (model_selection.rs:222-325): K nodes x n observations, eta and var
// row-major [K][n], normalised weights w[K]; observation i has family fam with a[i]
// (Gaussian precision / exposure / trials) and b[i] (NB size). cpo[i] / pit[i] are NAN
// where the reference returns None. Returns the failure count.
i64 orc_cpo_pit(int fam, int link, int K, i64 n, const double* eta, const double* var, const double* w,
                const double* y, const double* a, const double* b, double* cpo, double* pit) {
  i64 fails = 0;
  for (i64 i = 0; i < n; ++i) {
    cpo[i] = NAN;
    pit[i] = NAN;
    const double bi = b ? b[i] : 0.0;
    bool failed = false;
    for (int k = 0; k < K && !failed; ++k)
      failed = orc_cpo_failure(fam, link, y[i], a[i], bi, eta[(i64)k * n + i], var[(i64)k * n + i]) != 0;
    if (failed) {
      ++fails;
      continue;
    }
    std::vector<double> ll;
    bool bad = false;
    for (int k = 0; k < K; ++k) {
      const Eval e = eval_lik(fam, link, eta[(i64)k * n + i], y[i], a[i], bi);
      if (!e.ok) {
        bad = true;
        break;
      }
      ll.push_back(e.logp);
    }
    if (bad || ll.size() < 2) {
      ++fails;
      continue;
    }
    double max_neg = -INFINITY;
    for (double v : ll) max_neg = std::max(max_neg, -v);
    double inv = 0.0;
    for (int k = 0; k < K; ++k) inv += w[k] * std::exp(-ll[k] - max_neg);
    const double inv_scaled = g_plant == P_CPO ? inv : inv * std::exp(max_neg);
    if (!std::isfinite(inv_scaled) || inv_scaled <= 0.0) {
      ++fails;
      continue;
    }
    cpo[i] = 1.0 / inv_scaled;
    double ps = 0.0;
    bool ok = true;
    for (int k = 0; k < K && ok; ++k) {
      const double f = obs_cdf(fam, eta[(i64)k * n + i], var[(i64)k * n + i], y[i], a[i]);
      if (std::isnan(f)) ok = false;
      else ps += w[k] * f;
    }
    if (ok && std::isfinite(ps)) pit[i] = std::min(std::max(ps, 0.0), 1.0);
  }
  return fails;
}

void orc_eval_lik(int fam, int link, i64 m, const double* eta, const double* y, const double* a,
                  const double* b, double* logp, double* grad, double* hess) {
  for (i64 i = 0; i < m; ++i) {
    Eval e = eval_lik(fam, link, eta[i], y[i], a[i], b ? b[i] : 0.0);
    logp[i] = e.logp;
    grad[i] = e.grad;
    hess[i] = e.hess;
  }
}

}  // extern "C"

// --------------------------------------------------------------------------
// Laplace Newton loop (crates/inla_stats/src/inference.rs:776-975) on the
// oracle factorisation, with the assembly of crates/inla_math/src/ldlt.rs:373-453,
// design.rs:52-139 and constraints.rs:275-353.
// --------------------------------------------------------------------------
namespace {

struct Csc {
  i64 nr = 0, nc = 0;
  std::vector<i64> p, i;
  std::vector<double> x;
};

struct Trip {
  i64 r, c;
  double v;
};

// Triplets -> CSC with duplicates summed in insertion order (TriMat::to_csc).
Csc to_csc(i64 nr, i64 nc, std::vector<Trip>& t) {
  std::stable_sort(t.begin(), t.end(), [](const Trip& a, const Trip& b) {
    return a.c != b.c ? a.c < b.c : a.r < b.r;
  });
  Csc m;
  m.nr = nr;
  m.nc = nc;
  m.p.assign(nc + 1, 0);
  for (size_t k = 0; k < t.size();) {
    size_t e = k;
    double s = 0.0;
    while (e < t.size() && t[e].r == t[k].r && t[e].c == t[k].c) s += t[e++].v;
    m.i.push_back(t[k].r);
    m.x.push_back(s);
    m.p[t[k].c + 1]++;
    k = e;
  }
  for (i64 c = 0; c < nc; ++c) m.p[c + 1] += m.p[c];
  return m;
}

void push_all(const Csc& a, std::vector<Trip>& t, bool skip_zero) {
  for (i64 c = 0; c < a.nc; ++c)
    for (i64 q = a.p[c]; q < a.p[c + 1]; ++q)
      if (!skip_zero || a.x[q] != 0.0) t.push_back({a.i[q], c, a.x[q]});
}

std::vector<double> matvec(const Csc& a, const std::vector<double>& x) {  // design.rs:52-71
  std::vector<double> y(a.nr, 0.0);
  for (i64 c = 0; c < a.nc; ++c) {
    if (x[c] == 0.0) continue;
    for (i64 q = a.p[c]; q < a.p[c + 1]; ++q) y[a.i[q]] += a.x[q] * x[c];
  }
  return y;
}

std::vector<double> matvec_t(const Csc& a, const std::vector<double>& y) {  // design.rs:74-91
  std::vector<double> z(a.nc, 0.0);
  for (i64 c = 0; c < a.nc; ++c) {
    double s = 0.0;
    for (i64 q = a.p[c]; q < a.p[c + 1]; ++q) s += a.x[q] * y[a.i[q]];
    z[c] = s;
  }
  return z;
}

// design.rs:94-120 A^T diag(d) A through CSR rows of A.
void at_diag_a(const Csc& a, const std::vector<double>& d, std::vector<Trip>& t) {
  std::vector<i64> rp(a.nr + 1, 0);
  for (i64 q = 0; q < a.p[a.nc]; ++q) rp[a.i[q] + 1]++;
  for (i64 r = 0; r < a.nr; ++r) rp[r + 1] += rp[r];
  std::vector<i64> rc(a.p[a.nc]), fill(rp.begin(), rp.end() - 1);
  std::vector<double> rv(a.p[a.nc]);
  for (i64 c = 0; c < a.nc; ++c)
    for (i64 q = a.p[c]; q < a.p[c + 1]; ++q) {
      i64 k = fill[a.i[q]]++;
      rc[k] = c;
      rv[k] = a.x[q];
    }
  for (i64 r = 0; r < a.nr; ++r) {
    double di = d[r];
    if (di == 0.0 || !std::isfinite(di)) continue;
    for (i64 u = rp[r]; u < rp[r + 1]; ++u)
      for (i64 w = rp[r]; w < rp[r + 1]; ++w) {
        double v = di * rv[u] * rv[w];
        if (v != 0.0) t.push_back({rc[u], rc[w], v});
      }
  }
}

struct Model {
  Csc q, a;
  bool has_a = false;
  int fam = 0, link = 0;
  std::vector<double> y, pa, pb;
  i64 k = 0;
  std::vector<double> ca, ce;  // constraints, row-major k x n
};

bool evals_at(const Model& M, const std::vector<double>& x, std::vector<Eval>& ev) {
  std::vector<double> eta = M.has_a ? matvec(M.a, x) : x;
  ev.resize(eta.size());
  for (size_t i = 0; i < eta.size(); ++i) {
    ev[i] = eval_lik(M.fam, M.link, eta[i], M.y[i], M.pa[i], M.pb.empty() ? 0.0 : M.pb[i]);
    if (!ev[i].ok) return false;
  }
  return true;
}

// ldlt.rs:373-453
void newton_system(const Model& M, const Csc& qw, const std::vector<Eval>& ev,
                   const std::vector<double>& x, Csc& qpost, std::vector<double>& rhs) {
  i64 n = qw.nc;
  std::vector<double> qx = matvec(qw, x);
  std::vector<Trip> t;
  if (!M.has_a) {
    push_all(qw, t, false);
    for (i64 i = 0; i < n; ++i) {
      double h = -ev[i].hess;
      if (h != 0.0) t.push_back({i, i, h});
    }
    qpost = to_csc(n, n, t);
    rhs.resize(n);
    for (i64 i = 0; i < n; ++i) rhs[i] = ev[i].grad - (g_plant == P_NEWTON_RHS ? 0.0 : qx[i]);
  } else {
    std::vector<double> nh(ev.size()), g(ev.size());
    for (size_t i = 0; i < ev.size(); ++i) {
      nh[i] = -ev[i].hess;
      g[i] = ev[i].grad;
    }
    std::vector<Trip> tl;
    at_diag_a(M.a, nh, tl);
    Csc like = to_csc(n, n, tl);
    push_all(qw, t, false);
    push_all(like, t, false);
    qpost = to_csc(n, n, t);
    rhs = matvec_t(M.a, g);
    if (g_plant != P_NEWTON_RHS)
      for (i64 i = 0; i < n; ++i) rhs[i] -= qx[i];
  }
}

bool project(std::vector<double>& x, const Model& M) {  // constraints.rs:316-353
  i64 n = (i64)x.size(), k = M.k;
  std::vector<double> r(k), g(k * k);
  for (i64 a = 0; a < k; ++a) {
    double s = 0;
    for (i64 c = 0; c < n; ++c) s += M.ca[a * n + c] * x[c];
    r[a] = s - M.ce[a];
  }
  for (i64 a = 0; a < k; ++a)
    for (i64 b = 0; b < k; ++b) {
      double s = 0;
      for (i64 c = 0; c < n; ++c) s += M.ca[a * n + c] * M.ca[b * n + c];
      g[a * k + b] = s;
    }
  // constraints.rs:355-390 Gauss-Jordan with partial pivoting
  for (i64 i = 0; i < k; ++i) {
    i64 piv = i;
    for (i64 rr = i + 1; rr < k; ++rr)
      if (std::fabs(g[rr * k + i]) > std::fabs(g[piv * k + i])) piv = rr;
    if (std::fabs(g[piv * k + i]) < 1e-14) return false;
    if (piv != i) {
      for (i64 c = 0; c < k; ++c) std::swap(g[i * k + c], g[piv * k + c]);
      std::swap(r[i], r[piv]);
    }
    double dg = g[i * k + i];
    for (i64 c = i; c < k; ++c) g[i * k + c] /= dg;
    r[i] /= dg;
    for (i64 rr = 0; rr < k; ++rr) {
      if (rr == i) continue;
      double f = g[rr * k + i];
      for (i64 c = i; c < k; ++c) g[rr * k + c] -= f * g[i * k + c];
      r[rr] -= f * r[i];
    }
  }
  for (i64 c = 0; c < n; ++c) {
    double corr = 0;
    for (i64 a = 0; a < k; ++a) corr += M.ca[a * n + c] * r[a];
    x[c] += g_plant == P_PROJECT ? corr : -corr;
  }
  return true;
}

Handle* factor_csc(const Csc& m, int* st) {
  return (Handle*)orc_factor(m.nc, m.p.data(), m.i.data(), m.x.data(), nullptr, st);
}

}  // namespace

extern "C" {

// Laplace mode at one theta. Returns 0 ok; 1 singular; 3 Newton not converged /
// non-finite step; 4 invalid input. out_info = [log|Q_prior|, log|Q_post|,
// n_factorisations, n_solves]. *post receives the posterior factor handle
// (free with orc_free) when non-NULL.
int orc_find_mode(i64 n, const i64* qp, const i64* qi, const double* qx,
                  i64 nobs, const i64* ap, const i64* ai, const double* ax,
                  int fam, int link, const double* y, const double* pa, const double* pb,
                  i64 kc, const double* ca, const double* ce, int max_iter, double tol,
                  double* x_out, double* mlik_out, double* info, void** post) {
  Model M;
  M.q.nr = M.q.nc = n;
  M.q.p.assign(qp, qp + n + 1);
  M.q.i.assign(qi, qi + qp[n]);
  M.q.x.assign(qx, qx + qp[n]);
  M.has_a = ap != nullptr;
  if (M.has_a) {
    M.a.nr = nobs;
    M.a.nc = n;
    M.a.p.assign(ap, ap + n + 1);
    M.a.i.assign(ai, ai + ap[n]);
    M.a.x.assign(ax, ax + ap[n]);
  } else if (nobs != n) {
    return 4;
  }
  M.fam = fam;
  M.link = link;
  M.y.assign(y, y + nobs);
  M.pa.assign(pa, pa + nobs);
  if (pb) M.pb.assign(pb, pb + nobs);
  M.k = kc;
  if (kc) {
    M.ca.assign(ca, ca + kc * n);
    M.ce.assign(ce, ce + kc);
  }
  int nfact = 0, nsolve = 0;
  // working prior (inference.rs:817-827; constraints.rs:275-313)
  Csc qw;
  if (kc) {
    std::vector<Trip> t;
    push_all(M.q, t, true);
    const double kappa = 3.059023205018258e6 * (g_plant == P_KAPPA ? 0.5 : 1.0);
    for (i64 i = 0; i < n; ++i)
      for (i64 j = 0; j < n; ++j) {
        double s = 0;
        for (i64 r = 0; r < kc; ++r) s += M.ca[r * n + i] * M.ca[r * n + j];
        double v = kappa * s;
        if (v != 0.0) t.push_back({i, j, v});
      }
    qw = to_csc(n, n, t);
  } else {
    qw = M.q;
  }
  // prior log-det (inference.rs:829-840)
  double ld_prior;
  {
    Csc qj;
    if (!kc) {
      std::vector<Trip> t;
      push_all(qw, t, false);
      for (i64 i = 0; i < n; ++i) t.push_back({i, i, 1e-12});
      qj = to_csc(n, n, t);
    }
    int st;
    Handle* h = factor_csc(kc ? qw : qj, &st);
    ++nfact;
    if (st) { orc_free(h); return 1; }
    ld_prior = orc_log_abs_det(h);
    orc_free(h);
  }
  std::vector<double> x(n, 0.0);
  Handle* cur = nullptr;
  auto step_at = [&](const std::vector<double>& xx, std::vector<double>& step) -> int {
    std::vector<Eval> ev;
    if (!evals_at(M, xx, ev)) return 3;
    Csc qpost;
    std::vector<double> rhs;
    newton_system(M, qw, ev, xx, qpost, rhs);
    if (cur) orc_free(cur);
    int st;
    cur = factor_csc(qpost, &st);
    ++nfact;
    if (st) return 1;
    step.assign(n, 0.0);
    orc_solve(cur, rhs.data(), step.data());
    ++nsolve;
    return 0;
  };
  bool gauss_id = nobs > 0 && fam == 0 && link == 0;
  int rc = 0;
  if (gauss_id) {  // inference.rs:845-870
    if (max_iter == 0) rc = 3;
    std::vector<double> step;
    if (!rc) rc = step_at(x, step);
    if (!rc) {
      for (double s : step)
        if (!std::isfinite(s)) rc = 3;
    }
    if (!rc) {
      for (i64 i = 0; i < n; ++i) x[i] += step[i];
      if (kc && !project(x, M)) rc = 4;
    }
  } else {  // inference.rs:872-942
    bool conv = false;
    for (int it = 0; it < max_iter && !rc; ++it) {
      std::vector<double> step;
      rc = step_at(x, step);
      if (rc) break;
      double mx = 0;
      for (double s : step) {
        if (!std::isfinite(s)) rc = 3;
        mx = std::max(mx, std::fabs(s));
      }
      if (rc) break;
      double alpha = mx > 10.0 ? 10.0 / mx : 1.0;
      std::vector<double> xt(x);
      for (int bt = 0; bt < 12; ++bt) {
        for (i64 i = 0; i < n; ++i) xt[i] = x[i] + alpha * step[i];
        if (kc && !project(xt, M)) { rc = 4; break; }
        std::vector<Eval> ev;
        bool ok = evals_at(M, xt, ev);
        if (ok)
          for (auto& e : ev)
            if (!std::isfinite(e.logp)) ok = false;
        if (ok) break;
        alpha *= 0.5;
        if (alpha < 1e-8) { rc = 3; break; }
      }
      if (rc) break;
      double md = 0;
      for (i64 i = 0; i < n; ++i) {
        md = std::max(md, std::fabs(xt[i] - x[i]));
        x[i] = xt[i];
      }
      if (md < tol || mx < tol) {
        std::vector<double> s2;
        rc = step_at(x, s2);  // confirmation factorisation at x*
        conv = true;
        break;
      }
      if (it == max_iter - 1) rc = 3;
    }
    if (!rc && !conv) rc = 3;
  }
  if (rc) {
    if (cur) orc_free(cur);
    return rc;
  }
  if (kc && !project(x, M)) { orc_free(cur); return 4; }
  double ld_post = orc_log_abs_det(cur);
  std::vector<double> qxv(n, 0.0);
  for (i64 c = 0; c < n; ++c)
    for (i64 q = qw.p[c]; q < qw.p[c + 1]; ++q) qxv[qw.i[q]] += qw.x[q] * x[c];
  double quad = 0;
  for (i64 i = 0; i < n; ++i) quad += qxv[i] * x[i];
  std::vector<Eval> ev;
  evals_at(M, x, ev);
  double ll = 0;
  for (auto& e : ev) ll += e.logp;
  *mlik_out = ll - 0.5 * quad + (g_plant == P_MLIK ? 1.0 : 0.5) * ld_prior - 0.5 * ld_post;
  std::copy(x.begin(), x.end(), x_out);
  if (info) {
    info[0] = ld_prior;
    info[1] = ld_post;
    info[2] = nfact;
    info[3] = nsolve;
  }
  if (post) *post = cur;
  else orc_free(cur);
  return 0;
}

// ccd_design (crates/inla_math/src/integration.rs:1-62). points: (n_exp x m)
// row-major, capacity 1 + 2m + 2^m; returns n_exp or -1.
int orc_ccd_design(int m, double f0, double* points, double* weights) {
  if (m < 1 || m > 5) return -1;
  std::vector<std::vector<double>> pts;
  pts.push_back(std::vector<double>(m, 0.0));
  for (int i = 0; i < m; ++i) {
    std::vector<double> a(m, 0.0), b(m, 0.0);
    a[i] = f0;
    b[i] = -f0;
    pts.push_back(a);
    pts.push_back(b);
  }
  if (m >= 2 && m <= 4) {
    double v = f0 / std::sqrt((double)m);
    for (int k = 0; k < (1 << m); ++k) {
      std::vector<double> p(m);
      for (int i = 0; i < m; ++i) p[i] = ((k >> i) & 1) ? v : -v;
      pts.push_back(p);
    }
  } else if (m == 5) {
    double v = f0 / std::sqrt(5.0);
    for (int k = 0; k < 16; ++k) {
      std::vector<double> p(m);
      double sp = 1.0;
      for (int i = 0; i < 4; ++i) {
        double s = ((k >> i) & 1) ? 1.0 : -1.0;
        p[i] = s * v;
        sp *= s;
      }
      p[4] = (g_plant == P_CCD5 ? -sp : sp) * v;
      pts.push_back(p);
    }
  }
  int ne = (int)pts.size();
  const double f2 = g_plant == P_CCD_WEIGHT ? f0 : f0 * f0;
  double w = ne > 1 ? 1.0 / ((ne - 1) * (1.0 + std::exp(-0.5 * f0 * f0) * (f2 / m - 1.0))) : 1.0;
  for (int k = 0; k < ne; ++k) {
    for (int i = 0; i < m; ++i) points[k * m + i] = pts[k][i];
    weights[k] = w;
  }
  weights[0] = 1.0 - (ne - 1) * w;
  return ne;
}

// Test-only: switch on one deliberate mistake (see the Plant list); 0 switches it off.
void orc_plant(int code) { g_plant = code; }

// grid_design (crates/inla_math/src/integration.rs:64-156): the evaluator is the
// caller's callback g(z) (the log posterior at z). m = 1 or 2 (m > 2 is the CCD
// design with f0 = 2.5, handled by the caller). points receives up to cap rows of
// m; returns the number of points, -1 for m outside 1..2, -2 if cap is too small.
typedef double (*orc_eval_fn)(const double* z, void* ctx);
int orc_grid_design(int m, double step, double threshold, orc_eval_fn eval, void* ctx, double* points,
                    int cap) {
  if (m < 1 || m > 2) return -1;
  std::vector<std::vector<double>> pts;
  auto g = [&](std::vector<double> z) { return eval(z.data(), ctx); };
  const bool keep_beyond = g_plant == P_GRID;
  const double g0 = g(std::vector<double>(m, 0.0));
  if (m == 1) {
    pts.push_back({0.0});
    for (int dir = 0; dir < 2; ++dir) {
      for (int k = 1;; ++k) {
        std::vector<double> z{(dir == 0 ? 1.0 : -1.0) * (k * step)};
        const bool beyond = g0 - g(z) > threshold;
        if (beyond && !keep_beyond) break;
        pts.push_back(z);
        if (beyond) break;
      }
    }
  } else {
    auto extent = [&](double sx, double sy) {
      int e = 0;
      for (;;) {
        std::vector<double> z{sx * (e + 1) * step, sy * (e + 1) * step};
        if (g0 - g(z) > threshold) break;
        ++e;
      }
      return e;
    };
    const int max_i = extent(1, 0), min_i = extent(-1, 0), max_j = extent(0, 1), min_j = extent(0, -1);
    for (int i = -min_i; i <= max_i; ++i)
      for (int j = -min_j; j <= max_j; ++j) {
        std::vector<double> z{i * step, j * step};
        if (keep_beyond || g0 - g(z) <= threshold) pts.push_back(z);
      }
  }
  if ((int)pts.size() > cap) return -2;
  for (size_t k = 0; k < pts.size(); ++k)
    for (int i = 0; i < m; ++i) points[k * m + i] = pts[k][i];
  return (int)pts.size();
}

// invert_symmetric_matrix (integration.rs:158-226): Gauss-Jordan with partial
// pivoting on [A | I]; returns 1 for a pivot below 1e-14 ("Singular Hessian").
int orc_invert_symmetric(int m, const double* h, double* inv) {
  const int cols = 2 * m;
  std::vector<double> a((size_t)m * cols, 0.0);
  for (int i = 0; i < m; ++i) {
    for (int j = 0; j < m; ++j) a[i * cols + j] = h[i * m + j];
    a[i * cols + m + i] = 1.0;
  }
  for (int i = 0; i < m; ++i) {
    int mr = i;
    for (int r = i + 1; r < m; ++r)
      if (std::fabs(a[r * cols + i]) > std::fabs(a[mr * cols + i])) mr = r;
    if (std::fabs(a[mr * cols + i]) < 1e-14) return 1;
    if (mr != i)
      for (int c = 0; c < cols; ++c) std::swap(a[i * cols + c], a[mr * cols + c]);
    const double piv = a[i * cols + i];
    for (int c = 0; c < (g_plant == P_INVERT ? m : cols); ++c) a[i * cols + c] /= piv;
    for (int r = 0; r < m; ++r) {
      if (r == i) continue;
      const double f = a[r * cols + i];
      if (f == 0.0) continue;
      for (int c = 0; c < cols; ++c) a[r * cols + c] -= f * a[i * cols + c];
    }
  }
  for (int i = 0; i < m; ++i)
    for (int j = 0; j < m; ++j) inv[i * m + j] = a[i * cols + m + j];
  return 0;
}

// jacobi_eigen_scalar (integration.rs:249-318): cyclic Jacobi rotations on a
// row-major symmetric m x m matrix; eigenvalues in lam[m], eigenvectors as the
// columns of v (row-major m x m). Returns 1 when max_iter sweeps do not converge.
int orc_jacobi_eigen(int m, const double* mat, int max_iter, double* lam, double* v) {
  std::vector<double> a(mat, mat + (size_t)m * m);
  for (int i = 0; i < m * m; ++i) v[i] = 0.0;
  for (int i = 0; i < m; ++i) v[i * m + i] = 1.0;
  for (int sweep = 0; sweep < max_iter; ++sweep) {
    double off = 0.0;
    for (int i = 0; i < m; ++i)
      for (int j = i + 1; j < m; ++j) off += std::fabs(a[i * m + j]);
    if (off < 1e-12) {
      for (int i = 0; i < m; ++i) lam[i] = a[i * m + i];
      return 0;
    }
    for (int p = 0; p < m; ++p)
      for (int q = p + 1; q < m; ++q) {
        const double apq = a[p * m + q];
        if (std::fabs(apq) < 1e-15) continue;
        const double app = a[p * m + p], aqq = a[q * m + q];
        const double tau = (aqq - app) / (2.0 * apq);
        const double t = tau >= 0.0 ? 1.0 / (tau + std::sqrt(1.0 + tau * tau))
                                    : -1.0 / (-tau + std::sqrt(1.0 + tau * tau));
        const double c = 1.0 / std::sqrt(1.0 + t * t), sn = t * c;
        a[p * m + p] = app - t * apq;
        a[q * m + q] = aqq + t * apq;
        a[p * m + q] = a[q * m + p] = 0.0;
        for (int r = 0; r < m; ++r) {
          if (r == p || r == q) continue;
          const double arp = a[r * m + p], arq = a[r * m + q];
          a[r * m + p] = a[p * m + r] = c * arp - sn * arq;
          a[r * m + q] = a[q * m + r] = sn * arp + c * arq;
        }
        const double sv = g_plant == P_JACOBI ? -sn : sn;
        for (int r = 0; r < m; ++r) {
          const double vrp = v[r * m + p], vrq = v[r * m + q];
          v[r * m + p] = c * vrp - sv * vrq;
          v[r * m + q] = sv * vrp + c * vrq;
        }
      }
  }
  return 1;
}

}  // extern "C"
