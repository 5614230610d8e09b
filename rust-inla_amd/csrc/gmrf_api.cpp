// Host runtime behind include/gmrf_c.h: symbolic cache, device plans, batch
// workspaces, level-by-level launch sequences and the InlaSolver mirror.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/gmrf_c.h"
#include "plan.hpp"
#include "symbolic.hpp"

namespace gmrf {
hipError_t launch_factor_small(const DevPlan&, const int32_t*, int, int, int, const double*,
                               int64_t, double*, double*, hipStream_t);
hipError_t launch_factor_large(const DevPlan&, const int32_t*, int, int, const double*, int64_t,
                               double*, double*, hipStream_t);
hipError_t launch_logdet(const DevPlan&, int, const double*, double*, int32_t*, hipStream_t);
hipError_t launch_gather_perm(const DevPlan&, int, const double*, int64_t, double*, hipStream_t);
hipError_t launch_scatter_perm(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_diag_scale(const DevPlan&, int, const double*, double*, hipStream_t);
hipError_t launch_solve_fwd(const DevPlan&, const int32_t*, int, int, int, const double*, double*,
                            double*, hipStream_t);
hipError_t launch_solve_bwd(const DevPlan&, const int32_t*, int, int, int, const double*, double*,
                            hipStream_t);
hipError_t launch_selinv_small(const DevPlan&, const int32_t*, int, int, int, const double*,
                               double*, double*, hipStream_t);
hipError_t launch_selinv_large(const DevPlan&, const int32_t*, int, int, const double*, double*,
                               double*, double*, hipStream_t);
hipError_t launch_diag_out(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_pattern_out(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_mma_probe(const double*, const double*, int, double*, hipStream_t);
}  // namespace gmrf

using namespace gmrf;

namespace {

thread_local std::string g_err;
int g_device = 0;

gmrf_status fail(gmrf_status st, const std::string& msg) {
  g_err = msg;
  return st;
}

struct HipError : std::runtime_error {
  hipError_t code;
  HipError(hipError_t c, const char* what)
      : std::runtime_error(std::string(what) + ": " + hipGetErrorString(c)), code(c) {}
};

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw HipError(e, what);
}

gmrf_status from_exception() {
  try {
    throw;
  } catch (const HipError& e) {
    if (e.code == hipErrorOutOfMemory || e.code == hipErrorMemoryAllocation)
      return fail(GMRF_OUT_OF_MEMORY, e.what());
    return fail(GMRF_DEVICE_ERROR, e.what());
  } catch (const std::bad_alloc&) {
    return fail(GMRF_OUT_OF_MEMORY, "out of memory during sparse factorization");
  } catch (const std::exception& e) {
    return fail(GMRF_INVALID_ARGUMENT, e.what());
  } catch (...) {
    return fail(GMRF_DEVICE_ERROR, "unknown error");
  }
}

// Device allocation owned by a vector of raw pointers freed in the destructor.
struct DevArena {
  std::vector<void*> ptrs;
  ~DevArena() {
    for (void* p : ptrs) (void)hipFree(p);
  }
  template <typename T>
  T* alloc(size_t count) {
    void* p = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    check(hipMalloc(&p, bytes), "hipMalloc");
    ptrs.push_back(p);
    return (T*)p;
  }
  template <typename T>
  T* upload(const std::vector<T>& v) {
    T* d = alloc<T>(v.size());
    if (!v.empty()) check(hipMemcpy(d, v.data(), v.size() * sizeof(T), hipMemcpyHostToDevice), "upload");
    return d;
  }
};

struct LevelInfo {
  int64_t sb, se;     // small fronts [sb, se) in small list
  int64_t lb, le;     // large fronts [lb, le) in large list
  int64_t ab, ae;     // all fronts [ab, ae) in all list
  int32_t maxm_small, maxm_all;
};

}  // namespace

// Symbolic analysis plus its device plan; shared by batches and solvers.
struct SymData {
  Symbolic S;
  int device = 0;
  DevArena arena;
  DevPlan plan{};
  std::vector<LevelInfo> levels;
  int32_t* d_small = nullptr;
  int32_t* d_large = nullptr;
  int32_t* d_all = nullptr;
  int order_req = 0;
  std::vector<int64_t> given;  // given permutation (cache key)
  std::mutex dev_mu;
  bool dev_built = false;

  // Upload the plan on first use (symbolic analysis itself needs no device).
  void ensure_device() {
    std::lock_guard<std::mutex> lk(dev_mu);
    if (dev_built) return;
    build_device();
    dev_built = true;
  }

  void build_device() {
    check(hipSetDevice(device), "hipSetDevice");
    const int64_t ns = S.nsuper, n = S.n;
    if (n >= (int64_t)1 << 31 || S.nnz_full >= (int64_t)1 << 31)
      throw std::runtime_error("pattern too large for 32-bit device indices");
    std::vector<int32_t> sn_start(ns + 1), sn_r(ns), sn_parent(ns), child_ptr(ns + 1),
        child_idx(S.child_idx.size()), rows(S.sn_rows.size()), perm(n);
    for (int64_t s = 0; s <= ns; ++s) sn_start[s] = (int32_t)S.sn_start[s];
    for (int64_t s = 0; s < ns; ++s) {
      sn_r[s] = (int32_t)(S.sn_rows_ptr[s + 1] - S.sn_rows_ptr[s]);
      sn_parent[s] = (int32_t)S.sn_parent[s];
    }
    for (int64_t s = 0; s <= ns; ++s) child_ptr[s] = (int32_t)S.child_ptr[s];
    for (size_t q = 0; q < S.child_idx.size(); ++q) child_idx[q] = (int32_t)S.child_idx[q];
    for (size_t q = 0; q < S.sn_rows.size(); ++q) rows[q] = (int32_t)S.sn_rows[q];
    for (int64_t k = 0; k < n; ++k) perm[k] = (int32_t)S.perm[k];
    // input entries per supernode
    std::vector<int64_t> col2s(n);
    for (int64_t s = 0; s < ns; ++s)
      for (int64_t j = S.sn_start[s]; j < S.sn_start[s + 1]; ++j) col2s[j] = s;
    std::vector<int64_t> fa_ptr(ns + 1, 0);
    std::vector<int64_t> esn(S.nnz_full, -1);
    for (int64_t c = 0; c < n; ++c)
      for (int64_t p = S.indptr[c]; p < S.indptr[c + 1]; ++p) {
        if (S.amap[p] < 0) continue;
        int64_t fj = std::min(S.iperm[S.indices[p]], S.iperm[c]);
        esn[p] = col2s[fj];
        fa_ptr[esn[p] + 1]++;
      }
    for (int64_t s = 0; s < ns; ++s) fa_ptr[s + 1] += fa_ptr[s];
    std::vector<int32_t> fa_csc(fa_ptr[ns]), fa_slot(fa_ptr[ns]);
    {
      std::vector<int64_t> fill(fa_ptr.begin(), fa_ptr.end() - 1);
      for (int64_t p = 0; p < S.nnz_full; ++p) {
        if (esn[p] < 0) continue;
        int64_t s = esn[p], q = fill[s]++;
        fa_csc[q] = (int32_t)p;
        int64_t local = S.amap[p] - S.l_off[s];
        if (local < 0 || local >= ((int64_t)1 << 31)) throw std::runtime_error("front too large");
        fa_slot[q] = (int32_t)local;
      }
    }
    plan.n = (int32_t)n;
    plan.nsuper = (int32_t)ns;
    plan.nnz = S.nnz_full;
    plan.l_size = S.l_size;
    plan.u_size = S.u_size;
    plan.v_size = S.v_size;
    plan.sn_start = arena.upload(sn_start);
    plan.sn_r = arena.upload(sn_r);
    plan.sn_parent = arena.upload(sn_parent);
    std::vector<int64_t> loff(S.l_off.begin(), S.l_off.begin() + ns);
    std::vector<int64_t> uoff(S.u_off.begin(), S.u_off.begin() + ns);
    std::vector<int64_t> voff(S.v_off.begin(), S.v_off.begin() + ns);
    plan.l_off = arena.upload(loff);
    plan.u_off = arena.upload(uoff);
    plan.v_off = arena.upload(voff);
    plan.rows_ptr = arena.upload(S.sn_rows_ptr);
    plan.rows = arena.upload(rows);
    plan.rel = arena.upload(S.rel);
    plan.child_ptr = arena.upload(child_ptr);
    plan.child_idx = arena.upload(child_idx);
    plan.fa_ptr = arena.upload(fa_ptr);
    plan.fa_csc = arena.upload(fa_csc);
    plan.fa_slot = arena.upload(fa_slot);
    plan.perm = arena.upload(perm);
    plan.diag_slot = arena.upload(S.diag_slot);
    plan.slot_full = arena.upload(S.slot_full);
    // level classes
    std::vector<int32_t> small, large, all;
    levels.resize(S.nlevels);
    for (int l = 0; l < S.nlevels; ++l) {
      LevelInfo& L = levels[l];
      L.sb = (int64_t)small.size();
      L.lb = (int64_t)large.size();
      L.ab = (int64_t)all.size();
      L.maxm_small = 0;
      L.maxm_all = 0;
      for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
        int64_t s = S.level_sn[q];
        int32_t m = (int32_t)((S.sn_start[s + 1] - S.sn_start[s]) + sn_r[s]);
        all.push_back((int32_t)s);
        L.maxm_all = std::max(L.maxm_all, m);
        if (m <= kSmallFront) {
          small.push_back((int32_t)s);
          L.maxm_small = std::max(L.maxm_small, m);
        } else {
          large.push_back((int32_t)s);
        }
      }
      L.se = (int64_t)small.size();
      L.le = (int64_t)large.size();
      L.ae = (int64_t)all.size();
    }
    d_small = arena.upload(small);
    d_large = arena.upload(large);
    d_all = arena.upload(all);
  }
};

struct gmrf_symbolic {
  std::shared_ptr<SymData> d;
};

namespace {

// Process-wide pattern cache (the reference keeps one per solver; scratch.rs:92-118).
std::mutex g_cache_mu;
std::vector<std::shared_ptr<SymData>> g_cache;  // most recent last
constexpr size_t kCacheMax = 16;

std::shared_ptr<SymData> analyze_cached(int64_t n, const int64_t* indptr, const int64_t* indices,
                                        int order, const int64_t* given) {
  std::string err = validate_pattern(n, indptr, indices);
  if (!err.empty()) throw std::runtime_error(err);
  uint64_t h = hash_pattern(n, indptr, indices);
  {
    std::lock_guard<std::mutex> lk(g_cache_mu);
    for (size_t i = g_cache.size(); i-- > 0;) {
      auto& c = g_cache[i];
      if (c->S.pattern_hash != h || c->S.n != n || c->order_req != order || c->device != g_device)
        continue;
      if (!std::equal(indptr, indptr + n + 1, c->S.indptr.begin())) continue;
      if (!std::equal(indices, indices + indptr[n], c->S.indices.begin())) continue;
      if (order == GMRF_ORDER_GIVEN && !std::equal(given, given + n, c->given.begin())) continue;
      auto hit = c;
      g_cache.erase(g_cache.begin() + i);
      g_cache.push_back(hit);
      return hit;
    }
  }
  auto d = std::make_shared<SymData>();
  OrderRequest req = order == GMRF_ORDER_AMD       ? OrderRequest::AmdOnly
                     : order == GMRF_ORDER_NATURAL ? OrderRequest::Natural
                     : order == GMRF_ORDER_GIVEN   ? OrderRequest::Given
                                                   : OrderRequest::Chooser;
  d->S = analyze(n, indptr, indices, req, given);
  d->device = g_device;
  d->order_req = order;
  if (order == GMRF_ORDER_GIVEN) d->given.assign(given, given + n);
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache.push_back(d);
  if (g_cache.size() > kCacheMax) g_cache.erase(g_cache.begin());
  return d;
}

}  // namespace

// Numeric workspace for up to `cap` matrices of one pattern.
struct gmrf_batch {
  std::shared_ptr<SymData> sym;
  int cap = 0;
  hipStream_t stream = nullptr;  // stream in use (own or caller's)
  hipStream_t own = nullptr;
  DevArena arena;
  double *L = nullptr, *U = nullptr, *S = nullptr, *T = nullptr, *X = nullptr, *V = nullptr;
  double* d_logdet = nullptr;
  int32_t* d_status = nullptr;

  gmrf_batch(std::shared_ptr<SymData> s, int capacity) : sym(std::move(s)), cap(capacity) {
    check(hipSetDevice(sym->device), "hipSetDevice");
    sym->ensure_device();
    check(hipStreamCreateWithFlags(&own, hipStreamNonBlocking), "hipStreamCreate");
    stream = own;
    const Symbolic& Sy = sym->S;
    L = arena.alloc<double>((size_t)cap * Sy.l_size);
    U = arena.alloc<double>((size_t)cap * std::max<int64_t>(Sy.u_size, 1));
    S = arena.alloc<double>((size_t)cap * Sy.l_size);
    T = arena.alloc<double>((size_t)cap * Sy.l_size);
    X = arena.alloc<double>((size_t)cap * Sy.n);
    V = arena.alloc<double>((size_t)cap * std::max<int64_t>(Sy.v_size, 1));
    d_logdet = arena.alloc<double>(cap);
    d_status = arena.alloc<int32_t>(cap);
  }
  ~gmrf_batch() {
    if (own) {
      (void)hipStreamSynchronize(own);
      (void)hipStreamDestroy(own);
    }
  }

  void factorize(int count, const double* vals, int64_t ldv) {
    const DevPlan& P = sym->plan;
    for (const LevelInfo& lv : sym->levels) {
      check(launch_factor_small(P, sym->d_small + lv.sb, (int)(lv.se - lv.sb), lv.maxm_small, count,
                                vals, ldv, L, U, stream),
            "factor_small");
      check(launch_factor_large(P, sym->d_large + lv.lb, (int)(lv.le - lv.lb), count, vals, ldv, L,
                                U, stream),
            "factor_large");
    }
    check(launch_logdet(P, count, L, d_logdet, d_status, stream), "logdet");
  }

  void solve(int count, const double* rhs, int64_t ldr, double* x, int64_t ldx) {
    const DevPlan& P = sym->plan;
    check(launch_gather_perm(P, count, rhs, ldr, X, stream), "gather");
    for (const LevelInfo& lv : sym->levels)
      check(launch_solve_fwd(P, sym->d_all + lv.ab, (int)(lv.ae - lv.ab), lv.maxm_all, count, L, X,
                             V, stream),
            "solve_fwd");
    check(launch_diag_scale(P, count, L, X, stream), "diag_scale");
    for (size_t l = sym->levels.size(); l-- > 0;) {
      const LevelInfo& lv = sym->levels[l];
      check(launch_solve_bwd(P, sym->d_all + lv.ab, (int)(lv.ae - lv.ab), lv.maxm_all, count, L, X,
                             stream),
            "solve_bwd");
    }
    check(launch_scatter_perm(P, count, X, x, ldx, stream), "scatter");
  }

  void selinv(int count) {
    const DevPlan& P = sym->plan;
    for (size_t l = sym->levels.size(); l-- > 0;) {
      const LevelInfo& lv = sym->levels[l];
      check(launch_selinv_large(P, sym->d_large + lv.lb, (int)(lv.le - lv.lb), count, L, S, U, T,
                                stream),
            "selinv_large");
      check(launch_selinv_small(P, sym->d_small + lv.sb, (int)(lv.se - lv.sb), lv.maxm_small, count,
                                L, S, U, stream),
            "selinv_small");
    }
  }
};

// InlaSolver mirror: one matrix, host pointers, double-buffered so that a failed
// factorize leaves the previous factor usable (solver.rs:70-74 keeps it).
struct gmrf_solver {
  struct Slot {
    std::unique_ptr<gmrf_batch> bt;
    double logdet = 0.0;
    bool have_selinv = false;
  };
  Slot slot[2];
  int cur = -1;
  double* d_vals = nullptr;
  size_t d_vals_cap = 0;
  double* d_vec = nullptr;
  size_t d_vec_cap = 0;
  ~gmrf_solver() {
    if (d_vals) (void)hipFree(d_vals);
    if (d_vec) (void)hipFree(d_vec);
  }
  void ensure(double*& p, size_t& cap, size_t need) {
    if (cap >= need) return;
    if (p) (void)hipFree(p);
    p = nullptr;
    cap = 0;
    check(hipMalloc(&p, need * sizeof(double)), "hipMalloc");
    cap = need;
  }
};

extern "C" {

const char* gmrf_last_error(void) { return g_err.c_str(); }

gmrf_status gmrf_set_device(int32_t device) {
  int cnt = 0;
  if (hipGetDeviceCount(&cnt) != hipSuccess || device < 0 || device >= cnt)
    return fail(GMRF_INVALID_ARGUMENT, "invalid device index");
  g_device = device;
  return GMRF_OK;
}

gmrf_status gmrf_analyze(int64_t n, const int64_t* indptr, const int64_t* indices, int32_t order,
                         const int64_t* perm, gmrf_symbolic** out) {
  if (!out || !indptr || (n > 0 && !indices && indptr[n] > 0))
    return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (order < 0 || order > 3) return fail(GMRF_INVALID_ARGUMENT, "unknown ordering request");
  if (order == GMRF_ORDER_GIVEN && !perm)
    return fail(GMRF_INVALID_ARGUMENT, "GMRF_ORDER_GIVEN needs a permutation");
  try {
    auto d = analyze_cached(n, indptr, indices, order, perm);
    *out = new gmrf_symbolic{d};
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

void gmrf_symbolic_free(gmrf_symbolic* sym) { delete sym; }

gmrf_status gmrf_symbolic_get_stats(const gmrf_symbolic* sym, gmrf_symbolic_stats* o) {
  if (!sym || !o) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  const Symbolic& S = sym->d->S;
  o->n = S.n;
  o->nnz = S.nnz_full;
  o->nnz_q_lower = S.nnz_q_lower;
  o->order_kind = (int32_t)S.order_kind;
  o->nlevels = S.nlevels;
  o->nsuper = S.nsuper;
  o->nnz_l = S.nnz_l;
  o->l_size = S.l_size;
  o->u_size = S.u_size;
  o->max_front = S.max_front;
  o->flops_ldlt = S.flops_ldlt;
  o->flops_selinv = S.flops_selinv;
  o->flops_solve = S.flops_solve;
  o->sum_r2 = S.sum_r2;
  return GMRF_OK;
}

gmrf_status gmrf_symbolic_perm(const gmrf_symbolic* sym, int64_t* perm) {
  if (!sym || !perm) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  std::copy(sym->d->S.perm.begin(), sym->d->S.perm.end(), perm);
  return GMRF_OK;
}

gmrf_status gmrf_symbolic_etree(const gmrf_symbolic* sym, int64_t* parent) {
  if (!sym || !parent) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  std::copy(sym->d->S.parent.begin(), sym->d->S.parent.end(), parent);
  return GMRF_OK;
}

gmrf_status gmrf_symbolic_colcount(const gmrf_symbolic* sym, int64_t* cc) {
  if (!sym || !cc) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  std::copy(sym->d->S.colcount.begin(), sym->d->S.colcount.end(), cc);
  return GMRF_OK;
}

gmrf_status gmrf_symbolic_supernodes(const gmrf_symbolic* sym, int64_t* sn_start,
                                     int64_t* sn_parent) {
  if (!sym) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  const Symbolic& S = sym->d->S;
  if (sn_start) std::copy(S.sn_start.begin(), S.sn_start.end(), sn_start);
  if (sn_parent) std::copy(S.sn_parent.begin(), S.sn_parent.end(), sn_parent);
  return GMRF_OK;
}

int32_t gmrf_choose_order(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t* fwd) {
  std::vector<int64_t> f, inv;
  OrderKind k = choose_symmetric_order(n, indptr, indices, f, inv);
  if (k != OrderKind::Amd && fwd) std::copy(f.begin(), f.end(), fwd);
  return (int32_t)k;
}

gmrf_status gmrf_batch_create(gmrf_symbolic* sym, int32_t capacity, gmrf_batch** out) {
  if (!sym || !out || capacity <= 0) return fail(GMRF_INVALID_ARGUMENT, "invalid batch request");
  try {
    *out = new gmrf_batch(sym->d, capacity);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

void gmrf_batch_destroy(gmrf_batch* bt) { delete bt; }

gmrf_status gmrf_batch_factorize(gmrf_batch* bt, int32_t count, const double* values,
                                 int64_t ld_values) {
  if (!bt || !values) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_values < bt->sym->S.nnz_full) return fail(GMRF_INVALID_ARGUMENT, "ld_values < nnz");
  try {
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->factorize(count, values, ld_values);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_logdet_device(gmrf_batch* bt, int32_t count, double* logdet,
                                     int32_t* status) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  try {
    if (logdet)
      check(hipMemcpyAsync(logdet, bt->d_logdet, count * sizeof(double), hipMemcpyDeviceToDevice,
                           bt->stream),
            "copy logdet");
    if (status)
      check(hipMemcpyAsync(status, bt->d_status, count * sizeof(int32_t), hipMemcpyDeviceToDevice,
                           bt->stream),
            "copy status");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_logdet(gmrf_batch* bt, int32_t count, double* logdet, int32_t* status) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  try {
    std::vector<double> ld(count);
    std::vector<int32_t> st(count);
    check(hipMemcpyAsync(ld.data(), bt->d_logdet, count * sizeof(double), hipMemcpyDeviceToHost,
                         bt->stream),
          "copy logdet");
    check(hipMemcpyAsync(st.data(), bt->d_status, count * sizeof(int32_t), hipMemcpyDeviceToHost,
                         bt->stream),
          "copy status");
    check(hipStreamSynchronize(bt->stream), "sync");
    for (int b = 0; b < count; ++b) {
      if (logdet) logdet[b] = ld[b];
      if (status)
        status[b] = (st[b] & 2) ? GMRF_NOT_POSITIVE_DEFINITE : (st[b] & 1) ? GMRF_SINGULAR : GMRF_OK;
    }
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_solve(gmrf_batch* bt, int32_t count, const double* rhs, int64_t ld_rhs,
                             double* x, int64_t ld_x) {
  if (!bt || !rhs || !x) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_rhs < bt->sym->S.n || ld_x < bt->sym->S.n)
    return fail(GMRF_INVALID_ARGUMENT, "leading dimension < n");
  try {
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->solve(count, rhs, ld_rhs, x, ld_x);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_selinv(gmrf_batch* bt, int32_t count) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  try {
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->selinv(count);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_selinv_diag(gmrf_batch* bt, int32_t count, double* diag, int64_t ld_diag) {
  if (!bt || !diag) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_diag < bt->sym->S.n) return fail(GMRF_INVALID_ARGUMENT, "ld_diag < n");
  try {
    check(launch_diag_out(bt->sym->plan, count, bt->S, diag, ld_diag, bt->stream), "diag_out");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_selinv_pattern(gmrf_batch* bt, int32_t count, double* sigma,
                                      int64_t ld_sigma) {
  if (!bt || !sigma) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_sigma < bt->sym->S.nnz_full) return fail(GMRF_INVALID_ARGUMENT, "ld_sigma < nnz");
  try {
    check(launch_pattern_out(bt->sym->plan, count, bt->S, sigma, ld_sigma, bt->stream),
          "pattern_out");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_sync(gmrf_batch* bt) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  try {
    check(hipStreamSynchronize(bt->stream), "hipStreamSynchronize");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

void* gmrf_batch_stream(gmrf_batch* bt) { return bt ? (void*)bt->stream : nullptr; }

gmrf_status gmrf_batch_set_stream(gmrf_batch* bt, void* stream) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  bt->stream = stream ? (hipStream_t)stream : bt->own;
  return GMRF_OK;
}

gmrf_status gmrf_solver_create(gmrf_solver** out) {
  if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  *out = new gmrf_solver();
  return GMRF_OK;
}

void gmrf_solver_destroy(gmrf_solver* s) { delete s; }

int64_t gmrf_solver_n(const gmrf_solver* s) {
  if (!s || s->cur < 0) return 0;
  return s->slot[s->cur].bt->sym->S.n;
}

gmrf_status gmrf_solver_factorize(gmrf_solver* s, int64_t n, const int64_t* indptr,
                                  const int64_t* indices, const double* data) {
  if (!s || !indptr) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (n <= 0) return fail(GMRF_INVALID_ARGUMENT, "empty or invalid matrix: n must be > 0");
  try {
    check(hipSetDevice(g_device), "hipSetDevice");
    auto sym = analyze_cached(n, indptr, indices, GMRF_ORDER_CHOOSER, nullptr);
    const int target = s->cur < 0 ? 0 : 1 - s->cur;
    auto& sl = s->slot[target];
    if (!sl.bt || sl.bt->sym != sym) {
      sl.bt.reset();
      sl.bt.reset(new gmrf_batch(sym, 1));
    }
    const int64_t nnz = indptr[n];
    s->ensure(s->d_vals, s->d_vals_cap, (size_t)std::max<int64_t>(nnz, 1));
    check(hipMemcpyAsync(s->d_vals, data, nnz * sizeof(double), hipMemcpyHostToDevice,
                         sl.bt->stream),
          "upload values");
    sl.bt->factorize(1, s->d_vals, nnz);
    double ld;
    int32_t st;
    gmrf_status rc = gmrf_batch_logdet(sl.bt.get(), 1, &ld, &st);
    if (rc != GMRF_OK) return rc;
    if (st == GMRF_NOT_POSITIVE_DEFINITE)
      return fail(GMRF_NOT_POSITIVE_DEFINITE, "Matrix is not positive-definite during factorization");
    if (st == GMRF_SINGULAR)
      return fail(GMRF_SINGULAR, "Matrix is singular or structurally rank-deficient");
    sl.logdet = ld;
    sl.have_selinv = false;
    s->cur = target;
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_solve(gmrf_solver* s, const double* rhs, double* x) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->cur < 0) return fail(GMRF_SOLVE_FAILURE, "Linear system solve failed: Factorization not computed yet");
  if (!rhs || !x) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  try {
    auto& bt = s->slot[s->cur].bt;
    const int64_t n = bt->sym->S.n;
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    s->ensure(s->d_vec, s->d_vec_cap, (size_t)n);
    check(hipMemcpyAsync(s->d_vec, rhs, n * sizeof(double), hipMemcpyHostToDevice, bt->stream),
          "upload rhs");
    bt->solve(1, s->d_vec, n, s->d_vec, n);
    check(hipMemcpyAsync(x, s->d_vec, n * sizeof(double), hipMemcpyDeviceToHost, bt->stream),
          "download x");
    check(hipStreamSynchronize(bt->stream), "sync");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_diag_inv(gmrf_solver* s, double* out) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->cur < 0) return fail(GMRF_SOLVE_FAILURE, "Linear system solve failed: Factorization not computed yet");
  if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  try {
    auto& sl = s->slot[s->cur];
    auto& bt = sl.bt;
    const int64_t n = bt->sym->S.n;
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    if (!sl.have_selinv) {
      bt->selinv(1);
      sl.have_selinv = true;
    }
    s->ensure(s->d_vec, s->d_vec_cap, (size_t)n);
    check(launch_diag_out(bt->sym->plan, 1, bt->S, s->d_vec, n, bt->stream), "diag_out");
    check(hipMemcpyAsync(out, s->d_vec, n * sizeof(double), hipMemcpyDeviceToHost, bt->stream),
          "download diag");
    check(hipStreamSynchronize(bt->stream), "sync");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_log_abs_det(const gmrf_solver* s, double* out) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->cur < 0) return fail(GMRF_SOLVE_FAILURE, "Linear system solve failed: Factorization not computed yet");
  if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  *out = s->slot[s->cur].logdet;
  return GMRF_OK;
}

gmrf_status gmrf_probe_mma(const double* A, const double* B, int32_t K, double* T) {
  if (!A || !B || !T || K < 1 || K > 64) return fail(GMRF_INVALID_ARGUMENT, "invalid probe");
  double *dA = nullptr, *dB = nullptr, *dT = nullptr;
  try {
    check(hipSetDevice(g_device), "hipSetDevice");
    check(hipMalloc(&dA, 64 * 64 * sizeof(double)), "hipMalloc");
    check(hipMalloc(&dB, 64 * 64 * sizeof(double)), "hipMalloc");
    check(hipMalloc(&dT, 64 * 64 * sizeof(double)), "hipMalloc");
    check(hipMemcpy(dA, A, 64 * K * sizeof(double), hipMemcpyHostToDevice), "upload");
    check(hipMemcpy(dB, B, 64 * K * sizeof(double), hipMemcpyHostToDevice), "upload");
    check(launch_mma_probe(dA, dB, K, dT, 0), "probe");
    check(hipMemcpy(T, dT, 64 * 64 * sizeof(double), hipMemcpyDeviceToHost), "download");
    (void)hipFree(dA);
    (void)hipFree(dB);
    (void)hipFree(dT);
    return GMRF_OK;
  } catch (...) {
    if (dA) (void)hipFree(dA);
    if (dB) (void)hipFree(dB);
    if (dT) (void)hipFree(dT);
    return from_exception();
  }
}

}  // extern "C"
