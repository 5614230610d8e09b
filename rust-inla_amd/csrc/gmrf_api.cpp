// Host runtime behind include/gmrf_c.h: symbolic cache, device plans, batch
// workspaces, level-by-level launch sequences and the InlaSolver mirror.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <sstream>
#include <cstring>
#include <memory>
#include <mutex>
#include <stdexcept>
#include <string>
#include <vector>

#include "runtime.hpp"

namespace gmrf_rt {
thread_local std::string g_err;
thread_local int g_device = 0;
Profiler g_prof;
}  // namespace gmrf_rt

namespace {

// Workspace of gmrf_batch_solve_noalloc: a fixed array in device memory (one copy
// per GPU, held until the process ends) for callers that cannot allocate. In host
// code `scratch` only names the symbol; its device address on each GPU comes from
// hipGetSymbolAddress (scratch_on), and only that address is passed to kernels.
// Calls that overlap on one GPU share the array and race.
constexpr int64_t kScratchDoubles = (int64_t)4 << 20;  // 32 MiB
static __device__ __align__(256) double scratch[kScratchDoubles];
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kMaxDevices] = {};

// The array's address on `device` (looked up once, with that device current).
double* scratch_on(int device) {
  if (device < 0 || device >= kMaxDevices) throw std::runtime_error("device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[device]) {
    check(hipSetDevice(device), "hipSetDevice");
    void* p = nullptr;
    check(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)), "hipGetSymbolAddress(scratch)");
    if (!p) throw std::runtime_error("hipGetSymbolAddress(scratch) returned a null address");
    g_scratch_ptr[device] = (double*)p;
  }
  return g_scratch_ptr[device];
}

// Process-wide pattern cache (the reference keeps one per solver; scratch.rs:92-118).
std::mutex g_cache_mu;
std::vector<std::shared_ptr<SymData>> g_cache;  // most recent last
constexpr size_t kCacheMax = 16;

}  // namespace

OrderRequest order_request(int order) {
  switch (order) {
    case GMRF_ORDER_AMD: return OrderRequest::AmdOnly;
    case GMRF_ORDER_NATURAL: return OrderRequest::Natural;
    case GMRF_ORDER_GIVEN: return OrderRequest::Given;
    case GMRF_ORDER_ND: return OrderRequest::NestedDissection;
    case GMRF_ORDER_AUTO: return OrderRequest::Auto;
    default: return OrderRequest::Chooser;
  }
}

std::shared_ptr<SymData> analyze_uncached(int64_t n, const int64_t* indptr, const int64_t* indices,
                                          int order, const int64_t* given) {
  std::string err = validate_pattern(n, indptr, indices);
  if (!err.empty()) throw std::runtime_error(err);
  auto d = std::make_shared<SymData>();
  OrderRequest req = order_request(order);
  d->S = analyze(n, indptr, indices, req, given);
  d->pools = plan_pools(d->S);
  d->device = g_device;
  d->order_req = order;
  if (order == GMRF_ORDER_GIVEN) d->given.assign(given, given + n);
  return d;
}

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
  auto d = analyze_uncached(n, indptr, indices, order, given);
  std::lock_guard<std::mutex> lk(g_cache_mu);
  g_cache.push_back(d);
  if (g_cache.size() > kCacheMax) g_cache.erase(g_cache.begin());
  return d;
}


// Tuned schedules persist as text lines in the file named by GMRF_TUNE_FILE (if set):
//   gmrf-schedule 4 <pattern hash, hex> <n> <levels> <count> <fuse_below> <split_max>
//   <xcd_max> <factorisation merge masks> <selected-inversion merge masks> <stagger>
//   <pp_split_max> <trsm32_max> <gemm32_max>
// one hex digit per level for each mask string; a later line for the same key wins.
// Lines of earlier versions are read with the defaults for the fields they lack
// (version 1: no stagger; 2: no pp_split_max; 3: no trsm32_max / gemm32_max).
void SymData::load_schedules() {
  const char* path = std::getenv("GMRF_TUNE_FILE");
  if (!path) return;
  std::ifstream in(path);
  std::string line;
  while (std::getline(in, line)) {
    std::istringstream ss(line);
    std::string tag, hash, mf, ms;
    int ver = 0, count = 0;
    int64_t n = 0, nl = 0;
    Schedule sc;
    if (!(ss >> tag >> ver >> hash >> n >> nl >> count >> sc.fuse_below >> sc.split_max >> sc.xcd_max >> mf >> ms))
      continue;
    if (tag != "gmrf-schedule" || ver < 1 || ver > 4 || n != S.n || nl != (int64_t)S.nlevels || count < 1)
      continue;
    if (ver >= 2 && !(ss >> sc.stagger)) continue;
    if (ver >= 3 && !(ss >> sc.pp_split_max)) continue;
    if (ver >= 4 && !(ss >> sc.trsm32_max >> sc.gemm32_max)) continue;
    if (sc.stagger < -1 || sc.stagger >= nl) continue;
    if (std::strtoull(hash.c_str(), nullptr, 16) != S.pattern_hash) continue;
    if ((int64_t)mf.size() != nl || (int64_t)ms.size() != nl) continue;
    for (int64_t l = 0; l < nl; ++l) {
      sc.merge_fac.push_back((uint8_t)(std::strtol(std::string(1, mf[l]).c_str(), nullptr, 16) & 7));
      sc.merge_sel.push_back((uint8_t)(std::strtol(std::string(1, ms[l]).c_str(), nullptr, 16) & 7));
    }
    sched[count] = sc;
  }
}

namespace {

void save_schedule(const SymData& d, int count, const Schedule& sc) {
  const char* path = std::getenv("GMRF_TUNE_FILE");
  if (!path) return;
  std::FILE* f = std::fopen(path, "a");
  if (!f) return;
  std::string mf, ms;
  for (int l = 0; l < d.S.nlevels; ++l) {
    mf += "01234567"[sc.fac_mask(l) & 7];
    ms += "01234567"[sc.sel_mask(l) & 7];
  }
  std::fprintf(f, "gmrf-schedule 4 %016llx %lld %d %d %lld %lld %d %s %s %d %lld %lld %lld\n",
               (unsigned long long)d.S.pattern_hash, (long long)d.S.n, d.S.nlevels, count, (long long)sc.fuse_below,
               (long long)sc.split_max, sc.xcd_max, mf.c_str(), ms.c_str(), sc.stagger, (long long)sc.pp_split_max,
               (long long)sc.trsm32_max, (long long)sc.gemm32_max);
  std::fclose(f);
}

// Time (ms, best of `reps`) of factorisation + selected inversion of `count`
// matrices under schedule sc; with `marks`, the per-level small-front times of the
// last repetition go to fac_ms / sel_ms.
double time_schedule(gmrf_batch* bt, int count, const double* vals, int64_t ldv, const Schedule& sc, int reps,
                     gmrf_batch::TuneMarks* marks, std::vector<double>* fac_ms, std::vector<double>* sel_ms) {
  SymData& sd = *bt->sym;
  sd.set_schedule(count, sc);
  hipEvent_t e0, e1;
  check(hipEventCreate(&e0), "hipEventCreate");
  check(hipEventCreate(&e1), "hipEventCreate");
  double best = 1e300;
  try {
    for (int r = 0; r < reps; ++r) {
      bt->marks = (marks && r == reps - 1) ? marks : nullptr;
      check(hipEventRecord(e0, bt->stream), "hipEventRecord");
      const int split = bt->stagger_level(count);
      if (split >= 0) {
        bt->factor_selinv_staggered(count, vals, ldv, split);
      } else {
        bt->factorize(count, vals, ldv);
        bt->selinv(count);
      }
      check(hipEventRecord(e1, bt->stream), "hipEventRecord");
      check(hipEventSynchronize(e1), "hipEventSynchronize");
      float ms = 0.0f;
      check(hipEventElapsedTime(&ms, e0, e1), "hipEventElapsedTime");
      best = std::min(best, (double)ms);
    }
    bt->marks = nullptr;
    if (marks) {
      const size_t nl = sd.levels.size();
      fac_ms->assign(nl, 0.0);
      sel_ms->assign(nl, 0.0);
      for (size_t l = 0; l < nl; ++l) {
        float a = 0.0f, b = 0.0f;
        check(hipEventElapsedTime(&a, marks->fac_a[l], marks->fac_b[l]), "hipEventElapsedTime");
        check(hipEventElapsedTime(&b, marks->sel_a[l], marks->sel_b[l]), "hipEventElapsedTime");
        (*fac_ms)[l] = a;
        (*sel_ms)[l] = b;
      }
    }
  } catch (...) {
    bt->marks = nullptr;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    throw;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}

}  // namespace

// The autotuner behind gmrf_batch_autotune (and gmrf_model_autotune).
gmrf_status autotune_batch(gmrf_batch* bt, int count, const double* vals, int64_t ldv, double* default_ms,
                           double* tuned_ms) {
  try {
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->selinv_join();
    SymData& sd = *bt->sym;
    const size_t nl = sd.levels.size();
    const int reps = 3;
    const Schedule def;
    const Schedule prev = sd.schedule(count);
    const double t_def = time_schedule(bt, count, vals, ldv, def, reps, nullptr, nullptr, nullptr);
    // 1. small-front class grouping, level by level: every level under each of the 8
    //    masks, the per-level times from events around its small-front launches
    gmrf_batch::TuneMarks marks;
    for (auto* v : {&marks.fac_a, &marks.fac_b, &marks.sel_a, &marks.sel_b}) {
      v->resize(nl);
      for (auto& e : *v) check(hipEventCreate(&e), "hipEventCreate");
    }
    Schedule cur = def;
    std::vector<double> bf(nl, 1e300), bs(nl, 1e300), fm, sm;
    std::vector<uint8_t> mf(nl, 0), ms(nl, 0);
    try {
      for (int mask = 0; mask < 8; ++mask) {
        Schedule s = cur;
        s.merge_fac.assign(nl, (uint8_t)mask);
        s.merge_sel.assign(nl, (uint8_t)mask);
        time_schedule(bt, count, vals, ldv, s, 2, &marks, &fm, &sm);
        for (size_t l = 0; l < nl; ++l) {
          if (fm[l] < bf[l] * 0.97) bf[l] = fm[l], mf[l] = (uint8_t)mask;  // a clear gain, else the simpler mask
          if (sm[l] < bs[l] * 0.97) bs[l] = sm[l], ms[l] = (uint8_t)mask;
        }
      }
    } catch (...) {
      for (auto* v : {&marks.fac_a, &marks.fac_b, &marks.sel_a, &marks.sel_b})
        for (auto& e : *v) (void)hipEventDestroy(e);
      throw;
    }
    for (auto* v : {&marks.fac_a, &marks.fac_b, &marks.sel_a, &marks.sel_b})
      for (auto& e : *v) (void)hipEventDestroy(e);
    cur.merge_fac = mf;
    cur.merge_sel = ms;
    double t_cur = time_schedule(bt, count, vals, ldv, cur, reps, nullptr, nullptr, nullptr);
    // 2. the launch thresholds, one at a time
    for (int64_t v : {(int64_t)0, (int64_t)512, (int64_t)kFuseDiagBelow, (int64_t)8192, (int64_t)1 << 40}) {
      Schedule s = cur;
      s.fuse_below = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    for (int64_t v : {(int64_t)0, (int64_t)32, (int64_t)kGemmSplitMaxTasks, (int64_t)256, (int64_t)1024}) {
      Schedule s = cur;
      s.split_max = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    for (int32_t v : {0, 64, kXcdMaxTasks, 1 << 30}) {
      Schedule s = cur;
      s.xcd_max = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    for (int64_t v : {(int64_t)0, (int64_t)256, (int64_t)1024, (int64_t)4096, kPpSplitMaxTasks}) {
      Schedule s = cur;
      s.pp_split_max = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    for (int64_t v : {(int64_t)0, (int64_t)64, (int64_t)128, (int64_t)256, (int64_t)512, (int64_t)1 << 40}) {
      Schedule s = cur;
      s.trsm32_max = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    for (int64_t v : {(int64_t)0, (int64_t)64, (int64_t)128, (int64_t)256, (int64_t)512, (int64_t)1 << 40}) {
      Schedule s = cur;
      s.gemm32_max = v;
      const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
      if (t < t_cur) t_cur = t, cur = s;
    }
    // 3. with GMRF_TUNE_STAGGER set: two staggered pipelines of half the batch, from
    //    each level of the lower two thirds of the tree (the second half's lower
    //    levels beside the first's upper levels). Opt-in: each launch then covers half
    //    the matrices, and model evaluations with a staggered schedule run eagerly
    //    instead of from a captured graph.
    const bool try_stagger = std::getenv("GMRF_TUNE_STAGGER") != nullptr;
    if (try_stagger && count >= 2)
      for (int32_t v = 0; v < (int32_t)(2 * nl / 3); ++v) {
        Schedule s = cur;
        s.stagger = v;
        const double t = time_schedule(bt, count, vals, ldv, s, reps, nullptr, nullptr, nullptr);
        if (t < t_cur) t_cur = t, cur = s;
      }
    // keep the tuned schedule only if it beats the defaults on a fresh measurement
    const double t_def2 = time_schedule(bt, count, vals, ldv, def, reps, nullptr, nullptr, nullptr);
    const double t_best = time_schedule(bt, count, vals, ldv, cur, reps, nullptr, nullptr, nullptr);
    const double t_ref = std::min(t_def, t_def2);
    if (t_best < t_ref) {
      sd.set_schedule(count, cur);
      save_schedule(sd, count, cur);
    } else {
      sd.set_schedule(count, def);
    }
    (void)prev;
    if (default_ms) *default_ms = t_ref;
    if (tuned_ms) *tuned_ms = std::min(t_best, t_ref);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

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
  if (order < 0 || order > 5) return fail(GMRF_INVALID_ARGUMENT, "unknown ordering request");
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
  o->u_size = sym->d->pools.u_size;
  o->max_front = S.max_front;
  o->flops_ldlt = S.flops_ldlt;
  o->flops_selinv = S.flops_selinv;
  o->flops_solve = S.flops_solve;
  o->sum_r2 = S.sum_r2;
  o->pattern_hash = S.pattern_hash;
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
    char* pb = bt->pin.get((size_t)count * (sizeof(double) + sizeof(int32_t)));
    const double* ld = reinterpret_cast<const double*>(pb);
    const int32_t* st = reinterpret_cast<const int32_t*>(pb + (size_t)count * sizeof(double));
    check(hipMemcpyAsync(pb, bt->d_logdet, count * sizeof(double), hipMemcpyDeviceToHost, bt->stream),
          "copy logdet");
    check(hipMemcpyAsync(pb + (size_t)count * sizeof(double), bt->d_status, count * sizeof(int32_t),
                         hipMemcpyDeviceToHost, bt->stream),
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

gmrf_status gmrf_batch_solve_multi(gmrf_batch* bt, int32_t count, int32_t nrhs, const double* rhs,
                                   int64_t ld_rhs, double* x, int64_t ld_x) {
  if (!bt || !rhs || !x) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (nrhs < 1 || (int64_t)count * nrhs > 65535)
    return fail(GMRF_INVALID_ARGUMENT, "nrhs must be >= 1 and count * nrhs <= 65535");
  if (ld_rhs < bt->sym->S.n || ld_x < bt->sym->S.n)
    return fail(GMRF_INVALID_ARGUMENT, "leading dimension < n");
  try {
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->solve(count, rhs, ld_rhs, x, ld_x, 0, nrhs);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_solve_noalloc(gmrf_batch* bt, int32_t count, int32_t nrhs, const double* rhs,
                                     int64_t ld_rhs, double* x, int64_t ld_x) {
  if (!bt || !rhs || !x) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (nrhs < 1 || (int64_t)count * nrhs > 65535)
    return fail(GMRF_INVALID_ARGUMENT, "nrhs must be >= 1 and count * nrhs <= 65535");
  if (ld_rhs < bt->sym->S.n || ld_x < bt->sym->S.n)
    return fail(GMRF_INVALID_ARGUMENT, "leading dimension < n");
  const int64_t per = bt->solve_ws_per_vector();
  const int64_t need = (int64_t)count * nrhs * per;
  if (need > kScratchDoubles)
    return fail(GMRF_INVALID_ARGUMENT,
                "allocation-free solve workspace exceeded: count * nrhs = " + std::to_string((int64_t)count * nrhs) +
                    " vectors need " + std::to_string(need * 8) + " bytes, the fixed area holds " +
                    std::to_string(kScratchDoubles * 8) + " (at most " + std::to_string(kScratchDoubles / per) +
                    " vectors of this pattern)");
  try {
    double* ws = scratch_on(bt->sym->device);
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    bt->solve(count, rhs, ld_rhs, x, ld_x, 0, nrhs, 0, ws);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_batch_autotune(gmrf_batch* bt, int32_t count, const double* values, int64_t ld_values,
                                double* default_ms, double* tuned_ms) {
  if (!bt || !values) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > bt->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_values < bt->sym->S.nnz_full) return fail(GMRF_INVALID_ARGUMENT, "ld_values < nnz");
  return autotune_batch(bt, count, values, ld_values, default_ms, tuned_ms);
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

gmrf_status gmrf_batch_set_stream(gmrf_batch* bt, void* stream, int32_t use_given) {
  if (!bt) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  bt->stream = use_given ? (hipStream_t)stream : bt->own;
  return GMRF_OK;
}

gmrf_status gmrf_solver_create(gmrf_solver** out) {
  if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  *out = new gmrf_solver();
  return GMRF_OK;
}

void gmrf_solver_destroy(gmrf_solver* s) { delete s; }

int64_t gmrf_solver_n(const gmrf_solver* s) {
  if (!s || s->empty || s->cur < 0) return 0;
  return s->slot[s->cur].bt->sym->S.n;
}

}  // extern "C"

namespace {

// Numeric factorisation of `data` (values on sym's pattern) into the solver's spare
// slot; the slot becomes current only on success (solver.rs:70-74 keeps the old factor).
gmrf_status solver_factorize_sym(gmrf_solver* s, const std::shared_ptr<SymData>& sym,
                                 const double* data) {
  const int target = s->cur < 0 ? 0 : 1 - s->cur;
  auto& sl = s->slot[target];
  if (!sl.bt || sl.bt->sym != sym) {
    sl.bt.reset();
    sl.bt.reset(new gmrf_batch(sym, 1));
  }
  const int64_t nnz = sym->S.nnz_full;
  s->ensure(s->d_vals, s->d_vals_cap, (size_t)std::max<int64_t>(nnz, 1));
  // every earlier copy through the staging buffer has completed (each call synchronises)
  char* pb = s->pin.get((size_t)nnz * sizeof(double));
  std::memcpy(pb, data, (size_t)nnz * sizeof(double));
  check(hipMemcpyAsync(s->d_vals, pb, nnz * sizeof(double), hipMemcpyHostToDevice, sl.bt->stream),
        "upload values");
  sl.bt->factorize(1, s->d_vals, nnz);
  double ld;
  int32_t st;
  gmrf_status rc = gmrf_batch_logdet(sl.bt.get(), 1, &ld, &st);
  if (rc != GMRF_OK) return rc;
  if (st == GMRF_NOT_POSITIVE_DEFINITE)
    return fail(GMRF_NOT_POSITIVE_DEFINITE, "Matrix is not positive-definite during factorization");
  if (st == GMRF_SINGULAR) return fail(GMRF_SINGULAR, "Matrix is singular or structurally rank-deficient");
  sl.logdet = ld;
  sl.have_selinv = false;
  s->cur = target;
  s->empty = false;
  return GMRF_OK;
}

}  // namespace

extern "C" {

gmrf_status gmrf_solver_factorize(gmrf_solver* s, int64_t n, const int64_t* indptr,
                                  const int64_t* indices, const double* data) {
  if (!s || !indptr) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (n <= 0) return fail(GMRF_INVALID_ARGUMENT, "empty or invalid matrix: n must be > 0");
  gmrf_status rc;
  try {
    check(hipSetDevice(g_device), "hipSetDevice");
    // GMRF_SIMULATE_SPARSE_OOM set: report the sparse path out of memory (tests of the
    // fallback below, without exhausting the device)
    if (std::getenv("GMRF_SIMULATE_SPARSE_OOM")) {
      rc = fail(GMRF_OUT_OF_MEMORY, "sparse factorization out of memory (simulated)");
    } else {
      auto sym = analyze_cached(n, indptr, indices, GMRF_ORDER_CHOOSER, nullptr);
      rc = solver_factorize_sym(s, sym, data);
    }
  } catch (...) {
    rc = from_exception();
  }
  if (rc != GMRF_OUT_OF_MEMORY) return rc;
  // DefaultBackend::factorize (backend.rs:110-122): a sparse factorisation that runs out
  // of memory falls back to the dense one, on the matrix scattered row-major as
  // DenseBackend::factorize does (backend.rs:39-56); the dense factor needs 8 n^2 bytes
  std::vector<double> a;
  try {
    a.assign((size_t)n * (size_t)n, 0.0);
  } catch (const std::bad_alloc&) {
    return fail(GMRF_OUT_OF_MEMORY, "out of memory in the sparse factorization and for the dense fallback");
  }
  if (!indices || !data) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  for (int64_t c = 0; c < n; ++c)
    for (int64_t p = indptr[c]; p < indptr[c + 1]; ++p) {
      const int64_t r = indices[p];
      if (r < 0 || r >= n) return fail(GMRF_INVALID_ARGUMENT, "row index out of bounds");
      a[(size_t)r * n + c] = data[p];
    }
  return gmrf_solver_factorize_dense(s, n, a.data());
}

gmrf_status gmrf_solver_factorize_dense(gmrf_solver* s, int64_t n, const double* a) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (n < 0) return fail(GMRF_INVALID_ARGUMENT, "dense LDL^T matrix length: n must be >= 0");
  if (n == 0) {  // dense_faer.rs:33-39: an empty factor, not an error
    s->empty = true;
    return GMRF_OK;
  }
  if (!a) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  try {
    check(hipSetDevice(g_device), "hipSetDevice");
    // the full pattern's analysis is held by the solver's own slots (reused while n
    // stays the same), not by the process-wide pattern cache
    std::shared_ptr<SymData> sym;
    for (auto& sl : s->slot)
      if (sl.bt && sl.bt->sym->dense_n == n && sl.bt->sym->device == g_device) sym = sl.bt->sym;
    if (!sym) {
      std::vector<int64_t> ip(n + 1), ix((size_t)n * n);
      for (int64_t j = 0; j <= n; ++j) ip[j] = j * n;
      for (int64_t j = 0; j < n; ++j)
        for (int64_t i = 0; i < n; ++i) ix[(size_t)j * n + i] = i;
      sym = analyze_uncached(n, ip.data(), ix.data(), GMRF_ORDER_CHOOSER, nullptr);
      sym->dense_n = n;
    }
    // column j of the full pattern holds rows 0..n-1; entry (i, j) comes from the
    // lower triangle a[max(i, j) n + min(i, j)]
    std::vector<double> v((size_t)n * n);
    for (int64_t j = 0; j < n; ++j)
      for (int64_t i = 0; i < n; ++i) v[(size_t)j * n + i] = a[(size_t)std::max(i, j) * n + std::min(i, j)];
    const gmrf_status rc = solver_factorize_sym(s, sym, v.data());
    if (rc == GMRF_NOT_POSITIVE_DEFINITE)
      return fail(GMRF_SINGULAR, "Matrix is singular or structurally rank-deficient");
    return rc;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_solve(gmrf_solver* s, const double* rhs, double* x) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->empty) return GMRF_OK;  // 0 x 0 factor: nothing to write
  if (s->cur < 0) return fail(GMRF_SOLVE_FAILURE, "Linear system solve failed: Factorization not computed yet");
  if (!rhs || !x) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  try {
    auto& bt = s->slot[s->cur].bt;
    const int64_t n = bt->sym->S.n;
    check(hipSetDevice(bt->sym->device), "hipSetDevice");
    s->ensure(s->d_vec, s->d_vec_cap, (size_t)n);
    double* pb = reinterpret_cast<double*>(s->pin.get((size_t)n * sizeof(double)));
    std::memcpy(pb, rhs, n * sizeof(double));
    check(hipMemcpyAsync(s->d_vec, pb, n * sizeof(double), hipMemcpyHostToDevice, bt->stream), "upload rhs");
    bt->solve(1, s->d_vec, n, s->d_vec, n);
    check(hipMemcpyAsync(pb, s->d_vec, n * sizeof(double), hipMemcpyDeviceToHost, bt->stream), "download x");
    check(hipStreamSynchronize(bt->stream), "sync");
    std::memcpy(x, pb, n * sizeof(double));
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_diag_inv(gmrf_solver* s, double* out) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->empty) return GMRF_OK;  // 0 x 0 factor: nothing to write
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
    double* pb = reinterpret_cast<double*>(s->pin.get((size_t)n * sizeof(double)));
    check(hipMemcpyAsync(pb, s->d_vec, n * sizeof(double), hipMemcpyDeviceToHost, bt->stream),
          "download diag");
    check(hipStreamSynchronize(bt->stream), "sync");
    std::memcpy(out, pb, n * sizeof(double));
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_solver_log_abs_det(const gmrf_solver* s, double* out) {
  if (!s) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (s->empty) {
    if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
    *out = 0.0;  // empty sum of log|D|
    return GMRF_OK;
  }
  if (s->cur < 0) return fail(GMRF_SOLVE_FAILURE, "Linear system solve failed: Factorization not computed yet");
  if (!out) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  *out = s->slot[s->cur].logdet;
  return GMRF_OK;
}

gmrf_status gmrf_profile_enable(int32_t on) {
  try {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    if (on) g_prof.clear();
    g_prof.on = on != 0;
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_profile_summary(int32_t max_entries, char* names, int32_t name_len,
                                 int64_t* launches, double* total_ms, double* flops, double* bytes,
                                 double peak_flops, double peak_bytes, double* roof_ms,
                                 int32_t* n_entries) {
  if (!n_entries) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (roof_ms && !(peak_flops > 0.0 && peak_bytes > 0.0))
    return fail(GMRF_INVALID_ARGUMENT, "roof_ms needs positive peaks");
  try {
    std::lock_guard<std::mutex> lk(g_prof.mu);
    // GMRF_PROFILE_DUMP=<path>: every recorded launch (name, flops, bytes, ms) appended
    // there as one text line (a diagnostic of launch shapes against durations)
    const char* dump_path = std::getenv("GMRF_PROFILE_DUMP");
    std::FILE* dump = dump_path ? std::fopen(dump_path, "a") : nullptr;
    std::vector<std::string> keys;
    std::vector<int64_t> cnt;
    std::vector<double> ms, fl, by, rf;
    for (auto& r : g_prof.recs) {
      check(hipEventSynchronize(r.b), "hipEventSynchronize");
      float t = 0.0f;
      check(hipEventElapsedTime(&t, r.a, r.b), "hipEventElapsedTime");
      size_t k = std::find(keys.begin(), keys.end(), std::string(r.name)) - keys.begin();
      if (k == keys.size()) {
        keys.push_back(r.name);
        cnt.push_back(0);
        ms.push_back(0);
        fl.push_back(0);
        by.push_back(0);
        rf.push_back(0);
      }
      if (dump) std::fprintf(dump, "%s %.6e %.6e %.6f\n", r.name, r.flops, r.bytes, (double)t);
      cnt[k]++;
      ms[k] += t;
      fl[k] += r.flops;
      by[k] += r.bytes;
      if (roof_ms) rf[k] += 1e3 * std::max(r.flops / peak_flops, r.bytes / peak_bytes);
    }
    *n_entries = (int32_t)keys.size();
    for (int32_t k = 0; k < (int32_t)keys.size() && k < max_entries; ++k) {
      if (names && name_len > 0) {
        std::strncpy(names + (size_t)k * name_len, keys[k].c_str(), name_len - 1);
        names[(size_t)k * name_len + name_len - 1] = 0;
      }
      if (launches) launches[k] = cnt[k];
      if (total_ms) total_ms[k] = ms[k];
      if (flops) flops[k] = fl[k];
      if (bytes) bytes[k] = by[k];
      if (roof_ms) roof_ms[k] = rf[k];
    }
    if (dump) std::fclose(dump);
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
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
