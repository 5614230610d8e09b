// Internal runtime structures shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <iterator>
#include <map>
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
hipError_t launch_fl_assemble(const DevPlan&, const int32_t*, int, int, int, bool, const double*, int64_t, double*,
                              double*, double*,
                              hipStream_t);
hipError_t launch_fl_diag(const DevPlan&, const int32_t*, int, int, int, double*, double*,
                          hipStream_t);
hipError_t launch_fl_trsm(const DevPlan&, const int32_t*, int, int, int, bool, double*, const double*,
                          hipStream_t);
hipError_t launch_fl_update(const DevPlan&, const int32_t*, int, int, int, bool, double*, double*, double*,
                            hipStream_t);
hipError_t launch_sl_gather_lhat(const DevPlan&, const int32_t*, int, const int32_t*, int, int,
                                const double*, const double*, double*, const double*, const double*, double*,
                                hipStream_t);
hipError_t launch_sl_lhat(const DevPlan&, const int32_t*, int, int, int, const double*,
                          const double*, double*, hipStream_t);
hipError_t launch_sl_gemm(const DevPlan&, const int32_t*, int, int, int, int, bool, double*, const double*,
                          const double*, hipStream_t);
hipError_t launch_sl_pp_lhat(const DevPlan&, const int32_t*, int, const int32_t*, int, int, int, bool,
                             const double*, const double*, double*, double*, hipStream_t);
hipError_t launch_sl_pp_part(const DevPlan&, const int32_t*, int, int, int, const double*,
                             const double*, const double*, double*, hipStream_t);
hipError_t launch_sl_gemm_part(const DevPlan&, const int32_t*, int, int, int, const double*, const double*,
                               double*, hipStream_t);
hipError_t launch_sl_gemm_sum(const DevPlan&, const int32_t*, int, int, int, double*, const double*,
                              hipStream_t);
hipError_t launch_sl_pp_sum(const DevPlan&, const int32_t*, int, int, int, double*, const double*,
                            hipStream_t);
hipError_t launch_logdet(const DevPlan&, int, const double*, double*, uint32_t*, double*, int32_t*,
                         hipStream_t);
hipError_t launch_logdet_seg(const DevPlan&, int, const double*, const int64_t*, const int32_t*, int,
                             double*, int32_t*, hipStream_t);
hipError_t launch_gather_perm(const DevPlan&, int, const double*, int64_t, double*, hipStream_t,
                              int period);
hipError_t launch_scatter_perm(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_diag_scale(const DevPlan&, int, int, const double*, double*, hipStream_t);
hipError_t launch_solve_fwd(const DevPlan&, const int32_t*, int, int, int, int, int, const double*, double*,
                            double*, hipStream_t);
hipError_t launch_solve_bwd(const DevPlan&, const int32_t*, int, int, int, int, int, const double*, double*,
                            hipStream_t);
hipError_t launch_solve_fwd_tri(const DevPlan&, const int32_t*, int, int, int, int, bool,
                                const double*, const double*, double*, double*, hipStream_t);
hipError_t launch_solve_fwd_rows(const DevPlan&, const int32_t*, int, int, int, int, const double*,
                                 const double*, double*, hipStream_t);
hipError_t launch_solve_bwd_rows(const DevPlan&, const int32_t*, int, int, int, const double*,
                                 const double*, double*, hipStream_t);
hipError_t launch_solve_bwd_tri(const DevPlan&, const int32_t*, int, int, int, int, bool,
                                const double*, const double*, double*, const double*, hipStream_t);
hipError_t launch_selinv_small(const DevPlan&, const int32_t*, int, int, int, int, const double*,
                               double*, double*, const double*, hipStream_t);
hipError_t launch_diag_out(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_pattern_out(const DevPlan&, int, const double*, double*, int64_t, hipStream_t);
hipError_t launch_mma_probe(const double*, const double*, int, double*, hipStream_t);
}  // namespace gmrf

using namespace gmrf;

namespace gmrf_rt {

extern thread_local std::string g_err;
extern thread_local int g_device;

inline gmrf_status fail(gmrf_status st, const std::string& msg) {
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

inline gmrf_status from_exception() {
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

// Optional per-launch timing: HIP events around each launch on the launching
// stream, tagged with the launch's algorithmic flops and bytes.
struct ProfRec {
  const char* name;
  hipEvent_t a, b;
  double flops, bytes;
};
struct Profiler {
  bool on = false;
  std::mutex mu;
  std::vector<ProfRec> recs;
  std::vector<hipEvent_t> pool;
  hipEvent_t get() {
    if (!pool.empty()) {
      hipEvent_t e = pool.back();
      pool.pop_back();
      return e;
    }
    hipEvent_t e;
    check(hipEventCreate(&e), "hipEventCreate");
    return e;
  }
  void clear() {
    for (auto& r : recs) {
      pool.push_back(r.a);
      pool.push_back(r.b);
    }
    recs.clear();
  }
};
extern Profiler g_prof;

// Run `launch` (returning hipError_t) with profiling when enabled.
template <typename F>
inline void prof_launch(const char* name, double flops, double bytes, hipStream_t st, F launch) {
  if (!g_prof.on) {
    check(launch(), name);
    return;
  }
  std::lock_guard<std::mutex> lk(g_prof.mu);
  ProfRec r{name, g_prof.get(), g_prof.get(), flops, bytes};
  check(hipEventRecord(r.a, st), "hipEventRecord");
  check(launch(), name);
  check(hipEventRecord(r.b, st), "hipEventRecord");
  g_prof.recs.push_back(r);
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

// Pinned host staging, grown on demand: every copy between the host and the device
// goes through memory of the handle's own (no copies from or into pageable memory,
// which the runtime pins or stages per call, with many threads copying at once).
struct PinnedBuf {
  char* p = nullptr;
  size_t n = 0;
  // callers ensure no copy into or out of the buffer is in flight
  char* get(size_t bytes) {
    if (bytes > n) {
      if (p) (void)hipHostFree(p);
      p = nullptr;
      n = 0;
      check(hipHostMalloc((void**)&p, std::max<size_t>(bytes, 64), hipHostMallocDefault), "hipHostMalloc");
      n = std::max<size_t>(bytes, 64);
    }
    return p;
  }
  PinnedBuf() = default;
  PinnedBuf(const PinnedBuf&) = delete;
  PinnedBuf& operator=(const PinnedBuf&) = delete;
  ~PinnedBuf() {
    if (p) (void)hipHostFree(p);
  }
};

struct LevelInfo {
  int64_t sb, se;     // small fronts [sb, se) in small list, ordered by size class
  int64_t cls_b[kSmallClasses + 1];  // class q (m <= kClassM[q]) covers [cls_b[q], cls_b[q + 1])
  double fl_cls[kSmallClasses] = {}, sel_cls[kSmallClasses] = {}, by_cls[kSmallClasses] = {};
  int32_t maxw_cls[kSmallClasses] = {};  // widest supernode of each small class
  int64_t lb, le;     // large fronts [lb, le) in large list
  int64_t ab, ae;     // all fronts [ab, ae) in all list
  int32_t maxm_small, maxm_large, maxm_all, maxw_large;
  int32_t maxw_small = 0;  // widest small supernode of the level
  // algorithmic work of the level's launches (per matrix)
  double fl_small = 0, fl_large = 0, sel_small = 0, sel_large = 0, sol_small = 0, sol_large = 0;
  double by_small = 0, by_large = 0;
  // large-front task ranges (triples in SymData::d_tasks)
  struct Range {
    int64_t off = 0, cnt = 0;
    double flops = 0, bytes = 0;
  };
  Range asm_t, gather;
  int64_t ain_first = 0;  // the level's first assembly tile in the input-entry grouping
  Range frow, brow;  // solve: rows below the large supernodes (forward, backward)
  bool sol_split = false;  // tall fronts: the rows below are separate row-parallel launches
  double sol_tri = 0;  // solve work of the large supernodes' own blocks (per matrix)
  std::vector<Range> fdiag, ftrsm, fupd;   // per factor panel step
  std::vector<Range> lhat, sgemm, spp;     // per selinv panel step (from the right)
  std::vector<Range> sppart;  // tall fronts: Sigma[P,P] partial products (empty: single launch)
  std::vector<Range> sgpart;  // Sigma[B,P] split over K (used when a launch has few row blocks)
};

// First-fit allocator over one workspace, in doubles; freed ranges coalesce.
class FirstFit {
 public:
  int64_t alloc(int64_t len) {
    len = (len + 7) & ~int64_t(7);
    for (auto it = free_.begin(); it != free_.end(); ++it)
      if (it->second >= len) {
        const int64_t off = it->first, rest = it->second - len;
        free_.erase(it);
        if (rest > 0) free_[off + len] = rest;
        return off;
      }
    if (!free_.empty()) {
      auto last = std::prev(free_.end());
      if (last->first + last->second == top_) {
        const int64_t off = last->first;
        free_.erase(last);
        top_ = off + len;
        return off;
      }
    }
    const int64_t off = top_;
    top_ += len;
    return off;
  }
  void release(int64_t off, int64_t len) {
    len = (len + 7) & ~int64_t(7);
    auto it = free_.emplace(off, len).first;
    auto nx = std::next(it);
    if (nx != free_.end() && off + len == nx->first) {
      it->second += nx->second;
      free_.erase(nx);
    }
    if (it != free_.begin()) {
      auto pv = std::prev(it);
      if (pv->first + pv->second == it->first) {
        pv->second += it->second;
        free_.erase(it);
      }
    }
  }
  int64_t size() const { return top_; }

 private:
  std::map<int64_t, int64_t> free_;
  int64_t top_ = 0;
};

// Workspace offsets pooled by lifetime over the level schedule.
//  - factorisation update of s (r x r): written at level(s), read at
//    level(parent); freed once the parent's level has run;
//  - Sigma_RR of a small front s (r x r, selected inversion, levels walked from
//    the root): written at level(s), read by its children; freed after the
//    lowest child level (its own level for a leaf); shares the update buffer,
//    whose size is the larger of the two peaks;
//  - dense Sigma front of a large front s (m x m): lifetime as Sigma_RR;
//  - L_hat panel staging of a large front (m x 64): its own level only.
struct Pools {
  std::vector<int64_t> u_fac, u_sel, w_off, t_off, ts_off;
  int64_t u_size = 1, w_size = 1, t_size = 1, ts_size = 1;
};

inline Pools plan_pools(const Symbolic& S) {
  const int64_t ns = S.nsuper;
  Pools P;
  P.u_fac.assign(ns, 0);
  P.u_sel.assign(ns, 0);
  P.w_off.assign(ns, -1);
  P.t_off.assign(ns, -1);
  P.ts_off.assign(ns, -1);
  auto r_of = [&](int64_t s) { return S.sn_rows_ptr[s + 1] - S.sn_rows_ptr[s]; };
  auto m_of = [&](int64_t s) { return S.sn_start[s + 1] - S.sn_start[s] + r_of(s); };
  auto large = [&](int64_t s) { return m_of(s) > kSmallFront; };
  std::vector<int32_t> lev(ns, 0);
  for (int l = 0; l < S.nlevels; ++l)
    for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) lev[S.level_sn[q]] = l;
  FirstFit uf;
  for (int l = 0; l < S.nlevels; ++l) {
    for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int64_t s = S.level_sn[q], r = r_of(s);
      if (r > 0) P.u_fac[s] = uf.alloc(r * r);
    }
    for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int64_t s = S.level_sn[q];
      for (int64_t cq = S.child_ptr[s]; cq < S.child_ptr[s + 1]; ++cq) {
        const int64_t c = S.child_idx[cq], rc = r_of(c);
        if (rc > 0) uf.release(P.u_fac[c], rc * rc);
      }
    }
  }
  std::vector<std::vector<int64_t>> free_after(S.nlevels);
  for (int64_t s = 0; s < ns; ++s) {
    int fl = lev[s];
    for (int64_t cq = S.child_ptr[s]; cq < S.child_ptr[s + 1]; ++cq)
      fl = std::min(fl, lev[S.child_idx[cq]]);
    free_after[fl].push_back(s);
  }
  FirstFit us, wf;
  int64_t tmax = 0, tsmax = 0;
  for (int l = S.nlevels - 1; l >= 0; --l) {
    int64_t tl = 0, tsl = 0;
    for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
      const int64_t s = S.level_sn[q], r = r_of(s), m = m_of(s);
      if (large(s)) {
        P.w_off[s] = wf.alloc(m * m);
        P.t_off[s] = tl;
        tl += stage_size(m);
        P.ts_off[s] = tsl;
        tsl += (r + kSolvePart - 1) / kSolvePart * (m - r);
      } else if (r > 0) {
        P.u_sel[s] = us.alloc(r * r);
      }
    }
    tmax = std::max(tmax, tl);
    tsmax = std::max(tsmax, tsl);
    for (int64_t s : free_after[l]) {
      const int64_t r = r_of(s), m = m_of(s);
      if (large(s))
        wf.release(P.w_off[s], m * m);
      else if (r > 0)
        us.release(P.u_sel[s], r * r);
    }
  }
  P.u_size = std::max<int64_t>({uf.size(), us.size(), 1});
  P.w_size = std::max<int64_t>(wf.size(), 1);
  P.t_size = std::max<int64_t>(tmax, 1);
  P.ts_size = std::max<int64_t>(tsmax, 1);
  return P;
}

}  // namespace gmrf_rt

using namespace gmrf_rt;

// Launch schedule of a pattern for one batch size (matrices per launch): the
// autotuner's choice (gmrf_batch_autotune), the defaults otherwise.
//  - merge_fac / merge_sel, per level: bit c set runs small-front class c in one
//    launch with class c + 1 (the larger class's kernel, which handles any smaller
//    front): fewer launches for more padding;
//  - fuse_below: launches of fewer workgroups also factor the next diagonal block;
//  - split_max: Sigma[B,P] launches of at most this many row tiles split K;
//  - xcd_max: Sigma[B,P] launches with at most this many tasks per matrix keep a
//    matrix's tiles on one XCD;
//  - pp_split_max: Sigma[P,P] launches of at most this many tasks (fronts x
//    matrices) run each task as three 32 x 32 workgroups;
//  - trsm32_max / gemm32_max: panel-solve / Sigma[B,P] launches of at most this many
//    64 x 64 tiles (tiles x matrices) run each tile as four 32 x 32 workgroups;
//  - stagger (>= 0): the batch runs as two pipelines of half the matrices each
//    (factorisation, log determinants, selected inversion), the second on a stream of
//    its own that starts once the first has enqueued the factorisation of level
//    `stagger`, so that its lower levels (thousands of small fronts) fill the device
//    while the first walks its upper levels (a few large fronts, launches of few
//    workgroups). -1: one pipeline of all matrices.
struct Schedule {
  std::vector<uint8_t> merge_fac, merge_sel;
  int64_t fuse_below = kFuseDiagBelow;
  int64_t split_max = kGemmSplitMaxTasks;
  int32_t xcd_max = kXcdMaxTasks;
  int32_t stagger = -1;
  int64_t pp_split_max = kPpSplitMaxTasks;
  int64_t trsm32_max = kTrsm32MaxTasks;
  int64_t gemm32_max = kGemm32MaxTasks;
  uint8_t fac_mask(size_t l) const { return l < merge_fac.size() ? merge_fac[l] : 0; }
  uint8_t sel_mask(size_t l) const { return l < merge_sel.size() ? merge_sel[l] : 0; }
};

// Small-front launches of one level under a class-merge mask: f(c, e, hi) for each
// group of classes [c, e] holding fronts, hi the largest class with fronts.
template <typename F>
inline void small_groups(const int64_t* cls_b, uint8_t mask, F f) {
  int c = 0;
  while (c < kSmallClasses) {
    int e = c;
    while (e + 1 < kSmallClasses && ((mask >> e) & 1)) ++e;
    int hi = -1;
    for (int q = c; q <= e; ++q)
      if (cls_b[q + 1] > cls_b[q]) hi = q;
    if (hi >= 0) f(c, e, hi);
    c = e + 1;
  }
}

// Symbolic analysis plus its device plan; shared by batches and solvers.
struct SymData {
  Symbolic S;
  Pools pools;  // lifetime-pooled workspace offsets (host copy)
  int device = 0;
  DevArena arena;
  DevPlan plan{};
  std::vector<LevelInfo> levels;
  int32_t* d_small = nullptr;
  int32_t* d_large = nullptr;
  int32_t* d_all = nullptr;
  int32_t* d_tasks = nullptr;  // large-front task triples
  int order_req = 0;
  int64_t dense_n = -1;        // >= 0: the full n x n pattern of a dense input (not cached)
  std::vector<int64_t> given;  // given permutation (cache key)
  std::mutex dev_mu;
  bool dev_built = false;
  std::mutex sched_mu;
  std::map<int, Schedule> sched;  // by batch size
  bool sched_loaded = false;

  // The schedule for `count` matrices per launch (tuned, loaded from the tuning
  // file on first use, or the defaults).
  Schedule schedule(int count) {
    std::lock_guard<std::mutex> lk(sched_mu);
    if (!sched_loaded) {
      sched_loaded = true;
      load_schedules();
    }
    auto it = sched.find(count);
    return it != sched.end() ? it->second : Schedule{};
  }
  void set_schedule(int count, const Schedule& sc) {
    std::lock_guard<std::mutex> lk(sched_mu);
    sched_loaded = true;
    sched[count] = sc;
  }
  void load_schedules();  // gmrf_api.cpp

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
    plan.u_size = pools.u_size;
    plan.v_size = S.v_size;
    plan.sn_start = arena.upload(sn_start);
    plan.sn_r = arena.upload(sn_r);
    plan.sn_parent = arena.upload(sn_parent);
    std::vector<int64_t> loff(S.l_off.begin(), S.l_off.begin() + ns);
    std::vector<int64_t> voff(S.v_off.begin(), S.v_off.begin() + ns);
    plan.l_off = arena.upload(loff);
    plan.u_off = arena.upload(pools.u_fac);
    plan.u_off_sel = arena.upload(pools.u_sel);
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
      L.maxm_large = 0;
      L.maxw_large = 0;
      L.maxm_all = 0;
      std::vector<int32_t> cls[kSmallClasses];
      for (int64_t q = S.level_ptr[l]; q < S.level_ptr[l + 1]; ++q) {
        int64_t s = S.level_sn[q];
        int32_t m = (int32_t)((S.sn_start[s + 1] - S.sn_start[s]) + sn_r[s]);
        all.push_back((int32_t)s);
        L.maxm_all = std::max(L.maxm_all, m);
        double fl = 0, sel = 0, sol = 0;
        for (int64_t j = S.sn_start[s]; j < S.sn_start[s + 1]; ++j) {
          const double rj = (double)(S.colcount[j] - 1);
          fl += rj * (rj + 3.0);
          sel += 2.0 * rj * (rj + 1.0) + 1.0;
          sol += 4.0 * rj + 1.0;
        }
        const int64_t w = S.sn_start[s + 1] - S.sn_start[s];
        const double by = 8.0 * (double)m * (double)w + 12.0 * (double)(fa_ptr[s + 1] - fa_ptr[s]);
        if (m <= kSmallFront) {
          int c = 0;
          while (m > kClassM[c]) ++c;
          cls[c].push_back((int32_t)s);
          L.maxw_cls[c] = std::max(L.maxw_cls[c], (int32_t)w);
          L.maxw_small = std::max(L.maxw_small, (int32_t)w);
          L.sol_small += sol;
          L.maxm_small = std::max(L.maxm_small, m);
          L.fl_small += fl;
          L.sel_small += sel;
          L.by_small += by;
          L.fl_cls[c] += fl;
          L.sel_cls[c] += sel;
          L.by_cls[c] += by;
        } else {
          large.push_back((int32_t)s);
          L.sol_large += sol;
          L.sol_tri += 2.0 * (double)w * (double)w;
          L.maxm_large = std::max(L.maxm_large, m);
          L.maxw_large = std::max(L.maxw_large, (int32_t)w);
          L.fl_large += fl;
          L.sel_large += sel;
          L.by_large += by;
        }
      }
      for (int c = 0; c < kSmallClasses; ++c) {
        L.cls_b[c] = (int64_t)small.size();
        small.insert(small.end(), cls[c].begin(), cls[c].end());
      }
      L.cls_b[kSmallClasses] = (int64_t)small.size();
      L.se = (int64_t)small.size();
      L.le = (int64_t)large.size();
      L.ae = (int64_t)all.size();
    }
    d_small = arena.upload(small);
    d_large = arena.upload(large);
    d_all = arena.upload(all);
    build_large_plan(large, fa_ptr, fa_csc, fa_slot);
  }

  // Task lists, inverse child maps and Z offsets of the large fronts.
  void build_large_plan(const std::vector<int32_t>& large, const std::vector<int64_t>& fa_ptr,
                        const std::vector<int32_t>& fa_csc, const std::vector<int32_t>& fa_slot) {
    const int64_t ns = S.nsuper;
    auto mw = [&](int64_t s, int64_t& w, int64_t& r, int64_t& m) {
      w = S.sn_start[s + 1] - S.sn_start[s];
      r = S.sn_rows_ptr[s + 1] - S.sn_rows_ptr[s];
      m = w + r;
    };
    std::vector<int32_t> is_large(ns, 0);
    std::vector<int64_t> z_off(ns, -1);
    int64_t zsz = 0;
    for (int32_t s : large) {
      int64_t w, r, m;
      mw(s, w, r, m);
      is_large[s] = 1;
      z_off[s] = zsz;
      zsz += ((w + kTile - 1) / kTile) * kTile * kTile;
    }
    std::vector<int64_t> cinv_off(std::max<int64_t>((int64_t)S.child_idx.size(), 1), -1);
    std::vector<int32_t> cinv;
    for (int32_t pf : large) {
      int64_t w, r, m;
      mw(pf, w, r, m);
      for (int64_t q = S.child_ptr[pf]; q < S.child_ptr[pf + 1]; ++q) {
        const int64_t c = S.child_idx[q];
        cinv_off[q] = (int64_t)cinv.size();
        cinv.resize(cinv.size() + m, -1);
        int32_t* inv = cinv.data() + cinv_off[q];
        for (int64_t i = S.sn_rows_ptr[c]; i < S.sn_rows_ptr[c + 1]; ++i)
          inv[S.rel[i]] = (int32_t)(i - S.sn_rows_ptr[c]);
      }
    }
    plan.is_large = arena.upload(is_large);
    plan.w_off = arena.upload(pools.w_off);
    plan.w_size = pools.w_size;
    plan.t_off = arena.upload(pools.t_off);
    plan.t_size = pools.t_size;
    plan.ts_off = arena.upload(pools.ts_off);
    plan.ts_size = pools.ts_size;
    plan.z_off = arena.upload(z_off);
    plan.z_size = std::max<int64_t>(zsz, 1);
    plan.cinv_off = arena.upload(cinv_off);
    plan.cinv = arena.upload(cinv.empty() ? std::vector<int32_t>(1, -1) : cinv);
    // structural work per front (for the per-launch roofline accounting)
    auto struct_work = [&](int64_t s, double& fl, double& sel) {
      fl = sel = 0.0;
      for (int64_t j = S.sn_start[s]; j < S.sn_start[s + 1]; ++j) {
        const double rj = (double)(S.colcount[j] - 1);
        fl += rj * (rj + 3.0);
        sel += 2.0 * rj * (rj + 1.0) + 1.0;
      }
    };
    std::vector<int32_t> tk;
    auto add = [&](int64_t s, int64_t a, int64_t b) {
      tk.push_back((int32_t)s);
      tk.push_back((int32_t)a);
      tk.push_back((int32_t)b);
    };
    auto ceil64 = [](int64_t x) { return (x + kTile - 1) / kTile; };
    std::vector<int32_t> ain_ptr(1, 0), ain_pos, ain_csc;
    for (LevelInfo& L : levels) {
      std::vector<int64_t> fronts;
      for (int64_t q = L.lb; q < L.le; ++q) fronts.push_back(large[q]);
      if (fronts.empty()) continue;
      int64_t maxsteps = 0;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        maxsteps = std::max(maxsteps, ceil64(w));
      }
      L.fdiag.resize(maxsteps);
      L.ftrsm.resize(maxsteps);
      L.fupd.resize(maxsteps);
      L.lhat.resize(maxsteps);
      L.sgemm.resize(maxsteps);
      L.spp.resize(maxsteps);
      L.sppart.resize(maxsteps);
      L.sgpart.resize(maxsteps);
      // dense work shares of each front per launch kind and step
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        double F, G;
        struct_work(s, F, G);
        double Dfac = 0, Dsel = 0;
        std::vector<double> dd(maxsteps, 0), dt(maxsteps, 0), du(maxsteps, 0), dl(maxsteps, 0),
            dg(maxsteps, 0), dq(maxsteps, 0);
        // algorithmic bytes of each launch kind and step: every operand read and every
        // result written once (triangular blocks as triangles)
        const int64_t np = (w + kTile - 1) / kTile;
        for (int64_t k = 0; k < np; ++k) {
          const double c0 = (double)(k * kTile), nb = (double)std::min<int64_t>(kTile, w - k * kTile);
          const double T = (double)m - c0 - nb;
          dd[k] = nb * nb * nb / 3.0;
          dt[k] = nb * nb * T;
          du[k] = nb * T * (T + 1.0) / 2.0;
          Dfac += dd[k] + dt[k] + du[k];
          L.fdiag[k].bytes += 8.0 * (nb * (nb + 1.0) + 64.0 * nb);          // block in and out, Z out
          L.ftrsm[k].bytes += 8.0 * (2.0 * T * nb + 64.0 * nb);            // A21 in, Z in, L21 out
          L.fupd[k].bytes += 8.0 * (T * (T + 1.0) + T * nb);               // trailing lower in and out, panel in
          // selinv step k handles panel np - 1 - k
          const int64_t pk = np - 1 - k, sc0 = pk * kTile, sc1 = std::min<int64_t>(w, sc0 + kTile);
          const double nbs = (double)(sc1 - sc0), nB = (double)(m - sc1);
          dl[k] = nbs * nbs * nB / 2.0;
          dg[k] = nB * nB * nbs;
          dq[k] = nbs * nbs * nB + nbs * nbs * nbs / 6.0;
          Dsel += dl[k] + dg[k] + dq[k];
          L.lhat[k].bytes += 8.0 * (2.0 * nB * nbs + 64.0 * nbs);                   // L[B,P], Z in, L_hat out
          L.sgemm[k].bytes += 8.0 * (nB * (nB + 1.0) / 2.0 + 2.0 * nB * nbs);      // Sigma[B,B] lower, L_hat in, Sigma[B,P] out
          L.spp[k].bytes += 8.0 * (64.0 * nbs + 2.0 * nB * nbs + nbs * (nbs + 1.0) / 2.0);
        }
        for (int64_t k = 0; k < maxsteps; ++k) {
          if (Dfac > 0) {
            L.fdiag[k].flops += F * dd[k] / Dfac;
            L.ftrsm[k].flops += F * dt[k] / Dfac;
            L.fupd[k].flops += F * du[k] / Dfac;
          }
          if (Dsel > 0) {
            L.lhat[k].flops += G * dl[k] / Dsel;
            L.sgemm[k].flops += G * dg[k] / Dsel;
            L.spp[k].flops += G * dq[k] / Dsel;
          }
        }
        L.asm_t.bytes += 8.0 * (double)(m * w) + 12.0 * (double)(fa_ptr[s + 1] - fa_ptr[s]);
      }
      L.asm_t.off = (int64_t)tk.size() / 3;
      L.ain_first = (int64_t)ain_ptr.size() - 1;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        const int64_t nt = ceil64(m);
        // the front's input entries by tile, tiles in task order (J, then I >= J)
        auto tile_of = [&](int64_t I, int64_t J) { return J * nt - J * (J - 1) / 2 + (I - J); };
        std::vector<std::vector<std::pair<int32_t, int32_t>>> by_tile((size_t)(nt * (nt + 1) / 2));
        for (int64_t q = fa_ptr[s]; q < fa_ptr[s + 1]; ++q) {
          const int64_t slot = fa_slot[q], j = slot / m, i = slot % m;
          by_tile[(size_t)tile_of(i / kTile, j / kTile)].push_back(
              {(int32_t)((j % kTile) * (kTile + 1) + i % kTile), fa_csc[q]});
        }
        for (int64_t J = 0; J < nt; ++J)
          for (int64_t I = J; I < nt; ++I) {
            add(s, I, J);
            for (const auto& e : by_tile[(size_t)tile_of(I, J)]) {
              ain_pos.push_back(e.first);
              ain_csc.push_back(e.second);
            }
            if (ain_pos.size() >= ((size_t)1 << 31)) throw std::runtime_error("too many input entries");
            ain_ptr.push_back((int32_t)ain_pos.size());
          }
      }
      L.asm_t.cnt = (int64_t)tk.size() / 3 - L.asm_t.off;
      for (int64_t k = 0; k < maxsteps; ++k) {
        L.fdiag[k].off = (int64_t)tk.size() / 3;
        for (int64_t s : fronts) {
          int64_t w, r, m;
          mw(s, w, r, m);
          if (w > k * kTile) add(s, 0, 0);
        }
        L.fdiag[k].cnt = (int64_t)tk.size() / 3 - L.fdiag[k].off;
        L.ftrsm[k].off = (int64_t)tk.size() / 3;
        for (int64_t s : fronts) {
          int64_t w, r, m;
          mw(s, w, r, m);
          if (w <= k * kTile) continue;
          const int64_t c0 = k * kTile, nb = std::min<int64_t>(kTile, w - c0);
          for (int64_t I = 0; I < ceil64(m - c0 - nb); ++I) add(s, I, 0);
        }
        L.ftrsm[k].cnt = (int64_t)tk.size() / 3 - L.ftrsm[k].off;
        L.fupd[k].off = (int64_t)tk.size() / 3;
        for (int64_t s : fronts) {
          int64_t w, r, m;
          mw(s, w, r, m);
          if (w <= k * kTile) continue;
          const int64_t c0 = k * kTile, nb = std::min<int64_t>(kTile, w - c0), T = m - c0 - nb;
          const int64_t nt = ceil64(T);
          for (int64_t J = 0; J < nt; ++J)
            for (int64_t I = J; I < nt; ++I) add(s, I, J);
        }
        L.fupd[k].cnt = (int64_t)tk.size() / 3 - L.fupd[k].off;
      }
      L.gather.off = (int64_t)tk.size() / 3;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        L.gather.bytes += 8.0 * (double)r * (double)(r + 1);  // parent's Sigma_RR lower in, W lower out
        const int64_t nt = ceil64(r);
        for (int64_t J = 0; J < nt; ++J)
          for (int64_t I = J; I < nt; ++I) add(s, I, J);
      }
      L.gather.cnt = (int64_t)tk.size() / 3 - L.gather.off;
      int64_t rmax = 0;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        rmax = std::max(rmax, r);
      }
      L.sol_split = rmax >= kSolveSplitRows;
      L.frow.off = (int64_t)tk.size() / 3;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        for (int64_t c = 0; c < ceil64(r); ++c) add(s, c, 0);
      }
      L.frow.cnt = (int64_t)tk.size() / 3 - L.frow.off;
      L.brow.off = (int64_t)tk.size() / 3;
      for (int64_t s : fronts) {
        int64_t w, r, m;
        mw(s, w, r, m);
        const int64_t parts = (r + kSolvePart - 1) / kSolvePart;
        for (int64_t p = 0; p < parts; ++p)
          for (int64_t c = 0; c < ceil64(w); ++c) add(s, c, p);
      }
      L.brow.cnt = (int64_t)tk.size() / 3 - L.brow.off;
      for (int64_t k = 0; k < maxsteps; ++k) {
        auto rows_of = [&](int64_t s, int64_t& nB) {
          int64_t w, r, m;
          mw(s, w, r, m);
          const int64_t np = (w + kTile - 1) / kTile;
          if (k >= np) return false;
          const int64_t c1 = std::min<int64_t>(w, (np - 1 - k) * kTile + kTile);
          nB = m - c1;
          return true;
        };
        L.lhat[k].off = (int64_t)tk.size() / 3;
        for (int64_t s : fronts) {
          int64_t nB;
          if (!rows_of(s, nB)) continue;
          for (int64_t rb = 0; rb < ceil64(nB); ++rb) add(s, rb, 0);
        }
        L.lhat[k].cnt = (int64_t)tk.size() / 3 - L.lhat[k].off;
        L.sgemm[k].off = (int64_t)tk.size() / 3;
        for (int64_t s : fronts) {
          int64_t nB;
          if (!rows_of(s, nB)) continue;
          for (int64_t rb = 0; rb < ceil64(nB); ++rb) add(s, rb, 0);
        }
        L.sgemm[k].cnt = (int64_t)tk.size() / 3 - L.sgemm[k].off;
        L.sgpart[k].off = (int64_t)tk.size() / 3;
        bool longk = false;
        for (int64_t s : fronts) {
          int64_t nB;
          if (rows_of(s, nB)) longk |= nB > kGemmSplitK;
        }
        if (longk)
          for (int64_t s : fronts) {
            int64_t nB;
            if (!rows_of(s, nB)) continue;
            const int64_t parts = std::max<int64_t>(1, (nB + kGemmSplitK - 1) / kGemmSplitK);
            for (int64_t rb = 0; rb < ceil64(nB); ++rb)
              for (int64_t q = 0; q < parts; ++q) add(s, rb, q);
          }
        L.sgpart[k].cnt = (int64_t)tk.size() / 3 - L.sgpart[k].off;
        L.sgpart[k].flops = L.sgemm[k].flops;
        L.sgpart[k].bytes = L.sgemm[k].bytes;
        L.spp[k].off = (int64_t)tk.size() / 3;
        bool tall = false;
        for (int64_t s : fronts) {
          int64_t nB;
          if (!rows_of(s, nB)) continue;
          add(s, 0, 0);
          tall |= nB > kSelSplitK;
        }
        L.spp[k].cnt = (int64_t)tk.size() / 3 - L.spp[k].off;
        L.sppart[k].off = (int64_t)tk.size() / 3;
        if (tall)
          for (int64_t s : fronts) {
            int64_t nB;
            if (!rows_of(s, nB)) continue;
            for (int64_t q = 0; q < std::max<int64_t>(1, (nB + kSelSplitK - 1) / kSelSplitK); ++q) add(s, q, 0);
          }
        L.sppart[k].cnt = (int64_t)tk.size() / 3 - L.sppart[k].off;
        L.sppart[k].flops = L.spp[k].flops;
        L.sppart[k].bytes = L.spp[k].bytes;
      }
    }
    d_tasks = arena.upload(tk.empty() ? std::vector<int32_t>(3, 0) : tk);
    plan.ain_ptr = arena.upload(ain_ptr);
    plan.ain_pos = arena.upload(ain_pos.empty() ? std::vector<int32_t>(1, 0) : ain_pos);
    plan.ain_csc = arena.upload(ain_csc.empty() ? std::vector<int32_t>(1, 0) : ain_csc);
  }
};

struct gmrf_symbolic {
  std::shared_ptr<SymData> d;
};


// Numeric workspace for up to `cap` matrices of one pattern.
struct gmrf_batch {
  std::shared_ptr<SymData> sym;
  int cap = 0;
  hipStream_t stream = nullptr;  // stream in use (own or caller's)
  hipStream_t own = nullptr;
  DevArena arena;
  double *L = nullptr, *U = nullptr, *S = nullptr, *T = nullptr, *X = nullptr, *V = nullptr,
         *W = nullptr, *Z = nullptr;
  double* d_logdet = nullptr;
  int32_t* d_status = nullptr;
  PinnedBuf pin;                  // host staging of the log-determinant and status read-back
  double* d_ldpart = nullptr;     // per-matrix chunk partials of log|D|
  uint32_t* d_ldcount = nullptr;  // per-matrix chunk counters (zero between launches)
  // forward/backward workspaces for solves with several right-hand sides per
  // matrix (grown on demand; one vector = n + v_size + ts_size doubles)
  double* mrhs = nullptr;
  int64_t mrhs_vecs = 0;
  // selected inversion on a second stream beside the main stream's solves of the
  // same factors (selinv_async / selinv_join); the solves stage their partial sums
  // in TS (ts_size per vector), the selected inversion in T
  hipStream_t aux = nullptr;
  hipEvent_t ev_fork = nullptr, ev_join = nullptr;
  double* TS = nullptr;
  bool sel_inflight = false;
  // the second pipeline of a staggered batch (Schedule::stagger): its own stream,
  // the event it starts from (recorded mid-factorisation on the main stream) and the
  // event that ends it
  hipStream_t aux2 = nullptr;
  hipEvent_t ev_mid = nullptr, ev_p2 = nullptr;
  bool p2_inflight = false;
  // autotuner: events around each level's small-front launches (a[l] before, b[l]
  // after; factorisation in fac, selected inversion in sel), recorded when set
  struct TuneMarks {
    std::vector<hipEvent_t> fac_a, fac_b, sel_a, sel_b;
  }* marks = nullptr;

  // GMRF_LEVEL_TIMES set: the time of every level of a factorisation / selected
  // inversion / solve, from events at the level boundaries, printed on stderr when the phase
  // ends (synchronises, so the model evaluates eagerly, without graph capture; a
  // diagnostic of where the time goes in the tree).
  struct LevelClock {
    std::vector<hipEvent_t> ev;
    bool on = false;
    static bool wanted() {
      static const bool want = std::getenv("GMRF_LEVEL_TIMES") != nullptr;
      return want;
    }
    void start(size_t nl, hipStream_t st) {
      on = wanted();
      if (!on) return;
      if (ev.size() < nl + 1) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        ev.assign(nl + 1, nullptr);
        for (auto& e : ev) check(hipEventCreate(&e), "hipEventCreate");
      }
      check(hipEventRecord(ev[0], st), "hipEventRecord");
    }
    void mark(size_t i, hipStream_t st) {
      if (on) check(hipEventRecord(ev[i], st), "hipEventRecord");
    }
    // interval i covers level i (top_down: level nl - 1 - i)
    void report(const char* phase, size_t nl, int count, const std::vector<LevelInfo>& levels, bool top_down) {
      if (!on) return;
      check(hipEventSynchronize(ev[nl]), "hipEventSynchronize");
      for (size_t i = 0; i < nl; ++i) {
        const size_t l = top_down ? nl - 1 - i : i;
        float ms = 0.0f;
        check(hipEventElapsedTime(&ms, ev[i], ev[i + 1]), "hipEventElapsedTime");
        std::fprintf(stderr, "[gmrf] %s level %zu (%d matrices, %lld large fronts): %.4f ms\n", phase, l, count,
                     (long long)(levels[l].le - levels[l].lb), ms);
      }
    }
    ~LevelClock() {
      for (auto& e : ev) (void)hipEventDestroy(e);
    }
  } lclock, sclock;  // sclock: the level-scheduled solves (forward, then backward)

  gmrf_batch(std::shared_ptr<SymData> s, int capacity) : sym(std::move(s)), cap(capacity) {
    check(hipSetDevice(sym->device), "hipSetDevice");
    sym->ensure_device();
    check(hipStreamCreateWithFlags(&own, hipStreamNonBlocking), "hipStreamCreate");
    stream = own;
    const Symbolic& Sy = sym->S;
    L = arena.alloc<double>((size_t)cap * Sy.l_size);
    U = arena.alloc<double>((size_t)cap * sym->plan.u_size);
    S = arena.alloc<double>((size_t)cap * Sy.l_size);
    T = arena.alloc<double>((size_t)cap * sym->plan.t_size);
    TS = arena.alloc<double>((size_t)cap * sym->plan.ts_size);
    X = arena.alloc<double>((size_t)cap * Sy.n);
    V = arena.alloc<double>((size_t)cap * std::max<int64_t>(Sy.v_size, 1));
    W = arena.alloc<double>((size_t)cap * sym->plan.w_size);
    Z = arena.alloc<double>((size_t)cap * sym->plan.z_size);
    d_logdet = arena.alloc<double>(cap);
    d_status = arena.alloc<int32_t>(cap);
    d_ldpart = arena.alloc<double>((size_t)cap * 2 * 16);
    d_ldcount = arena.alloc<uint32_t>(cap);
    // the counters must be zero before the first k_logdet: cleared on the batch's own
    // (non-blocking) stream and waited for here. A clear on the null stream is not
    // ordered before work on a non-blocking stream, and a fresh allocation may reuse
    // memory another handle has just freed, so the first log determinant of a batch
    // could otherwise count from a stale value (and never be written).
    check(hipMemsetAsync(d_ldcount, 0, (size_t)cap * sizeof(uint32_t), own), "hipMemsetAsync");
    check(hipStreamSynchronize(own), "hipStreamSynchronize");
  }
  ~gmrf_batch() {
    if (aux) {
      (void)hipStreamSynchronize(aux);
      (void)hipStreamDestroy(aux);
      (void)hipEventDestroy(ev_fork);
      (void)hipEventDestroy(ev_join);
    }
    if (aux2) {
      (void)hipStreamSynchronize(aux2);
      (void)hipStreamDestroy(aux2);
      (void)hipEventDestroy(ev_mid);
      (void)hipEventDestroy(ev_p2);
    }
    if (own) {
      (void)hipStreamSynchronize(own);
      (void)hipStreamDestroy(own);
    }
    if (mrhs) (void)hipFree(mrhs);
  }

  // Selected inversion of matrices [first, first + count) on the second stream,
  // after everything enqueued so far on the main stream; selinv_join makes the main
  // stream wait for it. Until then the main stream may solve with the same factors
  // (L and Z are only read) but must not refactorise.
  void selinv_async(int count, int first = 0) {
    selinv_join();
    if (!aux) {
      check(hipStreamCreateWithFlags(&aux, hipStreamNonBlocking), "hipStreamCreate");
      check(hipEventCreateWithFlags(&ev_fork, hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&ev_join, hipEventDisableTiming), "hipEventCreate");
    }
    check(hipEventRecord(ev_fork, stream), "hipEventRecord");
    check(hipStreamWaitEvent(aux, ev_fork, 0), "hipStreamWaitEvent");
    hipStream_t main = stream;
    stream = aux;
    try {
      selinv(count, first);
    } catch (...) {
      stream = main;
      throw;
    }
    stream = main;
    check(hipEventRecord(ev_join, aux), "hipEventRecord");
    sel_inflight = true;
  }
  // The main stream waits for the work on the other streams: the selected inversion
  // of selinv_async and the second pipeline of a staggered batch.
  void selinv_join() {
    if (sel_inflight) {
      check(hipStreamWaitEvent(stream, ev_join, 0), "hipStreamWaitEvent");
      sel_inflight = false;
    }
    if (p2_inflight) {
      check(hipStreamWaitEvent(stream, ev_p2, 0), "hipStreamWaitEvent");
      p2_inflight = false;
    }
  }

  // Whether a batch of `count` matrices runs as two staggered pipelines, and from
  // which level (Schedule::stagger); never while level clocks or tuning marks
  // record, whose events serve one pipeline.
  // GMRF_STAGGER=<level> (read at each call; -1 for one pipeline) overrides the
  // schedule: a diagnostic for timing the levels against each other.
  int stagger_level(int count) const {
    if (count < 2 || marks || LevelClock::wanted()) return -1;
    const char* env = std::getenv("GMRF_STAGGER");
    const int lv = env ? std::atoi(env) : sym->schedule(count).stagger;
    return lv < (int)sym->levels.size() ? lv : -1;
  }

  // The second pipeline of a staggered batch: matrices [first, first + count) on
  // their own stream, from the point where the main stream's factorisation has
  // enqueued the marked level (factorize_levels with mark_level): factorisation,
  // log determinants and, with `inv`, selected inversion. selinv_join waits for it.
  // Must be enqueued before anything else joins (selinv_async joins first).
  void pipeline2_async(int count, const double* vals, int64_t ldv, int first, bool inv) {
    hipStream_t main = stream;
    stream = aux2;
    try {
      check(hipStreamWaitEvent(aux2, ev_mid, 0), "hipStreamWaitEvent");
      factorize_levels(count, vals, ldv, first, -1, false);
      logdet(count, first);
      if (inv) selinv(count, first, false);
      check(hipEventRecord(ev_p2, aux2), "hipEventRecord");
    } catch (...) {
      stream = main;
      throw;
    }
    stream = main;
    p2_inflight = true;
  }

  // Factorisation, log determinants and selected inversion of matrices [0, count)
  // as two staggered pipelines from level `split`: [0, c0) on the main stream,
  // [c0, count) beside it (c0 = ceil(count / 2)); the main stream waits for both.
  void factor_selinv_staggered(int count, const double* vals, int64_t ldv, int split) {
    const int c0 = (count + 1) / 2, c1 = count - c0;
    factorize_levels(c0, vals, ldv, 0, split);
    pipeline2_async(c1, vals + (int64_t)c0 * ldv, ldv, c0, true);
    logdet(c0, 0);
    selinv(c0, 0, false);
    selinv_join();
  }

  // Matrices [first, first + count) of the batch: the level loop, then the log
  // determinants and statuses.
  void factorize(int count, const double* vals, int64_t ldv, int first = 0) {
    factorize_levels(count, vals, ldv, first);
    logdet(count, first);
  }
  // The factors only (L, Z); logdet() completes the factorisation. A selected
  // inversion may be started in between (it reads L and Z only).
  // With mark_level >= 0 the event that starts a second pipeline (pipeline2_async)
  // is recorded once that level has been enqueued.
  void factorize_levels(int count, const double* vals, int64_t ldv, int first = 0, int mark_level = -1,
                        bool join = true) {
    if (join) selinv_join();
    if (mark_level >= 0 && !aux2) {
      check(hipStreamCreateWithFlags(&aux2, hipStreamNonBlocking), "hipStreamCreate");
      check(hipEventCreateWithFlags(&ev_mid, hipEventDisableTiming), "hipEventCreate");
      check(hipEventCreateWithFlags(&ev_p2, hipEventDisableTiming), "hipEventCreate");
    }
    const DevPlan& P = sym->plan;
    double* Lf = L + (int64_t)first * P.l_size;
    double* Uf = U + (int64_t)first * P.u_size;
    double* Zf = Z + (int64_t)first * P.z_size;
    const Schedule sc = sym->schedule(count);
    lclock.start(sym->levels.size(), stream);
    for (size_t li = 0; li < sym->levels.size(); ++li) {
      const LevelInfo& lv = sym->levels[li];
      if (marks) check(hipEventRecord(marks->fac_a[li], stream), "hipEventRecord");
      small_groups(lv.cls_b, sc.fac_mask(li), [&](int c, int e, int hi) {
        double fl = 0.0, by = 0.0;
        for (int q = c; q <= e; ++q) fl += lv.fl_cls[q], by += lv.by_cls[q];
        prof_launch("k_factor_small", fl * count, by * count, stream, [&] {
          return launch_factor_small(P, sym->d_small + lv.cls_b[c], (int)(lv.cls_b[e + 1] - lv.cls_b[c]),
                                     kClassM[hi], count, vals, ldv, Lf, Uf, stream);
        });
      });
      if (marks) check(hipEventRecord(marks->fac_b[li], stream), "hipEventRecord");
      if (lv.le > lv.lb) {
        const int32_t* tk = sym->d_tasks;
        // launches of few tiles also factor the next diagonal blocks (one launch
        // less per panel step); large launches keep their occupancy
        auto fuse = [&](int64_t tiles) { return tiles * count < sc.fuse_below; };
        bool diag_done = fuse(lv.asm_t.cnt);
        prof_launch("k_fl_assemble", diag_done && !lv.fdiag.empty() ? lv.fdiag[0].flops * count : 0.0,
                    (lv.asm_t.bytes + (diag_done && !lv.fdiag.empty() ? lv.fdiag[0].bytes : 0.0)) * count, stream,
                    [&] {
                      return launch_fl_assemble(P, tk + 3 * lv.asm_t.off, (int)lv.asm_t.cnt, (int)lv.ain_first,
                                                count, diag_done, vals, ldv, Lf, Uf, Zf, stream);
                    });
        for (size_t k = 0; k < lv.fdiag.size(); ++k) {
          if (!diag_done)
            prof_launch("k_fl_diag", lv.fdiag[k].flops * count, lv.fdiag[k].bytes * count, stream, [&] {
              return launch_fl_diag(P, tk + 3 * lv.fdiag[k].off, (int)lv.fdiag[k].cnt, (int)k, count,
                                    Lf, Zf, stream);
            });
          prof_launch("k_fl_trsm", lv.ftrsm[k].flops * count, lv.ftrsm[k].bytes * count, stream, [&] {
            return launch_fl_trsm(P, tk + 3 * lv.ftrsm[k].off, (int)lv.ftrsm[k].cnt, (int)k, count,
                                  lv.ftrsm[k].cnt * (int64_t)count <= sc.trsm32_max, Lf, Zf, stream);
          });
          const bool next = k + 1 < lv.fdiag.size();
          diag_done = next && fuse(lv.fupd[k].cnt);
          prof_launch("k_fl_update", (lv.fupd[k].flops + (diag_done ? lv.fdiag[k + 1].flops : 0.0)) * count,
                      (lv.fupd[k].bytes + (diag_done ? lv.fdiag[k + 1].bytes : 0.0)) * count, stream, [&] {
                        return launch_fl_update(P, tk + 3 * lv.fupd[k].off, (int)lv.fupd[k].cnt, (int)k, count,
                                                diag_done, Lf, Uf, Zf, stream);
                      });
        }
      }
      lclock.mark(li + 1, stream);
      if ((int)li == mark_level) check(hipEventRecord(ev_mid, stream), "hipEventRecord");
    }
    lclock.report("factorisation", sym->levels.size(), count, sym->levels, false);
  }
  void logdet(int count, int first = 0) {
    const DevPlan& P = sym->plan;
    const double* Lf = L + (int64_t)first * P.l_size;
    prof_launch("k_logdet", 0.0, 16.0 * P.n * count, stream, [&] {
      return launch_logdet(P, count, Lf, d_ldpart + (size_t)first * 2 * 16, d_ldcount + first,
                           d_logdet + first, d_status + first, stream);
    });
  }

  // Doubles of solve workspace per right-hand side (permuted vector, update
  // entries, split-K partial sums).
  int64_t solve_ws_per_vector() const {
    const DevPlan& P = sym->plan;
    return P.n + std::max<int64_t>(P.v_size, 1) + P.ts_size;
  }

  // Matrices [first, first + count), nr right-hand sides each: vector j of
  // matrix first + b is rhs + (b nr + j) ldr (rhs + ((b nr + j) % period) ldr when
  // period > 0: one set of vectors for every matrix), its solution x + (b nr + j) ldx.
  // ws (count nr solve_ws_per_vector() doubles), when given, holds the workspace
  // instead of the batch's own areas; nothing is allocated then.
  void solve(int count, const double* rhs, int64_t ldr, double* x, int64_t ldx, int first = 0,
             int nr = 1, int period = 0, double* ws = nullptr) {
    const DevPlan& P = sym->plan;
    const int64_t vsz = std::max<int64_t>(P.v_size, 1);
    double *Xf, *Vf, *Tf;
    if (ws) {
      const int64_t vecs = (int64_t)count * nr;
      Xf = ws;
      Vf = Xf + vecs * P.n;
      Tf = Vf + vecs * vsz;
    } else if (nr == 1) {
      Xf = X + (int64_t)first * P.n;
      Vf = V + (int64_t)first * vsz;
      Tf = TS + (int64_t)first * P.ts_size;
    } else {
      const int64_t vecs = (int64_t)count * nr;
      if (vecs > mrhs_vecs) {
        check(hipStreamSynchronize(stream), "sync");
        if (mrhs) check(hipFree(mrhs), "hipFree");
        mrhs = nullptr;
        mrhs_vecs = 0;
        check(hipMalloc(&mrhs, (size_t)vecs * (P.n + vsz + P.ts_size) * sizeof(double)), "hipMalloc");
        mrhs_vecs = vecs;
      }
      Xf = mrhs;
      Vf = Xf + vecs * P.n;
      Tf = Vf + vecs * vsz;
    }
    const double* Lf = L + (int64_t)first * P.l_size;
    const double* Zf = Z + (int64_t)first * P.z_size;
    const int32_t* tk = sym->d_tasks;
    check(launch_gather_perm(P, count * nr, rhs, ldr, Xf, stream, period), "gather");
    sclock.start(sym->levels.size(), stream);
    for (size_t li = 0; li < sym->levels.size(); ++li) {
      const LevelInfo& lv = sym->levels[li];
      if (lv.se > lv.sb)
        prof_launch("k_solve_fwd", 0.5 * lv.sol_small * count * nr, 2.0 * lv.sol_small * count * nr, stream, [&] {
          return launch_solve_fwd(P, sym->d_small + lv.sb, (int)(lv.se - lv.sb), lv.maxm_small, lv.maxw_small, count, nr,
                                  Lf, Xf, Vf, stream);
        });
      if (lv.le > lv.lb) {
        const bool sp = lv.sol_split;
        prof_launch("k_solve_fwd_tri", 0.5 * (sp ? lv.sol_tri : lv.sol_large) * count * nr,
                    2.0 * (sp ? lv.sol_tri : lv.sol_large) * count * nr, stream, [&] {
          return launch_solve_fwd_tri(P, sym->d_large + lv.lb, (int)(lv.le - lv.lb),
                                      sp ? lv.maxw_large : lv.maxm_large, count, nr, !sp, Lf, Zf, Xf, Vf,
                                      stream);
        });
        if (sp)
          prof_launch("k_solve_fwd_rows", 0.5 * (lv.sol_large - lv.sol_tri) * count * nr,
                      2.0 * (lv.sol_large - lv.sol_tri) * count * nr, stream, [&] {
            return launch_solve_fwd_rows(P, tk + 3 * lv.frow.off, (int)lv.frow.cnt, lv.maxw_large, count,
                                         nr, Lf, Xf, Vf, stream);
          });
      }
      sclock.mark(li + 1, stream);
    }
    sclock.report("forward solve", sym->levels.size(), count * nr, sym->levels, false);
    check(launch_diag_scale(P, count, nr, Lf, Xf, stream), "diag_scale");
    sclock.start(sym->levels.size(), stream);
    for (size_t l = sym->levels.size(); l-- > 0;) {
      const LevelInfo& lv = sym->levels[l];
      if (lv.le > lv.lb) {
        const bool sp = lv.sol_split;
        if (sp)
          prof_launch("k_solve_bwd_rows", 0.5 * (lv.sol_large - lv.sol_tri) * count * nr,
                      2.0 * (lv.sol_large - lv.sol_tri) * count * nr, stream, [&] {
            return launch_solve_bwd_rows(P, tk + 3 * lv.brow.off, (int)lv.brow.cnt, count, nr, Lf, Xf, Tf,
                                         stream);
          });
        prof_launch("k_solve_bwd_tri", 0.5 * (sp ? lv.sol_tri : lv.sol_large) * count * nr,
                    2.0 * (sp ? lv.sol_tri : lv.sol_large) * count * nr, stream, [&] {
          return launch_solve_bwd_tri(P, sym->d_large + lv.lb, (int)(lv.le - lv.lb),
                                      sp ? lv.maxw_large : lv.maxm_large, count, nr, !sp, Lf, Zf, Xf, Tf,
                                      stream);
        });
      }
      if (lv.se > lv.sb)
        prof_launch("k_solve_bwd", 0.5 * lv.sol_small * count * nr, 2.0 * lv.sol_small * count * nr, stream, [&] {
          return launch_solve_bwd(P, sym->d_small + lv.sb, (int)(lv.se - lv.sb), lv.maxm_small, lv.maxw_small, count, nr,
                                  Lf, Xf, stream);
        });
      sclock.mark(sym->levels.size() - l, stream);
    }
    sclock.report("backward solve", sym->levels.size(), count * nr, sym->levels, true);
    check(launch_scatter_perm(P, count * nr, Xf, x, ldx, stream), "scatter");
  }

  void selinv(int count, int first = 0, bool join = true) {
    if (join && stream != aux) selinv_join();
    const DevPlan& P = sym->plan;
    const double* Lf = L + (int64_t)first * P.l_size;
    double* Sf = S + (int64_t)first * P.l_size;
    double* Uf = U + (int64_t)first * P.u_size;
    double* Tf = T + (int64_t)first * P.t_size;
    double* Wf = W + (int64_t)first * P.w_size;
    const double* Zf = Z + (int64_t)first * P.z_size;
    const int32_t* tk = sym->d_tasks;
    const Schedule sc = sym->schedule(count);
    lclock.start(sym->levels.size(), stream);
    for (size_t l = sym->levels.size(); l-- > 0;) {
      const LevelInfo& lv = sym->levels[l];
      if (lv.le > lv.lb) {
        // L_hat of panel step k sits in buffer k % 2 of the staging area; the gather
        // launch forms the L_hat of step 0, the launch that closes step k that of k + 1
        const bool l0 = !lv.lhat.empty();
        prof_launch("k_sl_gather_lhat", l0 ? lv.lhat[0].flops * count : 0.0,
                    (lv.gather.bytes + (l0 ? lv.lhat[0].bytes : 0.0)) * count, stream, [&] {
          return launch_sl_gather_lhat(P, tk + 3 * lv.gather.off, (int)lv.gather.cnt,
                                       tk + 3 * (l0 ? lv.lhat[0].off : 0), l0 ? (int)lv.lhat[0].cnt : 0, count,
                                       Sf, Uf, Wf, Lf, Zf, Tf, stream);
        });
        for (size_t k = 0; k < lv.lhat.size(); ++k) {
          const bool next = k + 1 < lv.lhat.size() && lv.lhat[k + 1].cnt > 0;
          // few row blocks with a long K (the matrices of tall fronts, a small batch):
          // split K so that the product spreads over the device
          if (lv.sgpart[k].cnt > 0 && lv.sgemm[k].cnt * (int64_t)count <= sc.split_max) {
            prof_launch("k_sl_gemm_part", lv.sgpart[k].flops * count, lv.sgpart[k].bytes * count, stream, [&] {
              return launch_sl_gemm_part(P, tk + 3 * lv.sgpart[k].off, (int)lv.sgpart[k].cnt, (int)k, count,
                                         Sf, Wf, Tf, stream);
            });
            prof_launch("k_sl_gemm_sum", 0.0, 0.0, stream, [&] {
              return launch_sl_gemm_sum(P, tk + 3 * lv.sgemm[k].off, (int)lv.sgemm[k].cnt, (int)k, count,
                                        Sf, Tf, stream);
            });
          } else {
            prof_launch("k_sl_gemm", lv.sgemm[k].flops * count, lv.sgemm[k].bytes * count, stream, [&] {
              return launch_sl_gemm(P, tk + 3 * lv.sgemm[k].off, (int)lv.sgemm[k].cnt, (int)k, count, sc.xcd_max,
                                    lv.sgemm[k].cnt * (int64_t)count <= sc.gemm32_max, Sf, Wf, Tf, stream);
            });
          }
          if (lv.sppart[k].cnt > 0) {
            prof_launch("k_sl_pp_part", lv.sppart[k].flops * count, lv.sppart[k].bytes * count, stream, [&] {
              return launch_sl_pp_part(P, tk + 3 * lv.sppart[k].off, (int)lv.sppart[k].cnt, (int)k, count,
                                       Lf, Zf, Sf, Tf, stream);
            });
            prof_launch("k_sl_pp_sum", 0.0, 0.0, stream, [&] {
              return launch_sl_pp_sum(P, tk + 3 * lv.spp[k].off, (int)lv.spp[k].cnt, (int)k, count, Sf,
                                      Tf, stream);
            });
            if (next)
              prof_launch("k_sl_lhat", lv.lhat[k + 1].flops * count, lv.lhat[k + 1].bytes * count, stream, [&] {
                return launch_sl_lhat(P, tk + 3 * lv.lhat[k + 1].off, (int)lv.lhat[k + 1].cnt, (int)k + 1,
                                      count, Lf, Zf, Tf, stream);
              });
          } else {
            const int nl = next ? (int)lv.lhat[k + 1].cnt : 0;
            const int32_t* tl = tk + 3 * (next ? lv.lhat[k + 1].off : 0);
            prof_launch("k_sl_pp_lhat", (lv.spp[k].flops + (next ? lv.lhat[k + 1].flops : 0.0)) * count,
                        (lv.spp[k].bytes + (next ? lv.lhat[k + 1].bytes : 0.0)) * count, stream, [&] {
                          return launch_sl_pp_lhat(P, tk + 3 * lv.spp[k].off, (int)lv.spp[k].cnt, tl, nl, (int)k,
                                                   count, lv.spp[k].cnt * (int64_t)count <= sc.pp_split_max,
                                                   Lf, Zf, Sf, Tf, stream);
                        });
          }
        }
      }
      if (marks) check(hipEventRecord(marks->sel_a[l], stream), "hipEventRecord");
      small_groups(lv.cls_b, sc.sel_mask(l), [&](int c, int e, int hi) {
        double fl = 0.0, by = 0.0;
        int maxw = 0;
        for (int q = c; q <= e; ++q) fl += lv.sel_cls[q], by += lv.by_cls[q], maxw = std::max(maxw, lv.maxw_cls[q]);
        prof_launch("k_selinv_small", fl * count, by * count, stream, [&] {
          return launch_selinv_small(P, sym->d_small + lv.cls_b[c], (int)(lv.cls_b[e + 1] - lv.cls_b[c]),
                                     kClassM[hi], maxw, count, Lf, Sf, Uf, Wf, stream);
        });
      });
      if (marks) check(hipEventRecord(marks->sel_b[l], stream), "hipEventRecord");
      lclock.mark(sym->levels.size() - l, stream);
    }
    lclock.report("selected inversion", sym->levels.size(), count, sym->levels, true);
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
  bool empty = false;  // the 0 x 0 dense factor (dense_faer.rs:33-39): valid, nothing stored
  double* d_vals = nullptr;
  size_t d_vals_cap = 0;
  double* d_vec = nullptr;
  size_t d_vec_cap = 0;
  PinnedBuf pin;  // host staging of the values, right-hand side and results
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

// The analysis request of a GMRF_ORDER_* value.
OrderRequest order_request(int order);
std::shared_ptr<SymData> analyze_cached(int64_t n, const int64_t* indptr, const int64_t* indices,
                                        int order, const int64_t* given);
// Autotuner (gmrf_api.cpp): times candidate schedules of factorisation + selected
// inversion of `count` matrices on the given values and keeps the fastest.
gmrf_status autotune_batch(gmrf_batch* bt, int count, const double* vals, int64_t ldv, double* default_ms,
                           double* tuned_ms);

// The same analysis kept out of the process-wide cache (dense inputs: the solver
// holds the full pattern's analysis itself, keyed by n).
std::shared_ptr<SymData> analyze_uncached(int64_t n, const int64_t* indptr, const int64_t* indices,
                                          int order, const int64_t* given);
