// Heterogeneous batches: k independent matrices with different patterns
// factorised as the block-diagonal union, so that every level launch covers the
// fronts of all k models. Each pattern keeps the ordering its own analysis
// chooses (crates/inla_math/src/ordering.rs:8-162 per model); the union is
// analysed with the concatenation of those orderings, which keeps every
// block's elimination tree, supernodes and fill exactly as for the block alone.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <memory>
#include <vector>

#include "runtime.hpp"

struct gmrf_multi {
  int32_t k = 0;
  int32_t cap = 0;
  std::vector<int64_t> n_off, nnz_off;
  std::vector<int32_t> order_kind;
  std::shared_ptr<SymData> sym;
  std::unique_ptr<gmrf_batch> bt;
  DevArena arena;
  int64_t* d_seg_ptr = nullptr;  // [k + 1] factor positions of each block
  int32_t* d_seg_pos = nullptr;  // [n_total]
  double* d_blk_logdet = nullptr;
  int32_t* d_blk_status = nullptr;
};

extern "C" {

gmrf_status gmrf_multi_analyze(int32_t k, const int64_t* n, const int64_t* const* indptr,
                               const int64_t* const* indices, int32_t capacity, gmrf_multi** out) {
  if (!out || !n || !indptr || !indices || k < 1 || capacity < 1)
    return fail(GMRF_INVALID_ARGUMENT, "invalid argument");
  *out = nullptr;
  try {
    auto M = std::make_unique<gmrf_multi>();
    M->k = k;
    M->cap = capacity;
    M->n_off.assign(k + 1, 0);
    M->nnz_off.assign(k + 1, 0);
    M->order_kind.resize(k);
    for (int32_t q = 0; q < k; ++q) {
      if (!indptr[q] || !indices[q] || n[q] <= 0) throw std::runtime_error("empty pattern in batch");
      M->n_off[q + 1] = M->n_off[q] + n[q];
      M->nnz_off[q + 1] = M->nnz_off[q] + indptr[q][n[q]];
    }
    const int64_t nt = M->n_off[k];
    std::vector<int64_t> ip(nt + 1, 0), ix(M->nnz_off[k]), perm(nt);
    for (int32_t q = 0; q < k; ++q) {
      const int64_t n0 = M->n_off[q], e0 = M->nnz_off[q];
      for (int64_t c = 0; c < n[q]; ++c) ip[n0 + c + 1] = e0 + indptr[q][c + 1];
      for (int64_t e = 0; e < indptr[q][n[q]]; ++e) ix[e0 + e] = n0 + indices[q][e];
      Symbolic Sq = analyze(n[q], indptr[q], indices[q], OrderRequest::Chooser, nullptr);
      M->order_kind[q] = (int32_t)Sq.order_kind;
      for (int64_t p = 0; p < n[q]; ++p) perm[n0 + p] = n0 + Sq.perm[p];
    }
    M->sym = analyze_cached(nt, ip.data(), ix.data(), GMRF_ORDER_GIVEN, perm.data());
    M->bt.reset(new gmrf_batch(M->sym, capacity));
    // factor positions grouped by block (positions ascending inside a block)
    const Symbolic& S = M->sym->S;
    std::vector<int64_t> seg_ptr(k + 1, 0);
    std::vector<int32_t> blk(nt);
    for (int64_t p = 0; p < nt; ++p) {
      const int64_t o = S.perm[p];
      blk[p] = (int32_t)(std::upper_bound(M->n_off.begin(), M->n_off.end(), o) - M->n_off.begin() - 1);
      seg_ptr[blk[p] + 1]++;
    }
    for (int32_t q = 0; q < k; ++q) seg_ptr[q + 1] += seg_ptr[q];
    std::vector<int32_t> pos(nt);
    {
      std::vector<int64_t> fill(seg_ptr.begin(), seg_ptr.end() - 1);
      for (int64_t p = 0; p < nt; ++p) pos[fill[blk[p]]++] = (int32_t)p;
    }
    M->d_seg_ptr = M->arena.upload(seg_ptr);
    M->d_seg_pos = M->arena.upload(pos);
    M->d_blk_logdet = M->arena.alloc<double>((size_t)capacity * k);
    M->d_blk_status = M->arena.alloc<int32_t>((size_t)capacity * k);
    *out = M.release();
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

void gmrf_multi_free(gmrf_multi* M) { delete M; }

gmrf_status gmrf_multi_info(const gmrf_multi* M, int64_t* n_off, int64_t* nnz_off, int32_t* order_kind) {
  if (!M) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (n_off) std::copy(M->n_off.begin(), M->n_off.end(), n_off);
  if (nnz_off) std::copy(M->nnz_off.begin(), M->nnz_off.end(), nnz_off);
  if (order_kind) std::copy(M->order_kind.begin(), M->order_kind.end(), order_kind);
  return GMRF_OK;
}

gmrf_symbolic* gmrf_multi_symbolic(gmrf_multi* M) {
  if (!M) return nullptr;
  return new gmrf_symbolic{M->sym};
}

gmrf_batch* gmrf_multi_batch(gmrf_multi* M) { return M ? M->bt.get() : nullptr; }

gmrf_status gmrf_multi_factorize(gmrf_multi* M, int32_t count, const double* values, int64_t ld_values) {
  if (!M || !values) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > M->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  if (ld_values < M->nnz_off[M->k]) return fail(GMRF_INVALID_ARGUMENT, "ld_values < total nnz");
  try {
    check(hipSetDevice(M->sym->device), "hipSetDevice");
    gmrf_batch* bt = M->bt.get();
    bt->factorize(count, values, ld_values);
    check(launch_logdet_seg(M->sym->plan, count, bt->L, M->d_seg_ptr, M->d_seg_pos, M->k,
                            M->d_blk_logdet, M->d_blk_status, bt->stream),
          "logdet_seg");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

gmrf_status gmrf_multi_logdet(gmrf_multi* M, int32_t count, double* logdet, int32_t* status) {
  if (!M) return fail(GMRF_INVALID_ARGUMENT, "null argument");
  if (count < 1 || count > M->cap) return fail(GMRF_INVALID_ARGUMENT, "count exceeds batch capacity");
  try {
    hipStream_t st = M->bt->stream;
    const size_t nk = (size_t)count * M->k;
    if (logdet)
      check(hipMemcpyAsync(logdet, M->d_blk_logdet, nk * sizeof(double), hipMemcpyDefault, st), "copy");
    if (status)
      check(hipMemcpyAsync(status, M->d_blk_status, nk * sizeof(int32_t), hipMemcpyDefault, st), "copy");
    check(hipStreamSynchronize(st), "sync");
    return GMRF_OK;
  } catch (...) {
    return from_exception();
  }
}

}  // extern "C"
