// Device-side view of a symbolic analysis, shared by every matrix of a batch.
#pragma once

#include <cstdint>

namespace gmrf {

// All pointers are device pointers. Offsets are in doubles, per matrix.
struct DevPlan {
  int32_t n;
  int32_t nsuper;
  int64_t nnz;     // entries of the input CSC (both triangles)
  int64_t l_size;  // doubles of the L layout per matrix
  int64_t u_size;  // doubles of the update / Sigma_RR workspace per matrix
  int64_t v_size;  // doubles of the forward-solve update vectors per matrix

  const int32_t* sn_start;   // [nsuper+1]
  const int32_t* sn_r;       // [nsuper] rows below the supernode
  const int32_t* sn_parent;  // [nsuper] -1 for roots
  const int64_t* l_off;      // [nsuper]
  const int64_t* u_off;      // [nsuper] update block during factorisation (pooled by lifetime)
  const int64_t* u_off_sel;  // [nsuper] Sigma_RR block during selected inversion (pooled)
  const int64_t* v_off;      // [nsuper]
  const int64_t* rows_ptr;   // [nsuper+1]
  const int32_t* rows;       // R_s (factor indices)
  const int32_t* rel;        // position of R_s[i] in the parent front
  const int32_t* child_ptr;  // [nsuper+1]
  const int32_t* child_idx;
  const int64_t* fa_ptr;     // [nsuper+1] input entries assembled into supernode s
  const int32_t* fa_csc;     // CSC entry index
  const int32_t* fa_slot;    // slot j*m+i inside the L block of s
  // input entries of the large fronts grouped by assembly tile: tile t of the
  // assembly task list holds entries ain_ptr[t] .. ain_ptr[t + 1], each at LDS
  // position ain_pos = 65 (j - 64 J) + (i - 64 I) of the tile, value vals[ain_csc]
  const int32_t* ain_ptr;
  const int32_t* ain_pos;
  const int32_t* ain_csc;
  const int32_t* perm;       // [n] original index of factor position k
  const int64_t* diag_slot;  // [n] L-layout slot of D_k
  const int64_t* slot_full;  // [nnz] L-layout slot of the (symmetric) CSC entry

  // Large fronts (m > kSmallFront): tile-parallel kernels.
  const int32_t* is_large;    // [nsuper]
  const int64_t* w_off;       // [nsuper] dense m x m Sigma front (large fronts, pooled), -1 otherwise
  int64_t w_size;             // doubles of the Sigma-front pool per matrix
  const int64_t* t_off;       // [nsuper] selected-inversion staging (large fronts, per level)
  int64_t t_size;             // doubles of the selected-inversion staging per matrix
  const int64_t* ts_off;      // [nsuper] backward-solve partial sums (large fronts, per level)
  int64_t ts_size;            // doubles of the solve staging per right-hand side
  const int64_t* z_off;       // [nsuper] per-panel 64 x 64 blocks Z = L11^-1 (large fronts)
  int64_t z_size;             // doubles of the Z workspace per matrix
  const int64_t* cinv_off;    // [n children] per (front, child) slot: offset into cinv
  const int32_t* cinv;        // parent front position -> child row index, -1 if absent
};

// Independent blocks of a Laplace model (a heterogeneous batch of models in one
// block-diagonal union): block q owns the latent indices [lseg[q], lseg[q + 1]), the
// observations [oseg[q], oseg[q + 1]) (the model's internal order) and the pattern
// entries [eseg[q], eseg[q + 1]); lblk / oblk / eblk give each element's block. A
// per-point quantity of the model becomes per (point b, block q), index b nb + q.
// nb = 1 with null maps and segments is a single model.
struct LapBlocks {
  int32_t nb = 1;
  const int32_t* lblk = nullptr;
  const int32_t* oblk = nullptr;
  const int32_t* eblk = nullptr;
  const int64_t* lseg = nullptr;
  const int64_t* oseg = nullptr;
  const int64_t* eseg = nullptr;
};

constexpr int kSmallFront = 64;  // fronts with m <= 64 run one workgroup each, in LDS
// small-front classes by front size: one kernel instance per class, 4 m / 16 waves
constexpr int kSmallClasses = 4;
constexpr int kClassM[kSmallClasses] = {16, 32, 48, 64};
constexpr int kPanel = 32;       // K chunk of the MFMA tile engine
constexpr int kTile = 64;        // panel width / tile edge of the large-front kernels
constexpr int kSolvePart = 1024;  // rows of R_s per partial sum in the backward solve
constexpr int kSolveSplitRows = 1024;  // levels whose tallest front has r >= this split the solve
constexpr int64_t kLongGather = 512;  // Laplace gather lists longer than this get a workgroup each
constexpr int kFuseDiagBelow = 2048;  // factor launches of fewer workgroups also factor the next diagonal block
constexpr int kSelSplitK = 512;  // rows below a panel per partial product of Sigma[P,P] (tall fronts)
constexpr int kGemmSplitK = 512;  // K per partial product of Sigma[B,P] when a level has few row blocks
constexpr int kGemmSplitMaxTasks = 96;  // split Sigma[B,P] when a launch has at most this many row blocks
constexpr int kXcdMaxTasks = 256;  // XCD-grouped launches: at most this many tasks per matrix
constexpr int64_t kPpSplitMaxTasks = (int64_t)1 << 30;  // Sigma[P,P] launches of at most this many tasks split each in three
constexpr int64_t kTrsm32MaxTasks = 64;   // panel-solve launches of at most this many tiles split each in four
constexpr int64_t kGemm32MaxTasks = 512;  // Sigma[B,P] launches of at most this many tiles split each in four

// Staging area of a large front (T, per level): L_hat of two panels (m x 64 each,
// panel step k in buffer k % 2, so that the next panel's L_hat can be formed while
// the current one is still read), the Sigma[P,P] partial products, then the
// Sigma[B,P] partial tiles. The backward solves stage elsewhere (ts_off: ceil(r /
// kSolvePart) x w per large front and right-hand side).
constexpr int64_t stage_lhat_off(int64_t m, int step) { return (step & 1) * m * 64; }
constexpr int64_t stage_pp_off(int64_t m) { return 2 * m * 64; }
constexpr int64_t stage_gemm_off(int64_t m) {
  return stage_pp_off(m) + ((m + kSelSplitK - 1) / kSelSplitK) * 64 * 64;
}
constexpr int64_t stage_size(int64_t m) {
  return stage_gemm_off(m) + ((m + 63) / 64) * ((m + kGemmSplitK - 1) / kGemmSplitK) * 64 * 64;
}

}  // namespace gmrf
