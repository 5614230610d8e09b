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
  const int64_t* u_off;      // [nsuper]
  const int64_t* v_off;      // [nsuper]
  const int64_t* rows_ptr;   // [nsuper+1]
  const int32_t* rows;       // R_s (factor indices)
  const int32_t* rel;        // position of R_s[i] in the parent front
  const int32_t* child_ptr;  // [nsuper+1]
  const int32_t* child_idx;
  const int64_t* fa_ptr;     // [nsuper+1] input entries assembled into supernode s
  const int32_t* fa_csc;     // CSC entry index
  const int32_t* fa_slot;    // slot j*m+i inside the L block of s
  const int32_t* perm;       // [n] original index of factor position k
  const int64_t* diag_slot;  // [n] L-layout slot of D_k
  const int64_t* slot_full;  // [nnz] L-layout slot of the (symmetric) CSC entry
};

// Host-side level lists for one launch class: supernode ids on the device plus
// per-level [begin, end) into that array and the largest front of the level.
struct LevelClass {
  int32_t* d_list = nullptr;  // device
  std::int64_t* begin = nullptr;
  std::int64_t* end = nullptr;
  int32_t* maxm = nullptr;
  int32_t* maxw = nullptr;
};

constexpr int kSmallFront = 64;  // fronts with m <= 64 run one wavefront each, in LDS
constexpr int kPanel = 32;       // panel width of the large-front kernels

}  // namespace gmrf
