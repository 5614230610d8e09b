// Symbolic analysis of a symmetric sparsity pattern for the supernodal LDL^T.
//
// Runs once per CSC pattern on the host. The result fixes the factor ordering,
// the column elimination tree, the relaxed supernode partition, the level
// schedule the GPU kernels walk (leaves -> roots for factor / forward solve,
// roots -> leaves for selected inversion / backward solve), and every index map
// the kernels need (CSC entry -> L slot, child front -> parent front).
//
// Reference interface replaced: faer::sparse::linalg::cholesky::
// factorize_symbolic_cholesky as called from
// crates/inla_math/src/sparse_ldlt.rs:159-186, with the ordering chosen by
// crates/inla_math/src/ordering.rs:149-162.
#pragma once

#include <cstdint>
#include <string>
#include <vector>

namespace gmrf {

// Which branch of the ordering chooser produced the factor ordering.
enum class OrderKind : int {
  Amd = 0,        // fill-reducing approximate minimum degree
  TimeMajor = 1,  // mixture-major -> time-major Kronecker relabelling
  Rcm = 2,        // reverse Cuthill-McKee
  Natural = 3,    // identity (explicit request only)
  Given = 4,      // caller-supplied permutation (explicit request only)
  NestedDissection = 5,  // level-structure nested dissection (explicit request or Auto)
};

// Ordering request for analyze(); Chooser mirrors the reference decision.
// Auto: the chooser's ordering or nested dissection, whichever has the lower
// level-schedule cost (level_schedule_cost).
enum class OrderRequest : int { Chooser = 0, AmdOnly = 1, Natural = 2, Given = 3, NestedDissection = 4, Auto = 5 };

struct Symbolic {
  int64_t n = 0;
  int64_t nnz_full = 0;  // stored entries of the input CSC (both triangles)
  std::vector<int64_t> indptr, indices;  // copy of the pattern (cache key)
  uint64_t pattern_hash = 0;

  OrderKind order_kind = OrderKind::Amd;
  std::vector<int64_t> perm;   // perm[k]  = original index at factor position k
  std::vector<int64_t> iperm;  // iperm[i] = factor position of original index i

  std::vector<int64_t> parent;    // column elimination tree (factor order), -1 = root
  std::vector<int64_t> colcount;  // structural nnz of column j of L incl. the diagonal

  // Relaxed supernodes, numbered in postorder (children before parents).
  int64_t nsuper = 0;
  std::vector<int64_t> sn_start;    // [nsuper+1] first column of supernode s
  std::vector<int64_t> sn_parent;   // supernode tree, -1 = root
  std::vector<int64_t> sn_rows_ptr; // [nsuper+1] into sn_rows
  std::vector<int64_t> sn_rows;     // R_s: sorted rows strictly below the supernode
  std::vector<int64_t> child_ptr;   // [nsuper+1] into child_idx
  std::vector<int64_t> child_idx;   // children of s, ascending
  std::vector<int32_t> rel;         // [sn_rows.size()] position of R_s[i] in parent front

  // Storage plan (doubles, per matrix of a batch).
  std::vector<int64_t> l_off;  // L block of s: m_s x w_s column-major, ld = m_s
  int64_t l_size = 0;
  std::vector<int64_t> u_off;  // update / Sigma_RR block of s: r_s x r_s column-major
  int64_t u_size = 0;
  std::vector<int64_t> v_off;  // update vector of s (r_s) for the forward solve
  int64_t v_size = 0;

  // Level schedule: level(s) = 1 + max level(children), leaves at 0.
  int nlevels = 0;
  std::vector<int64_t> level_ptr;  // [nlevels+1]
  std::vector<int64_t> level_sn;   // supernodes of each level

  // CSC entry e (row >= col in the ORIGINAL numbering, as the reference's
  // csc_to_faer keeps) -> slot in the L layout; -1 for entries above the diagonal.
  std::vector<int64_t> amap;
  // Every CSC entry -> slot of its symmetric position in the L layout (-1 if
  // outside the structure of L, possible only for upper entries without a
  // lower counterpart).
  std::vector<int64_t> slot_full;
  // Diagonal slot of every factor column j in the L layout.
  std::vector<int64_t> diag_slot;

  // Algorithmic counts (structural, relaxation zeros excluded).
  int64_t nnz_l = 0;          // sum colcount
  double sum_r2 = 0;          // sum_j r_j^2, r_j = colcount_j - 1
  double flops_ldlt = 0;      // sum_j r_j (r_j + 3)
  double flops_selinv = 0;    // sum_j 2 r_j (r_j + 1) + n
  double flops_solve = 0;     // 4 sum_j r_j + n
  int64_t nnz_q_lower = 0;    // entries with row >= col
  int64_t max_front = 0;      // max m_s
};

// Ordering-chooser pieces restated from crates/inla_math/src/ordering.rs.
int64_t lower_bandwidth(int64_t n, const int64_t* indptr, const int64_t* indices);
int64_t bandwidth_with_fwd(int64_t n, const int64_t* indptr, const int64_t* indices,
                           const std::vector<int64_t>& fwd);
bool try_time_major_kronecker(int64_t n, const int64_t* indptr, const int64_t* indices,
                              std::vector<int64_t>& fwd, std::vector<int64_t>& inv);
void reverse_cuthill_mckee(int64_t n, const int64_t* indptr, const int64_t* indices,
                           std::vector<int64_t>& fwd, std::vector<int64_t>& inv);
// Returns the chooser's branch; fwd[old] = new when the branch is not Amd.
OrderKind choose_symmetric_order(int64_t n, const int64_t* indptr, const int64_t* indices,
                                 std::vector<int64_t>& fwd, std::vector<int64_t>& inv);

// Approximate minimum degree on the symmetric pattern; order[k] = k-th pivot.
void amd_order(int64_t n, const int64_t* indptr, const int64_t* indices,
               std::vector<int64_t>& order);

// Nested dissection; order[k] = k-th pivot. Subgraphs of at most `leaf` vertices are
// not split further.
void nested_dissection_order(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t leaf,
                             std::vector<int64_t>& order);
constexpr int64_t kNdLeaf = 48;

// Estimated seconds of a batched factorisation and selected inversion under this
// analysis' level schedule (the chain of dependent level launches plus the flops).
double level_schedule_cost(const Symbolic& S);

// Validate a CSC pattern (square, sorted unique rows in range). Empty on success.
std::string validate_pattern(int64_t n, const int64_t* indptr, const int64_t* indices);

// Full analysis. `given_perm` (perm[k] = original index) is used for Given.
// Throws std::runtime_error with a message on invalid input.
Symbolic analyze(int64_t n, const int64_t* indptr, const int64_t* indices, OrderRequest req,
                 const int64_t* given_perm);

uint64_t hash_pattern(int64_t n, const int64_t* indptr, const int64_t* indices);

}  // namespace gmrf
