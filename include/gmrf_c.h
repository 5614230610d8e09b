/*
 * gmrf_c.h — C ABI of the MI355X (gfx950) numeric GMRF core.
 *
 * Replaces the numeric path behind rust-inla's solver interface:
 *   trait InlaSolver            crates/inla_math/src/solver.rs:29-44
 *   FaerCpuSolver               crates/inla_math/src/solver.rs:48-99
 *   factorize_sparse / refactorize_numeric_arc / symbolic_pattern /
 *   sparse_solve_in_place / sparse_diagonal_inverse
 *                               crates/inla_math/src/sparse_ldlt.rs:196-517
 * and adds what the reference has no interface for: batched numeric
 * factorisation over one symbolic analysis, batched solves and selected
 * inversion, and the selected inverse on the pattern of Q.
 *
 * Conventions
 *  - Matrices are square CSC with 64-bit indptr/indices, rows sorted and unique
 *    within each column, both triangles stored (as sprs::CsMat in the reference,
 *    crates/inla_math/src/sparse.rs:3). Only entries with row >= col are read
 *    numerically, as crates/inla_math/src/sparse_ldlt.rs:74-102 does.
 *  - Every function returns a gmrf_status; on failure gmrf_last_error() returns
 *    a thread-local message. Status codes map 1:1 onto the reference's
 *    SolverError (solver.rs:10-17) / MathError (error.rs:7-37).
 *  - gmrf_solver_* take host pointers (like the reference's slices).
 *    gmrf_batch_* take DEVICE pointers on the library's device and enqueue
 *    work on the batch's stream; gmrf_batch_sync() waits for it.
 *  - A handle is used by one thread at a time; distinct handles may be used
 *    concurrently (the reference's Send + Sync with &mut self).
 */
#ifndef GMRF_C_H
#define GMRF_C_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef enum gmrf_status {
  GMRF_OK = 0,
  GMRF_SINGULAR = 1,              /* SolverError::Singular            */
  GMRF_NOT_POSITIVE_DEFINITE = 2, /* SolverError::NotPositiveDefinite */
  GMRF_SOLVE_FAILURE = 3,         /* SolverError::SolveFailure(msg)   */
  GMRF_INVALID_ARGUMENT = 4,      /* MathError::{NotSquare, DimensionMismatch, InvalidMatrix, ...} */
  GMRF_OUT_OF_MEMORY = 5,         /* MathError::OutOfMemory           */
  GMRF_DEVICE_ERROR = 6           /* HIP runtime failure (no reference counterpart) */
} gmrf_status;

/* Ordering request for gmrf_analyze. CHOOSER mirrors choose_symmetric_order
 * (crates/inla_math/src/ordering.rs:149-162): time-major Kronecker relabelling,
 * else RCM when it halves the envelope, else approximate minimum degree. */
typedef enum gmrf_order {
  GMRF_ORDER_CHOOSER = 0,
  GMRF_ORDER_AMD = 1,
  GMRF_ORDER_NATURAL = 2,
  GMRF_ORDER_GIVEN = 3, /* perm[k] = original index at factor position k */
  /* No reference counterpart: orderings for the level-synchronous device schedule.
   * ND: nested dissection by level-structure separators (short, wide elimination
   * trees: fewer dependent level launches). AUTO: the CHOOSER's ordering or ND,
   * whichever has the lower estimated schedule cost (dependent panel steps of the
   * levels plus flops). Results agree with the CHOOSER's to rounding. */
  GMRF_ORDER_ND = 4,
  GMRF_ORDER_AUTO = 5
} gmrf_order;

typedef struct gmrf_symbolic gmrf_symbolic;
typedef struct gmrf_batch gmrf_batch;
typedef struct gmrf_solver gmrf_solver;

typedef struct gmrf_symbolic_stats {
  int64_t n;
  int64_t nnz;           /* stored entries of the input CSC */
  int64_t nnz_q_lower;   /* entries with row >= col */
  int32_t order_kind;    /* 0 AMD, 1 time-major, 2 RCM, 3 natural, 4 given, 5 nested dissection */
  int32_t nlevels;       /* height of the supernodal tree */
  int64_t nsuper;
  int64_t nnz_l;         /* structural nnz(L) incl. the diagonal */
  int64_t l_size;        /* doubles of the supernodal L storage (incl. relaxation zeros) */
  int64_t u_size;        /* doubles of the update workspace (pooled by front lifetime) */
  int64_t max_front;
  double flops_ldlt;     /* sum_j r_j (r_j + 3) */
  double flops_selinv;   /* sum_j 2 r_j (r_j + 1) + n */
  double flops_solve;    /* 4 sum_j r_j + n */
  double sum_r2;
  uint64_t pattern_hash; /* hash of (n, indptr, indices): the key of tuned schedules */
} gmrf_symbolic_stats;

/* Thread-local message of the last failure on this thread ("" if none). */
const char* gmrf_last_error(void);

/* Select the HIP device used by handles created afterwards (default 0). */
gmrf_status gmrf_set_device(int32_t device);

/* ---- symbolic analysis (once per sparsity pattern) -------------------------
 * Replaces factorize_symbolic / symbolic_pattern (sparse_ldlt.rs:159-186,473-477)
 * and the pattern cache of crates/inla_math/src/scratch.rs:92-118; the cache
 * here is process-wide, keyed by the exact (indptr, indices) arrays. */
gmrf_status gmrf_analyze(int64_t n, const int64_t* indptr, const int64_t* indices,
                         int32_t order, const int64_t* perm, gmrf_symbolic** out);
void gmrf_symbolic_free(gmrf_symbolic* sym);
gmrf_status gmrf_symbolic_get_stats(const gmrf_symbolic* sym, gmrf_symbolic_stats* out);
/* perm[k] = original index at factor position k (orig_of_factor, sparse_ldlt.rs:272-277) */
gmrf_status gmrf_symbolic_perm(const gmrf_symbolic* sym, int64_t* perm);
/* column elimination tree in factor order, -1 for roots */
gmrf_status gmrf_symbolic_etree(const gmrf_symbolic* sym, int64_t* parent);
/* structural column counts of L (incl. diagonal) in factor order */
gmrf_status gmrf_symbolic_colcount(const gmrf_symbolic* sym, int64_t* colcount);
/* supernode partition: sn_start[nsuper+1] and supernodal parents[nsuper] */
gmrf_status gmrf_symbolic_supernodes(const gmrf_symbolic* sym, int64_t* sn_start,
                                     int64_t* sn_parent);
/* Ordering-chooser restatement (ordering.rs:8-162) for inspection:
 * returns the branch (0 AMD, 1 time-major, 2 RCM) and, when not AMD, fwd[old] = new. */
int32_t gmrf_choose_order(int64_t n, const int64_t* indptr, const int64_t* indices, int64_t* fwd);

/* ---- batched numeric phase (device pointers) --------------------------------
 * A batch holds up to `capacity` numeric factorisations of one pattern. */
gmrf_status gmrf_batch_create(gmrf_symbolic* sym, int32_t capacity, gmrf_batch** out);
void gmrf_batch_destroy(gmrf_batch* bt);
/* Numeric LDL^T of `count` matrices; values[b * ld_values + e] is entry e of the
 * CSC of matrix b (device). Replaces factorize_numeric_ldlt as called from
 * sparse_ldlt.rs:196-241,480-517 — no pivoting, no regularisation. */
gmrf_status gmrf_batch_factorize(gmrf_batch* bt, int32_t count, const double* values,
                                 int64_t ld_values);
/* log|Q_b| = sum log|D| (ldlt.rs:64-67) and per-matrix status (0 ok,
 * GMRF_SINGULAR, GMRF_NOT_POSITIVE_DEFINITE). Both outputs are host pointers;
 * the call synchronises the batch stream. */
gmrf_status gmrf_batch_logdet(gmrf_batch* bt, int32_t count, double* logdet, int32_t* status);
/* Device-side variant (no synchronisation): logdet/status are device pointers. */
gmrf_status gmrf_batch_logdet_device(gmrf_batch* bt, int32_t count, double* logdet,
                                     int32_t* status);
/* x_b = Q_b^-1 rhs_b, original order (sparse_ldlt.rs:243-265), device pointers;
 * x may alias rhs. */
gmrf_status gmrf_batch_solve(gmrf_batch* bt, int32_t count, const double* rhs, int64_t ld_rhs,
                             double* x, int64_t ld_x);
/* nrhs right-hand sides per matrix in one launch sequence: vector j of matrix b
 * is rhs + (b nrhs + j) ld_rhs, its solution x + (b nrhs + j) ld_x (device
 * pointers, x may alias rhs). The reference solves one vector per call
 * (sparse_ldlt.rs:243-265); several columns serve low-rank corrections. */
gmrf_status gmrf_batch_solve_multi(gmrf_batch* bt, int32_t count, int32_t nrhs, const double* rhs,
                                   int64_t ld_rhs, double* x, int64_t ld_x);
/* No counterpart in the reference. For callers that cannot allocate device memory
 * (code under HIP graph capture, serving loops, harnesses that forbid hipMalloc):
 * gmrf_batch_solve_multi with the same arguments and the same results, whose
 * workspace (count * nrhs vectors of n + update + staging doubles; the batch's own
 * multi-vector area is grown with hipMalloc on demand) is instead a fixed 32 MiB
 * array inside the library, one per GPU, held until the process ends. Requests
 * that do not fit return GMRF_INVALID_ARGUMENT (the message gives the limit for the
 * pattern) before any work is enqueued; nothing is allocated. Every call on one GPU
 * uses that GPU's array: calls that overlap on one GPU (several streams or threads)
 * race; calls ordered on one stream do not. */
gmrf_status gmrf_batch_solve_noalloc(gmrf_batch* bt, int32_t count, int32_t nrhs, const double* rhs,
                                     int64_t ld_rhs, double* x, int64_t ld_x);
/* Autotuner (no reference counterpart). Times factorisation + selected inversion of
 * `count` matrices of this pattern (device values as for gmrf_batch_factorize) under
 * candidate launch schedules and keeps the fastest for the pattern and batch size:
 * small-front size classes run per level separately or merged into one launch
 * (chosen level by level, for factorisation and selected inversion separately), the
 * workgroup threshold below which a trailing-update launch also factors the next
 * diagonal block, the split of Sigma[B,P] over K, and XCD grouping of its tiles. The
 * tuned schedule is kept only if it beats the defaults; it is also appended to the
 * file named by the GMRF_TUNE_FILE environment variable (if set), from which later
 * processes load it by pattern hash. default_ms / tuned_ms (host, may be NULL) receive
 * the measured times. The batch's factors are overwritten: factorise again after. */
gmrf_status gmrf_batch_autotune(gmrf_batch* bt, int32_t count, const double* values, int64_t ld_values,
                                double* default_ms, double* tuned_ms);
/* Selected inversion on the supernodal pattern of L (Takahashi recurrence). */
gmrf_status gmrf_batch_selinv(gmrf_batch* bt, int32_t count);
/* diag(Q_b^-1) in original order after gmrf_batch_selinv (sparse_ldlt.rs:421-431). */
gmrf_status gmrf_batch_selinv_diag(gmrf_batch* bt, int32_t count, double* diag, int64_t ld_diag);
/* (Q_b^-1)_e for every stored CSC entry e after gmrf_batch_selinv (device). */
gmrf_status gmrf_batch_selinv_pattern(gmrf_batch* bt, int32_t count, double* sigma,
                                      int64_t ld_sigma);
gmrf_status gmrf_batch_sync(gmrf_batch* bt);
/* The batch's HIP stream (hipStream_t) for callers that enqueue around it. */
void* gmrf_batch_stream(gmrf_batch* bt);
/* use_given != 0: enqueue subsequent batch work on `stream` (hipStream_t, not
 * owned; NULL is the default stream), keeping order with the caller's work
 * without host synchronisation. use_given == 0: back to the batch's own stream. */
gmrf_status gmrf_batch_set_stream(gmrf_batch* bt, void* stream, int32_t use_given);

/* ---- heterogeneous batches (several patterns, one launch sequence) ----------
 * The reference factorises independent models one solver at a time
 * (inference.rs:1336-1337); gmrf_factorize_multi of the drop-in contract is
 * this group. k patterns (full symmetric CSC each) are analysed separately,
 * with the ordering chooser of ordering.rs:8-162 per pattern, and factorised
 * as their block-diagonal union: every level launch covers all k models.
 * Values, right-hand sides and outputs use the union layout: the k CSC value
 * arrays (or vectors) concatenated in pattern order. */
typedef struct gmrf_multi gmrf_multi;
gmrf_status gmrf_multi_analyze(int32_t k, const int64_t* n, const int64_t* const* indptr,
                               const int64_t* const* indices, int32_t capacity, gmrf_multi** out);
void gmrf_multi_free(gmrf_multi* m);
/* n_off[k + 1], nnz_off[k + 1]: block offsets in the union vectors / values;
 * order_kind[k]: the chooser branch of each pattern (0 AMD, 1 time-major, 2 RCM). */
gmrf_status gmrf_multi_info(const gmrf_multi* m, int64_t* n_off, int64_t* nnz_off,
                            int32_t* order_kind);
/* A new handle on the union analysis (free with gmrf_symbolic_free). */
gmrf_symbolic* gmrf_multi_symbolic(gmrf_multi* m);
/* The union's batch, owned by m: use gmrf_batch_solve / _selinv / _selinv_diag /
 * _selinv_pattern / _set_stream on it; do not destroy it. */
gmrf_batch* gmrf_multi_batch(gmrf_multi* m);
/* Numeric LDL^T of `count` union matrices (device values, ld_values >= total
 * nnz) plus the per-block log|Q_q| and status. */
gmrf_status gmrf_multi_factorize(gmrf_multi* m, int32_t count, const double* values,
                                 int64_t ld_values);
/* logdet[b * k + q], status[b * k + q] (host or device pointers); synchronises. */
gmrf_status gmrf_multi_logdet(gmrf_multi* m, int32_t count, double* logdet, int32_t* status);

/* ---- InlaSolver mirror (host pointers) --------------------------------------
 * gmrf_solver_factorize = InlaSolver::factorize (solver.rs:70-74): symbolic
 * analysis on the first call or when the pattern changes, numeric otherwise.
 * On failure the previously stored factor is kept, as in the reference. When the
 * sparse factorisation runs out of device memory, the matrix is factorised densely
 * instead (DefaultBackend::factorize, backend.rs:110-122; 8 n^2 bytes, as
 * gmrf_solver_factorize_dense); GMRF_OUT_OF_MEMORY only if that fails too.
 * solve / diag_inv / log_abs_det before any successful factorize return
 * GMRF_SOLVE_FAILURE "Factorization not computed yet" (solver.rs:77-79). */
gmrf_status gmrf_solver_create(gmrf_solver** out);
void gmrf_solver_destroy(gmrf_solver* s);
gmrf_status gmrf_solver_factorize(gmrf_solver* s, int64_t n, const int64_t* indptr,
                                  const int64_t* indices, const double* data);
/* Dense input (LdltBackend::factorize_dense, crates/inla_math/src/backend.rs:14-20,59-69;
 * dense_faer.rs:25-75): a is n x n row-major, of which the lower triangle is read; it
 * is factorised as one dense front (the full pattern), after which solve / diag_inv /
 * log_abs_det apply as for gmrf_solver_factorize. A failed factorisation returns
 * GMRF_SINGULAR (the reference maps every dense failure to MathError::Singular) and
 * keeps the previous factor. n == 0 stores the empty factor (dense_faer.rs:33-39):
 * solve and diag_inv then write nothing and log_abs_det is 0; n < 0 is
 * GMRF_INVALID_ARGUMENT. The full pattern's analysis stays with the solver (reused
 * while n is unchanged), outside the process-wide pattern cache. */
gmrf_status gmrf_solver_factorize_dense(gmrf_solver* s, int64_t n, const double* a);
gmrf_status gmrf_solver_solve(gmrf_solver* s, const double* rhs, double* x);
gmrf_status gmrf_solver_diag_inv(gmrf_solver* s, double* out);
gmrf_status gmrf_solver_log_abs_det(const gmrf_solver* s, double* out);
int64_t gmrf_solver_n(const gmrf_solver* s);

/* ---- batched Laplace evaluation -------------------------------------------
 * A latent Gaussian model with an affine precision
 *   Q(theta) = sum_k coef_k(theta) Q_k   (all Q_k on one full symmetric CSC pattern
 *                                         that also contains the pattern of A^T A),
 * observations y_i with linear predictor eta = A x (A: n_obs x n CSC; NULL means
 * the identity) and one likelihood family. gmrf_model_eval runs, for a batch of
 * theta-points, what the reference runs per point in
 * find_latent_mode_a_with_solver (crates/inla_stats/src/inference.rs:776-975):
 * prior factorisation (Q + 1e-12 I, or Q itself when constraints are given —
 * the caller then includes kappa A_c^T A_c as a component with a constant
 * coefficient, constraints.rs:275-313), the Newton loop with step cap and
 * backtracking, the confirmation factorisation, the Laplace marginal likelihood
 * (inference.rs:972), diag(Q_post^-1) by selected inversion, and the gradient of
 * the marginal likelihood w.r.t. theta (no reference counterpart). */
enum { GMRF_FAMILY_GAUSSIAN = 0, GMRF_FAMILY_POISSON = 1, GMRF_FAMILY_BINOMIAL = 2,
       GMRF_FAMILY_NBINOMIAL = 3 };
enum { GMRF_LINK_IDENTITY = 0, GMRF_LINK_LOG = 1, GMRF_LINK_LOGIT = 2 };

typedef struct gmrf_model gmrf_model;

typedef struct gmrf_model_desc {
  int64_t n;                 /* latent dimension */
  const int64_t* indptr;     /* full symmetric CSC pattern (host) */
  const int64_t* indices;
  int32_t n_comp;            /* K */
  const double* comps;       /* K x nnz values of Q_1..Q_K on the pattern (host) */
  int64_t n_obs;
  const int64_t* a_indptr;   /* A: n_obs x n CSC (host), NULL for eta = x */
  const int64_t* a_indices;
  const double* a_values;
  int32_t family;            /* GMRF_FAMILY_* */
  int32_t link;              /* GMRF_LINK_* */
  const double* y;           /* n_obs responses */
  const double* obs_a;       /* n_obs: Gaussian precision scale / Poisson exposure /
                                binomial trials / NB exposure (NULL = 1) */
  const double* obs_b;       /* n_obs: NB size (NULL otherwise) */
  int32_t n_constr;          /* k hard constraints A_c x = e (projection), 0 for none */
  const double* constr_a;    /* k x n row-major (host) */
  const double* constr_e;    /* k */
  int32_t capacity;          /* max theta-points per gmrf_model_eval */
  double constr_kappa;       /* 0: any kappa A_c^T A_c augmentation is one of the components
                                (constraints.rs:275-313 forms it explicitly). > 0: the
                                augmentation Q + kappa A_c^T A_c is applied to the prior and
                                the posterior as an exact low-rank correction of the grounded
                                precision, so long constraint rows do not densify the pattern
                                (no reference counterpart; k <= 32) */
  int32_t order;             /* GMRF_ORDER_CHOOSER (the reference's choice), GMRF_ORDER_AMD,
                                GMRF_ORDER_NATURAL (the caller's order), GMRF_ORDER_ND or
                                GMRF_ORDER_AUTO */
  int32_t n_blocks;          /* 0 or 1: one model. nb > 1: nb independent models (a heterogeneous
                                batch, e.g. many areal models of uneven size) as one
                                block-diagonal union: block q owns the latent indices
                                [block_ptr[q], block_ptr[q + 1]), the pattern entries of those
                                columns (the pattern must be block diagonal) and the
                                observations that load them (each observation loads one block
                                and every block has observations). Each block is ordered on its
                                own pattern; the factorisations, solves and selected inversions
                                run on the union in one launch sequence, while the Newton steps,
                                convergence, status, mlik and gradient are per (point, block):
                                in gmrf_eval_args every per-point quantity (coefs, lpar, dcoefs,
                                dlog_lpar, mlik, grad, status, n_iter) gains a block index,
                                [point][block][...]; x and var stay count x n. No constraints.
                                No reference counterpart: the reference fits such models one
                                at a time (crates/inla_stats/src/inference.rs:776-975 each) */
  const int64_t* block_ptr;  /* nb + 1 latent offsets (host), block_ptr[0] = 0, [nb] = n */
  int32_t lean;              /* 0: the prior and posterior factors of the capacity points share one
                                batch of 2 capacity matrices. 1: the prior is factorised (and, for the
                                gradient, inverted) first and alone in the posterior's slots, keeping
                                only its log determinant, status and traces tr(Sigma_prior Q_k); the
                                batch then holds capacity matrices (half the device memory), for
                                models whose factor fills most of the device. Same results; no graph
                                capture. No reference counterpart (the reference factorises the
                                prior in its own solver, inference.rs:827-836) */
} gmrf_model_desc;

typedef struct gmrf_eval_args {
  int32_t count;             /* theta-points in this call (<= capacity) */
  int32_t n_theta;           /* m, length of theta (for the gradient) */
  const double* coefs;       /* device, count x K (NULL in values mode, qvals below) */
  const double* lpar;        /* device, count: Gaussian precision multiplier (NULL = 1) */
  const double* dcoefs;      /* device, count x m x K: d coef_k / d theta_l (NULL: no gradient) */
  const double* dlog_lpar;   /* device, count x m: d log(lpar) / d theta_l (NULL = 0) */
  int32_t max_iter;          /* Newton iterations (the reference uses 200) */
  double tol;                /* Newton tolerance (the reference uses 1e-5) */
  /* outputs, device pointers (NULL to skip x / var / grad) */
  double* mlik;              /* count */
  double* x;                 /* count x n, posterior mode (original order) */
  double* var;               /* count x n, diag(Q_post^-1) */
  double* grad;              /* count x m, d mlik / d theta */
  /* outputs, host pointers */
  int32_t* status;           /* count: GMRF_OK or the failure of that point */
  int32_t* n_iter;           /* count: Newton iterations used */
  /* outputs, device pointers, count x n_obs in the caller's observation order (NULL to
     skip): the linear predictor at the mode, eta = A x* (x* without A), and its variance
     in the reference's diagonal form sum_j A_ij^2 diag(Q_post^-1)_j (x's variances
     without A): the cond_eta / cond_eta_var of crates/inla_stats/src/inference.rs:1371-1379
     (predictor_variances_diag, crates/inla_math/src/design.rs:198-210) that DIC, WAIC and
     CPO/PIT take. Requesting them runs the general (uncaptured) evaluation. */
  double* eta;
  double* eta_var;
  /* values mode (NULL: the coefficient form above). Device, count x nnz: the entries of
     Q(theta_b) on the model's pattern, used instead of coefs x components, for callers
     whose Q(theta) is not affine in fixed matrices (the reference's build_prior callback,
     inference.rs:1246-1300); dqvals (device, count x m x nnz, NULL: no gradient): the
     entries of dQ/dtheta_l on the pattern, used instead of dcoefs. Not with blocks or
     the low-rank constraint form. */
  const double* qvals;
  const double* dqvals;
} gmrf_eval_args;

gmrf_status gmrf_model_create(const gmrf_model_desc* desc, gmrf_model** out);
void gmrf_model_destroy(gmrf_model* model);
gmrf_status gmrf_model_eval(gmrf_model* model, gmrf_eval_args* args);
/* The model's stream (or the caller's, after set_stream). gmrf_model_eval orders its
 * work after what is already enqueued there and leaves its device outputs ordered
 * before what follows; inside, the selected inversion of the final factors runs on a
 * second stream of the model (default priority) that forks from and joins back into
 * this one with events. */
void* gmrf_model_stream(gmrf_model* model);

/* Per-observation log-likelihood log pi(y_i | eta) for count predictor vectors
 * (eta[b * n_obs + i], device; y, obs_a, obs_b per observation and lpar per vector,
 * device; lpar NULL = 1, obs_b only for the negative binomial): eval_likelihood(..).logp
 * of crates/inla_stats/src/inference.rs:170-311 as summed by deviance_at_config /
 * compute_waic (crates/inla_stats/src/model_selection.rs:82-91,152-199), on `stream`
 * (NULL: the null stream); logp[b * n_obs + i] (device). The call synchronises the
 * stream and returns GMRF_INVALID_ARGUMENT when an input is one on which
 * eval_likelihood returns an error (binomial y > n or y < 0, a negative count or
 * exposure, a Gaussian precision <= 0, a negative-binomial size <= 0 or mean <= 0,
 * a non-finite eta); a zero Poisson rate with y > 0 is logp = -inf, not an error. */
gmrf_status gmrf_obs_loglik(int32_t family, int32_t link, int32_t count, int64_t n_obs,
                            const double* eta, const double* y, const double* obs_a,
                            const double* obs_b, const double* lpar, double* logp, void* stream);
/* gmrf_obs_loglik with the derivatives eval_likelihood returns: grad = d log pi / d eta
 * and neg_hess = -d2 log pi / d eta2 (the Newton weight of the Laplace evaluation; Fisher
 * scoring where the reference scores, inference.rs:170-311), as
 * check_cpo_failure_for_obs takes them (crates/inla_stats/src/model_selection.rs:328-396).
 * Any of logp, grad, neg_hess may be NULL. Same errors as gmrf_obs_loglik. */
gmrf_status gmrf_obs_eval(int32_t family, int32_t link, int32_t count, int64_t n_obs, const double* eta,
                          const double* y, const double* obs_a, const double* obs_b, const double* lpar,
                          double* logp, double* grad, double* neg_hess, void* stream);
gmrf_status gmrf_model_set_stream(gmrf_model* model, void* stream, int32_t use_given);
/* gmrf_batch_autotune for the model's batch (the prior and posterior matrices of the
 * last gmrf_model_eval, 2 count matrices) on that evaluation's values; call after an
 * evaluation. Later evaluations of the same size use the tuned schedule (the model's
 * captured graphs are rebuilt). */
gmrf_status gmrf_model_autotune(gmrf_model* model, double* default_ms, double* tuned_ms);

/* ---- diagnostics -------------------------------------------------------------
 * Per-launch timing of the level kernels with HIP events on the launching
 * stream (no reference counterpart; used by the benchmark's roofline).
 * gmrf_profile_enable(1) clears and starts recording; gmrf_profile_summary
 * waits for the recorded events and returns per-kernel launch counts, total
 * milliseconds, the algorithmic flops / bytes of the recorded launches and
 * roof_ms, the sum over launches of max(flops / peak_flops, bytes / peak_bytes)
 * (peaks in flop/s and byte/s): the time the launches would take at the roofline;
 * `names` receives max_entries strings of name_len bytes. */
gmrf_status gmrf_profile_enable(int32_t on);
gmrf_status gmrf_profile_summary(int32_t max_entries, char* names, int32_t name_len,
                                 int64_t* launches, double* total_ms, double* flops, double* bytes,
                                 double peak_flops, double peak_bytes, double* roof_ms,
                                 int32_t* n_entries);

/* T = A B^T for one 64x64 tile through the FP64 MFMA tile engine (A, B are
 * 64 x K column-major host arrays, K <= 64). No reference counterpart: pins the
 * MFMA lane layout in the tests. */
gmrf_status gmrf_probe_mma(const double* A, const double* B, int32_t K, double* T);

#ifdef __cplusplus
}
#endif

#endif /* GMRF_C_H */
