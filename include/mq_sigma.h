/*
 * mq_sigma.h -- C-ABI of the MI355X (gfx950) sigma-vector / boundary-propagation engine.
 *
 * The entry points replace the execution layer of QCMaquis' contraction engine
 * (sekelle/maquis-dmrg).  Every function returns an int status (MQ_OK == 0); on failure
 * mq_last_error() returns a message.  No entry point ever calls exit().
 * Reference interfaces replaced (paths relative to dmrg/framework/dmrg/ of the reference):
 *
 *   create_contraction_schedule     mp_tensors/contractions/create_schedule/site_hamil_schedule.hpp:40-93
 *   super_hamil_mv                  solver/super_hamil_mv.hpp:100-155   (ietl::mult, solver/ietl_interface_dv.h:112-118)
 *   overlap_mpo_left_step           mp_tensors/contractions/boundary_ops/move_boundary.hpp:147-279
 *   overlap_mpo_right_step          mp_tensors/contractions/boundary_ops/move_boundary.hpp:281-404
 *   solve<double> (Jacobi-Davidson) solver/solver.cpp:14-34, ietl/jacobi.h:386-474
 *   DavidsonVector ops              solver/davidson_vector.cpp:197-254, solver/ietl_interface_dv.h:33-69
 *
 * Data conventions (identical to the reference):
 *   - all matrices are column-major doubles;
 *   - an MPS site tensor is "right-paired": one block per left-index sector lb (left_i order),
 *     ls(lb) rows, columns = concatenation over physical sectors s (phys_i order) of
 *     d_s * rs(lc+s) columns (sub-index major), i.e. ProductBasis(phys_i, right_i, fuse(-a,b));
 *   - "left-paired": one block per right-index sector rb (right_i order), rows = concatenation
 *     over s of d_s * ls(rc-s) rows, rs(rb) columns, i.e. ProductBasis(phys_i, left_i);
 *   - a Davidson vector is the concatenation of the right-paired blocks, each padded to a
 *     multiple of MQ_BUFFER_ALIGNMENT elements (solver/constants.h:6, davidson_vector.cpp:43-55);
 *   - a boundary is a set of cohorts (bra sector, ket sector); cohort ci stores n_blocks(ci)
 *     column-major ls x rs blocks back to back; offsets[ci][b] is the element offset of MPO
 *     bond index b, -1 when absent (mp_tensors/boundary.h:76-234).
 */
#ifndef MQ_SIGMA_H
#define MQ_SIGMA_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MQ_BUFFER_ALIGNMENT 128u
#define MQ_NONE 0xffffffffu

enum mq_status {
    MQ_OK = 0,
    MQ_ERR_INVALID = 1,   /* bad argument, shape or descriptor */
    MQ_ERR_HIP = 2,       /* a HIP runtime call failed */
    MQ_ERR_NOMEM = 3,     /* device or host allocation failed */
    MQ_ERR_CAPACITY = 4,  /* a size exceeds a fixed capacity */
    MQ_ERR_INTERNAL = 5
};

enum mq_symmetry {
    MQ_SYM_U1 = 1,     /* charge q[0]                 (block_matrix/symmetry/u1.h)    */
    MQ_SYM_TWOU1 = 2,  /* charge (q[0], q[1])         (block_matrix/symmetry/2u1.h)   */
    MQ_SYM_SU2U1 = 3   /* charge (N, 2S)              (block_matrix/symmetry/su2u1.h) */
};

typedef struct { int32_t q[2]; } mq_charge;

/* Index<SymmGroup>: (charge, size) pairs, in the caller's order (block_matrix/indexing_stable.hpp:97-258). */
typedef struct {
    uint32_t n;
    const mq_charge* charges;
    const uint32_t* sizes;
} mq_index_desc;

/* One site operator with its SparseOperator entries (block_matrix/sparse_operator.h:35-233).
 * Block k maps physical sector in_charge[k] (W.basis().left_charge) to out_charge[k]
 * (W.basis().right_charge); entry e: sigma(out, col[e]) += coeff[e] * psi(in, row[e]). */
typedef struct {
    int32_t spin;                 /* 2S of the operator (SU2U1), 0 for abelian groups */
    uint32_t n_blocks;
    const mq_charge* in_charge;
    const mq_charge* out_charge;
    const uint32_t* entry_begin;  /* n_blocks + 1 */
    const uint32_t* row;
    const uint32_t* col;
    const uint8_t* row_spin;      /* SU2 site spins of the entry (0, 1 or 2); NULL => 0 */
    const uint8_t* col_spin;
    const double* coeff;
} mq_op_desc;

/* MPOTensor (mp_tensors/mpotensor.h:45-135): terms (b1, b2, op, scale) sorted by (b1, b2);
 * several ops per (b1, b2) are allowed (term_descriptor).  Bond properties: 2S per bond index
 * and the Hermitian-conjugate pairing (mp_tensors/mpotensor_detail.h:141-187). */
typedef struct {
    uint32_t row_dim, col_dim;
    const int32_t* left_spin;     /* row_dim entries, may be NULL (all 0) */
    const int32_t* right_spin;    /* col_dim entries, may be NULL */
    const uint32_t* left_herm;    /* row_dim, conjugate bond index or MQ_NONE; may be NULL */
    const int32_t* left_phase;
    const uint32_t* right_herm;   /* col_dim */
    const int32_t* right_phase;
    uint32_t n_ops;
    const mq_op_desc* ops;
    uint32_t n_terms;
    const uint32_t* term_b1;
    const uint32_t* term_b2;
    const uint32_t* term_op;
    const double* term_scale;
} mq_mpo_desc;

/* Site tensor indices (MPSTensor row_dim / site_dim / col_dim). */
typedef struct {
    int32_t sym;
    mq_index_desc left_i, phys_i, right_i;
} mq_site_desc;

/* BoundaryIndex<double, SymmGroup> (mp_tensors/boundary.h:76-234 and boundary_index_rt.hpp:37-86). */
typedef struct {
    mq_index_desc bra, ket;
    uint32_t n_cohorts;
    uint32_t aux_dim;
    const uint32_t* bra_pos;        /* n_cohorts: lbrb_ci row  */
    const uint32_t* ket_pos;        /* n_cohorts: lbrb_ci column */
    const uint8_t* tr;              /* n_cohorts */
    const uint32_t* n_blocks;       /* n_cohorts (BoundaryIndexRT::n_blocks) */
    const int64_t* offsets;         /* n_cohorts * aux_dim, -1 == absent */
    const double* conj_scales;      /* n_cohorts * aux_dim */
    const uint8_t* transposes;      /* n_cohorts * aux_dim, may be NULL */
} mq_bindex_desc;

/* Flattened ScheduleNew<double> (solver/tasks.h:50-348): the reference's own plan format.
 * MPSBlock m owns T-tuples [t_begin[m], t_begin[m+1]) and cohorts [coh_begin[m], coh_begin[m+1]).
 * Within a cohort, units are (phys-out sector, sub-index) pairs in sfold order; a unit owns
 * lines (b1 row id, number of tasks) and tasks (ti, col, alpha) in line order. */
typedef struct {
    uint32_t n_blocks;
    const uint64_t* t_begin;
    const uint32_t* t_mps_offset;
    const uint32_t* t_ci;
    const uint32_t* t_ci_eff;
    const uint32_t* t_lb_ket;
    const uint64_t* coh_begin;
    const uint32_t* coh_lb;
    const uint32_t* coh_rb;
    const uint32_t* coh_ls;
    const uint32_t* coh_rs;
    const uint32_t* coh_ci;
    const uint32_t* coh_ci_eff;
    const uint32_t* coh_nSrows;
    const uint32_t* coh_stripe;
    const uint64_t* coh_unit_begin;
    const uint32_t* unit_offset;
    const uint32_t* unit_ms;
    const uint64_t* unit_line_begin;
    const uint32_t* line_b1;
    const uint32_t* line_count;
    const uint64_t* unit_task_begin;
    const uint32_t* task_ti;
    const uint32_t* task_col;
    const double* task_alpha;
    /* cohort -> new-boundary block offsets (boundary steps only, else NULL):
       coh_mpo_offset_begin[c] .. +aux_dim, -1 == absent (Cohort::get_offsets) */
    uint32_t aux_dim;
    const int64_t* coh_mpo_offsets;
} mq_schedule_desc;

typedef struct mq_ctx mq_ctx;
typedef struct mq_boundary mq_boundary;
typedef struct mq_plan mq_plan;
typedef struct mq_schedule mq_schedule;

/* ---------------------------------------------------------------- context */
const char* mq_last_error(void);
int mq_ctx_create(int device, mq_ctx** out);
/* Plans and boundaries made on a context must be destroyed before it. */
int mq_ctx_destroy(mq_ctx* ctx);
int mq_ctx_synchronize(mq_ctx* ctx);
/* Stream the context launches on (a hipStream_t); lets callers time with events on it. */
void* mq_ctx_stream(mq_ctx* ctx);

/* ---------------------------------------------------------------- layouts */
/* Number of elements and per-block offsets of the padded Davidson-vector layout of the
 * right-paired tensor (n_blocks == left_i.n).  block_offsets may be NULL. */
int mq_site_vector_layout(const mq_site_desc* site, uint64_t* n_elements, uint64_t* block_offsets);
/* Same for the left-paired layout (one block per right_i sector). */
int mq_site_left_layout(const mq_site_desc* site, uint64_t* n_elements, uint64_t* block_offsets);
/* Reshape between the right-paired and left-paired layouts on the device
 * (mp_tensors/reshapes.h:197-242 and its inverse).  src/dst are device pointers. */
int mq_reshape_right_to_left(mq_ctx* ctx, const mq_site_desc* site, const double* src, double* dst);
int mq_reshape_left_to_right(mq_ctx* ctx, const mq_site_desc* site, const double* src, double* dst);

/* ---------------------------------------------------------------- boundaries */
/* Boundary(ud, ld, ad) constructor (mp_tensors/boundary.h:258-283): one cohort per sector of
 * idx, aux_dim blocks each, every element 1.0. */
int mq_boundary_create_trivial(mq_ctx* ctx, const mq_index_desc* idx, uint32_t aux_dim, mq_boundary** out);
/* Upload a boundary: host_cohorts[ci] points at n_blocks(ci)*ls*rs doubles (NULL for tr cohorts). */
int mq_boundary_create(mq_ctx* ctx, const mq_bindex_desc* idx, const double* const* host_cohorts,
                       mq_boundary** out);
int mq_boundary_destroy(mq_boundary* b);
/* Export the index.  Arrays returned stay valid until the boundary is destroyed. */
int mq_boundary_index(const mq_boundary* b, mq_bindex_desc* out);
/* Copy cohort ci to the host (n_blocks*ls*rs doubles). */
int mq_boundary_download(const mq_boundary* b, uint32_t ci, double* host);
/* Device pointer of cohort ci. */
int mq_boundary_cohort_ptr(const mq_boundary* b, uint32_t ci, const double** dev);
/* Residency of a boundary, in place of the reference's storage::gpu / storage::disk boundary
 * cache (utils/storage.h:87-200 evict / prefetch / fetch, 271-539 disk spill; driven from
 * optimize/ts_optimize.hpp:153-182,307-345).
 * mq_boundary_evict: the device copy is freed after queued work on the context's stream has run;
 * the first eviction copies the data to pinned host memory (a boundary never changes after it is
 * made, so that copy stays valid).  With path != NULL the host copy is also written to that file
 * (once) and freed.  A boundary that a live sigma plan reads cannot be evicted (MQ_ERR_INVALID).
 * mq_boundary_prefetch: allocates the device copy and starts the host -> device copy (reading the
 * spill file first if there is no host copy) on the context's copy stream; it returns at once,
 * and work that reads the boundary afterwards waits for the copy on the device.  Every entry
 * point that reads an evicted boundary (plan creation, steps, download, traces, cohort_ptr)
 * prefetches it first.  mq_boundary_residency reports where the data is and the bytes held on the
 * device and in pinned host memory.  Spill files are deleted with the boundary. */
enum { MQ_RESIDENT_DEVICE = 0, MQ_RESIDENT_HOST = 1, MQ_RESIDENT_DISK = 2 };
int mq_boundary_evict(mq_ctx* ctx, mq_boundary* b, const char* path);
int mq_boundary_prefetch(mq_ctx* ctx, mq_boundary* b);
int mq_boundary_residency(const mq_boundary* b, int* where, uint64_t* device_bytes, uint64_t* host_bytes);
/* Boundary::traces() (mp_tensors/boundary.h:339-350): per bond index b, the sum of all elements
 * of every block b; out has aux_dim entries. */
int mq_boundary_traces(const mq_boundary* b, double* out);

/* ---------------------------------------------------------------- SU2 couplings */
/* SU2::set_coupling (block_matrix/symmetry/gsl_coupling.h:72-92): the four 9j-based coupling
 * factors (cases 0..3 = site spins (0,0), (2,0), (0,2), (2,2)) for doubled spins a..i, times init. */
int mq_su2_coupling(int a, int b, int c, int d, int e, int f, int g, int h, int i, double init, double* out4);
/* Wigner 9j symbol of doubled spins (gsl_sf_coupling_9j semantics). */
double mq_wigner9j(int a, int b, int c, int d, int e, int f, int g, int h, int i);

/* ---------------------------------------------------------------- schedules */
/* Schedule builders run on the host only (no device needed); they take the boundary indices.
 * create_contraction_schedule (site_hamil_schedule.hpp:40-93): right_tasks + site_hamil_tasks. */
int mq_schedule_build_sigma(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                            const mq_bindex_desc* right, mq_schedule** out);
/* Schedule of overlap_mpo_left_step (move_boundary.hpp:147-213; left_tasks.hpp:38-172). */
int mq_schedule_build_left(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                           int symmetric, mq_schedule** out);
/* Schedule of overlap_mpo_right_step (move_boundary.hpp:281-337; right_tasks.hpp:39-167). */
int mq_schedule_build_right(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* right,
                            int symmetric, mq_schedule** out);
/* Index-only forms of the two step schedules: the result holds the new boundary's index
 * (mq_schedule_out_index, identical to the full builders') and no task lists.  They let a caller
 * lay out boundaries along a chain without building the steps' schedules.  No reference
 * counterpart (the reference always builds the full schedule, move_boundary.hpp:147-404). */
int mq_schedule_build_left_index(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* left,
                                 int symmetric, mq_schedule** out);
int mq_schedule_build_right_index(const mq_site_desc* site, const mq_mpo_desc* mpo, const mq_bindex_desc* right,
                                  int symmetric, mq_schedule** out);
/* Index of the boundary a left/right-step schedule produces (BoundaryIndex after add_cohort and
 * complement_transpose, move_boundary.hpp:202-209, 331-337). */
int mq_schedule_out_index(const mq_schedule* s, mq_bindex_desc* out);
int mq_schedule_export(const mq_schedule* s, mq_schedule_desc* out);
/* Flop count of the schedule with the reference's model (tasks.cpp:388-398, 777-796). */
int mq_schedule_flops(const mq_schedule* s, uint64_t* flops);
int mq_schedule_destroy(mq_schedule* s);

/* ---------------------------------------------------------------- sigma */
/* super_hamil_mv replacement: plan for H_eff * psi at one site.  The plan keeps pointers to the
 * device data of left/right; they must outlive the plan. */
int mq_sigma_plan_create(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                         const mq_boundary* left, const mq_boundary* right, mq_plan** out);
/* Plan from a caller-built ScheduleNew (the drop-in seam: the caller keeps its own
 * create_contraction_schedule).  ket_rows[lb] = ls of psi block lb; ket_cols[lb] = its columns. */
int mq_sigma_plan_create_from_schedule(mq_ctx* ctx, const mq_schedule_desc* sched,
                                       uint32_t n_ket_blocks, const uint32_t* ket_rows,
                                       const uint32_t* ket_cols, const mq_boundary* left,
                                       const mq_boundary* right, mq_plan** out);
/* Sigma algorithms.  SCHEDULE executes the reference's ScheduleNew tables (T GEMMs, alpha-list
 * gather, L GEMMs).  GROUPED groups sector pairs by charge transfer (and, for SU2U1, by the spin
 * of the pair's left sector, which fixes every 9j coupling of the group) and performs the MPO-term
 * sum as one dense MFMA GEMM per group; same result, no per-sector alpha lists.  AUTO = GROUPED. */
enum mq_sigma_algorithm { MQ_SIGMA_AUTO = 0, MQ_SIGMA_SCHEDULE = 1, MQ_SIGMA_GROUPED = 2 };
/* OR-ed into the algorithm of mq_sigma_plan_create_ex: keep what mq_sigma_adjoint needs (the map
 * from MPO terms to W_D entries and the adjoint task tables).  GROUPED plans only. */
enum { MQ_SIGMA_ADJOINT = 0x100 };
int mq_sigma_plan_create_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                            const mq_boundary* left, const mq_boundary* right, int algorithm, mq_plan** out);
/* sigma = H_eff * psi; psi and sigma are device pointers in the padded Davidson-vector layout. */
int mq_sigma_apply(mq_plan* plan, const double* psi, double* sigma);
/* Arithmetic of the sigma contraction.  psi, sigma, boundaries stay FP64 in memory in every mode.
 *   F64     FP64 MFMA throughout (mq_sigma_apply).
 *   F32     T = psi R, S = T W and sigma = L S on FP32 MFMA with FP32 accumulation; T and S are
 *           held in FP32, the final accumulation into sigma is FP64.
 *   BF16X3  as F32, each product formed from the bf16 split x = hi + lo as hi*hi + hi*lo + lo*hi.
 * Reduced precisions need a GROUPED plan.  No reference counterpart: the reference's sigma is
 * FP64 only (solver/tasks.cpp). */
enum mq_precision { MQ_PREC_F64 = 0, MQ_PREC_F32 = 1, MQ_PREC_BF16X3 = 2 };
int mq_sigma_apply_ex(mq_plan* plan, const double* psi, double* sigma, int precision);
/* Graph replay of applications (default on).  Each precision has one HIP graph of the whole
 * application over plan-owned operand slots: psi is copied into the input slot, the graph is
 * replayed, and the output slot is copied to sigma, so every application (any operands, e.g. each
 * Jacobi-Davidson iteration) replays it.  The first application of a precision runs eagerly, the
 * second is captured on the context's stream.  enable = 0 issues every launch directly on the
 * caller's operands and drops the graphs.  No reference counterpart (the reference issues cuBLAS
 * calls per block, solver/tasks.cpp). */
int mq_plan_set_graphs(mq_plan* plan, int enable);
/* Adjoint of the sigma contraction with respect to the MPO coefficients (the gradient with respect
 * to the reference's alpha list, tasks.cpp:466-494 and op_iterate.hpp:58-63, taken per MPO term):
 *   grad[t] = d(psi . H_eff psi) / d term_scale[t] = psi . (L_b1 (x) op_t (x) R_b2) psi
 * for every term t of the plan's mq_mpo_desc, in its order.  For a normalized eigenvector of H_eff
 * this is dE/d scale_t (Hellmann-Feynman); summed over the sites where an integral enters the MPO
 * it gives dE/dh_ij and dE/d(ij|kl), the 1- and 2-RDM elements.  psi: device pointer (padded
 * layout); grad: host array of n_terms doubles.  FP64.  The plan must have been created with
 * MQ_SIGMA_ADJOINT. */
int mq_sigma_adjoint(mq_plan* plan, const double* psi, double* grad);
/* Autotune the GEMM tile configurations of the plan's T and final stages for `precision`.  Tasks
 * fall into row-size classes (M <= 16, 32, 64, 128, larger); for each class present, candidate tile
 * shapes are timed on the plan's own stage (one warm-up and three repetitions) and the fastest is
 * kept.  Decisions are cached in the context, keyed by stage, precision and the stage's flop
 * histogram over the classes (log2 buckets): a plan whose histogram was tuned before applies the
 * cached choice without timing.
 * In FP32 and BF16x3 the S-stage library GEMMs (hipBLASLt) are tuned too: up to 8 library
 * candidates per GEMM are timed and the fastest kept, cached by exact shape and, as a library
 * solution index, by shape class (half-octave buckets of M, N, K).  Any later GEMM of that class,
 * in any plan of the context, adopts the class's solution when the library reports that it
 * supports the shape; such GEMMs are not timed again.
 * precision | MQ_AUTOTUNE_LIBRARY_ONLY tunes only the library GEMMs not already covered by the
 * caches (the cheap per-site form for sweeps: after the first sites nearly every GEMM is covered).
 * timed (may be NULL) receives the number of stage timings made.  Captured graphs of the plan are
 * dropped.  No reference counterpart. */
enum { MQ_AUTOTUNE_LIBRARY_ONLY = 0x10 };
int mq_plan_autotune(mq_plan* plan, int precision, int* timed);
/* n independent applications (plans of one context, e.g. the sites of independent chains) issued
 * as one batch: each instance on one of four side streams forked from and joined back to the
 * context's stream, so small instances run concurrently.  The batch of a given (plans, precision)
 * runs eagerly on first use and is captured into one HIP graph on the second, replayed after.
 * psi[i] / sigma[i] are device vectors of plans[i].  Destroying a plan drops the batch graphs that
 * contain it.  No reference counterpart (the reference applies one site Hamiltonian at a time). */
int mq_sigma_apply_batched(mq_ctx* ctx, int n, mq_plan* const* plans, const double* const* psi,
                           double* const* sigma, int precision);
/* Measurement: runs `reps` sigma applications stage by stage with HIP events on the context's
 * stream and returns the average milliseconds of each stage: [0] S zero-fill, [1] T-stage GEMM,
 * [2] S gather, [3] final GEMM.  flops_out (may be NULL) receives the algorithmic flops of the
 * three compute stages (T GEMM, gather, final GEMM) and gather_bytes (may be NULL) the
 * algorithmic bytes of the gather (T slab reads + S writes). */
int mq_plan_profile(mq_plan* plan, const double* psi, double* sigma, int reps, int precision,
                    double* stage_ms, uint64_t* flops_out, uint64_t* gather_bytes);
int mq_plan_vector_size(const mq_plan* plan, uint64_t* n_elements);
int mq_plan_flops(const mq_plan* plan, uint64_t* flops);
/* Bytes of device workspace held by the plan (T and S buffers and tables). */
int mq_plan_workspace_bytes(const mq_plan* plan, uint64_t* bytes);
int mq_plan_destroy(mq_plan* plan);

/* ---------------------------------------------------------------- boundary steps */
/* overlap_mpo_left_step: bra_lp is the bra in the left-paired layout, ket_rp the ket in the
 * right-paired layout (both device pointers of the same site tensor shape). */
int mq_boundary_left_step(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                          const double* bra_lp, const double* ket_rp, const mq_boundary* left,
                          int symmetric, mq_boundary** out);
/* overlap_mpo_right_step: bra_rp and ket_rp both right-paired. */
int mq_boundary_right_step(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                           const double* bra_rp, const double* ket_rp, const mq_boundary* right,
                           int symmetric, mq_boundary** out);
/* As mq_boundary_left_step / mq_boundary_right_step with the algorithm chosen (mq_sigma_algorithm).
 * GROUPED (abelian groups): the new boundary's index comes from the index-only builder; the step
 * is T (psi R, or L psi), S_D = T_D W_D (one library GEMM per charge-transfer group) and the
 * contraction with the bra, as in the grouped sigma, without the per-sector alpha lists; for SU2U1
 * the pairs are grouped by charge transfer and left-sector spin and each W_D carries the 9j
 * couplings.  The plain entry points use AUTO (= GROUPED). */
int mq_boundary_left_step_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                             const double* bra_lp, const double* ket_rp, const mq_boundary* left,
                             int symmetric, int algorithm, mq_boundary** out);
int mq_boundary_right_step_ex(mq_ctx* ctx, const mq_site_desc* site, const mq_mpo_desc* mpo,
                              const double* bra_rp, const double* ket_rp, const mq_boundary* right,
                              int symmetric, int algorithm, mq_boundary** out);

/* ---------------------------------------------------------------- Davidson */
/* Device-resident vector kernels (DavidsonVector ops, solver/davidson_vector.cpp:197-254). */
int mq_vec_dot(mq_ctx* ctx, uint64_t n, const double* x, const double* y, double* result);
int mq_vec_nrm2(mq_ctx* ctx, uint64_t n, const double* x, double* result);
int mq_vec_axpy(mq_ctx* ctx, uint64_t n, double a, const double* x, double* y);
int mq_vec_scal(mq_ctx* ctx, uint64_t n, double a, double* x);
int mq_vec_copy(mq_ctx* ctx, uint64_t n, const double* x, double* y);
/* Batched dot products result[j] = V_j . x, j < k, with V_j = V + j * ldv (k device vectors in one
 * buffer, as the Jacobi-Davidson basis is stored); result is host memory.  Replaces the k separate
 * ietl::dot calls of one Gram-Schmidt pass or one subspace-matrix column (ietl/jacobi.h:408-427).
 * The partial sums live in a context buffer that grows with k.  Fixed reduction order. */
int mq_vec_multi_dot(mq_ctx* ctx, uint64_t n, int k, const double* V, uint64_t ldv, const double* x,
                     double* result);
/* Allocation-free form of mq_vec_multi_dot.  No counterpart in the reference: it is for callers
 * that cannot allocate device memory (code under HIP graph capture, tight serving loops, harnesses
 * that forbid hipMalloc).  It calls no hipMalloc / hipFree and no allocating library routine, and
 * does not synchronise: result is DEVICE memory (k doubles), written in stream order on the
 * context's stream.  Its partial sums live in a static per-GPU scratch of MQ_NOALLOC_MAXK x 512
 * doubles (256 KiB, held by every GPU the library runs on until the process ends); k above
 * MQ_NOALLOC_MAXK is rejected with MQ_ERR_INVALID.  All calls on one GPU share that scratch, so calls
 * that overlap in time on one GPU (other contexts' streams, other threads) race; calls ordered on
 * one stream do not.  Same kernels and reduction order as mq_vec_multi_dot: identical results. */
#define MQ_NOALLOC_MAXK 64
int mq_vec_multi_dot_noalloc(mq_ctx* ctx, uint64_t n, int k, const double* V, uint64_t ldv, const double* x,
                             double* result);

/* Jacobi-Davidson smallest eigenpair of H_eff (ietl/jacobi.h:386-516, solver/ietl_jacobi_davidson.h:35-76).
 * init and out are device vectors in the plan's layout (may alias).  gmres_iter == 0 uses the
 * plain correction t = -r + (r.u)/(u.u) u (ietl/jacobi.h:150-156); gmres_iter > 0 refines it with
 * GMRES (at most gmres_iter Krylov vectors, ietl/gmres.h:78-140) on the projected correction
 * equation (1 - uu^T)(H - theta)(1 - uu^T) t = -r to the tolerance 2^-(k+1) at step k
 * (jcd_gmres_solver, ietl/jacobi.h:126-164; the reference's parameter ietl_jcd_gmres). */
int mq_jd_solve(mq_plan* plan, const double* init, double* out, double tol, int max_iter,
                int gmres_iter, double* energy, int* iterations);
/* mq_jd_solve with ortho vectors: the search space is kept orthogonal to the n_ortho device vectors
 * ortho[0..n_ortho) (plan layout), as the reference's solve(..., ortho_vecs, ...)
 * (solver/solver.h:65-69, solver/ietl_jacobi_davidson.h:36-47) with DavidsonVS::project
 * (solver/ietl_interface_dv.h:94-99): the vectors are Gram-Schmidt'ed among themselves (internal
 * copies; the caller's are not modified), then projected out of the start vector, of every new
 * search vector and of H u before the residual.  Finds the lowest eigenpair in their complement
 * (excited states).  At most vec_size - 1 of them are used. */
int mq_jd_solve_ex(mq_plan* plan, const double* init, double* out, double tol, int max_iter,
                   int gmres_iter, int n_ortho, const double* const* ortho, double* energy, int* iterations);
/* Jacobi-Davidson as iterative refinement over a reduced-precision sigma.  Each outer step runs
 * the Jacobi-Davidson iteration above with sigma in `precision` (the first search vector of an
 * outer step after the first carries its exact FP64 product), then recomputes the Ritz vector's
 * product, Rayleigh quotient and residual H u - E u with the FP64 sigma; convergence is tested on
 * that FP64 residual with the rule of ietl/iteration.h:56-58.  Orthogonalisation, subspace
 * matrix and eigenproblem stay FP64.  max_iter bounds the reduced-precision sigma calls of one
 * outer step, max_outer the outer steps.  counts (may be NULL) receives {reduced-precision sigma
 * calls, FP64 sigma calls, outer steps}; residual (may be NULL) the final FP64 residual norm.
 * precision == MQ_PREC_F64 gives the plain solver with one extra FP64 check per outer step. */
int mq_jd_solve_refined(mq_plan* plan, const double* init, double* out, double tol, int max_iter,
                        int precision, int max_outer, double* energy, int* counts, double* residual);

#ifdef __cplusplus
}
#endif

#endif /* MQ_SIGMA_H */
