/*
 * doug.h -- C-ABI of the MI355X (gfx950) implementation of DOUG's solve path:
 * PCG preconditioned by two-level additive Schwarz with a smoothed-aggregation
 * coarse grid (doug_aggr with solver 2, method 1, levels 2, fine_method 1,
 * coarse_method 1).  One context = one DOUG process (numprocs == 1) with all
 * of its subdomains resident on one GPU.
 *
 * Conventions (kept from the reference, src/datatypes/SpMtx/SpMtx_class.F90
 * and src/utils/globals.F90):
 *   - scalars are int32 / double; index arrays are int32 with a caller-chosen
 *     base (1 = Fortran numbering as in DOUG's files, 0 = C numbering);
 *   - the caller owns every host array; the library copies what it keeps;
 *   - every entry point returns DOUG_OK (0) or an error code; the message is
 *     available from doug_last_error().  DOUG's own behaviour on any error is
 *     DOUG_abort -> MPI_ABORT (src/utils/DOUG_utils.F90:254-286): a caller
 *     wanting that behaviour aborts on a non-zero status;
 *   - all GPU work runs on the context's stream (default priority); the
 *     *_device entry points take device pointers and return after the work
 *     has completed on that stream.
 */
#ifndef DOUG_H
#define DOUG_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  DOUG_OK = 0,
  DOUG_E_ARG = 1,      /* invalid argument */
  DOUG_E_STATE = 2,    /* call out of order (e.g. solve before setup) */
  DOUG_E_HIP = 3,      /* HIP runtime error */
  DOUG_E_NUMERIC = 4,  /* non-positive pivot in a subdomain or coarse factor */
  DOUG_E_NOMEM = 5,    /* device allocation failed */
  DOUG_E_UNSUPPORTED = 6
};

/* Precision of the subdomain factors.  The outer PCG (residuals, dot
 * products, SpMV), the restriction/prolongation and the coarse factor are
 * FP64 in both forms; with DOUG_PREC_FP32_FACTORS the subdomain Cholesky
 * factors are computed and applied in FP32 (matrix cores, FP32
 * accumulation) and the FP64 Krylov iteration refines the result to the
 * requested FP64 tolerance. */
enum { DOUG_PREC_FP64 = 0, DOUG_PREC_FP32_FACTORS = 1 };

typedef struct doug_ctx doug_ctx;

/* Control values on the path.  Field names and the meaning of -1 ("not set")
 * follow DOUG.ctl (src/utils/globals.F90:85-116, src/utils/controls.F90). */
typedef struct {
  int32_t radius1, radius2, minasize1, minasize2, maxasize1, maxasize2;
  double strong1, strong2;
  int32_t smoothers, overlap, levels, method, fine_method;
  int32_t symmstruct, symmnumeric;
  /* prolongator damping: DOUG hard-codes 0.33333333333333
   * (src/coarse/CoarseMtx.F90:64); other values are a documented extension */
  double omega;
  /* symmetric Gauss-Seidel sweeps of fine_method 2 (ctl num_iters; -1 or
   * less than 1 means 3, src/main/aggr.F90:220-222) */
  int32_t num_iters, reserved;
} doug_params;

/* Sizes and timings of the last doug_setup (seconds).  factor_bytes and
 * factor_tiles count the stored subdomain factor tiles (64x64, inside each
 * factor's envelope); factor_flops is the matrix-core work of their
 * factorisation. */
typedef struct {
  int32_t n, nagr_fine, nagr_coarse, nsub, ncoarse, overlap, smoothers;
  int64_t nnz, sub_total, factor_bytes, coarse_factor_bytes;
  double t_aggregate, t_coarse_build, t_subdomains, t_upload, t_factor, t_coarse_factor;
  double t_total;
  int64_t factor_tiles;
  double factor_flops;
  /* device bytes the context keeps after the set-up for the next one (the
   * coarse build's temporaries; 0 with DOUG_OPT_KEEP_SETUP_TEMPS 0) */
  int64_t temp_bytes;
} doug_info;

/* Context.  Replaces the process-wide state DOUG_Init sets up
 * (src/utils/DOUG_utils.F90:367-429); `device` is a HIP device ordinal, or
 * -1 for a host-only context that can run doug_setup_host and the getters
 * but no GPU work. */
int doug_ctx_create(int device, doug_ctx **out);
int doug_ctx_destroy(doug_ctx *ctx);
const char *doug_last_error(const doug_ctx *ctx);

/* Defaults identical to an empty DOUG.ctl for the keys above, plus
 * levels 2, method 1, fine_method 1, omega 0.33333333333333. */
void doug_params_default(doug_params *p);
int doug_set_params(doug_ctx *ctx, const doug_params *p);

/* Assembled matrix as COO triples, duplicates allowed (SpMV adds them).
 * Replaces ReadInSparseAssembled + Distribution_NewInit for one process
 * (src/datatypes/SpMtx/SpMtx_class.F90:410-453,
 *  src/components/Distribution.F90:50-75). */
int doug_set_matrix_coo(doug_ctx *ctx, int32_t n, int64_t nnz, const int32_t *indi,
                        const int32_t *indj, const double *val, int32_t index_base);

/* Optional non-overlapping partition of the unknowns into subdomains; takes
 * the place of METIS in Partitionings_metis_InitCoarse
 * (src/components/Partitioning_metis.F90:34-86).  Without it the subdomains
 * are the coarse aggregates expanded to fine nodes, DOUG's one-process rule
 * (src/components/Partitioning_aggr.F90:120-193).  starts has nparts+1
 * entries (starts[0] == index_base); nodes lists each part's unknowns. */
int doug_set_partition(doug_ctx *ctx, int32_t nparts, const int32_t *starts,
                       const int32_t *nodes, int32_t index_base);

/* Subdomains from `nsub` parts of the fine-aggregate adjacency graph, the
 * subdomain source of DOUG's multi-process runs: Partitionings_metis_InitCoarse
 * (src/components/Partitioning_metis.F90:34-86) over SpMtx_buildAggrAdjncy
 * (src/datatypes/SpMtx/SpMtx_arrangement.F90:2558-2639), with the project
 * partitioner (doug_partition_graph) in place of METIS; each subdomain lists
 * its nodes in node order.  Used when no partition is set and nsub > 1;
 * 0 or 1 restores the one-process rule. */
int doug_set_num_subdomains(doug_ctx *ctx, int32_t nsub);

/* Fine aggregates given by the caller in place of Partitionings_aggr_InitFine's
 * (src/components/Partitioning_aggr.F90:33-117), e.g. the aggregates a DOUG
 * run lists with `verbose` 5 or more (src/datatypes/Aggregate_mod.F90:115-120):
 * num[i] is the aggregate (1..nagr) of
 * unknown i, or 0 for an unknown no aggregate holds; every aggregate must
 * hold an unknown.  The full aggregates add the others one by one in node
 * order (src/datatypes/SpMtx/SpMtx_aggregation.F90:562-575).  The smoothed
 * restriction, the coarse matrix and (without other subdomains) the coarse
 * aggregation follow from them.  nagr <= 0 restores the aggregation. */
int doug_set_aggregates(doug_ctx *ctx, int32_t nagr, const int32_t *num);

/* Strength flags of the fine matrix given by the caller in place of
 * SpMtx_find_strong's (src/datatypes/SpMtx/SpMtx_aggregation.F90:829-899,
 * the A%strong a DOUG run holds after it): one flag per COO entry of
 * doug_set_matrix_coo, in its input order, nonzero = strong.  They select
 * the aggregation's strong graph and the entries the prolongator smoother
 * keeps (src/coarse/CoarseMtx.F90:164-220).  Not with doug_set_ranks, whose
 * ranks compute their own.  strong == NULL restores the computed flags. */
int doug_set_strong_mask(doug_ctx *ctx, int64_t nnz, const uint8_t *strong);

/* Restriction R (nagr x n) given by the caller as COO triples in place of the
 * smoothed aggregation's (IntRestBuild,
 * src/coarse/CoarseMtx.F90:50-474): sorted by (row, column) with a repeated
 * pair summed in input order, as SpMtx_consolidate does
 * (src/datatypes/SpMtx/SpMtx_arrangement.F90:258-292).  The coarse matrix
 * is R A R^T (CoarseMtxBuild, src/coarse/CoarseMtx.F90:476-516).  Subdomains
 * from the coarse aggregation need nagr equal to the fine aggregate count
 * (its coarse unknowns are the fine aggregates); with a partition, given
 * subdomains or doug_set_num_subdomains any nagr works.  Not with
 * doug_set_ranks.  nagr <= 0 restores the built restriction. */
int doug_set_restriction_coo(doug_ctx *ctx, int32_t nagr, int64_t nnz, const int32_t *ri,
                             const int32_t *rj, const double *rv, int32_t index_base);

/* Final subdomain index sets, overlap included, in place of
 * FinePreconditioner_Init's (src/components/Preconditioner_base.F90:96-123),
 * e.g. Decomposition%subd(:)%inds of a DOUG run: starts has nsub+1 entries
 * (starts[0] == index_base), inds lists each set (non-empty, an unknown at
 * most once per set; sets may overlap).  Takes precedence over
 * doug_set_partition and doug_set_num_subdomains.  nsub <= 0 restores them. */
int doug_set_subdomains(doug_ctx *ctx, int32_t nsub, const int32_t *starts, const int32_t *inds,
                        int32_t index_base);

/* Virtual ranks: the decomposition of a DOUG run on `nranks` MPI processes,
 * restated on one host and solved on one GPU.  rank_of_node[i] (index_base
 * 0 or 1) is the rank that owns unknown i, M%eptnmap of
 * SpMtx_DistributeAssembled (src/components/Distribution_assm.F90:105-270);
 * NULL partitions the matrix graph with the project partitioner, in place of
 * METIS on SpMtx_buildAdjncy (Distribution_assm.F90:191-197).  At setup each
 * rank builds DOUG's local numbering (SpMtx_localize/SpMtx_build_ghost/
 * SpMtx_Build_lggl, src/datatypes/SpMtx/SpMtx_distribution.F90:139-290,
 * SpMtx_arrangement.F90:744-1710), aggregates its own rows
 * (Partitioning_aggr.F90:81-92), splits them into doug_set_num_subdomains
 * parts (one when unset) and adds the overlap layers over its local matrix;
 * aggregates and subdomains are then numbered rank by rank.  A negative
 * radius1 is negated as the reference does with several processes.
 * Requires max(overlap, smoothers) >= 1 and no doug_set_partition.
 * nranks <= 0 restores the one-process setup. */
int doug_set_ranks(doug_ctx *ctx, int32_t nranks, const int32_t *rank_of_node,
                   int32_t index_base);
/* After setup with virtual ranks: rank `rank`'s (0-based) local numbering,
 * M%lg_fmap (1-based global numbers, *nlocal of them, the first *ninner its
 * own).  lg may be NULL to query the sizes. */
int doug_get_rank_map(const doug_ctx *ctx, int32_t rank, int32_t *ninner, int32_t *nlocal,
                      int32_t *lg);

/* The project graph partitioner, standing in for METIS_PartGraphKway
 * (called from src/datatypes/Graph.f90:295-315): recursive bisection by
 * breadth-first growth from a pseudo-peripheral vertex plus boundary
 * refinement, balancing the vertex weights (vwgt may be NULL) to within 2 %
 * of a part's share.  xadj/adjncy: symmetric adjacency without self loops,
 * indices in `index_base`; part[v] receives a part number in index_base ..
 * index_base + nparts - 1.  Host only, deterministic. */
int doug_partition_graph(int32_t n, const int32_t *xadj, const int32_t *adjncy,
                         const int32_t *vwgt, int32_t nparts, int32_t index_base,
                         int32_t *part);

/* Builds the preconditioner: Partitionings_aggr_InitFine,
 * CoarsePreconditioner_smooth_Init, Partitionings_aggr_InitCoarse,
 * FinePreconditioner_Init and the factorisations of Factorise_subdomains and
 * of the coarse matrix (src/main/aggr.F90:131-219,
 * src/solvers/subsolvers.F90:78-163).  precision: DOUG_PREC_FP64 or
 * DOUG_PREC_FP32_FACTORS. */
int doug_setup(doug_ctx *ctx, int32_t precision);
/* The host half of doug_setup alone: aggregates, restriction, coarse matrix
 * and subdomain index sets, without any device allocation or factorisation. */
int doug_setup_host(doug_ctx *ctx);
int doug_get_info(const doug_ctx *ctx, doug_info *info);

/* Setup products, 1-based as DOUG stores them (caller allocates):
 * aggregates (Aggregate_mod.F90:39-51): num[n], starts[nagr+1], nodes[...]. */
int doug_get_fine_aggregates(const doug_ctx *ctx, int32_t *num, int32_t *starts, int32_t *nodes);
int doug_get_coarse_aggregates(const doug_ctx *ctx, int32_t *num, int32_t *starts,
                               int32_t *nodes);
/* Decomposition%subd(:)%inds (Preconditioner_base.F90:96-123) */
int doug_get_subdomains(const doug_ctx *ctx, int32_t *starts, int32_t *inds);
/* restriction R (CoarseMtx.F90:50-474) as CSR, rows ascending in column;
 * coarse matrix A_c (CoarseMtx.F90:476-516) as CSR in DOUG's entry order */
int doug_get_restriction(const doug_ctx *ctx, int32_t *rowptr, int32_t *cols, double *vals);
int doug_get_coarse_matrix(const doug_ctx *ctx, int32_t *rowptr, int32_t *cols, double *vals);
int64_t doug_restriction_nnz(const doug_ctx *ctx);
int64_t doug_coarse_matrix_nnz(const doug_ctx *ctx);

/* The *_device entry points take device pointers and run on the context's
 * own stream, returning when it has drained; that stream does not wait for
 * work on other streams, so the inputs must be complete when they are called. */

/* y = A x (SpMtx_Ax / Distribution_pmvm, SpMtx_op_Ax.f90:44-101) */
int doug_spmv(doug_ctx *ctx, const double *x, double *y);
int doug_spmv_device(doug_ctx *ctx, const double *dx, double *dy);

/* z = M^{-1} r, the two-level additive Schwarz preconditioner
 * (`preconditioner` + prec2Level, src/solvers/pcg.F90:54-270) */
int doug_prec_apply(doug_ctx *ctx, const double *r, double *z);
int doug_prec_apply_device(doug_ctx *ctx, const double *dr, double *dz);

/* doug_prec_apply_device without any device allocation at or after the
 * call: the apply's work vectors (the subdomain vectors, the coarse vectors
 * and their partial sums) live in one static device array per GPU of
 * 4194304 doubles (32 MiB), held until the process ends; a set-up needing
 * more is refused with DOUG_E_ARG.  The results equal
 * doug_prec_apply_device's bit for bit.  Every call on one GPU shares that
 * array, so calls that overlap on one GPU (several threads, or contexts on
 * the same device) race.  No counterpart in the reference: it serves callers
 * that cannot allocate (code under HIP graph capture, allocation-free
 * serving loops). */
int doug_prec_apply_device_noalloc(doug_ctx *ctx, const double *dr, double *dz);

/* pcg_weigs (src/solvers/pcg.F90:289-462): x holds x0 on entry (DOUG's
 * driver passes zero) and the solution on return.  Stops when
 * ||r_k||^2 / ||r_0||^2 <= tol^2 or after maxit iterations (maxit == -1 means
 * 100000).  cond receives the Lanczos condition estimate; res_hist (optional,
 * maxit entries) the residual norm of every iteration, for the first 2^20
 * iterations.  The stopping test runs on the GPU; an iteration queued behind
 * it after it passed leaves x unchanged. */
int doug_pcg(doug_ctx *ctx, const double *b, double *x, double tol, int32_t maxit, int32_t *it,
             double *cond, double *res_hist);
int doug_pcg_device(doug_ctx *ctx, const double *db, double *dx, double tol, int32_t maxit,
                    int32_t *it, double *cond, double *res_hist);

/* Timing of one operator over `reps` back-to-back launches on the context's
 * stream (HIP events); ms receives the average milliseconds per launch.
 * which: 0 = SpMV, 1 = subdomain forward/backward solves, 2 = whole
 * preconditioner apply, 3 = copy of one n-vector (a known-size streaming
 * kernel for calibrating hardware byte counters), 4 = preconditioner apply
 * to 32 right-hand sides at once (needs doug_prepare_batch). */
int doug_time_operator(doug_ctx *ctx, int32_t which, int32_t reps, double *ms);

/* Launch-shape autotuner (no counterpart in the reference, whose solve path
 * has no launch shapes).  Times, with HIP events on the context's stream,
 * every SpMV shape -- 128, 256 or 512 rows per workgroup, each with and
 * without XCD-contiguous block order --, the dispatch order of the
 * subdomain solves (as partitioned, or largest first when sizes differ) and,
 * after doug_prepare_batch, the tiles per step of the batched solve,
 * `reps` launches each (<= 0: 20), and keeps the fastest.  All choices give
 * bit-identical results.  The choice stays with the context across later
 * doug_setup calls.  `out` may be NULL. */
typedef struct doug_tuning {
  int32_t spmv_shape;     /* 0..5: rows = 128 << (shape / 2); odd = XCD order */
  int32_t spmv_rows;      /* rows per workgroup of the chosen shape */
  int32_t spmv_xcd_order; /* 1: consecutive row blocks kept on one XCD */
  int32_t trsv_order;     /* 0 as partitioned, 1 largest subdomain first */
  double spmv_ms[6];      /* measured average per shape */
  double trsv_ms[2];      /* measured per order; 0 where not timed */
  int32_t trsm_tiles_per_step; /* batched envelope solve: 1 or 2 tiles staged per step */
  int32_t trsm_col_groups;     /* its workgroups per subdomain for 32 columns: 1 or 2 */
  double trsm_ms[2];      /* 32-column preconditioner apply per tiles-per-step choice, at
                             the chosen column groups; 0 where not timed (timed only
                             after doug_prepare_batch) */
} doug_tuning;
int doug_autotune(doug_ctx *ctx, int32_t reps, doug_tuning *out);
int doug_get_tuning(const doug_ctx *ctx, doug_tuning *out);

/* Execution options.  DOUG_OPT_ITERATION_GRAPH (default 1): from the second
 * PCG iteration on, replay the iteration's kernels (preconditioner, dots,
 * updates, SpMV, scalar read-back) from one captured HIP graph.
 * DOUG_OPT_SPMV_SHAPE (0..5) and DOUG_OPT_TRSV_ORDER (0/1) set what
 * doug_autotune would choose, see doug_tuning.  DOUG_OPT_LOCAL_ORDER
 * (default 1) places each subdomain's unknowns in its factor in reverse
 * Cuthill-McKee order, which narrows the stored tile envelope; 0 keeps the
 * index-set order.  It needs a new doug_setup and changes rounding only.
 * DOUG_OPT_DEVICE_COARSE (default 1): build the smoothed restriction and the
 * Galerkin coarse matrix on the GPU (same bits as the host build, which 0
 * selects); applies to smoothers = 1 at the next doug_setup.
 * DOUG_OPT_BATCH_TILES (1 or 2, default 2): off-diagonal tiles the batched
 * envelope solve stages per step (what doug_autotune sets).
 * DOUG_OPT_DEVICE_BLOCKS (default 1): take the subdomain factors' tile
 * envelopes and entries on the GPU from the uploaded matrix (same tiles as
 * the host scatter map, which 0 selects; a matrix holding a duplicate
 * (i, j) pair always uses the host map); at the next doug_setup.
 * DOUG_OPT_KEEP_SETUP_TEMPS (default 1): keep the coarse build's device
 * temporaries between set-ups (doug_info.temp_bytes), so that a repeated
 * set-up of the same size neither allocates nor frees them; 0 frees them
 * now and at the end of every set-up.
 * DOUG_OPT_BATCH_COLUMN_GROUPS (1, 2 or 3, default 1): layout of the
 * 32-column batched envelope solve: 1, one workgroup of four wavefronts per
 * subdomain; 2, two such workgroups of 16 columns each; 3, one workgroup of
 * eight wavefronts whose two halves take 16 columns each of one staged tile
 * (doug_autotune picks the fastest); the bits are the same in every layout. */
enum {
  DOUG_OPT_ITERATION_GRAPH = 1,
  DOUG_OPT_SPMV_SHAPE = 2,
  DOUG_OPT_TRSV_ORDER = 3,
  DOUG_OPT_LOCAL_ORDER = 4,
  DOUG_OPT_DEVICE_COARSE = 5,
  DOUG_OPT_BATCH_TILES = 6,
  DOUG_OPT_DEVICE_BLOCKS = 7,
  DOUG_OPT_KEEP_SETUP_TEMPS = 8,
  DOUG_OPT_BATCH_COLUMN_GROUPS = 9
};
int doug_set_option(doug_ctx *ctx, int32_t option, int64_t value);

/* Adjoint of the solve with respect to the matrix values.  For a loss L with
 * x = A^{-1} b and lambda = A^{-T} dL/dx (a second doug_pcg with A^T = A),
 * dL/da_ij = -lambda_i x_j.  dgrad (device, nnz entries) receives that value
 * for every entry of the doug_set_matrix_coo input, in the input's order
 * (duplicates each receive the full value).  No counterpart in the reference:
 * it serves differentiable callers (torch.autograd). */
int doug_value_grad_device(doug_ctx *ctx, const double *dlam, const double *dx, double *dgrad);

/* Several right-hand sides in one pass.  After doug_setup, doug_prepare_batch
 * allocates the multi-vector workspaces; the subdomain factors are the
 * single-column path's tile envelopes (L_i below the diagonal, inv(L_ii) on
 * it).  Each preconditioner apply is then one multi-vector restriction, two
 * chunked GEMMs on the coarse inverse factor, one launch of the batched
 * forward/backward envelope solve for every subdomain and all columns (tile x
 * panel products on the matrix cores, each factor tile read once per apply
 * whatever the number of columns) and one overlap sum.  Multi-vectors
 * are node-major: entry (i, r) at i*nrhs + r; 1 <= nrhs <= 32.  Every column
 * runs pcg_weigs's recurrence (pcg.F90:387-435) and stops on its own test;
 * its[r] receives the iteration count of column r, res_hist (optional,
 * maxit*nrhs) the residual norms. */
int doug_prepare_batch(doug_ctx *ctx);
int doug_prec_apply_batch_device(doug_ctx *ctx, int32_t nrhs, const double *dr, double *dz);
int doug_pcg_batch(doug_ctx *ctx, int32_t nrhs, const double *B, double *X, double tol,
                   int32_t maxit, int32_t *its, double *res_hist);
int doug_pcg_batch_device(doug_ctx *ctx, int32_t nrhs, const double *dB, double *dX, double tol,
                          int32_t maxit, int32_t *its, double *res_hist);

#ifdef __cplusplus
}
#endif
#endif
