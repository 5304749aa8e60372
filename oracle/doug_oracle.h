/*
 * doug_oracle.h -- CPU restatement of DOUG's two-level Schwarz PCG solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the checker that the MI355X library is
 * compared against; it is never linked into, loaded by or called from the
 * product library.  Only tests/, __graft_entry__.smoke() and the cpu_baseline
 * leg of bench.py may use it.
 *
 * All index arguments are 1-based, exactly as in the reference (Fortran
 * default INTEGER); all arithmetic is FP64 with no fused multiply-add, as a
 * default x86-64 build of the reference computes it.
 *
 * Pinning status of each part is recorded in the header of doug_oracle.c.
 */
#ifndef DOUG_ORACLE_H
#define DOUG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* Control values on the path (names and defaults as src/utils/globals.F90:85-116;
 * -1 means "not set in DOUG.ctl"). */
typedef struct {
  int radius1, radius2, minasize1, minasize2, maxasize1, maxasize2;
  double strong1, strong2;
  int smoothers, overlap, levels, method, fine_method;
  int symmstruct, symmnumeric;
  /* prolongator damping; the reference hard-codes 0.33333333333333
   * (src/coarse/CoarseMtx.F90:64) */
  double omega;
  /* symmetric Gauss-Seidel sweeps of fine_method 2 (ctl num_iters; <= 0
   * means 3, src/main/aggr.F90:220-222) */
  int num_iters;
} or_params;

typedef struct or_state or_state;

void or_params_default(or_params *p);

/* Reset the variables the reference keeps across calls inside one process:
 * `allmind` (src/datatypes/Aggregate_mod.F90:348) and the initialised local
 * `beta` of SpMtx_aggregate (SpMtx_aggregation.F90:83). */
void or_reset_process(void);

/* doug_aggr setup for one process (src/main/aggr.F90:129-219, numprocs==1):
 * fine aggregation, smoothed restriction, Galerkin coarse matrix, coarse
 * aggregation (or an injected partition when nparts>0), overlapping subdomain
 * index sets and the exact subdomain / coarse factorisations.
 * Returns NULL on error (message via or_last_error()). */
or_state *or_setup(int n, long nnz, const int *indi, const int *indj,
                   const double *val, const or_params *p, int nparts,
                   const int *pstarts, const int *pnodes);
/* Several processes: the per-rank decomposition of a DOUG run on nranks
 * MPI processes (rank_of_node[i] in 1..nranks, the METIS part of unknown i),
 * each rank aggregating its own rows with fresh process state and keeping
 * one subdomain (nsub_per_rank <= 1); the solve then uses the assembled
 * matrix with the rank-numbered aggregates and subdomains.  When lg_out is
 * not NULL, lg_out[r] receives a malloc'd copy of rank r's local-to-global
 * map (nlf_out[r] entries, ninner_out[r] of them its own). */
or_state *or_setup_ranks(int n, long nnz, const int *indi, const int *indj, const double *val,
                         const or_params *p, int nranks, const int *rank_of_node,
                         int nsub_per_rank, int **lg_out, int *nlf_out, int *ninner_out);
/* For the next or_setup of this thread only: strength flags per input entry
 * (nnz of them, NULL = computed) replacing SpMtx_find_strong's, and a
 * restriction given as rnnz 1-based triples with rnagr rows (rnagr = 0 =
 * built) replacing the smoothed aggregation's.  No counterpart in the
 * reference; it checks the library's injection entry points.  The arrays
 * must stay alive until or_setup returns. */
void or_set_injection(long nnz, const unsigned char *strong, int rnagr, long rnnz,
                      const int *ri, const int *rj, const double *rv);
void or_free(or_state *s);
/* The fine strength flags of a one-process or_setup, one per input entry in
 * input order (SpMtx_find_strong's A%strong, or the injected flags); returns
 * their count (0 for a several-process state).  strong may be NULL. */
long or_get_strong(const or_state *s, unsigned char *strong);
void or_free_ptr(void *q);
const char *or_last_error(void);
/* worker threads of the parallel loops (ranks, subdomain factorisations and
 * solves; OMP_NUM_THREADS) */
int or_num_threads(void);

/* sizes */
int or_n(const or_state *s);
int or_nagr_fine(const or_state *s);       /* fAggr%inner%nagr */
int or_nagr_coarse(const or_state *s);     /* cAggr%full%nagr (0 if injected) */
int or_nsub(const or_state *s);
long or_sub_total(const or_state *s);      /* sum of subdomain sizes */
long or_R_nnz(const or_state *s);
long or_AC_nnz(const or_state *s);
int or_overlap(const or_state *s);         /* ol actually used */
int or_smoothers(const or_state *s);       /* smoothers after clamping */

/* getters (caller allocates) -- all 1-based */
void or_get_fine_aggr(const or_state *s, int *num, int *starts, int *nodes);
void or_get_coarse_aggr(const or_state *s, int *num, int *starts, int *nodes);
void or_get_subdomains(const or_state *s, int *starts, int *inds);
/* R in its final arrangement order (by columns) */
void or_get_R(const or_state *s, int *indi, int *indj, double *val);
/* A_c in its final order (row arranged, first-touch column order) */
void or_get_AC(const or_state *s, int *indi, int *indj, double *val);
/* the system matrix after the scale/unscale round trip, in row-arranged order */
long or_A_nnz(const or_state *s);
void or_get_A(const or_state *s, int *indi, int *indj, double *val);

/* operators */
void or_spmv(const or_state *s, const double *x, double *y);
void or_prec_apply(or_state *s, const double *r, double *z);
/* pcg_weigs: returns iterations; x holds x0 on entry (zero in the driver) */
int or_pcg(or_state *s, const double *b, double *x, double tol, int maxit,
           double *cond, double *res_hist);

/* stand-alone pieces used for pinning */
/* SpMtx_find_strong + SpMtx_aggregate (SpMtx_aggregation.F90:49-641,829-899)
 * on a given 1-based COO matrix, inside the current process state (see
 * or_reset_process): has_fine != 0 makes it a coarse-level call.  symm
 * selects symmetric scaling.  num_inner/num_full receive the n aggregate
 * numbers (0 = none); returns the inner aggregate count. */
int or_aggregate(int n, long nnz, const int *indi, const int *indj, const double *val,
                 int symm, int radius, int minasize, int maxasize, double alpha, int has_fine,
                 int *num_inner, int *num_full);
/* CSR SpMV summing each row in stored order, y_i = ((0 + a1*x1) + a2*x2) ... */
void or_spmv_csr(int n, const int *rowptr0, const int *col0, const double *val,
                 const double *x, double *y);
/* Lanczos estimate from CG coefficients (pcg.F90:465-497) followed by EISPACK
 * tql1; alpha/beta are 1-based arrays of length it (beta(1) unused);
 * eig receives the `it` eigenvalues in ascending order; returns ierr. */
int or_lanczos_eigs(int it, const double *alpha, const double *beta, double *eig);
/* tql1 alone on a symmetric tridiagonal matrix (d[0..n-1], e[1..n-1] as
 * subdiagonal with e[0] ignored); returns ierr */
int or_tql1(int n, double *d, double *e);
/* Add_layers (src/datatypes/Decomposition.f90:56-99) on a 1-based CSR graph;
 * nodes must have room for the result; returns the new count. */
int or_add_layers(const int *bounds, const int *adj, int n, int *nodes,
                  int nnodes, int nlayers);

#ifdef __cplusplus
}
#endif
#endif
