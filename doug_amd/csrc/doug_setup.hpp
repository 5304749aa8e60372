// Host-side setup of DOUG's two-level Schwarz preconditioner (numprocs == 1):
// strength of connection, greedy aggregation, smoothed restriction, Galerkin
// coarse matrix, subdomain index sets.  Sequential, order-dependent graph work
// that decides *which* unknowns go together; every FP64 contraction that runs
// per iteration lives in the HIP kernels.
#pragma once
#include <climits>
#include <cstdint>
#include <memory>
#include <string>
#include <vector>

namespace doug {

// CSR with 0-based indices; within a row the entries keep the order the
// reference's stable arrangement gives them (src/datatypes/SpMtx/
// SpMtx_arrangement.F90:58-252).
struct Csr {
  int nrows = 0, ncols = 0;
  std::vector<int> ptr, col;
  std::vector<double> val;
  std::vector<int> src;  // input position of each entry (coo_to_csr only)
  int64_t nnz() const { return (int64_t)col.size(); }
};

struct Params {
  int radius1 = -1, radius2 = -1, minasize1 = -1, minasize2 = -1;
  int maxasize1 = -1, maxasize2 = -1;
  double strong1 = -1.0, strong2 = -1.0;
  int smoothers = -1, overlap = -1, levels = 2, method = 1, fine_method = 1;
  int symmstruct = 0, symmnumeric = 0;
  double omega = 0.33333333333333;  // CoarseMtx.F90:64
  int num_iters = -1;               // fine_method 2 sweeps (aggr.F90:220-222)
};

// Aggregates (Aggregate_mod.F90:39-51): num is 1-based aggregate number per
// node (0 = none); starts/nodes is the 0-based CSR listing nodes ascending.
struct Aggregates {
  int nagr = 0;
  std::vector<int> num, starts, nodes;
};

// Variables the reference keeps alive across calls within one process:
// `allmind` (Aggregate_mod.F90:348) and the initialised local `beta` of
// SpMtx_aggregate (SpMtx_aggregation.F90:83).
struct ProcessState {
  int allmind = INT_MAX;
  double beta = 1.0e-5;
};

// COO (1-based or 0-based) -> CSR by a stable counting sort on rows.
Csr coo_to_csr(int n, int64_t nnz, const int32_t* indi, const int32_t* indj, const double* val,
               int base, std::string& err);

// SpMtx_scale (SpMtx_arrangement.F90:298-351): returns the scaled values and
// the stored diagonal; symmetric scaling if `symm`, row scaling otherwise.
void scale_values(const Csr& A, bool symm, std::vector<double>& scaled, std::vector<double>& diag);
// scale_values on `input` (A's values), and the values after the reference's
// scale/unscale round trip, in one pass: scaled and round_trip are what
// scale_values followed by unscale_values give.
void scale_round_trip(const Csr& A, const std::vector<double>& input, bool symm,
                      std::vector<double>& scaled, std::vector<double>& diag,
                      std::vector<double>& round_trip);
// SpMtx_unscale (SpMtx_arrangement.F90:355-388) applied to scaled values.
void unscale_values(const Csr& A, bool symm, const std::vector<double>& diag,
                    std::vector<double>& vals);

// SpMtx_find_strong (SpMtx_aggregation.F90:829-899)
std::vector<uint8_t> find_strong(const Csr& A, const std::vector<double>& scaled,
                                 const std::vector<double>& diag, double alpha);

// SpMtx_aggregate (SpMtx_aggregation.F90:49-641) and its helpers.
void aggregate(const Csr& A, const std::vector<double>& scaled, const std::vector<uint8_t>& strong,
               int neighood, int minasize, int maxasize, double alpha, bool coarse_level,
               ProcessState& ps, Aggregates& inner, Aggregates& full);

// IntRestBuild (CoarseMtx.F90:50-474) for smoothers 0 to 8 (above 8 the
// reference falls back to 2).  Returned as CSR with each row sorted by
// column: CoarseMtxBuild arranges R by columns in place, and every later
// use of R (the Galerkin product, prec2Level) sees that order.
// ghost (optional, per row): rows of a process's A_ghost, smoothed with
// their diagonal entry in place of the filtered diagonal (CoarseMtx.F90:164-220).
Csr restriction(const Csr& A, const std::vector<uint8_t>& strong, const Aggregates& inner,
                int smoothers, double omega, const std::vector<uint8_t>* ghost = nullptr);

// CoarseMtxBuild (CoarseMtx.F90:476-516): A_c = R (A R^T), rows in the
// reference's first-touch column order; Rt = transpose(R).
Csr galerkin(const Csr& A, const Csr& R, const Csr& Rt);

// SpMtx_AB (SpMtx_op_AB.f90:39-200) on row-arranged operands: output rows
// in first-touch column order, products summed in traversal order.
Csr first_touch_product(const Csr& X, const Csr& Y);

// Transpose keeping ascending source-row order inside each output row.
Csr transpose(const Csr& A);

// Get_aggregate_nodes (Aggregate_mod.F90:162-190) for every coarse
// aggregate: the np=1 subdomain partition (Partitioning_aggr.F90:176-189).
void coarse_partition(const Aggregates& cfull, const Aggregates& ffull, int n,
                      std::vector<int>& starts, std::vector<int>& nodes);

// Add_layers (Decomposition.f90:56-99) for every part: appends `ol` BFS
// layers over the row graph of A.
void add_layers_all(const Csr& A, const std::vector<int>& pstarts, const std::vector<int>& pnodes,
                    int ol, std::vector<int>& sstarts, std::vector<int>& sinds);

// Worker threads of the row-parallel setup loops started from the calling
// thread: min(hardware threads, 16), or fewer after set_setup_thread_cap(c)
// with c > 0 (c = 0 lifts the cap).  The cap is per calling thread; results
// never depend on the thread count.
int setup_threads();
void set_setup_thread_cap(int cap);

// Reverse Cuthill-McKee order of each block of `inds` over the graph of A
// restricted to the block: out holds the same blocks, each a permutation of
// its input.  Independent of the thread count.
// `ws`, when given, lends its workers' node-indexed scratch.
struct RankWorkspaces;
void block_rcm(const Csr& A, const std::vector<int>& starts, const std::vector<int>& inds,
               std::vector<int>& out, RankWorkspaces* ws = nullptr);
// block_rcm on the calling thread with the caller's scratch `floc`
// (A.nrows entries, all -1 on entry and on exit).
void block_rcm_serial(const Csr& A, const std::vector<int>& starts, const std::vector<int>& inds,
                      std::vector<int>& out, std::vector<int>& floc);

// Project graph partitioner (stands in for METIS): recursive bisection of a
// symmetric adjacency (0-based CSR, no self loops) into nparts parts of
// balanced vertex weight (vwgt may be null: unit weights).  part[v] in
// [0, nparts).  Deterministic.
void partition_graph(int n, const int* xadj, const int* adjncy, const int* vwgt, int nparts,
                     std::vector<int>& part);

// Adjacency of the fine aggregates (SpMtx_buildAggrAdjncy,
// SpMtx_arrangement.F90:2558-2639) over `full.nagr` vertices.
void aggregate_graph(const Csr& A, const Aggregates& inner, const Aggregates& full,
                     std::vector<int>& xadj, std::vector<int>& adjncy);

// ---- virtual ranks: DOUG's decomposition over several MPI processes ----
//
// One rank's view after SpMtx_distributeWithOverlap / SpMtx_build_ghost /
// SpMtx_Build_lggl (SpMtx_arrangement.F90:744-1710, SpMtx_distribution.F90:
// 139-290): local freedoms ordered [sent to a neighbour | rest of the inner
// overlap | other inner | outer overlap | received], each ascending in global
// number; `A` is the rank's local matrix (own rows and the rows of outside
// nodes nearer than the outer layer) in local numbering, each row's entries
// ordered as the reference's blocks leave them: columns of sent nodes first,
// then the others, each group in ascending global column order.
struct RankLocal {
  int ninner = 0;
  std::vector<int> lg;  // local -> global (0-based)
  // nrows = ncols = lg.size(); rows of the outer layer are empty.  Kept by
  // virtual_ranks with `keep_local` only: then the outer layer's rows hold
  // A_ghost's entries (values as read) and A.src gives every entry's
  // position in the assembled matrix the caller arranged.
  Csr A;
  int nsent = 0;        // own nodes sent to a neighbour: the first nsent local nodes
  int nrecv = 0;        // the outer layer: the last nrecv local nodes
  Csr G;                // A_ghost rows during the set-up
  Aggregates agg;       // own aggregates in local numbering (keep_local)
};

// Per-rank results of the decomposition: each rank aggregates its own rows
// (Partitioning_aggr.F90:81-92) with a fresh process state and splits its
// inner nodes into `nsub_per_rank` parts of its aggregate graph
// (Partitioning_metis.F90:34-86, project partitioner for more than one
// part), then adds `ol_sub` layers over its local matrix
// (Preconditioner_base.F90:96-123).  Aggregates are numbered rank by rank
// (CoarseAllgathers.F90:531-540); subdomains are listed rank by rank.
struct RankSetup {
  std::vector<RankLocal> ranks;       // numbering only: each local matrix is released after use
  Aggregates inner, full;             // global numbering
  std::vector<int> sstarts, sinds;    // subdomains with their layers, global ids
  std::vector<int> dinds;             // sinds, each subdomain in block_rcm order (if asked)
  // every rank's own nodes in its local order, rank after rank (own_off[r]
  // is rank r's first position): the process-major node order
  std::vector<int> order, own_off;
};

// `As`: the assembled matrix pattern with each row sorted by column
// (stable); `val`: its values as read, in As's entry order.  `eptn[i]` in
// [0, nranks).  `ol_dist` = max(overlap, smoothers) >= 1; `ol_sub` <= ol_dist.
// Per-worker scratch of virtual_ranks (node-indexed arrays of the whole
// matrix), kept by the caller so that later set-ups reuse it.
struct RankScratch;
struct RankWorkspaces {
  std::vector<std::unique_ptr<RankScratch>> w;
  RankWorkspaces();
  ~RankWorkspaces();
  void clear();
};

// With `rcm_graph`, each rank also orders its subdomains by block_rcm over
// that graph (out.dinds), inside the rank's worker.  With `keep_local` every
// rank keeps its local matrix with A_ghost's rows and its own aggregates
// (ranks_coarse_space's input); `pos` (optional) gives the position of each
// of As's entries in the caller's matrix, recorded in RankLocal::A.src (As's
// own positions without it), e.g. sort_rows_by_column's src.
void virtual_ranks(const Csr& As, const std::vector<double>& val, const std::vector<int>& eptn,
                   int nranks, int ol_dist, int ol_sub,
                   int nsub_per_rank, int radius1, int minasize1, int maxasize1, double strong1,
                   RankSetup& out, const Csr* rcm_graph = nullptr, RankWorkspaces* ws = nullptr,
                   bool keep_local = false, const std::vector<int>* pos = nullptr);

// The coarse space of a DOUG run on several processes
// (CoarsePreconditioner_smooth_Init, CoarsePreconditioner_smooth.F90:47-89,
// and the coarse matrix exchange, pcg.F90:78-93, CoarseAllgathers.F90:
// 175-257), from virtual_ranks' kept local matrices:
//  - each aggregate's row of R as its owner builds it (IntRestBuild with
//    A_ghost, CoarseMtx.F90:50-474): rows of the rank's matrix smooth with
//    the filtered diagonal summed in the rank's entry order and the values
//    after the scaling round trip, A_ghost's rows with their diagonal entry
//    and the values as read; the nodes of an aggregate in the owner's local
//    order;
//  - A_c as the sum over the ranks, in rank order, of R(:, own) (A R^T)(own, :)
//    (CoarseMtxBuild, CoarseMtx.F90:476-516, then SpMtx_consolidate),
//    each rank's products in its entry orders; rows sorted by column.
// val_rt / val_in: the caller's matrix values after the round trip and as
// read, strong: its strength flags (symmetric scaling), all indexed by the
// positions in RankLocal::A.src.  R rows ascending in column, Rt = R^T.
void ranks_coarse_space(const RankSetup& rs, int n, const std::vector<double>& val_rt,
                        const std::vector<double>& val_in, const std::vector<uint8_t>& strong,
                        int smoothers, double omega, Csr& R, Csr& Rt, Csr& AC);

// Stable per-row sort by column (SpMtx_arrange with sort, SpMtx_arrangement.F90:159-227),
// carrying `val`; src receives each entry's position in A.
Csr sort_rows_by_column(const Csr& A, const std::vector<double>& val);
// Whether every row of A already lists its columns in non-decreasing order.
bool rows_sorted_by_column(const Csr& A);

}  // namespace doug
