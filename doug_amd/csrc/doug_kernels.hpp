// Launch wrappers for the gfx950 kernels of the Schwarz-PCG solve path.
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>

namespace doug {

constexpr int kTile = 64;              // dense tile edge (one wavefront wide)
constexpr int kTileElems = kTile * kTile;
constexpr int kSpmvThreads = 256;      // CSR-stream block
constexpr int kSpmvChunk = 2048;       // products staged in LDS per block
constexpr int kMaxTrsvTiles = 256;     // subdomain size limit of the one-block solve (16384)
constexpr int kRedBlocks = 1024;       // fixed partial count of the grid reductions
constexpr int kRedBlocksWide = 8192;   // grid of the column-pair reductions (partials per column)
constexpr int kVecChunk = 1024;        // entries per workgroup of the PCG vector kernels
// partial sums the single-vector reductions of an n-vector need
constexpr int64_t vec_partials(int64_t n) { return (n + 255) / 256 + kRedBlocks; }
constexpr int kMaxRhs = 32;            // right-hand sides per batched solve

// Tile row i of a block's lower envelope: tiles (i, fc..i) are stored, the
// ones left of fc are zero.  fc never decreases with i, so the tiles of
// column j are rows j..lr.  Cholesky fill stays inside a row envelope, so
// the factor needs no other tiles.
struct TileRow {
  int32_t off;  // tile index of (i, fc) within the block's storage
  int32_t fc;   // first stored tile column
  int32_t lr;   // last tile row whose envelope reaches column i
  int32_t pad;
};

// Symmetric block in tile-envelope lower storage: tile (i,k), fc_i <= k <= i,
// starts at base + (tr[i].off + k - tr[i].fc) * kTileElems; each tile is
// 64x64 row-major.  After factorisation the off-diagonal tiles hold L and the
// diagonal tiles hold inv(L_ii).  A dense block has fc = 0 on every row.
struct BlockDesc {
  int64_t base;        // element offset of the block's tile storage
  int64_t vec_off;     // offset into the concatenated subdomain vectors
  int64_t woff;        // first row of the block in the padded multi-vector workspaces
  const TileRow* tr;   // device pointer to the block's nt tile rows
  int32_t nt;          // tiles per dimension
  int32_t s;           // true size (rows), <= nt*64
};

// ---- dense batched factorisation (T = double, or float for FP32 factors) ----
// Scatter A entries (converted to T) into zeroed tile storage; ones on the
// padded diagonal.
template <typename T>
void launch_scatter_entries(T* tiles, const int64_t* dst, const int32_t* src, const double* val,
                            int64_t nentries, hipStream_t st);
// Positions that receive several entries (duplicate (i, j) pairs): group g
// stores the FP64 sum of val[src[seg[g]..seg[g+1])], in that order, at
// tiles[dst[g]].  Runs after launch_scatter_entries on the same stream.
template <typename T>
void launch_scatter_sums(T* tiles, const int64_t* dst, const int64_t* seg, const int32_t* src,
                         const double* val, int64_t ngroups, hipStream_t st);
// Envelope and entries of index blocks from the device copies of A (CSR),
// the blocks' unknown lists `inds` and each unknown's ascending occurrence
// positions in them (occ_ptr/occ_pos).  launch_block_envelope lowers fc
// (one entry per tile row, block b's rows from row0[b], holding the tile row
// index on entry) to each tile row's first stored tile column;
// launch_scatter_blocks writes the lower entries into zeroed tile storage.
// A must have no duplicate (i, j) pair.
void launch_block_envelope(const int32_t* starts, const int32_t* row0, int nblocks, const int32_t* inds,
                           const int32_t* occ_ptr, const int32_t* occ_pos, const int32_t* a_ptr,
                           const int32_t* a_col, int32_t* fc, hipStream_t st);
template <typename T>
void launch_scatter_blocks(T* tiles, const BlockDesc* d, int nblocks, const int32_t* inds,
                           const int32_t* occ_ptr, const int32_t* occ_pos, const int32_t* a_ptr,
                           const int32_t* a_col, const double* a_val, hipStream_t st);

template <typename T>
void launch_pad_identity(T* tiles, const BlockDesc* d, int nblocks, int maxpad, hipStream_t st);
// Left-looking tile Cholesky, step j: diagonal tiles (SYRK + POTRF + inverse)
// then the tiles below them (GEMM + TRSM), on the matrix cores in T.  `flag`
// receives the first failing block+1 on a non-positive pivot.  `maxcount`
// bounds lr_j - j over the blocks.
// With left = false the kernels skip the left-looking sums (the trailing
// updates were applied by launch_rl_update).
template <typename T>
void launch_chol_diag(T* tiles, const BlockDesc* d, int nblocks, int j, int* flag,
                      hipStream_t st, bool left = true);
template <typename T>
void launch_chol_offdiag(T* tiles, const BlockDesc* d, int nblocks, int j, int maxcount,
                         hipStream_t st, bool left = true);
// Dense single block d[0]: right-looking trailing update after step j,
// A_ik -= L_ij L_kj^T for j < k <= i <= last, where `last` is the last tile
// row whose envelope reaches column j (L_ij = 0 below it; nt - 1: all).
void launch_rl_update(double* tiles, const BlockDesc* d, int nt, int j, int last, hipStream_t st);
// W = L^{-1} of a dense factor (diagonal tiles inv(L_ii)): launch_inv_init,
// then launch_inv_step for m = 0 .. nt-1 (right-looking substitution), with
// `last` the last tile row of L's column m inside the envelope.  Step m reads
// block column m of L alone: it may start once the factorisation's step m
// has solved that column, while the later steps go on elsewhere.
void launch_inv_init(double* W, const BlockDesc* d, int nt, hipStream_t st);
void launch_inv_step(const double* L, double* W, const BlockDesc* d, int nt, int m, int last,
                     hipStream_t st);
// The values after the scale/unscale round trip (val) and the strength
// flags of the symmetric- (symm) or row-scaled matrix, from the values as
// read; diag and sv are n-entry work arrays.  Bit-identical to the host's
// scale_round_trip + find_strong.
void launch_values_strong(int n, const int32_t* ptr, const int32_t* col, const double* vin, int symm,
                          double alpha, double* diag, double* sv, double* val, uint8_t* strong,
                          hipStream_t st);
// idx[p] = map[idx[p]] (renumbering of column indices)
void launch_map_index(int32_t* idx, int64_t n, const int32_t* map, hipStream_t st);

// ---- per-iteration operators ----
// y = A x, CSR-stream: each row summed in stored order, products rounded
// before the add (bit-identical to a sequential row loop without FMA).
// `shape` (0..kSpmvShapes-1) picks rows per block spmv_shape_rows(shape)
// and, for odd shapes, the XCD-contiguous block order; the row blocks must
// have been built for that row count (at most 8 entries per row on average).
constexpr int kSpmvShapes = 6;
constexpr int kSpmvDefaultShape = 2;
constexpr int spmv_shape_rows(int shape) { return 128 << (shape >> 1); }
void launch_spmv(int nblocks_rows, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                 const double* val, const double* x, double* y, hipStream_t st,
                 int shape = kSpmvDefaultShape);
// Batched forward/backward substitution, one block per subdomain: gathers
// r(inds), solves L L^T u = r in the factor's precision, writes u to sol
// (FP64, concatenated layout).
// maxnt: the largest block's tile count (sets the dynamic LDS); call
// prepare_trsv<T>(maxnt) once before the first launch for that size.
template <typename T>
void prepare_trsv(int maxnt);
template <typename T>
// skip (device, optional): the launch does nothing when *skip != 0.
void launch_trsv(const T* tiles, const BlockDesc* d, int nblocks, int maxnt, const int32_t* inds,
                 const double* r, double* sol, hipStream_t st, const double* skip = nullptr);
// u = W c and c' = W^T u for the coarse inverse factor; `part` holds
// nt * gemv_chunks(nt) * 64 doubles of partial sums.
constexpr int kGemvChunk = 8;  // tiles per workgroup
constexpr int gemv_chunks(int nt) { return (nt + kGemvChunk - 1) / kGemvChunk; }
// perm (nt*64 entries, or null): u = W c(perm), W's rows and columns in the
// factor's order, c in the original one.  W in FP64 or FP32 (products and
// sums in FP64).
template <typename TW>
void launch_gemv_w(const TW* W, const BlockDesc* d, int nt, const double* c, double* u,
                   double* part, hipStream_t st, const int32_t* perm = nullptr);
template <typename TW>
void launch_gemv_wt(const TW* W, const BlockDesc* d, int nt, const double* u, double* c,
                    double* part, hipStream_t st);
// dst[0..n) = (float)src
void launch_to_float(const double* src, float* dst, int64_t n, hipStream_t st);
// y = A x with one wavefront per row (rows of hundreds of entries), each row
// summed in a fixed order
void launch_spmv_wave_rows(int nrows, const int32_t* ptr, const int32_t* col, const double* val,
                           const double* x, double* y, hipStream_t st);
// z[i] = sum over occurrences (ascending subdomain) of sol, then
//        + sum_a Rt(i,a) csol[a]  (coarse prolongation), fine part first.
// Symmetric Gauss-Seidel sweeps over each subdomain's own matrix (the
// fine_method 2 preconditioner), one workgroup per subdomain with its
// iterate in LDS.  Rows are level-scheduled: levptr[lf0 .. lf0 + nlf] (and
// [lb0 .. lb0 + nlb] for the backward sweep) delimit each level's rows in
// lrows[2 off ..] (forward, m rows) and lrows[2 off + m ..] (backward).  Row
// r's off-diagonal entries are ecol/eval[e0 + rbeg[off + r] ..] (rlen[off + r]
// of them, ascending column, local numbering).
constexpr int kSgsMaxRows = 16384;  // LDS holds the iterate
struct SgsDesc {
  int64_t off;  // first position in the subdomain-vector layout
  int64_t e0;   // first entry
  int32_t m;    // rows (the subdomain's unknowns, ascending)
  int32_t nlf, nlb;
  int32_t lf0, lb0;
  int32_t pad;
};
void prepare_sgs(int mmax);
// out[p] = x[idx[p]] for p < n
void launch_gather_positions(int64_t n, const int32_t* idx, const double* x, double* out,
                             hipStream_t st);
// sub holds the right-hand side on entry and the iterate after `iters`
// forward and backward sweeps on return
void launch_sgs(const SgsDesc* d, int nsub, int mmax, const int32_t* levptr, const int32_t* lrows,
                const int32_t* rbeg, const int32_t* rlen, const int32_t* ecol, const double* eval,
                const double* diag, double* sub, int iters, hipStream_t st);

void launch_overlap_sum(int n, const int32_t* occ_ptr, const int32_t* occ_pos, const double* sol,
                        const int32_t* rt_ptr, const int32_t* rt_col, const double* rt_val,
                        const double* csol, double* z, hipStream_t st);

// The same z, with r.z reduced into scal (rho and beta, as launch_dot's mode
// 1); `partial` holds vec_partials(n) doubles.
void launch_overlap_sum_dot(int n, const int32_t* occ_ptr, const int32_t* occ_pos, const double* sol,
                            const int32_t* rt_ptr, const int32_t* rt_col, const double* rt_val,
                            const double* csol, double* z, const double* r, double* partial,
                            double* scal, hipStream_t st);

// ---- PCG vector kernels with deterministic two-level reductions ----
// scal layout (device doubles): see doug_capi.hip.  `partial` holds
// vec_partials(n) doubles.
void launch_dot(const double* a, const double* b, int n, double* partial, double* out, int mode,
                double* scal, hipStream_t st);
void launch_p_update(int n, const double* z, double* p, const double* scal, int first,
                     hipStream_t st);
// With hist, the residual's finish keeps the iteration count, the
// (alpha, beta, residual) history and the convergence flag on the device
// (scal slots S_DONE.. of the kernels); x and r stay unchanged once set.
void launch_xr_update(int n, double* x, double* r, const double* p, const double* q,
                      double* partial, double* scal, hipStream_t st, double* hist = nullptr);
void launch_copy(double* dst, const double* src, int n, hipStream_t st);
// dst(i, c) = src(i, c) for c < ncols, 0 for ncols <= c < ldd (n rows;
// leading dimensions ldd and lds)
void launch_repack(double* dst, int ldd, const double* src, int lds, int64_t n, int ncols,
                   hipStream_t st);
// dst[0..n) = 0 (a kernel, so that captured graphs hold no memset node)
void launch_zero(double* dst, int64_t n, hipStream_t st);
void launch_residual(int n, const double* b, const double* ax, double* r, hipStream_t st);

// ---- several right-hand sides at once ----
// Multi-vectors are node-major: entry (i, r) of an n x nrhs block sits at
// i*ld + r.  Subdomain workspaces hold nt*64 rows of NR (16 or 32) columns
// per block, starting at row woff.
// Forward then backward substitution with the envelope factor of every
// block for NR (16 or 32) columns at once, one workgroup per block, tile
// products on the matrix cores: U(block rows) <- L^{-T} L^{-1} rhs(inds, :),
// then out[(vec_off + p) * ldo + r] = U for p < s, r < nrhs.  rhs has
// leading dimension ldr; U is T workspace of nt*64 rows of NR from woff.
// kb: off-diagonal tiles staged per step (1 or 2).  cg = 2 with NR = 32:
// two workgroups per block, 16 columns each (U then holds the two groups'
// nt*64 x 16 panels one after the other in the block's 32-column space).
template <typename T>
void launch_trsm_multi(const T* L, const BlockDesc* d, int nblocks, const int32_t* inds,
                       const double* rhs, int ldr, T* U, double* out, int ldo, int nrhs, int NR,
                       int kb, int cg, hipStream_t st);
// Dense coarse inverse (block d[0], nt tiles) times NR columns, chunked with
// partial panels in `part` (nt * gemv_chunks(nt) * 64 * NR doubles):
// U_i = sum_{k<=i} W_ik G_k (all nt*64 rows), and
// out(p, c) = (sum_{i>=j} W_ij^T U_i)(p, c) for p < rows, c < nrhs (leading dimension ldo).
void launch_wgemm_fwd(const double* W, const BlockDesc* d, int nt, const double* G, double* U,
                      double* part, int NR, hipStream_t st, const int32_t* perm = nullptr);
void launch_wgemm_bwd(const double* W, const BlockDesc* d, int nt, int rows, const double* U,
                      double* out, int ldo, int nrhs, double* part, int NR, hipStream_t st);
// y(:, r) = A x(:, r) for r < nrhs, each row summed in stored order
void launch_spmv_multi(int nrb, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                       const double* val, const double* x, int ldx, double* y, int ldy, int nrhs,
                       hipStream_t st);
// The same for long rows (the restriction): one workgroup per row, each
// row's entries split over 256/(nrhs/2) threads per column pair and added in
// a fixed tree order.  nrhs, ldx, ldy even, nrhs <= 32.
void launch_spmv_multi_rows(int nrows, const int32_t* ptr, const int32_t* col, const double* val,
                            const double* x, int ldx, double* y, int ldy, int nrhs, hipStream_t st);
void launch_overlap_sum_multi(int n, int nrhs, const int32_t* occ_ptr, const int32_t* occ_pos,
                              const double* sol, const int32_t* rt_ptr, const int32_t* rt_col,
                              const double* rt_val, const double* csol, int ldc, double* z,
                              hipStream_t st);
// per-column dot products; mode as launch_dot, scalars at scal[r*8 + k]
void launch_dot_multi(const double* a, const double* b, int n, int nrhs, double* partial,
                      unsigned* counter, int mode, double* scal, hipStream_t st);
void launch_p_update_multi(int n, int nrhs, const double* z, double* p, const double* scal,
                           int first, hipStream_t st);
void launch_xr_update_multi(int n, int nrhs, double* x, double* r, const double* p,
                            const double* q, double* partial, unsigned* counter, double* scal,
                            hipStream_t st);
// fused forms (nrhs even): overlap sum with rho = r.z (mode 1), and
// p_new = z + beta p_old, q = A p_new with p.q (mode 2).  Column-pair
// reductions take max(kRedBlocksWide, nrb) partial slots of kMaxRhs.
void launch_overlap_sum_dot_multi2(int n, int nrhs, const int32_t* occ_ptr, const int32_t* occ_pos,
                                   const double* sol, const int32_t* rt_ptr, const int32_t* rt_col,
                                   const double* rt_val, const double* csol, int ldc, double* z,
                                   const double* r, double* partial, unsigned* counter, double* scal,
                                   hipStream_t st);
// The SpMV form also applies the previous iteration's x += alpha p_old (not
// on the first), which the column-pair x/r update leaves to it;
// launch_x_final_multi2 adds the last iteration's after the loop.
void launch_spmv_p_multi2(int nrb, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                          const double* val, const double* z, const double* pold, double* pnew, double* q,
                          int nrhs, int first, double* partial, unsigned* counter, double* scal, double* x,
                          hipStream_t st);
void launch_x_final_multi2(int n, int nrhs, double* x, const double* p, const double* scal, hipStream_t st);

// ---- coarse space on the device (setup; bit-identical to the host build) ----
// Rows of R^T = S T^T into slots at ptr[j] (capacity = row length of A),
// cnt[j] = entries written.  strong: per entry of A; num: 1-based aggregate
// of each node (0 = none).
// the coarse space of a DOUG run on several processes with overlap 1
// (doug_coarse.hip): rows in their owner's entry order with the filtered and
// the A_ghost diagonals; R^T rows in process order; entry ids in process
// order for the sort that gives R in process order; A_c with each process's
// part summed first, rows sorted by column
void launch_mp_layout(int n, const int32_t* order, const int32_t* owner, const int32_t* off,
                      const int32_t* nsent, int32_t* proc, uint8_t* sent, hipStream_t st);
void launch_mp_own_rows(int n, const int32_t* ptr, const int32_t* col, const double* val,
                        const double* valin, const uint8_t* strong, const int32_t* owner,
                        const uint8_t* sent, int32_t* ocol, double* oval, double* dfilt,
                        double* dghost, hipStream_t st);
void launch_mp_rt_rows(int n, const int32_t* order, const int32_t* ptr, const int32_t* col,
                       const double* val, const double* valin, const uint8_t* strong,
                       const int32_t* num, const int32_t* owner, const int32_t* proc,
                       const double* dfilt, const double* dghost, double omega, int32_t* slot_col,
                       double* slot_val, int32_t* cnt, hipStream_t st);
void launch_mp_row_lengths(int n, const int32_t* order, const int32_t* rtptr, int32_t* len,
                           hipStream_t st);
void launch_mp_proc_entries(int n, const int32_t* order, const int32_t* rtptr, const int32_t* rtcol,
                            const int32_t* off, int32_t* ids, int32_t* keys, hipStream_t st);
void launch_mp_galerkin_wave(int nrows, const int32_t* xptr, const int32_t* xcol, const double* xval,
                             const int32_t* owner, const int32_t* yptr, const int32_t* ycol,
                             const double* yval, const int64_t* slot_off, int32_t* slot_col,
                             double* slot_val, int32_t* cnt, int* overflow, hipStream_t st);
void launch_rt_rows(int n, const int32_t* ptr, const int32_t* col, const double* val,
                    const uint8_t* strong, const int32_t* num, double omega, int32_t* slot_col,
                    double* slot_val, int32_t* cnt, hipStream_t st);
// ub[i] = number of products of row i of X*Y
void launch_product_bound(int nrows, const int32_t* xptr, const int32_t* xcol, const int32_t* yptr,
                          int64_t* ub, hipStream_t st);
// rows of X*Y in first-touch column order into slots at slot_off[i]
// one wavefront per row; *overflow set (and the rows incomplete) when a row
// has more than 128 distinct columns
void launch_first_touch_wave(int nrows, const int32_t* xptr, const int32_t* xcol,
                             const double* xval, const int32_t* yptr, const int32_t* ycol,
                             const double* yval, const int64_t* slot_off, int32_t* slot_col,
                             double* slot_val, int32_t* cnt, int* overflow, hipStream_t st);
void launch_first_touch(int nrows, const int32_t* xptr, const int32_t* xcol, const double* xval,
                        const int32_t* yptr, const int32_t* ycol, const double* yval,
                        const int64_t* slot_off, int32_t* slot_col, double* slot_val, int32_t* cnt,
                        hipStream_t st);
// slots -> CSR with row pointers out_ptr
void launch_compact(int nrows, const int32_t* slot_off, const int32_t* out_ptr,
                    const int32_t* slot_col, const double* slot_val, int32_t* out_col,
                    double* out_val, hipStream_t st);
void launch_compact(int nrows, const int64_t* slot_off, const int32_t* out_ptr,
                    const int32_t* slot_col, const double* slot_val, int32_t* out_col,
                    double* out_val, hipStream_t st);
void launch_row_ids(int nrows, const int32_t* ptr, int32_t* rowid, hipStream_t st);
void launch_iota(int64_t n, int32_t* v, hipStream_t st);
void launch_key_histogram(int64_t n, const int32_t* keys, int32_t* counts, hipStream_t st);
// out[e] = (rowid[order[e]], val[order[e]])
void launch_gather_entries(int64_t n, const int32_t* order, const int32_t* rowid, const double* val,
                           int32_t* out_col, double* out_val, hipStream_t st);
// hipcub prefix sums and stable radix sort; tmp == nullptr queries `bytes`
hipError_t exclusive_scan(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int n,
                          hipStream_t st);
hipError_t exclusive_scan(void* tmp, size_t& bytes, const int64_t* in, int64_t* out, int n,
                          hipStream_t st);
hipError_t stable_sort_pairs(void* tmp, size_t& bytes, const int32_t* keys_in, int32_t* keys_out,
                             const int32_t* vals_in, int32_t* vals_out, int n, int end_bit,
                             hipStream_t st);

// grad[src[p]] = -lam[row(p)] * x[col[p]] over the CSR entries of A
void launch_value_grad(int n, const int32_t* ptr, const int32_t* col, const int32_t* src,
                       const double* lam, const double* x, double* grad, hipStream_t st);

}  // namespace doug
