// C-ABI of the MI355X Schwarz-PCG solve path (include/doug.h).
//
// Host side: ingest, aggregation, index sets and the per-subdomain scatter
// maps (doug_setup.cpp).  Device side: every factorisation and every
// per-iteration operator (doug_kernels.hip).  There is no CPU fallback: a
// missing GPU is an error.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <functional>
#include <mutex>
#include <numeric>
#include <string>
#include <thread>
#include <vector>

#include "../../include/doug.h"
#include "doug_kernels.hpp"
#include "doug_setup.hpp"

using namespace doug;

namespace {

enum { S_RHO = 0, S_RHO_PREV = 1, S_BETA = 2, S_PQ = 3, S_ALPHA = 4, S_RES = 5 };
// the device's PCG convergence test (as in the kernels), then the slot count
enum { S_DONE = 8, S_ITER = 9, S_THR = 10, S_INIT = 11, S_MAXIT = 12, S_HCAP = 13, S_N = 16 };

double now() {
  return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// A host thread that enqueues one part of a set-up on another of the
// context's streams names it here; every `ctx->stream` read on that thread
// then yields it (the set-up's subdomain work runs on a second host thread
// beside the coarse space's construction).
thread_local hipStream_t t_stream_override = nullptr;

struct StreamRef {
  hipStream_t s = nullptr;
  operator hipStream_t() const { return t_stream_override ? t_stream_override : s; }
};

struct StreamOverride {
  explicit StreamOverride(hipStream_t st) { t_stream_override = st; }
  ~StreamOverride() { t_stream_override = nullptr; }
  StreamOverride(const StreamOverride&) = delete;
  StreamOverride& operator=(const StreamOverride&) = delete;
};

}  // namespace

struct doug_ctx {
  int device = 0;
  StreamRef stream;  // the main stream (or this thread's override)
  // guards allocs, err and the reaper while two host threads set up
  std::mutex mu;
  hipStream_t stream2 = nullptr;  // subdomain factorisation beside the coarse one
  hipStream_t stream3 = nullptr;  // coarse inverse one step behind the coarse factorisation
  std::vector<hipEvent_t> cevents;  // coarse factorisation step ends, then the inverse's end
  // temporaries of the device coarse build, kept between set-ups
  std::vector<std::pair<void*, size_t>> scratch_slots;
  std::thread reaper;  // frees the buffers of the previous set-up
  std::string err;
  Params p;
  doug_info info{};

  // host problem
  bool have_matrix = false, ready = false, host_ready = false;
  int dup_state = -1;   // A holds a duplicate (i, j) pair: 1, none: 0, not yet known: -1
  Csr A;                // values after the scale/unscale round trip once set up
  std::vector<double> A_input_val;
  std::vector<double> scaled;  // strength-test values (kept to reuse the allocation)
  std::vector<uint8_t> strong;
  std::vector<int> part_starts, part_nodes;
  bool have_partition = false;
  int num_subdomains = 0;  // > 1: parts of the fine-aggregate graph (doug_set_num_subdomains)
  // virtual ranks (doug_set_ranks): rank of each unknown, and the per-rank
  // decomposition of the last setup
  int nranks = 0;
  std::vector<int> eptn;
  // caller aggregates (doug_set_aggregates) and subdomains (doug_set_subdomains)
  int inj_nagr = 0;
  // caller's strength flags (per input entry, input order) and restriction
  std::vector<uint8_t> inj_strong;
  Csr inj_R;
  bool have_inj_R = false;
  std::vector<int> inj_num;
  std::vector<int> inj_sstarts, inj_sinds;
  int ol_dist = 0;
  RankSetup rsetup;
  RankWorkspaces rank_ws;  // the rank workers' scratch, reused by later setups
  ProcessState ps;
  Aggregates fin, ffull, cin, cfull;
  Csr R, Rt, AC;
  int ol = 0, smoothers = 0;
  std::vector<int> sstarts, sinds;
  bool has_coarse = false;
  bool coarse_on_device = false;   // R, R^T built on the device (not uploaded)
  int64_t nnz_R = 0;               // entries of R; with coarse_on_device R's columns and
                                   // values stay on the device until a getter asks
  uint8_t* d_strong = nullptr;     // strength mask, uploaded while the aggregation runs
  double* d_valin = nullptr;       // values as read (several processes, overlap 1)
  int32_t* d_eptn = nullptr;       // owner of each unknown (with d_valin)
  // several processes: place of each node in process order, the inverse,
  // and whether its owner sends it (staging of the device coarse build)
  std::vector<int32_t> mp_nsent;  // nodes each rank sends (its first own nodes)
  bool ranks_kept_local = true;    // rsetup.ranks hold the local matrices
  // whether the host's A.val and strong hold the round-trip values and the
  // strength flags (a several-process set-up forms them on the device and
  // copies them back only when host code needs them)
  bool host_values = true;
  bool use_device_coarse = true;   // DOUG_OPT_DEVICE_COARSE
  bool use_device_blocks = true;   // DOUG_OPT_DEVICE_BLOCKS
  bool keep_setup_temps = true;    // DOUG_OPT_KEEP_SETUP_TEMPS

  // device: matrix
  std::vector<void*> allocs;
  int32_t *d_ptr = nullptr, *d_col = nullptr, *d_rowblk = nullptr, *d_src = nullptr;
  double* d_val = nullptr;
  int nrb = 0;  // row blocks of kSpmvThreads rows (multi-vector SpMV)
  // the single-vector SpMV runs with the tuned shape over its own row blocks
  int spmv_shape = kSpmvDefaultShape;
  int32_t* d_rowblk_a = nullptr;
  int nrb_a = 0;
  // restriction (CSR rows = aggregates), prolongation (CSR rows = nodes)
  int32_t *d_Rptr = nullptr, *d_Rcol = nullptr, *d_Rrowblk = nullptr;
  double* d_Rval = nullptr;
  int nrbR = 0;
  int32_t *d_Rtptr = nullptr, *d_Rtcol = nullptr;
  double* d_Rtval = nullptr;
  // subdomains
  int nsub = 0;
  BlockDesc* d_sub = nullptr;
  std::vector<BlockDesc> h_sub;
  int trsv_order = 0;               // 1: largest subdomain dispatched first
  int trsm_kb = 2;                  // tiles per step of the batched envelope solve
  int trsm_cg = 1;                  // column layout of that solve (1..3, DOUG_OPT_BATCH_COLUMN_GROUPS)
  int local_order = 1;              // 1: reverse Cuthill-McKee inside each subdomain
  BlockDesc* d_sub_trsv = nullptr;  // descriptors in dispatch order
  doug_tuning tuning{};
  double* d_tiles = nullptr;   // FP64 factors
  float* d_tiles32 = nullptr;  // FP32 factors (DOUG_PREC_FP32_FACTORS)
  int precision = DOUG_PREC_FP64;
  int32_t *d_inds = nullptr, *d_occ_ptr = nullptr, *d_occ_pos = nullptr;
  double* d_subsol = nullptr;
  // fine_method 2: symmetric Gauss-Seidel on each subdomain's matrix
  SgsDesc* d_sgs = nullptr;
  int32_t *d_sgs_levptr = nullptr, *d_sgs_lrows = nullptr, *d_sgs_rbeg = nullptr;
  int32_t *d_sgs_rlen = nullptr, *d_sgs_col = nullptr;
  double *d_sgs_val = nullptr, *d_sgs_diag = nullptr;
  int sgs_mmax = 0, sgs_iters = 3;
  // coarse
  int nc = 0, cnt = 0;
  BlockDesc* d_cdesc = nullptr;
  double *d_ctiles = nullptr, *d_cW = nullptr;
  // coarse unknown at each position of the coarse factor's order (reverse
  // Cuthill-McKee of A_c), padded with the identity to cnt * 64 entries
  int32_t* d_cperm = nullptr;
  float* d_cW32 = nullptr;  // the coarse inverse rounded to FP32 (FP32-factor form)
  double* d_cpart = nullptr;  // partial sums of the chunked coarse GEMVs
  double* d_cPm = nullptr;    // partial panels of the chunked coarse products (batch)
  double *d_crhs = nullptr, *d_cu = nullptr, *d_csol = nullptr;
  // PCG
  int n = 0;
  double *d_x = nullptr, *d_r = nullptr, *d_z = nullptr, *d_p = nullptr, *d_q = nullptr;
  double *d_b = nullptr, *d_scal = nullptr, *d_partial = nullptr, *d_tmp = nullptr;
  unsigned* d_counter = nullptr;
  int* d_flag = nullptr;
  double* h_scal = nullptr;  // pinned, 2 * S_N: the states of two iterations in flight
  double* d_hist = nullptr;  // PCG (alpha, beta, residual) per iteration, hist_cap entries
  int64_t hist_cap = 0;
  // several right-hand sides (doug_prepare_batch): W = L^{-1} per subdomain,
  // padded workspaces, node-major multi-vectors of kMaxRhs columns
  bool batch_ready = false;
  int maxnt = 0;
  int64_t work_rows = 0;

  void* d_U = nullptr;
  double *d_cG = nullptr, *d_cU = nullptr, *d_cS = nullptr, *d_subsolm = nullptr;
  double *d_Xm = nullptr, *d_Rm = nullptr, *d_Zm = nullptr, *d_Pm = nullptr, *d_Qm = nullptr;
  double* d_Pm2 = nullptr;  // second p buffer of the fused even-column iteration
  double *d_Bm = nullptr, *d_Tm = nullptr, *d_scalm = nullptr, *d_partialm = nullptr;
  unsigned* d_counterm = nullptr;
  double* h_scalm = nullptr;  // pinned, kMaxRhs * 8
  // one PCG iteration captured as a HIP graph (iterations >= 2)
  bool use_graph = true;
  hipGraphExec_t iter_graph = nullptr;
  // batched iteration, for batch_graph_nrp columns; two graphs when the
  // iteration alternates between the two p buffers
  hipGraphExec_t batch_graph[2] = {nullptr, nullptr};
  int batch_graph_nrp = 0;
};

#define HIPCHK(expr)                                                              \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      std::lock_guard<std::mutex> g_(ctx->mu);                                    \
      ctx->err = std::string(#expr) + ": " + hipGetErrorString(e_);               \
      return DOUG_E_HIP;                                                          \
    }                                                                             \
  } while (0)

static int fail(doug_ctx* ctx, int code, const std::string& msg) {
  std::lock_guard<std::mutex> g(ctx->mu);
  ctx->err = msg;
  return code;
}

static void join_reaper(doug_ctx* ctx);

// hipMalloc; when the device is full while the previous set-up's buffers are
// still being freed on the reaper thread, waits for that thread and tries
// once more.
static hipError_t device_malloc(doug_ctx* ctx, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    std::lock_guard<std::mutex> g(ctx->mu);
    if (ctx->reaper.joinable()) {
      (void)hipGetLastError();
      join_reaper(ctx);
      e = hipMalloc(p, bytes);
    }
  }
  return e;
}

template <class T>
static int dalloc(doug_ctx* ctx, T** out, size_t count) {
  void* p = nullptr;
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  hipError_t e = device_malloc(ctx, &p, bytes);
  std::lock_guard<std::mutex> g(ctx->mu);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    ctx->err = "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e);
    return DOUG_E_NOMEM;
  }
  ctx->allocs.push_back(p);
  *out = (T*)p;
  return DOUG_OK;
}

template <class T>
static int upload(doug_ctx* ctx, T** out, const std::vector<T>& h) {
  int rc = dalloc(ctx, out, h.size());
  if (rc) return rc;
  if (!h.empty()) HIPCHK(hipMemcpyAsync(*out, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice, ctx->stream));
  return DOUG_OK;
}

static void drop_iteration_graph(doug_ctx* ctx) {
  if (ctx->iter_graph) {
    (void)hipGraphExecDestroy(ctx->iter_graph);
    ctx->iter_graph = nullptr;
  }
  for (hipGraphExec_t& g : ctx->batch_graph)
    if (g) {
      (void)hipGraphExecDestroy(g);
      g = nullptr;
    }
  ctx->batch_graph_nrp = 0;
}

static void join_reaper(doug_ctx* ctx) {
  if (ctx->reaper.joinable()) ctx->reaper.join();
}

// The coarse build's device temporaries (kept between set-ups unless
// DOUG_OPT_KEEP_SETUP_TEMPS is 0).
static void free_scratch_slots(doug_ctx* ctx) {
  for (auto& q : ctx->scratch_slots)
    if (q.first) (void)hipFree(q.first);
  ctx->scratch_slots.clear();
}

static int64_t scratch_slot_bytes(const doug_ctx* ctx) {
  int64_t b = 0;
  for (const auto& q : ctx->scratch_slots) b += (int64_t)q.second;
  return b;
}

// Drops every device buffer of the last set-up.  In the background (a new
// set-up) a host thread frees them while the set-up's host work goes on:
// hipFree waits for the device and unmaps, milliseconds for the factors.
static void free_device(doug_ctx* ctx, bool background = false) {
  drop_iteration_graph(ctx);
  ctx->d_valin = nullptr;
  ctx->d_eptn = nullptr;
  ctx->d_cperm = nullptr;
  ctx->d_cW32 = nullptr;
  ctx->d_hist = nullptr;
  ctx->hist_cap = 0;
  join_reaper(ctx);
  if (background && !ctx->allocs.empty()) {
    std::vector<void*> old;
    old.swap(ctx->allocs);
    const int dev = ctx->device;
    ctx->reaper = std::thread([dev, old = std::move(old)]() {
      (void)hipSetDevice(dev);
      for (void* p : old) (void)hipFree(p);
    });
  } else {
    for (void* p : ctx->allocs) (void)hipFree(p);
    ctx->allocs.clear();
  }
  ctx->ready = false;
  ctx->batch_ready = false;
}

// CSR-stream row blocks: consecutive rows with at most kSpmvThreads rows and
// kSpmvChunk nonzeros; a longer row gets a block of its own.
static std::vector<int32_t> row_blocks(const Csr& M, int rows = kSpmvThreads) {
  const int64_t chunk = 8 * (int64_t)rows;
  std::vector<int32_t> rb{0};
  int r = 0;
  while (r < M.nrows) {
    int r1 = r;
    int64_t nz = 0;
    while (r1 < M.nrows && r1 - r < rows) {
      int64_t rn = M.ptr[r1 + 1] - M.ptr[r1];
      if (nz + rn > chunk && r1 > r) break;
      nz += rn;
      r1++;
      if (nz > chunk) break;
    }
    rb.push_back(r1);
    r = r1;
  }
  return rb;
}

// element offset of (li, lj), li >= lj inside the envelope, from a block's base
static int64_t tile_pos(const TileRow* tr, int li, int lj) {
  const int ti = li / kTile, tj = lj / kTile;
  return (int64_t)(tr[ti].off + tj - tr[ti].fc) * kTileElems + (int64_t)(li % kTile) * kTile +
         (lj % kTile);
}

// ---------------------------------------------------------------------------
extern "C" {

void doug_params_default(doug_params* p) {
  Params d;
  p->radius1 = d.radius1;
  p->radius2 = d.radius2;
  p->minasize1 = d.minasize1;
  p->minasize2 = d.minasize2;
  p->maxasize1 = d.maxasize1;
  p->maxasize2 = d.maxasize2;
  p->strong1 = d.strong1;
  p->strong2 = d.strong2;
  p->smoothers = d.smoothers;
  p->overlap = d.overlap;
  p->levels = d.levels;
  p->method = d.method;
  p->fine_method = d.fine_method;
  p->symmstruct = d.symmstruct;
  p->symmnumeric = d.symmnumeric;
  p->omega = d.omega;
  p->num_iters = d.num_iters;
  p->reserved = 0;
}

int doug_ctx_create(int device, doug_ctx** out) {
  if (!out) return DOUG_E_ARG;
  *out = nullptr;
  if (device == -1) {  // host-only context: setup products, no GPU work
    doug_ctx* ctx = new doug_ctx();
    ctx->device = -1;
    *out = ctx;
    return DOUG_OK;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    (void)hipGetLastError();
    return DOUG_E_HIP;  // no GPU: there is no CPU path
  }
  if (device < 0 || device >= ndev) return DOUG_E_ARG;
  doug_ctx* ctx = new doug_ctx();
  ctx->device = device;
  if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&ctx->stream.s, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream2, hipStreamNonBlocking) != hipSuccess ||
      hipStreamCreateWithFlags(&ctx->stream3, hipStreamNonBlocking) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_scal, 2 * S_N * sizeof(double), hipHostMallocDefault) != hipSuccess ||
      hipHostMalloc((void**)&ctx->h_scalm, kMaxRhs * 8 * sizeof(double), hipHostMallocDefault) != hipSuccess) {
    if (ctx->h_scal) (void)hipHostFree(ctx->h_scal);
    for (hipStream_t st : {ctx->stream.s, ctx->stream2, ctx->stream3})
      if (st) (void)hipStreamDestroy(st);
    delete ctx;
    return DOUG_E_HIP;
  }
  *out = ctx;
  return DOUG_OK;
}

int doug_ctx_destroy(doug_ctx* ctx) {
  if (!ctx) return DOUG_OK;
  if (ctx->device < 0) {
    delete ctx;
    return DOUG_OK;
  }
  (void)hipSetDevice(ctx->device);
  for (hipStream_t st : {ctx->stream.s, ctx->stream2, ctx->stream3})
    if (st) (void)hipStreamSynchronize(st);
  free_device(ctx);
  free_scratch_slots(ctx);
  if (ctx->h_scal) (void)hipHostFree(ctx->h_scal);
  if (ctx->h_scalm) (void)hipHostFree(ctx->h_scalm);
  for (hipEvent_t e : ctx->cevents) (void)hipEventDestroy(e);
  for (hipStream_t st : {ctx->stream.s, ctx->stream2, ctx->stream3})
    if (st) (void)hipStreamDestroy(st);
  delete ctx;
  return DOUG_OK;
}

const char* doug_last_error(const doug_ctx* ctx) { return ctx ? ctx->err.c_str() : "null context"; }

int doug_set_params(doug_ctx* ctx, const doug_params* q) {
  if (!ctx || !q) return DOUG_E_ARG;
  Params& p = ctx->p;
  p.radius1 = q->radius1;
  p.radius2 = q->radius2;
  p.minasize1 = q->minasize1;
  p.minasize2 = q->minasize2;
  p.maxasize1 = q->maxasize1;
  p.maxasize2 = q->maxasize2;
  p.strong1 = q->strong1;
  p.strong2 = q->strong2;
  p.smoothers = q->smoothers;
  p.overlap = q->overlap;
  p.levels = q->levels;
  p.method = q->method;
  p.fine_method = q->fine_method;
  p.symmstruct = q->symmstruct;
  p.symmnumeric = q->symmnumeric;
  p.omega = q->omega;
  p.num_iters = q->num_iters;
  ctx->ready = false;
  return DOUG_OK;
}

int doug_set_matrix_coo(doug_ctx* ctx, int32_t n, int64_t nnz, const int32_t* indi,
                        const int32_t* indj, const double* val, int32_t base) {
  if (!ctx) return DOUG_E_ARG;
  if (n <= 0 || nnz <= 0 || !indi || !indj || !val || (base != 0 && base != 1))
    return fail(ctx, DOUG_E_ARG, "doug_set_matrix_coo: bad arguments");
  if (nnz > INT32_MAX) return fail(ctx, DOUG_E_UNSUPPORTED, "nnz above 2^31-1 is not supported");
  std::string err;
  Csr A = coo_to_csr(n, nnz, indi, indj, val, base, err);
  if (!err.empty()) return fail(ctx, DOUG_E_ARG, err);
  ctx->A = std::move(A);
  ctx->A_input_val = ctx->A.val;
  ctx->n = n;
  ctx->have_matrix = true;
  ctx->dup_state = -1;
  ctx->ready = false;
  return DOUG_OK;
}

int doug_set_num_subdomains(doug_ctx* ctx, int32_t nsub) {
  if (!ctx) return DOUG_E_ARG;
  if (nsub < 0) return fail(ctx, DOUG_E_ARG, "number of subdomains must be >= 0");
  ctx->num_subdomains = nsub;
  ctx->ready = false;
  return DOUG_OK;
}

int doug_set_aggregates(doug_ctx* ctx, int32_t nagr, const int32_t* num) {
  if (!ctx) return DOUG_E_ARG;
  ctx->ready = ctx->host_ready = false;
  if (nagr <= 0) {
    ctx->inj_nagr = 0;
    ctx->inj_num.clear();
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_aggregates before doug_set_matrix_coo");
  if (!num) return fail(ctx, DOUG_E_ARG, "num is NULL");
  std::vector<int> cnt((size_t)nagr + 1, 0);
  for (int i = 0; i < ctx->n; i++) {
    if (num[i] < 0 || num[i] > nagr) return fail(ctx, DOUG_E_ARG, "aggregate number out of range");
    cnt[num[i]]++;
  }
  for (int a = 1; a <= nagr; a++)
    if (cnt[a] == 0) return fail(ctx, DOUG_E_ARG, "aggregate " + std::to_string(a) + " holds no unknown");
  ctx->inj_num.assign(num, num + ctx->n);
  ctx->inj_nagr = nagr;
  return DOUG_OK;
}

int doug_set_strong_mask(doug_ctx* ctx, int64_t nnz, const uint8_t* strong) {
  if (!ctx) return DOUG_E_ARG;
  ctx->ready = ctx->host_ready = false;
  if (!strong) {
    ctx->inj_strong.clear();
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_strong_mask before doug_set_matrix_coo");
  if (nnz != (int64_t)ctx->A.src.size())
    return fail(ctx, DOUG_E_ARG, "the strength mask needs one flag per matrix entry (" +
                                     std::to_string(ctx->A.src.size()) + ")");
  ctx->inj_strong.assign(strong, strong + nnz);
  return DOUG_OK;
}

int doug_set_restriction_coo(doug_ctx* ctx, int32_t nagr, int64_t nnz, const int32_t* ri,
                             const int32_t* rj, const double* rv, int32_t base) {
  if (!ctx) return DOUG_E_ARG;
  ctx->ready = ctx->host_ready = false;
  if (nagr <= 0) {
    ctx->have_inj_R = false;
    ctx->inj_R = Csr();
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_restriction_coo before doug_set_matrix_coo");
  if (nnz < 0 || (nnz > 0 && (!ri || !rj || !rv)) || (base != 0 && base != 1))
    return fail(ctx, DOUG_E_ARG, "bad restriction triples");
  // sorted by (row, column), a repeated pair summed in input order
  // (SpMtx_consolidate, SpMtx_arrangement.F90:258-292)
  std::vector<int64_t> ord((size_t)nnz);
  for (int64_t t = 0; t < nnz; t++) {
    const int64_t r = (int64_t)ri[t] - base, c = (int64_t)rj[t] - base;
    if (r < 0 || r >= nagr || c < 0 || c >= ctx->n)
      return fail(ctx, DOUG_E_ARG, "restriction entry " + std::to_string(t) + " out of range");
    ord[(size_t)t] = t;
  }
  std::stable_sort(ord.begin(), ord.end(), [&](int64_t a, int64_t b) {
    return ri[a] != ri[b] ? ri[a] < ri[b] : rj[a] < rj[b];
  });
  Csr R;
  R.nrows = nagr;
  R.ncols = ctx->n;
  R.ptr.assign((size_t)nagr + 1, 0);
  for (int64_t t : ord) {
    const int r = ri[t] - base, c = rj[t] - base;
    if (!R.col.empty() && R.ptr[r + 1] > 0 && R.col.back() == c) {
      R.val.back() = R.val.back() + rv[t];
      continue;
    }
    R.col.push_back(c);
    R.val.push_back(rv[t]);
    R.ptr[r + 1]++;
  }
  for (int r = 0; r < nagr; r++) R.ptr[r + 1] += R.ptr[r];
  ctx->inj_R = std::move(R);
  ctx->have_inj_R = true;
  return DOUG_OK;
}

int doug_set_subdomains(doug_ctx* ctx, int32_t nsub, const int32_t* starts, const int32_t* inds,
                        int32_t base) {
  if (!ctx) return DOUG_E_ARG;
  ctx->ready = ctx->host_ready = false;
  if (nsub <= 0) {
    ctx->inj_sstarts.clear();
    ctx->inj_sinds.clear();
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_subdomains before doug_set_matrix_coo");
  if (!starts || !inds || starts[0] != base || (base != 0 && base != 1))
    return fail(ctx, DOUG_E_ARG, "bad subdomain starts");
  std::vector<int> st((size_t)nsub + 1), in;
  for (int q = 0; q <= nsub; q++) st[q] = starts[q] - base;
  std::vector<int> seen((size_t)ctx->n, -1);
  for (int q = 0; q < nsub; q++) {
    if (st[q + 1] <= st[q]) return fail(ctx, DOUG_E_ARG, "subdomain " + std::to_string(q + base) + " is empty");
    for (int k = st[q]; k < st[q + 1]; k++) {
      const int v = inds[k] - base;
      if (v < 0 || v >= ctx->n) return fail(ctx, DOUG_E_ARG, "subdomain index out of range");
      if (seen[v] == q) return fail(ctx, DOUG_E_ARG, "an unknown appears twice in one subdomain");
      seen[v] = q;
      in.push_back(v);
    }
  }
  ctx->inj_sstarts.swap(st);
  ctx->inj_sinds.swap(in);
  return DOUG_OK;
}

int doug_set_ranks(doug_ctx* ctx, int32_t nranks, const int32_t* rank_of_node, int32_t base) {
  if (!ctx) return DOUG_E_ARG;
  if (nranks <= 0) {
    ctx->nranks = 0;
    ctx->eptn.clear();
    ctx->rsetup = RankSetup{};
    ctx->rank_ws.clear();
    ctx->ready = ctx->host_ready = false;
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_ranks before doug_set_matrix_coo");
  if (base != 0 && base != 1) return fail(ctx, DOUG_E_ARG, "base must be 0 or 1");
  const int n = ctx->n;
  std::vector<int> eptn((size_t)n);
  if (rank_of_node) {
    for (int i = 0; i < n; i++) {
      eptn[i] = rank_of_node[i] - base;
      if (eptn[i] < 0 || eptn[i] >= nranks) return fail(ctx, DOUG_E_ARG, "rank of unknown " + std::to_string(i + base) + " out of range");
    }
  } else {
    // SpMtx_buildAdjncy + METIS on the matrix graph (Distribution_assm.F90:
    // 191-197): here the project partitioner on the symmetrised pattern
    const Csr& A = ctx->A;
    std::vector<std::vector<int>> adj((size_t)n);
    for (int i = 0; i < n; i++)
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
        const int j = A.col[p];
        if (j == i) continue;
        adj[i].push_back(j);
        adj[j].push_back(i);
      }
    std::vector<int> xadj((size_t)n + 1, 0), adjncy;
    for (int i = 0; i < n; i++) {
      std::sort(adj[i].begin(), adj[i].end());
      adj[i].erase(std::unique(adj[i].begin(), adj[i].end()), adj[i].end());
      adjncy.insert(adjncy.end(), adj[i].begin(), adj[i].end());
      xadj[i + 1] = (int)adjncy.size();
    }
    partition_graph(n, xadj.data(), adjncy.data(), nullptr, nranks, eptn);
  }
  std::vector<int> cnt((size_t)nranks, 0);
  for (int r : eptn) cnt[r]++;
  for (int r = 0; r < nranks; r++)
    if (cnt[r] == 0) return fail(ctx, DOUG_E_ARG, "rank " + std::to_string(r + base) + " owns no unknown");
  ctx->eptn.swap(eptn);
  ctx->nranks = nranks;
  ctx->ready = ctx->host_ready = false;
  return DOUG_OK;
}

int doug_get_rank_map(const doug_ctx* ctx, int32_t rank, int32_t* ninner, int32_t* nlocal, int32_t* lg) {
  if (!ctx) return DOUG_E_ARG;
  if (!ctx->host_ready || ctx->nranks <= 0) return DOUG_E_STATE;
  if (rank < 0 || rank >= (int)ctx->rsetup.ranks.size()) return DOUG_E_ARG;
  const RankLocal& R = ctx->rsetup.ranks[rank];
  if (ninner) *ninner = R.ninner;
  if (nlocal) *nlocal = (int32_t)R.lg.size();
  if (lg)
    for (size_t k = 0; k < R.lg.size(); k++) lg[k] = R.lg[k] + 1;
  return DOUG_OK;
}

int doug_partition_graph(int32_t n, const int32_t* xadj, const int32_t* adjncy, const int32_t* vwgt,
                         int32_t nparts, int32_t base, int32_t* part) {
  if (n < 0 || nparts < 1 || !xadj || !part || (base != 0 && base != 1)) return DOUG_E_ARG;
  if (n == 0) return DOUG_OK;
  std::vector<int> x((size_t)n + 1), a((size_t)(xadj[n] - xadj[0]));
  for (int i = 0; i <= n; i++) x[i] = xadj[i] - xadj[0];
  for (size_t p = 0; p < a.size(); p++) {
    a[p] = adjncy[p] - base;
    if (a[p] < 0 || a[p] >= n) return DOUG_E_ARG;
  }
  std::vector<int> out;
  partition_graph(n, x.data(), a.data(), vwgt, nparts, out);
  for (int i = 0; i < n; i++) part[i] = out[i] + base;
  return DOUG_OK;
}

int doug_set_partition(doug_ctx* ctx, int32_t nparts, const int32_t* starts, const int32_t* nodes,
                       int32_t base) {
  if (!ctx) return DOUG_E_ARG;
  if (nparts <= 0) {
    ctx->have_partition = false;
    ctx->part_starts.clear();
    ctx->part_nodes.clear();
    return DOUG_OK;
  }
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "doug_set_partition before doug_set_matrix_coo");
  if (!starts || !nodes || starts[0] != base) return fail(ctx, DOUG_E_ARG, "bad partition");
  int64_t tot = (int64_t)starts[nparts] - base;
  if (tot != ctx->n) return fail(ctx, DOUG_E_ARG, "partition must list every unknown once");
  std::vector<char> seen((size_t)ctx->n, 0);
  ctx->part_starts.assign((size_t)nparts + 1, 0);
  ctx->part_nodes.resize((size_t)tot);
  for (int i = 0; i <= nparts; i++) {
    ctx->part_starts[i] = starts[i] - base;
    if (i > 0 && ctx->part_starts[i] < ctx->part_starts[i - 1]) return fail(ctx, DOUG_E_ARG, "partition starts must be non-decreasing");
  }
  for (int64_t t = 0; t < tot; t++) {
    int v = nodes[t] - base;
    if (v < 0 || v >= ctx->n || seen[v]) return fail(ctx, DOUG_E_ARG, "partition must list every unknown once");
    seen[v] = 1;
    ctx->part_nodes[t] = v;
  }
  ctx->have_partition = true;
  ctx->ready = false;
  return DOUG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Tile storage for a set of index blocks: descriptors, tile-row tables,
// scatter map and factorisation on the device.  With `envelope` each block
// keeps only the tiles inside the lower envelope of M(inds, inds) (its row
// order is the order of `inds`); otherwise every lower tile.  `offcount[j]`
// receives max over blocks of lr_j - j.  Everything is enqueued on the
// context stream, so that host setup work can overlap the factorisation.
template <typename T>
static int build_blocks(doug_ctx* ctx, const Csr& M, const double* d_vals,
                        const std::vector<int>& starts, const std::vector<int>& inds,
                        bool one_block_solve, bool envelope, std::vector<BlockDesc>& desc,
                        BlockDesc** d_desc, T** d_tiles, int64_t* total_elems,
                        std::vector<int>& offcount, double* chol_flops = nullptr) {
  const int nb = (int)starts.size() - 1;
  desc.resize((size_t)nb);
  std::vector<int64_t> row0((size_t)nb + 1, 0);
  int ntmax = 0;
  for (int b = 0; b < nb; b++) {
    int s = starts[b + 1] - starts[b];
    int nt = (s + kTile - 1) / kTile;
    if (nt > kMaxTrsvTiles && one_block_solve)
      return fail(ctx, DOUG_E_UNSUPPORTED,
                  "subdomain of " + std::to_string(s) + " unknowns exceeds the one-block solve limit " +
                      std::to_string(kMaxTrsvTiles * kTile));
    row0[b + 1] = row0[b] + nt;
    ntmax = std::max(ntmax, nt);
  }
  std::vector<TileRow> rows((size_t)row0[nb]);
  const int nth = setup_threads();
  auto for_blocks = [&](auto&& body) {
    std::vector<std::thread> th;
    for (int t = 0; t < nth; t++)
      th.emplace_back([&, t]() {
        std::vector<int> floc((size_t)M.nrows, -1);
        for (int b = t; b < nb; b += nth) body(t, b, floc);
      });
    for (auto& x : th) x.join();
  };
  // tile-row envelopes: fc from the first lower entry of each row, made
  // non-decreasing so that every column's tiles are contiguous
  for_blocks([&](int, int b, std::vector<int>& floc) {
    const int s0 = starts[b], s = starts[b + 1] - s0, nt = (s + kTile - 1) / kTile;
    TileRow* tr = rows.data() + row0[b];
    for (int i = 0; i < nt; i++) tr[i] = TileRow{0, envelope ? i : 0, i, 0};
    if (envelope) {
      for (int li = 0; li < s; li++) floc[inds[s0 + li]] = li;
      for (int li = 0; li < s; li++) {
        int g = inds[s0 + li], f = li;
        for (int p = M.ptr[g]; p < M.ptr[g + 1]; p++) {
          int lj = floc[M.col[p]];
          if (lj >= 0 && lj < f) f = lj;
        }
        TileRow& r = tr[li / kTile];
        r.fc = std::min(r.fc, f / kTile);
      }
      for (int li = 0; li < s; li++) floc[inds[s0 + li]] = -1;
      for (int i = nt - 2; i >= 0; i--) tr[i].fc = std::min(tr[i].fc, tr[i + 1].fc);
    }
    int off = 0;
    for (int i = 0; i < nt; i++) {
      tr[i].off = off;
      off += i - tr[i].fc + 1;
      for (int j = tr[i].fc; j <= i; j++) tr[j].lr = std::max(tr[j].lr, i);
    }
  });
  int rc;
  TileRow* d_rows;
  if ((rc = upload(ctx, &d_rows, rows))) return rc;
  int64_t base = 0, woff = 0;
  offcount.assign((size_t)ntmax, 0);
  for (int b = 0; b < nb; b++) {
    const int s = starts[b + 1] - starts[b], nt = (int)(row0[b + 1] - row0[b]);
    const TileRow* tr = rows.data() + row0[b];
    desc[b] = BlockDesc{base, (int64_t)starts[b], woff, d_rows + row0[b], nt, s};
    base += (int64_t)(tr[nt - 1].off + nt - tr[nt - 1].fc) * kTileElems;
    woff += (int64_t)nt * kTile;
    for (int j = 0; j < nt; j++) offcount[j] = std::max(offcount[j], tr[j].lr - j);
  }
  *total_elems = base;
  if (chol_flops) {
    // matrix-core work of the tile factorisation: per step j, (j - fc_j)
    // SYRK tiles, one 64x64 POTRF and inverse, and for each tile (i, j) below
    // the diagonal (j - fc_i) GEMM tiles plus one triangular product
    const double g = 2.0 * kTile * kTile * kTile;
    double f = 0.0;
    for (int b = 0; b < nb; b++) {
      const TileRow* tr = rows.data() + row0[b];
      for (int j = 0; j < desc[b].nt; j++) {
        f += g * (j - tr[j].fc) + 2.0 * g / 3.0;
        for (int i = j + 1; i <= tr[j].lr; i++) f += g * (j - tr[i].fc + 1);
      }
    }
    *chol_flops = f;
  }
  // scatter map: lower-triangle entries of M(inds, inds), blocks in parallel.
  // A pair (i, j) that M stores several times is summed, as the reference's
  // factorisation does (umf4triplet2col in Fact.F90:102-110 adds duplicate
  // triplets): its first occurrence stays in the plain scatter and the group
  // of all its occurrences, in stored order, goes to the summing pass.
  std::vector<std::vector<int64_t>> dsts((size_t)nth), gdsts((size_t)nth), gcnts((size_t)nth);
  std::vector<std::vector<int32_t>> srcs((size_t)nth), gsrcs((size_t)nth);
  for_blocks([&](int t, int b, std::vector<int>& floc) {
    std::vector<int64_t> chk;
    std::vector<char> rep;
    const TileRow* tr = rows.data() + row0[b];
    const int s0 = starts[b], s = starts[b + 1] - s0;
    for (int li = 0; li < s; li++) floc[inds[s0 + li]] = li;
    for (int li = 0; li < s; li++) {
      int g = inds[s0 + li];
      const size_t r0 = dsts[t].size();
      for (int p = M.ptr[g]; p < M.ptr[g + 1]; p++) {
        int lj = floc[M.col[p]];
        if (lj < 0 || lj > li) continue;
        dsts[t].push_back(desc[b].base + tile_pos(tr, li, lj));
        srcs[t].push_back(p);
      }
      const size_t r1 = dsts[t].size();
      if (r1 - r0 < 2) continue;
      chk.assign(dsts[t].begin() + r0, dsts[t].end());
      std::sort(chk.begin(), chk.end());
      if (std::adjacent_find(chk.begin(), chk.end()) == chk.end()) continue;
      // this row holds a repeated pair: group the occurrences, keep firsts
      rep.assign(r1 - r0, 0);
      size_t w = r0;
      for (size_t a = r0; a < r1; a++) {
        if (rep[a - r0]) continue;
        int64_t cnt = 1;
        for (size_t c = a + 1; c < r1; c++)
          if (dsts[t][c] == dsts[t][a]) { rep[c - r0] = 1; cnt++; }
        if (cnt > 1) {
          gdsts[t].push_back(dsts[t][a]);
          gcnts[t].push_back(cnt);
          for (size_t c = a; c < r1; c++)
            if (dsts[t][c] == dsts[t][a]) gsrcs[t].push_back(srcs[t][c]);
        }
        dsts[t][w] = dsts[t][a];
        srcs[t][w] = srcs[t][a];
        w++;
      }
      dsts[t].resize(w);
      srcs[t].resize(w);
    }
    for (int li = 0; li < s; li++) floc[inds[s0 + li]] = -1;
  });
  std::vector<int64_t> dst, gdst, gseg(1, 0);
  std::vector<int32_t> src, gsrc;
  for (int t = 0; t < nth; t++) {
    dst.insert(dst.end(), dsts[t].begin(), dsts[t].end());
    src.insert(src.end(), srcs[t].begin(), srcs[t].end());
    gdst.insert(gdst.end(), gdsts[t].begin(), gdsts[t].end());
    gsrc.insert(gsrc.end(), gsrcs[t].begin(), gsrcs[t].end());
    for (int64_t c : gcnts[t]) gseg.push_back(gseg.back() + c);
  }
  if ((rc = dalloc(ctx, d_tiles, (size_t)base))) return rc;
  if ((rc = upload(ctx, d_desc, desc))) return rc;
  int64_t* d_dst;
  int32_t* d_src;
  if ((rc = upload(ctx, &d_dst, dst))) return rc;
  if ((rc = upload(ctx, &d_src, src))) return rc;
  HIPCHK(hipMemsetAsync(*d_tiles, 0, (size_t)base * sizeof(T), ctx->stream));
  int maxpad = 0;
  for (auto& d : desc) maxpad = std::max(maxpad, d.nt * kTile - d.s);
  launch_pad_identity(*d_tiles, *d_desc, nb, maxpad, ctx->stream);
  launch_scatter_entries(*d_tiles, d_dst, d_src, d_vals, (int64_t)dst.size(), ctx->stream);
  if (!gdst.empty()) {
    int64_t *d_gdst, *d_gseg;
    int32_t* d_gsrc;
    if ((rc = upload(ctx, &d_gdst, gdst)) || (rc = upload(ctx, &d_gseg, gseg)) ||
        (rc = upload(ctx, &d_gsrc, gsrc)))
      return rc;
    launch_scatter_sums(*d_tiles, d_gdst, d_gseg, d_gsrc, d_vals, (int64_t)gdst.size(), ctx->stream);
  }
  // the host copies above must stay alive until the stream consumed them
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipGetLastError());
  return DOUG_OK;
}

// Whether some row of A holds the same column twice.
static bool has_duplicate_entries(const Csr& A) {
  std::atomic<bool> dup{false};
  const int nth = setup_threads();
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back([&, t]() {
      std::vector<int> row;
      const int b = (int)((int64_t)A.nrows * t / nth), e = (int)((int64_t)A.nrows * (t + 1) / nth);
      for (int i = b; i < e && !dup.load(std::memory_order_relaxed); i++) {
        row.assign(A.col.begin() + A.ptr[i], A.col.begin() + A.ptr[i + 1]);
        std::sort(row.begin(), row.end());
        if (std::adjacent_find(row.begin(), row.end()) != row.end()) dup = true;
      }
    });
  for (auto& x : th) x.join();
  return dup;
}

// build_blocks for the subdomain factors (envelope storage, one-block solve
// limit) with the envelope and the entries taken on the device from the
// uploaded matrix, unknown lists (ctx->d_inds) and occurrence lists, so the
// host builds no scatter map.  A must have no duplicate (i, j) pair.
template <typename T>
static int build_blocks_device(doug_ctx* ctx, const std::vector<int>& starts, std::vector<BlockDesc>& desc,
                               BlockDesc** d_desc, T** d_tiles, int64_t* total_elems,
                               std::vector<int>& offcount, double* chol_flops) {
  const int nb = (int)starts.size() - 1;
  desc.resize((size_t)nb);
  std::vector<int32_t> row0((size_t)nb + 1, 0), st32(starts.begin(), starts.end());
  int ntmax = 0;
  for (int b = 0; b < nb; b++) {
    int s = starts[b + 1] - starts[b];
    int nt = (s + kTile - 1) / kTile;
    if (nt > kMaxTrsvTiles)
      return fail(ctx, DOUG_E_UNSUPPORTED,
                  "subdomain of " + std::to_string(s) + " unknowns exceeds the one-block solve limit " +
                      std::to_string(kMaxTrsvTiles * kTile));
    row0[b + 1] = row0[b] + nt;
    ntmax = std::max(ntmax, nt);
  }
  std::vector<int32_t> fc((size_t)row0[nb]);
  for (int b = 0; b < nb; b++)
    for (int i = 0; i < row0[b + 1] - row0[b]; i++) fc[row0[b] + i] = i;
  int rc;
  int32_t *d_st, *d_r0, *d_fc;
  if ((rc = upload(ctx, &d_st, st32)) || (rc = upload(ctx, &d_r0, row0)) || (rc = upload(ctx, &d_fc, fc)))
    return rc;
  launch_block_envelope(d_st, d_r0, nb, ctx->d_inds, ctx->d_occ_ptr, ctx->d_occ_pos, ctx->d_ptr,
                        ctx->d_col, d_fc, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(fc.data(), d_fc, fc.size() * sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  // tile-row tables as build_blocks makes them from the first columns
  std::vector<TileRow> rows((size_t)row0[nb]);
  for (int b = 0; b < nb; b++) {
    const int nt = row0[b + 1] - row0[b];
    TileRow* tr = rows.data() + row0[b];
    for (int i = 0; i < nt; i++) tr[i] = TileRow{0, fc[row0[b] + i], i, 0};
    for (int i = nt - 2; i >= 0; i--) tr[i].fc = std::min(tr[i].fc, tr[i + 1].fc);
    int off = 0;
    for (int i = 0; i < nt; i++) {
      tr[i].off = off;
      off += i - tr[i].fc + 1;
      for (int j = tr[i].fc; j <= i; j++) tr[j].lr = std::max(tr[j].lr, i);
    }
  }
  TileRow* d_rows;
  if ((rc = upload(ctx, &d_rows, rows))) return rc;
  int64_t base = 0, woff = 0;
  offcount.assign((size_t)ntmax, 0);
  for (int b = 0; b < nb; b++) {
    const int s = starts[b + 1] - starts[b], nt = row0[b + 1] - row0[b];
    const TileRow* tr = rows.data() + row0[b];
    desc[b] = BlockDesc{base, (int64_t)starts[b], woff, d_rows + row0[b], nt, s};
    base += (int64_t)(tr[nt - 1].off + nt - tr[nt - 1].fc) * kTileElems;
    woff += (int64_t)nt * kTile;
    for (int j = 0; j < nt; j++) offcount[j] = std::max(offcount[j], tr[j].lr - j);
  }
  *total_elems = base;
  if (chol_flops) {
    const double g = 2.0 * kTile * kTile * kTile;
    double f = 0.0;
    for (int b = 0; b < nb; b++) {
      const TileRow* tr = rows.data() + row0[b];
      for (int j = 0; j < desc[b].nt; j++) {
        f += g * (j - tr[j].fc) + 2.0 * g / 3.0;
        for (int i = j + 1; i <= tr[j].lr; i++) f += g * (j - tr[i].fc + 1);
      }
    }
    *chol_flops = f;
  }
  if ((rc = dalloc(ctx, d_tiles, (size_t)base))) return rc;
  if ((rc = upload(ctx, d_desc, desc))) return rc;
  HIPCHK(hipMemsetAsync(*d_tiles, 0, (size_t)base * sizeof(T), ctx->stream));
  int maxpad = 0;
  for (auto& d : desc) maxpad = std::max(maxpad, d.nt * kTile - d.s);
  launch_pad_identity(*d_tiles, *d_desc, nb, maxpad, ctx->stream);
  launch_scatter_blocks(*d_tiles, *d_desc, nb, ctx->d_inds, ctx->d_occ_ptr, ctx->d_occ_pos, ctx->d_ptr,
                        ctx->d_col, ctx->d_val, ctx->stream);
  // the host tables above must stay alive until the stream consumed them
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipGetLastError());
  return DOUG_OK;
}

// Enqueue the left-looking tile Cholesky of every block; a non-positive
// pivot sets *flag (checked once the stream has drained).
template <typename T>
static int enqueue_factor(doug_ctx* ctx, T* tiles, const BlockDesc* d_desc,
                          const std::vector<BlockDesc>& desc, const std::vector<int>& offcount,
                          int* flag) {
  for (int j = 0; j < (int)offcount.size(); j++) {
    launch_chol_diag(tiles, d_desc, (int)desc.size(), j, flag, ctx->stream);
    launch_chol_offdiag(tiles, d_desc, (int)desc.size(), j, offcount[j], ctx->stream);
  }
  HIPCHK(hipGetLastError());
  return DOUG_OK;
}

extern "C" {

// ---- setup phases --------------------------------------------------------
// (1) values and strength: SpMtx_scale / SpMtx_find_strong / SpMtx_unscale
//     (SpMtx_arrangement.F90:298-388, SpMtx_aggregation.F90:829-899)
static void phase_values(doug_ctx* ctx, std::vector<double>& scaled, std::vector<uint8_t>& strong) {
  const Params& p = ctx->p;
  Csr& A = ctx->A;
  const bool symm = p.symmstruct && p.symmnumeric;
  const double strong_conn1 = p.strong1 != 0.0 ? p.strong1 : 0.67;
  std::vector<double> diag;
  // the reference continues with the values after the scale/unscale round trip
  scale_round_trip(A, ctx->A_input_val, symm, scaled, diag, A.val);
  strong = find_strong(A, scaled, diag, strong_conn1);
  if (!ctx->inj_strong.empty())  // the caller's flags, moved to the row order
    for (size_t q = 0; q < strong.size(); q++) strong[q] = ctx->inj_strong[(size_t)A.src[q]] != 0;
}

// virtual_ranks for the context's decomposition and aggregation parameters
static void run_virtual_ranks(doug_ctx* ctx, RankSetup& out, const Csr* rcm, bool keep_local) {
  const Params& p = ctx->p;
  const double strong_conn1 = p.strong1 != 0.0 ? p.strong1 : 0.67;
  const int radius1 = p.radius1 > 0 ? p.radius1 : 2;
  const int min_asize1 = p.minasize1 > 0 ? p.minasize1 : (int)(0.5 * (double)((2 * radius1 + 1) * (2 * radius1 + 1)));
  const int max_asize1 = p.maxasize1 > 0 ? p.maxasize1 : (2 * radius1 + 1) * (2 * radius1 + 1);
  if (rows_sorted_by_column(ctx->A)) {
    virtual_ranks(ctx->A, ctx->A_input_val, ctx->eptn, ctx->nranks, ctx->ol_dist, ctx->ol,
                  ctx->num_subdomains, radius1, min_asize1, max_asize1, strong_conn1, out, rcm,
                  &ctx->rank_ws, keep_local);
  } else {
    Csr As = sort_rows_by_column(ctx->A, ctx->A_input_val);
    virtual_ranks(As, As.val, ctx->eptn, ctx->nranks, ctx->ol_dist, ctx->ol, ctx->num_subdomains,
                  radius1, min_asize1, max_asize1, strong_conn1, out, rcm, &ctx->rank_ws, keep_local,
                  &As.src);
  }
}

// (2) fine aggregates, smoothed restriction and Galerkin coarse matrix
//     (Partitioning_aggr.F90:33-117, CoarsePreconditioner_smooth.F90:47-89)
static void phase_aggregate(doug_ctx* ctx, const std::vector<double>& scaled,
                            const std::vector<uint8_t>& strong) {
  const Params& p = ctx->p;
  double t0 = now();
  const double strong_conn1 = p.strong1 != 0.0 ? p.strong1 : 0.67;
  const int radius1 = p.radius1 > 0 ? p.radius1 : 2;
  const int min_asize1 = p.minasize1 > 0 ? p.minasize1 : (int)(0.5 * (double)((2 * radius1 + 1) * (2 * radius1 + 1)));
  const int max_asize1 = p.maxasize1 > 0 ? p.maxasize1 : (2 * radius1 + 1) * (2 * radius1 + 1);
  if (ctx->inj_nagr > 0) {
    // the caller's aggregates; the full ones add the unaggregated unknowns
    // one by one in node order (SpMtx_aggregation.F90:562-575)
    const int n = ctx->A.nrows;
    std::vector<int> full(ctx->inj_num);
    int nf = ctx->inj_nagr;
    for (int i = 0; i < n; i++)
      if (full[i] == 0) full[i] = ++nf;
    auto form = [n](int nagr, const std::vector<int>& num, Aggregates& a) {
      a.nagr = nagr;
      a.num = num;
      a.starts.assign((size_t)nagr + 1, 0);
      for (int x : num)
        if (x > 0) a.starts[x]++;
      for (int c = 0; c < nagr; c++) a.starts[c + 1] += a.starts[c];
      a.nodes.assign((size_t)a.starts[nagr], 0);
      std::vector<int> fill(a.starts.begin(), a.starts.end() - 1);
      for (int i = 0; i < n; i++)
        if (num[i] > 0) a.nodes[fill[num[i] - 1]++] = i;
    };
    form(ctx->inj_nagr, ctx->inj_num, ctx->fin);
    form(nf, full, ctx->ffull);
  } else if (ctx->nranks > 0) {
    // each rank aggregates its own rows of the values as read
    // with at least one rank per worker thread the ranks also put their
    // subdomains in the factor order of the device copy (reverse
    // Cuthill-McKee over A's pattern); with fewer ranks the threaded
    // block_rcm of the subdomain preparation keeps every thread busy.
    // When the subdomain work is prepared beside the coarse build (doug_setup)
    // that order is left to it too: only the aggregates are needed before
    // the coarse space can start.
    const bool beside = p.fine_method == 1 && ctx->ol_dist == 1 && ctx->smoothers == 1 &&
                        ctx->use_device_coarse && ctx->inj_nagr == 0 && ctx->has_coarse;
    const Csr* rcm = (p.fine_method == 1 && ctx->local_order == 1 && ctx->device >= 0 &&
                      ctx->nranks >= setup_threads() && !beside)
                         ? &ctx->A
                         : nullptr;
    // the device coarse build forms the several-process coarse space from
    // the whole matrix: the ranks' local matrices (and their ghost rows) are
    // then not kept; the host build, should it be needed, forms them again
    ctx->ranks_kept_local = !beside || ctx->device < 0;
    run_virtual_ranks(ctx, ctx->rsetup, rcm, ctx->ranks_kept_local);
    ctx->fin = std::move(ctx->rsetup.inner);
    ctx->ffull = std::move(ctx->rsetup.full);
  } else {
    aggregate(ctx->A, scaled, strong, radius1, min_asize1, max_asize1, strong_conn1, false, ctx->ps,
              ctx->fin, ctx->ffull);
  }
  ctx->info.t_aggregate = now() - t0;
}

// several processes: each aggregate's row of R as its owner builds it and
// A_c as the sum of the processes' parts (ranks_coarse_space), from the
// local matrices the rank workers kept
static void phase_coarse_ranks(doug_ctx* ctx, const std::vector<uint8_t>& strong) {
  double t1 = now();
  ranks_coarse_space(ctx->rsetup, ctx->A.nrows, ctx->A.val, ctx->A_input_val, strong,
                     ctx->smoothers, ctx->p.omega, ctx->R, ctx->Rt, ctx->AC);
  for (RankLocal& L : ctx->rsetup.ranks) {
    L.A = Csr();
    L.agg = Aggregates();
  }
  ctx->coarse_on_device = false;
  ctx->info.t_coarse_build = now() - t1;
}

// the caller's R (doug_set_restriction_coo): R^T and A_c = R A R^T on the
// host, as phase_coarse_host forms them
static int phase_coarse_given(doug_ctx* ctx) {
  double t1 = now();
  const bool coarse_parts = ctx->inj_sstarts.empty() && !ctx->have_partition && ctx->num_subdomains <= 1;
  if (coarse_parts && ctx->inj_R.nrows != ctx->fin.nagr)
    return fail(ctx, DOUG_E_ARG, "the restriction has " + std::to_string(ctx->inj_R.nrows) +
                                     " rows and the fine aggregation " + std::to_string(ctx->fin.nagr) +
                                     " aggregates: the subdomains from the coarse aggregation need them "
                                     "equal (or give a partition)");
  ctx->R = ctx->inj_R;
  ctx->Rt = transpose(ctx->R);
  ctx->AC = galerkin(ctx->A, ctx->R, ctx->Rt);
  ctx->coarse_on_device = false;
  ctx->info.t_coarse_build = now() - t1;
  return DOUG_OK;
}

// whether the last aggregation was the virtual ranks' (their local matrices
// are kept for the coarse space)
static bool rank_coarse(const doug_ctx* ctx) { return ctx->nranks > 0 && ctx->inj_nagr == 0; }

// host build of R, R^T and A_c (the host-only setup and smoothers = 0)
static void phase_coarse_host(doug_ctx* ctx, const std::vector<uint8_t>& strong) {
  double t1 = now();
  ctx->R = restriction(ctx->A, strong, ctx->fin, ctx->smoothers, ctx->p.omega);
  ctx->Rt = transpose(ctx->R);
  ctx->AC = galerkin(ctx->A, ctx->R, ctx->Rt);
  ctx->coarse_on_device = false;
  ctx->info.t_coarse_build = now() - t1;
}

}  // extern "C"

// Device build of the smoothed restriction and the Galerkin matrix (the
// same bits as phase_coarse_host): R^T row by row, R by a stable sort of
// R^T's entries on their aggregate, T' = A R^T and A_c = R T' by
// first-touch products.  R and R^T stay on the device for the
// preconditioner; R and A_c are copied to the host (getters, coarse factor).
// Temporaries of one device build, in slots that the context keeps between
// set-ups: the k-th request takes slot k, grown when too small, so that a
// repeated set-up of the same size neither allocates nor frees (hipFree
// waits for the device).  The set-up ends with the device idle, so the next
// one may overwrite them.
struct DevTemp {
  doug_ctx* ctx;
  std::vector<std::pair<void*, size_t>>& slots;
  size_t next = 0;
  explicit DevTemp(doug_ctx* c) : ctx(c), slots(c->scratch_slots) {}
  template <typename T>
  hipError_t alloc(T** q, size_t count) {
    const size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    if (next == slots.size()) slots.emplace_back(nullptr, 0);
    std::pair<void*, size_t>& sl = slots[next++];
    if (sl.second < bytes) {
      if (sl.first) (void)hipFree(sl.first);
      sl = {nullptr, 0};
      hipError_t e = device_malloc(ctx, &sl.first, bytes);
      if (e != hipSuccess) {
        sl.first = nullptr;
        return e;
      }
      sl.second = bytes;
    }
    *q = (T*)sl.first;
    return hipSuccess;
  }
};

// With `ranks` (virtual ranks, overlap 1): DOUG's several-process coarse
// space -- the same bits as ranks_coarse_space, see doug_coarse.hip.  Returns
// DOUG_E_UNSUPPORTED without side effects on the results when a row of A_c
// has more columns than the wavefront kernel holds (the caller then builds
// on the host).
static int coarse_build_device(doug_ctx* ctx, const std::vector<uint8_t>& strong, bool ranks = false) {
  double t1 = now();
  const Csr& A = ctx->A;
  const int n = A.nrows, nagr = ctx->fin.nagr;
  const int64_t nnz = A.nnz();
  hipStream_t st = ctx->stream;
  DevTemp tmp(ctx);
  int rc;
  void* scan_tmp = nullptr;
  size_t scan_cap = 0;
  // hipcub temporary storage, grown on demand
  auto tmp_bytes = [&](size_t bytes) -> int {
    if (bytes <= scan_cap) return DOUG_OK;
    HIPCHK(tmp.alloc((char**)&scan_tmp, bytes));
    scan_cap = bytes;
    return DOUG_OK;
  };
  auto scan32 = [&](const int32_t* in, int32_t* out, int cnt) -> int {
    size_t bytes = 0;
    HIPCHK(exclusive_scan(nullptr, bytes, in, out, cnt, st));
    if (int r = tmp_bytes(bytes)) return r;
    HIPCHK(exclusive_scan(scan_tmp, bytes, in, out, cnt, st));
    return DOUG_OK;
  };
  auto scan64 = [&](const int64_t* in, int64_t* out, int cnt) -> int {
    size_t bytes = 0;
    HIPCHK(exclusive_scan(nullptr, bytes, in, out, cnt, st));
    if (int r = tmp_bytes(bytes)) return r;
    HIPCHK(exclusive_scan(scan_tmp, bytes, in, out, cnt, st));
    return DOUG_OK;
  };
  auto read_back = [&](const auto* dptr, auto& value) -> int {
    HIPCHK(hipMemcpyAsync(&value, dptr, sizeof(value), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    return DOUG_OK;
  };
  uint8_t* d_strong;
  int32_t* d_num;
  if (ctx->d_strong) {
    d_strong = ctx->d_strong;
  } else {
    HIPCHK(tmp.alloc(&d_strong, (size_t)nnz));
    HIPCHK(hipMemcpyAsync(d_strong, strong.data(), (size_t)nnz, hipMemcpyHostToDevice, st));
  }
  HIPCHK(tmp.alloc(&d_num, (size_t)n));
  HIPCHK(hipMemcpyAsync(d_num, ctx->fin.num.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  // R^T
  int32_t *s_col, *cnt;
  double* s_val;
  HIPCHK(tmp.alloc(&s_col, (size_t)nnz));
  HIPCHK(tmp.alloc(&s_val, (size_t)nnz));
  HIPCHK(tmp.alloc(&cnt, (size_t)n + 1));
  HIPCHK(hipMemsetAsync(cnt, 0, ((size_t)n + 1) * sizeof(int32_t), st));
  // several processes: the process layout, each row in its owner's order
  int32_t *d_owner = nullptr, *d_proc = nullptr, *d_order = nullptr, *a_col = nullptr;
  uint8_t* d_sent = nullptr;
  double *a_val = nullptr, *d_dfilt = nullptr, *d_dghost = nullptr;
  if (ranks) {
    // the ranks' workers listed their own nodes in local order, rank after
    // rank (the process-major order); the position of each node in it and
    // whether its owner sends it are formed on the device
    const RankSetup& rs = ctx->rsetup;
    const int nr = (int)rs.ranks.size();
    if ((int)rs.order.size() != n || (int)rs.own_off.size() != nr + 1 || rs.own_off[nr] != n)
      return fail(ctx, DOUG_E_STATE, "the ranks' node order is missing");
    ctx->mp_nsent.resize((size_t)nr);
    for (int r = 0; r < nr; r++) ctx->mp_nsent[r] = rs.ranks[r].nsent;
    int32_t *d_off, *d_nsent;
    HIPCHK(tmp.alloc(&d_proc, (size_t)n));
    HIPCHK(tmp.alloc(&d_order, (size_t)n));
    HIPCHK(tmp.alloc(&d_sent, (size_t)n));
    HIPCHK(tmp.alloc(&d_off, (size_t)nr + 1));
    HIPCHK(tmp.alloc(&d_nsent, (size_t)nr));
    HIPCHK(tmp.alloc(&a_col, (size_t)nnz));
    HIPCHK(tmp.alloc(&a_val, (size_t)nnz));
    HIPCHK(tmp.alloc(&d_dfilt, (size_t)n));
    HIPCHK(tmp.alloc(&d_dghost, (size_t)n));
    if (ctx->d_eptn) {
      d_owner = ctx->d_eptn;  // uploaded with the matrix
    } else {
      HIPCHK(tmp.alloc(&d_owner, (size_t)n));
      HIPCHK(hipMemcpyAsync(d_owner, ctx->eptn.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    }
    HIPCHK(hipMemcpyAsync(d_order, rs.order.data(), (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_off, rs.own_off.data(), ((size_t)nr + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(d_nsent, ctx->mp_nsent.data(), (size_t)nr * sizeof(int32_t), hipMemcpyHostToDevice, st));
    launch_mp_layout(n, d_order, d_owner, d_off, d_nsent, d_proc, d_sent, st);
    launch_mp_own_rows(n, ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_valin, d_strong, d_owner, d_sent,
                       a_col, a_val, d_dfilt, d_dghost, st);
    launch_mp_rt_rows(n, d_order, ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_valin, d_strong, d_num,
                      d_owner, d_proc, d_dfilt, d_dghost, ctx->p.omega, s_col, s_val, cnt, st);
  } else {
    launch_rt_rows(n, ctx->d_ptr, ctx->d_col, ctx->d_val, d_strong, d_num, ctx->p.omega, s_col, s_val,
                   cnt, st);
  }
  if ((rc = dalloc(ctx, &ctx->d_Rtptr, (size_t)n + 1)) || (rc = scan32(cnt, ctx->d_Rtptr, n + 1)))
    return rc;
  int32_t nnz_rt = 0;
  if ((rc = read_back(ctx->d_Rtptr + n, nnz_rt))) return rc;
  if ((rc = dalloc(ctx, &ctx->d_Rtcol, (size_t)nnz_rt)) || (rc = dalloc(ctx, &ctx->d_Rtval, (size_t)nnz_rt)))
    return rc;
  launch_compact(n, ctx->d_ptr, ctx->d_Rtptr, s_col, s_val, ctx->d_Rtcol, ctx->d_Rtval, st);
  // R = transpose(R^T), rows listing nodes in ascending order
  int32_t *rowid, *iota, *keys, *order, *rcount;
  HIPCHK(tmp.alloc(&rowid, (size_t)nnz_rt));
  HIPCHK(tmp.alloc(&iota, (size_t)nnz_rt));
  HIPCHK(tmp.alloc(&keys, (size_t)nnz_rt));
  HIPCHK(tmp.alloc(&order, (size_t)nnz_rt));
  HIPCHK(tmp.alloc(&rcount, (size_t)nagr + 1));
  launch_row_ids(n, ctx->d_Rtptr, rowid, st);
  launch_iota(nnz_rt, iota, st);
  int end_bit = 1;
  while (end_bit < 31 && (1 << end_bit) <= nagr) end_bit++;
  {
    size_t bytes = 0;
    HIPCHK(stable_sort_pairs(nullptr, bytes, ctx->d_Rtcol, keys, iota, order, nnz_rt, end_bit, st));
    if ((rc = tmp_bytes(bytes))) return rc;
    HIPCHK(stable_sort_pairs(scan_tmp, bytes, ctx->d_Rtcol, keys, iota, order, nnz_rt, end_bit, st));
  }
  HIPCHK(hipMemsetAsync(rcount, 0, ((size_t)nagr + 1) * sizeof(int32_t), st));
  launch_key_histogram(nnz_rt, ctx->d_Rtcol, rcount, st);
  if ((rc = dalloc(ctx, &ctx->d_Rptr, (size_t)nagr + 1)) || (rc = scan32(rcount, ctx->d_Rptr, nagr + 1)))
    return rc;
  if ((rc = dalloc(ctx, &ctx->d_Rcol, (size_t)nnz_rt)) || (rc = dalloc(ctx, &ctx->d_Rval, (size_t)nnz_rt)))
    return rc;
  launch_gather_entries(nnz_rt, order, rowid, ctx->d_Rtval, ctx->d_Rcol, ctx->d_Rval, st);
  // X*Y with first-touch rows: bound, slots, products, compaction
  auto product = [&](int nrows, const int32_t* xp, const int32_t* xc, const double* xv,
                     const int32_t* yp, const int32_t* yc, const double* yv, int32_t** op,
                     int32_t** oc, double** ov, int32_t* onnz, bool keep, bool wave) -> int {
    int64_t *ub, *off;
    int32_t *pc, *ccount;
    double* pv;
    HIPCHK(tmp.alloc(&ub, (size_t)nrows + 1));
    HIPCHK(tmp.alloc(&off, (size_t)nrows + 1));
    HIPCHK(hipMemsetAsync(ub, 0, ((size_t)nrows + 1) * sizeof(int64_t), st));
    launch_product_bound(nrows, xp, xc, yp, ub, st);
    if (int r = scan64(ub, off, nrows + 1)) return r;
    int64_t total = 0;
    if (int r = read_back(off + nrows, total)) return r;
    HIPCHK(tmp.alloc(&pc, (size_t)total));
    HIPCHK(tmp.alloc(&pv, (size_t)total));
    HIPCHK(tmp.alloc(&ccount, (size_t)nrows + 1));
    HIPCHK(hipMemsetAsync(ccount, 0, ((size_t)nrows + 1) * sizeof(int32_t), st));
    bool done = false;
    if (wave) {  // long rows: a wavefront per row, unless some row has > 128 columns
      int* ovf;
      HIPCHK(tmp.alloc(&ovf, 1));
      HIPCHK(hipMemsetAsync(ovf, 0, sizeof(int), st));
      launch_first_touch_wave(nrows, xp, xc, xv, yp, yc, yv, off, pc, pv, ccount, ovf, st);
      int o = 0;
      if (int r = read_back(ovf, o)) return r;
      done = o == 0;
      if (!done) HIPCHK(hipMemsetAsync(ccount, 0, ((size_t)nrows + 1) * sizeof(int32_t), st));
    }
    if (!done) launch_first_touch(nrows, xp, xc, xv, yp, yc, yv, off, pc, pv, ccount, st);
    if (keep) {
      if (int r = dalloc(ctx, op, (size_t)nrows + 1)) return r;
    } else {
      HIPCHK(tmp.alloc(op, (size_t)nrows + 1));
    }
    if (int r = scan32(ccount, *op, nrows + 1)) return r;
    if (int r = read_back(*op + nrows, *onnz)) return r;
    if (*onnz < 0) return fail(ctx, DOUG_E_UNSUPPORTED, "coarse product above 2^31-1 entries");
    HIPCHK(tmp.alloc(oc, (size_t)*onnz));
    HIPCHK(tmp.alloc(ov, (size_t)*onnz));
    launch_compact(nrows, off, *op, pc, pv, *oc, *ov, st);
    HIPCHK(hipGetLastError());
    return DOUG_OK;
  };
  int32_t *tp_ptr, *tp_col, *ac_ptr, *ac_col;
  double *tp_val, *ac_val;
  int32_t nnz_tp = 0, nnz_ac = 0;
  if (ranks) {
    // R with each row in process order: R^T's entries listed row by row in
    // process order, stably sorted on their aggregate
    int32_t *len, *poff, *ids, *pkeys, *pkeys_out, *porder, *rp_col;
    double* rp_val;
    HIPCHK(tmp.alloc(&len, (size_t)n + 1));
    HIPCHK(tmp.alloc(&poff, (size_t)n + 1));
    HIPCHK(tmp.alloc(&ids, (size_t)nnz_rt));
    HIPCHK(tmp.alloc(&pkeys, (size_t)nnz_rt));
    HIPCHK(tmp.alloc(&pkeys_out, (size_t)nnz_rt));
    HIPCHK(tmp.alloc(&porder, (size_t)nnz_rt));
    HIPCHK(tmp.alloc(&rp_col, (size_t)nnz_rt));
    HIPCHK(tmp.alloc(&rp_val, (size_t)nnz_rt));
    HIPCHK(hipMemsetAsync(len, 0, ((size_t)n + 1) * sizeof(int32_t), st));
    launch_mp_row_lengths(n, d_order, ctx->d_Rtptr, len, st);
    if ((rc = scan32(len, poff, n + 1))) return rc;
    launch_mp_proc_entries(n, d_order, ctx->d_Rtptr, ctx->d_Rtcol, poff, ids, pkeys, st);
    {
      size_t bytes = 0;
      HIPCHK(stable_sort_pairs(nullptr, bytes, pkeys, pkeys_out, ids, porder, nnz_rt, end_bit, st));
      if ((rc = tmp_bytes(bytes))) return rc;
      HIPCHK(stable_sort_pairs(scan_tmp, bytes, pkeys, pkeys_out, ids, porder, nnz_rt, end_bit, st));
    }
    launch_gather_entries(nnz_rt, porder, rowid, ctx->d_Rtval, rp_col, rp_val, st);
    // T = A R^T with each row in its owner's entry order
    if ((rc = product(n, ctx->d_ptr, a_col, a_val, ctx->d_Rtptr, ctx->d_Rtcol, ctx->d_Rtval, &tp_ptr,
                      &tp_col, &tp_val, &nnz_tp, false, false)))
      return rc;
    // A_c: each process's part, then their sum in process order
    int64_t *ub, *off;
    int32_t *pc, *ccount;
    double* pv;
    HIPCHK(tmp.alloc(&ub, (size_t)nagr + 1));
    HIPCHK(tmp.alloc(&off, (size_t)nagr + 1));
    HIPCHK(hipMemsetAsync(ub, 0, ((size_t)nagr + 1) * sizeof(int64_t), st));
    launch_product_bound(nagr, ctx->d_Rptr, rp_col, tp_ptr, ub, st);
    if ((rc = scan64(ub, off, nagr + 1))) return rc;
    int64_t total = 0;
    if ((rc = read_back(off + nagr, total))) return rc;
    HIPCHK(tmp.alloc(&pc, (size_t)total));
    HIPCHK(tmp.alloc(&pv, (size_t)total));
    HIPCHK(tmp.alloc(&ccount, (size_t)nagr + 1));
    HIPCHK(hipMemsetAsync(ccount, 0, ((size_t)nagr + 1) * sizeof(int32_t), st));
    int* ovf;
    HIPCHK(tmp.alloc(&ovf, 1));
    HIPCHK(hipMemsetAsync(ovf, 0, sizeof(int), st));
    launch_mp_galerkin_wave(nagr, ctx->d_Rptr, rp_col, rp_val, d_owner, tp_ptr, tp_col, tp_val, off,
                            pc, pv, ccount, ovf, st);
    int o = 0;
    if ((rc = read_back(ovf, o))) return rc;
    if (o) return fail(ctx, DOUG_E_UNSUPPORTED, "a coarse matrix row above 128 entries");
    HIPCHK(tmp.alloc(&ac_ptr, (size_t)nagr + 1));
    if ((rc = scan32(ccount, ac_ptr, nagr + 1))) return rc;
    if ((rc = read_back(ac_ptr + nagr, nnz_ac))) return rc;
    HIPCHK(tmp.alloc(&ac_col, (size_t)nnz_ac));
    HIPCHK(tmp.alloc(&ac_val, (size_t)nnz_ac));
    launch_compact(nagr, off, ac_ptr, pc, pv, ac_col, ac_val, st);
    HIPCHK(hipGetLastError());
  } else {
    if ((rc = product(n, ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_Rtptr, ctx->d_Rtcol, ctx->d_Rtval,
                      &tp_ptr, &tp_col, &tp_val, &nnz_tp, false, false)))
      return rc;
    if ((rc = product(nagr, ctx->d_Rptr, ctx->d_Rcol, ctx->d_Rval, tp_ptr, tp_col, tp_val, &ac_ptr,
                      &ac_col, &ac_val, &nnz_ac, false, true)))
      return rc;
  }
  // host copies: R's row pointers (row blocks; its entries are fetched by
  // doug_get_restriction) and A_c (coarse factor)
  Csr& R = ctx->R;
  R.nrows = nagr;
  R.ncols = n;
  R.ptr.resize((size_t)nagr + 1);
  R.col.clear();
  R.val.clear();
  ctx->nnz_R = nnz_rt;
  Csr& AC = ctx->AC;
  AC.nrows = AC.ncols = nagr;
  AC.ptr.resize((size_t)nagr + 1);
  AC.col.resize((size_t)nnz_ac);
  AC.val.resize((size_t)nnz_ac);
  HIPCHK(hipMemcpyAsync(R.ptr.data(), ctx->d_Rptr, R.ptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(AC.ptr.data(), ac_ptr, AC.ptr.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(AC.col.data(), ac_col, AC.col.size() * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  HIPCHK(hipMemcpyAsync(AC.val.data(), ac_val, AC.val.size() * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  HIPCHK(hipGetLastError());
  ctx->Rt = Csr{};
  ctx->Rt.nrows = n;
  ctx->Rt.ncols = nagr;
  ctx->coarse_on_device = true;
  ctx->info.t_coarse_build = now() - t1;
  return DOUG_OK;
}

extern "C" {

// (3) subdomain index sets: injected partition, or coarse aggregates
//     (Partitioning_aggr.F90:120-193), plus `ol` layers (Preconditioner_base.F90:96-123)
static void phase_subdomains(doug_ctx* ctx) {
  const Params& p = ctx->p;
  Csr& A = ctx->A;
  const int n = A.nrows;
  double t2 = now();
  std::vector<int> pstarts, pnodes;
  ctx->cin = Aggregates{};
  ctx->cfull = Aggregates{};
  if (!ctx->inj_sstarts.empty()) {  // the caller's index sets, overlap included
    ctx->sstarts = ctx->inj_sstarts;
    ctx->sinds = ctx->inj_sinds;
    ctx->nsub = (int)ctx->sstarts.size() - 1;
    ctx->info.t_subdomains = now() - t2;
    return;
  }
  if (ctx->nranks > 0) {  // the ranks' subdomains, layers included
    ctx->sstarts = std::move(ctx->rsetup.sstarts);
    ctx->sinds = std::move(ctx->rsetup.sinds);
    ctx->nsub = (int)ctx->sstarts.size() - 1;
    ctx->info.t_subdomains = now() - t2;
    return;
  }
  if (ctx->have_partition) {
    pstarts = ctx->part_starts;
    pnodes = ctx->part_nodes;
  } else if (ctx->num_subdomains > 1) {
    // Partitionings_metis_InitCoarse (Partitioning_metis.F90:34-86) with the
    // project partitioner: parts of the fine-aggregate graph, each listing
    // its nodes in node order
    std::vector<int> xadj, adjncy, part;
    aggregate_graph(A, ctx->fin, ctx->ffull, xadj, adjncy);
    partition_graph(ctx->ffull.nagr, xadj.data(), adjncy.data(), nullptr, ctx->num_subdomains, part);
    const int k = ctx->num_subdomains;
    pstarts.assign((size_t)k + 1, 0);
    for (int i = 0; i < n; i++) pstarts[part[ctx->ffull.num[i] - 1] + 1]++;
    for (int q = 0; q < k; q++) pstarts[q + 1] += pstarts[q];
    pnodes.assign((size_t)n, 0);
    std::vector<int> fill(pstarts.begin(), pstarts.end() - 1);
    for (int i = 0; i < n; i++) pnodes[fill[part[ctx->ffull.num[i] - 1]]++] = i;
    // parts the partitioner left empty are dropped
    std::vector<int> ps{0};
    for (int q = 0; q < k; q++)
      if (pstarts[q + 1] > pstarts[q]) ps.push_back(pstarts[q + 1]);
    pstarts.swap(ps);
  } else {
    const double strong_conn1 = p.strong1 != 0.0 ? p.strong1 : 0.67;
    const int radius1 = p.radius1 > 0 ? p.radius1 : 2;
    const double strong_conn2 = p.strong2 > 0 ? p.strong2 : strong_conn1 / 2.0;
    Csr& AC = ctx->AC;
    std::vector<double> cscaled, cdiag;
    scale_values(AC, false, cscaled, cdiag);  // SpMtx_AB results are not flagged symmetric
    std::vector<uint8_t> cstrong = find_strong(AC, cscaled, cdiag, strong_conn2);
    int r2;
    if (p.radius2 > 0) {
      r2 = p.radius2;
    } else {
      int nn = (int)std::sqrt(1.0 * (double)n);
      r2 = (int)std::lround(3 * std::sqrt((double)nn) / (2 * radius1 + 1) - 1);
    }
    int min2 = p.minasize2 > 0 ? p.minasize2
                               : (p.radius2 > 0 ? 2 * p.radius2 + 1 : (int)(0.5 * (double)((2 * r2 + 1) * (2 * r2 + 1))));
    int max2 = p.maxasize2 > 0 ? p.maxasize2 : (2 * r2 + 1) * (2 * r2 + 1);
    aggregate(AC, cscaled, cstrong, r2, min2, max2, strong_conn2, true, ctx->ps, ctx->cin, ctx->cfull);
    AC.val = cscaled;
    unscale_values(AC, false, cdiag, AC.val);
    coarse_partition(ctx->cfull, ctx->ffull, n, pstarts, pnodes);
  }
  add_layers_all(A, pstarts, pnodes, ctx->ol, ctx->sstarts, ctx->sinds);
  ctx->nsub = (int)ctx->sstarts.size() - 1;
  ctx->info.t_subdomains = now() - t2;
}

static int check_params(doug_ctx* ctx) {
  if (!ctx->have_matrix) return fail(ctx, DOUG_E_STATE, "setup before doug_set_matrix_coo");
  const Params& p = ctx->p;
  if (p.fine_method < 0 || p.fine_method > 2)
    return fail(ctx, DOUG_E_UNSUPPORTED, "fine_method must be 0 (none), 1 (exact solves) or 2 (SGS)");
  // method 5 ("multiplicative on the fine level, additive with the coarse")
  // is method 1 with one process: the coarse correction is restricted from
  // the same residual and added to the same sum (pcg.F90:145-150,205-208).
  // With one process methods 2-4 restrict a residual vector that the
  // preconditioner never computes (pcg.F90:146-147,253-255), so they have no
  // defined result; with several processes the restriction takes r
  // (pcg.F90:130-141) and they are defined (enqueue_prec).
  if (p.method < 0 || p.method > 5 || (ctx->nranks <= 0 && p.method >= 2 && p.method <= 4))
    return fail(ctx, DOUG_E_UNSUPPORTED,
                "method must be 0 (none), 1 or 5 (additive Schwarz); 2-4 need virtual ranks");
  // a partition or rank map given for an earlier matrix
  if (ctx->have_partition && (int64_t)ctx->part_nodes.size() != ctx->n)
    return fail(ctx, DOUG_E_STATE, "the partition was set for another matrix: set it again");
  if (ctx->nranks > 0 && (int64_t)ctx->eptn.size() != ctx->n)
    return fail(ctx, DOUG_E_STATE, "the rank map was set for another matrix: call doug_set_ranks again");
  if (ctx->inj_nagr > 0 && (int64_t)ctx->inj_num.size() != ctx->n)
    return fail(ctx, DOUG_E_STATE, "the aggregates were set for another matrix: set them again");
  if (!ctx->inj_sinds.empty()) {
    for (int v : ctx->inj_sinds)
      if (v >= ctx->n)
        return fail(ctx, DOUG_E_STATE, "the subdomains were set for another matrix: set them again");
  }
  if (ctx->nranks > 0 && (ctx->inj_nagr > 0 || !ctx->inj_sstarts.empty()))
    return fail(ctx, DOUG_E_ARG, "virtual ranks make their own aggregates and subdomains");
  if (ctx->nranks > 0 && (ctx->have_inj_R || !ctx->inj_strong.empty()))
    return fail(ctx, DOUG_E_ARG, "virtual ranks make their own strength flags and restriction");
  if (!ctx->inj_strong.empty() && ctx->inj_strong.size() != ctx->A.src.size())
    return fail(ctx, DOUG_E_STATE, "the strength mask was set for another matrix: set it again");
  if (ctx->have_inj_R && ctx->inj_R.ncols != ctx->n)
    return fail(ctx, DOUG_E_STATE, "the restriction was set for another matrix: set it again");
  if (ctx->nranks > 0) {
    // several processes: a negative radius1 selects the graph partition of
    // the matrix and is negated (Distribution_assm.F90:127-134); the
    // distribution ships max(overlap, smoothers) layers (:138), before the
    // smoothers are reduced to radius1
    if (ctx->have_partition)
      return fail(ctx, DOUG_E_ARG, "virtual ranks choose their own subdomains: clear the partition");
    if (ctx->p.radius1 < 0) ctx->p.radius1 = -ctx->p.radius1;
    ctx->ol_dist = std::max(p.overlap, p.smoothers);
    if (ctx->ol_dist < 1)
      return fail(ctx, DOUG_E_UNSUPPORTED, "virtual ranks need max(overlap, smoothers) >= 1");
  }
  int smoothers = p.smoothers;
  if (smoothers > p.radius1) smoothers = p.radius1;  // CoarseMtx.F90:75-78
  if (smoothers < 0)  // -1 selects an exact smoother the reference has commented out
    return fail(ctx, DOUG_E_UNSUPPORTED, "smoothers must resolve to 0 or more (got " + std::to_string(smoothers) + ")");
  ctx->smoothers = smoothers;
  ctx->ol = p.overlap < 0 ? std::max(smoothers, 0) : p.overlap;  // aggr.F90:207-212
  ctx->has_coarse = p.levels > 1;
  ctx->ps = ProcessState{};
  ctx->ready = false;
  ctx->host_ready = false;
  ctx->info = doug_info{};
  return DOUG_OK;
}

static void fill_info(doug_ctx* ctx) {
  doug_info& info = ctx->info;
  info.n = ctx->A.nrows;
  info.nnz = ctx->A.nnz();
  info.nagr_fine = ctx->fin.nagr;
  info.nagr_coarse = ctx->cfull.nagr;
  info.nsub = ctx->nsub;
  info.ncoarse = ctx->R.nrows;
  info.overlap = ctx->ol;
  info.smoothers = ctx->smoothers;
  info.sub_total = (int64_t)ctx->sinds.size();
}

static int setup_host(doug_ctx* ctx) {
  int rc = check_params(ctx);
  if (rc) return rc;
  double t0 = now();
  std::vector<double> scaled;
  std::vector<uint8_t> strong;
  phase_values(ctx, scaled, strong);
  phase_aggregate(ctx, scaled, strong);
  if (rank_coarse(ctx))
    phase_coarse_ranks(ctx, strong);
  else if (ctx->have_inj_R)
    rc = phase_coarse_given(ctx);
  else
    phase_coarse_host(ctx, strong);
  if (rc) return rc;
  phase_subdomains(ctx);
  fill_info(ctx);
  ctx->info.t_total = now() - t0;
  ctx->host_ready = true;
  return DOUG_OK;
}

// Dispatch order of the subdomain solves: 0 as partitioned, 1 largest first
// (longest-processing-time order, which shortens the tail when sizes vary).
static int set_trsv_order(doug_ctx* ctx, int order) {
  ctx->trsv_order = order;
  ctx->d_sub_trsv = ctx->d_sub;
  if (order == 0 || ctx->h_sub.empty()) return DOUG_OK;
  std::vector<BlockDesc> sorted(ctx->h_sub);
  std::stable_sort(sorted.begin(), sorted.end(),
                   [](const BlockDesc& a, const BlockDesc& b) { return a.nt > b.nt; });
  int rc = upload(ctx, &ctx->d_sub_trsv, sorted);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(ctx->stream));  // `sorted` goes out of scope
  return DOUG_OK;
}

// Matrix, vectors, subdomain data and the subdomain factorisation (enqueued).
// The matrix (values after the scale/unscale round trip) and its SpMV row
// blocks on the device.
// Without `values` d_val is only allocated (the device forms the values).
static int upload_matrix(doug_ctx* ctx, bool values = true) {
  const Csr& A = ctx->A;
  double t = now();
  int rc;
  ctx->d_src = nullptr;  // uploaded by the first doug_value_grad_device
  if ((rc = upload(ctx, &ctx->d_ptr, A.ptr)) || (rc = upload(ctx, &ctx->d_col, A.col)) ||
      (rc = values ? upload(ctx, &ctx->d_val, A.val) : dalloc(ctx, &ctx->d_val, A.col.size())))
    return rc;
  // several processes with overlap 1: A_ghost's rows smooth with the
  // values as read (the device coarse build)
  if (rank_coarse(ctx) && ctx->ol_dist == 1 && ctx->smoothers == 1 && ctx->use_device_coarse &&
      ((rc = upload(ctx, &ctx->d_valin, ctx->A_input_val)) || (rc = upload(ctx, &ctx->d_eptn, ctx->eptn))))
    return rc;
  std::vector<int32_t> rb = row_blocks(A);
  ctx->nrb = (int)rb.size() - 1;
  if ((rc = upload(ctx, &ctx->d_rowblk, rb))) return rc;
  ctx->d_rowblk_a = ctx->d_rowblk;
  ctx->nrb_a = ctx->nrb;
  std::vector<int32_t> rba;
  if (spmv_shape_rows(ctx->spmv_shape) != kSpmvThreads) {
    rba = row_blocks(A, spmv_shape_rows(ctx->spmv_shape));
    ctx->nrb_a = (int)rba.size() - 1;
    if ((rc = upload(ctx, &ctx->d_rowblk_a, rba))) return rc;
  }
  HIPCHK(hipStreamSynchronize(ctx->stream));  // staging vectors go out of scope
  ctx->info.t_upload = now() - t;
  return DOUG_OK;
}

// FinePreconditioner_sgs_Init (FinePreconditioner_sgs.F90:32-64): each
// subdomain keeps the entries of the consolidated matrix (SpMtx_addAll,
// SpMtx_operation.F90:728-759: rows and columns sorted, duplicates summed in
// stored order) with both ends in its index set; the sweeps visit its rows
// ascending and each row's entries by ascending column.  For k_sgs the rows
// are levelled: in the forward sweep row i comes after every row j < i that
// shares an entry with it (i reads the new x_j, or j reads the old x_i) and
// gets level 1 + the largest such level; the backward sweep mirrors this.
static int build_sgs(doug_ctx* ctx, const std::vector<int>& dinds) {
  const Csr& A = ctx->A;
  const int n = A.nrows;
  const int nb = (int)ctx->sstarts.size() - 1;
  struct Piece {
    std::vector<int32_t> ptr{0}, col, levf, levb, rowsf, rowsb;
    std::vector<double> val, diag;
    int nlf = 0, nlb = 0;
  };
  std::vector<Piece> pieces((size_t)nb);
  int mmax = 0;
  for (int b = 0; b < nb; b++) mmax = std::max(mmax, ctx->sstarts[b + 1] - ctx->sstarts[b]);
  if (mmax > kSgsMaxRows)
    return fail(ctx, DOUG_E_UNSUPPORTED, "subdomain of " + std::to_string(mmax) +
                                             " unknowns exceeds the Gauss-Seidel limit " +
                                             std::to_string(kSgsMaxRows));
  const int nth = std::max(1, std::min(setup_threads(), nb));
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back([&, t]() {
      std::vector<int> floc((size_t)n, -1), lev, cnt;
      std::vector<std::pair<int, double>> ent;
      std::vector<int32_t> sptr, sadj;
      for (int b = t; b < nb; b += nth) {
        Piece& pc = pieces[b];
        const int s0 = ctx->sstarts[b], m = ctx->sstarts[b + 1] - s0;
        for (int r = 0; r < m; r++) floc[dinds[s0 + r]] = r;
        pc.diag.assign((size_t)m, 0.0);
        for (int r = 0; r < m; r++) {
          const int g = dinds[s0 + r];
          ent.clear();
          for (int q = A.ptr[g]; q < A.ptr[g + 1]; q++)
            if (floc[A.col[q]] >= 0) ent.push_back({A.col[q], A.val[q]});
          std::stable_sort(ent.begin(), ent.end(),
                           [](const auto& x, const auto& y) { return x.first < y.first; });
          size_t w = 0;  // consolidate: duplicates summed in stored order
          for (size_t q = 0; q < ent.size(); q++) {
            if (w > 0 && ent[w - 1].first == ent[q].first)
              ent[w - 1].second = ent[w - 1].second + ent[q].second;
            else
              ent[w++] = ent[q];
          }
          ent.resize(w);
          for (auto& e : ent) {
            if (e.first == g) {
              pc.diag[r] = e.second;
            } else {
              pc.col.push_back(floc[e.first]);
              pc.val.push_back(e.second);
            }
          }
          pc.ptr.push_back((int32_t)pc.col.size());
        }
        for (int r = 0; r < m; r++) floc[dinds[s0 + r]] = -1;
        // symmetrised coupling graph
        sptr.assign((size_t)m + 1, 0);
        for (int r = 0; r < m; r++)
          for (int q = pc.ptr[r]; q < pc.ptr[r + 1]; q++) {
            sptr[r + 1]++;
            sptr[pc.col[q] + 1]++;
          }
        for (int r = 0; r < m; r++) sptr[r + 1] += sptr[r];
        sadj.resize((size_t)sptr[m]);
        cnt.assign(sptr.begin(), sptr.end() - 1);
        for (int r = 0; r < m; r++)
          for (int q = pc.ptr[r]; q < pc.ptr[r + 1]; q++) {
            sadj[cnt[r]++] = pc.col[q];
            sadj[cnt[pc.col[q]]++] = r;
          }
        // levels, then rows grouped by level (ascending rows inside a level)
        auto levels = [&](bool fwd, std::vector<int32_t>& rows, int& nl) {
          lev.assign((size_t)m, 0);
          nl = 0;
          for (int k = 0; k < m; k++) {
            const int r = fwd ? k : m - 1 - k;
            int l = 0;
            for (int q = sptr[r]; q < sptr[r + 1]; q++) {
              const int j = sadj[q];
              if (fwd ? j < r : j > r) l = std::max(l, lev[j] + 1);
            }
            lev[r] = l;
            nl = std::max(nl, l + 1);
          }
          std::vector<int32_t> start((size_t)nl + 1, 0);
          for (int r = 0; r < m; r++) start[lev[r] + 1]++;
          for (int l = 0; l < nl; l++) start[l + 1] += start[l];
          rows.resize((size_t)m);
          std::vector<int32_t> fill(start.begin(), start.end() - 1);
          for (int r = 0; r < m; r++) rows[fill[lev[r]]++] = r;
          return start;
        };
        pc.levf = levels(true, pc.rowsf, pc.nlf);
        pc.levb = levels(false, pc.rowsb, pc.nlb);
      }
    });
  for (auto& x : th) x.join();
  std::vector<SgsDesc> desc((size_t)nb);
  std::vector<int32_t> levptr, lrows, rbeg, rlen, col;
  std::vector<double> val, diag;
  rbeg.reserve(dinds.size());
  rlen.reserve(dinds.size());
  diag.reserve(dinds.size());
  lrows.reserve(2 * dinds.size());
  for (int b = 0; b < nb; b++) {
    const Piece& pc = pieces[b];
    const int m = (int)pc.diag.size();
    desc[b].off = ctx->sstarts[b];
    desc[b].e0 = (int64_t)col.size();
    desc[b].m = m;
    desc[b].nlf = pc.nlf;
    desc[b].pad = 0;
    desc[b].lf0 = (int32_t)levptr.size();
    levptr.insert(levptr.end(), pc.levf.begin(), pc.levf.end());
    desc[b].nlb = pc.nlb;
    desc[b].lb0 = (int32_t)levptr.size();
    levptr.insert(levptr.end(), pc.levb.begin(), pc.levb.end());
    lrows.insert(lrows.end(), pc.rowsf.begin(), pc.rowsf.end());
    lrows.insert(lrows.end(), pc.rowsb.begin(), pc.rowsb.end());
    for (int r = 0; r < m; r++) {
      rbeg.push_back(pc.ptr[r]);
      rlen.push_back(pc.ptr[r + 1] - pc.ptr[r]);
    }
    col.insert(col.end(), pc.col.begin(), pc.col.end());
    val.insert(val.end(), pc.val.begin(), pc.val.end());
    diag.insert(diag.end(), pc.diag.begin(), pc.diag.end());
  }
  int rc;
  if ((rc = upload(ctx, &ctx->d_sgs, desc)) || (rc = upload(ctx, &ctx->d_sgs_levptr, levptr)) ||
      (rc = upload(ctx, &ctx->d_sgs_lrows, lrows)) || (rc = upload(ctx, &ctx->d_sgs_rbeg, rbeg)) ||
      (rc = upload(ctx, &ctx->d_sgs_rlen, rlen)) || (rc = upload(ctx, &ctx->d_sgs_col, col)) ||
      (rc = upload(ctx, &ctx->d_sgs_val, val)) || (rc = upload(ctx, &ctx->d_sgs_diag, diag)))
    return rc;
  ctx->sgs_mmax = mmax;
  ctx->sgs_iters = ctx->p.num_iters > 0 ? ctx->p.num_iters : 3;  // aggr.F90:220-222
  prepare_sgs(mmax);
  return DOUG_OK;
}

// the strength mask for the device build of R (read by the aggregation too,
// never written once set up)
static int upload_strong(doug_ctx* ctx, const std::vector<uint8_t>& strong) {
  ctx->d_strong = nullptr;
  if (!(ctx->smoothers == 1 && ctx->use_device_coarse && ctx->has_coarse)) return DOUG_OK;
  return upload(ctx, &ctx->d_strong, strong);
}

// phase_values on the device from the uploaded values as read (d_valin):
// d_val and d_strong; the host copies follow on demand (host_values_now)
static int device_values(doug_ctx* ctx) {
  const Params& p = ctx->p;
  const int n = ctx->A.nrows;
  double *d_diag, *d_sv;
  int rc;
  if ((rc = dalloc(ctx, &d_diag, (size_t)n)) || (rc = dalloc(ctx, &d_sv, (size_t)n)) ||
      (rc = dalloc(ctx, &ctx->d_strong, ctx->A.col.size())))
    return rc;
  const double strong_conn1 = p.strong1 != 0.0 ? p.strong1 : 0.67;
  launch_values_strong(n, ctx->d_ptr, ctx->d_col, ctx->d_valin, (p.symmstruct && p.symmnumeric) ? 1 : 0,
                       strong_conn1, d_diag, d_sv, ctx->d_val, ctx->d_strong, ctx->stream);
  HIPCHK(hipGetLastError());
  ctx->host_values = false;
  return DOUG_OK;
}

// the device's round-trip values and strength flags copied to the host
static int host_values_now(doug_ctx* ctx) {
  if (ctx->host_values) return DOUG_OK;
  ctx->strong.resize(ctx->A.col.size());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  HIPCHK(hipMemcpy(ctx->A.val.data(), ctx->d_val, ctx->A.val.size() * sizeof(double), hipMemcpyDeviceToHost));
  HIPCHK(hipMemcpy(ctx->strong.data(), ctx->d_strong, ctx->strong.size(), hipMemcpyDeviceToHost));
  ctx->host_values = true;
  return DOUG_OK;
}

static int device_fine(doug_ctx* ctx, hipEvent_t ev0, hipEvent_t ev1) {
  const Params& p = ctx->p;
  doug_info& info = ctx->info;
  Csr& A = ctx->A;
  const int n = A.nrows;
  double t3 = now();
  int rc;
  // device order of each subdomain's unknowns (the reported index sets keep
  // the reference's order): reverse Cuthill-McKee narrows the factor's
  // tile envelope
  std::vector<int> dinds;
  if (p.fine_method == 1 && ctx->local_order == 1) {
    if (ctx->nranks > 0 && ctx->rsetup.dinds.size() == ctx->sinds.size())
      dinds = std::move(ctx->rsetup.dinds);  // ordered by the ranks' workers
    else
      block_rcm(A, ctx->sstarts, ctx->sinds, dinds, &ctx->rank_ws);
  } else {
    dinds = ctx->sinds;
    if (p.fine_method == 2)  // Gauss-Seidel sweeps the rows in ascending order
      for (size_t b = 0; b + 1 < ctx->sstarts.size(); b++)
        std::sort(dinds.begin() + ctx->sstarts[b], dinds.begin() + ctx->sstarts[b + 1]);
  }
  // occurrence lists for the deterministic overlap sum (a node's
  // occurrences in ascending subdomain order), formed on the device: counts
  // per node, their prefix sum, and the positions stably sorted on their node
  if ((rc = upload(ctx, &ctx->d_inds, dinds))) return rc;
  const int64_t total = (int64_t)dinds.size();
  if ((rc = dalloc(ctx, &ctx->d_occ_ptr, (size_t)n + 1))) return rc;
  hipStream_t st = ctx->stream;
  if (p.fine_method >= 1) {
    int32_t *cnt, *keys_out, *iota;
    if ((rc = dalloc(ctx, &cnt, (size_t)n + 1)) || (rc = dalloc(ctx, &ctx->d_occ_pos, (size_t)total)) ||
        (rc = dalloc(ctx, &keys_out, (size_t)total)) || (rc = dalloc(ctx, &iota, (size_t)total)))
      return rc;
    HIPCHK(hipMemsetAsync(cnt, 0, ((size_t)n + 1) * sizeof(int32_t), st));
    launch_key_histogram(total, ctx->d_inds, cnt, st);
    launch_iota(total, iota, st);
    int end_bit = 1;
    while (end_bit < 31 && (1 << end_bit) <= n) end_bit++;
    size_t scan_bytes = 0, sort_bytes = 0;
    HIPCHK(exclusive_scan(nullptr, scan_bytes, cnt, ctx->d_occ_ptr, n + 1, st));
    HIPCHK(stable_sort_pairs(nullptr, sort_bytes, ctx->d_inds, keys_out, iota, ctx->d_occ_pos, (int)total,
                             end_bit, st));
    char* tmp;
    if ((rc = dalloc(ctx, &tmp, std::max(scan_bytes, sort_bytes)))) return rc;
    HIPCHK(exclusive_scan(tmp, scan_bytes, cnt, ctx->d_occ_ptr, n + 1, st));
    HIPCHK(stable_sort_pairs(tmp, sort_bytes, ctx->d_inds, keys_out, iota, ctx->d_occ_pos, (int)total,
                             end_bit, st));
  } else {  // no fine solves: z starts as r (FinePreconditioner_Apply, type NONE)
    if ((rc = dalloc(ctx, &ctx->d_occ_pos, (size_t)n))) return rc;
    launch_iota((int64_t)n + 1, ctx->d_occ_ptr, st);
    launch_iota(n, ctx->d_occ_pos, st);
  }
  if ((rc = dalloc(ctx, &ctx->d_subsol, ctx->sinds.size()))) return rc;
  for (double** v : {&ctx->d_x, &ctx->d_r, &ctx->d_z, &ctx->d_p, &ctx->d_q, &ctx->d_b, &ctx->d_tmp})
    if ((rc = dalloc(ctx, v, (size_t)n))) return rc;
  if ((rc = dalloc(ctx, &ctx->d_scal, S_N)) || (rc = dalloc(ctx, &ctx->d_partial, (size_t)vec_partials(n))) ||
      (rc = dalloc(ctx, &ctx->d_counter, 1)))
    return rc;
  HIPCHK(hipMemsetAsync(ctx->d_counter, 0, sizeof(unsigned), ctx->stream));
  HIPCHK(hipMemsetAsync(ctx->d_scal, 0, S_N * sizeof(double), ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));  // host staging vectors go out of scope
  info.t_upload += now() - t3;

  if (p.fine_method == 2 && ((rc = host_values_now(ctx)) || (rc = build_sgs(ctx, dinds)))) return rc;
  // ---- Factorise_subdomains (subsolvers.F90:78-163) ----
  int64_t elems = 0;
  size_t esize = ctx->precision == DOUG_PREC_FP32_FACTORS ? sizeof(float) : sizeof(double);
  std::vector<int> offcount;
  if (p.fine_method == 1) {
    // duplicate (i, j) pairs would race in the device scatter: the host
    // scatter map finds whether one falls inside a subdomain (and is the
    // one DOUG_OPT_DEVICE_BLOCKS 0 asks for)
    if (ctx->dup_state < 0) ctx->dup_state = has_duplicate_entries(A) ? 1 : 0;
    const bool dup = ctx->dup_state == 1 || !ctx->use_device_blocks;
    if (ctx->precision == DOUG_PREC_FP32_FACTORS) {
      if ((rc = dup ? build_blocks(ctx, A, ctx->d_val, ctx->sstarts, dinds, true, true, ctx->h_sub,
                                   &ctx->d_sub, &ctx->d_tiles32, &elems, offcount, &info.factor_flops)
                    : build_blocks_device(ctx, ctx->sstarts, ctx->h_sub, &ctx->d_sub, &ctx->d_tiles32,
                                          &elems, offcount, &info.factor_flops)))
        return rc;
      HIPCHK(hipEventRecord(ev0, ctx->stream));
      if ((rc = enqueue_factor(ctx, ctx->d_tiles32, ctx->d_sub, ctx->h_sub, offcount, ctx->d_flag)))
        return rc;
    } else {
      if ((rc = dup ? build_blocks(ctx, A, ctx->d_val, ctx->sstarts, dinds, true, true, ctx->h_sub,
                                   &ctx->d_sub, &ctx->d_tiles, &elems, offcount, &info.factor_flops)
                    : build_blocks_device(ctx, ctx->sstarts, ctx->h_sub, &ctx->d_sub, &ctx->d_tiles,
                                          &elems, offcount, &info.factor_flops)))
        return rc;
      HIPCHK(hipEventRecord(ev0, ctx->stream));
      if ((rc = enqueue_factor(ctx, ctx->d_tiles, ctx->d_sub, ctx->h_sub, offcount, ctx->d_flag)))
        return rc;
    }
  } else {
    HIPCHK(hipEventRecord(ev0, ctx->stream));
  }
  HIPCHK(hipEventRecord(ev1, ctx->stream));
  info.factor_bytes = elems * (int64_t)esize;
  info.factor_tiles = elems / kTileElems;
  if ((rc = set_trsv_order(ctx, ctx->trsv_order))) return rc;
  ctx->maxnt = 0;
  ctx->work_rows = 0;
  for (auto& d : ctx->h_sub) {
    ctx->maxnt = std::max(ctx->maxnt, d.nt);
    ctx->work_rows += (int64_t)d.nt * kTile;
  }
  if (ctx->precision == DOUG_PREC_FP32_FACTORS)
    prepare_trsv<float>(ctx->maxnt);
  else
    prepare_trsv<double>(ctx->maxnt);
  return DOUG_OK;
}

// Restriction/prolongation, coarse factor and its explicit inverse (enqueued).
static int device_coarse(doug_ctx* ctx, hipEvent_t ev2, hipEvent_t ev3) {
  int rc;
  ctx->nc = ctx->R.nrows;
  HIPCHK(hipEventRecord(ev2, ctx->stream));
  if (ctx->has_coarse) {
    const Csr& R = ctx->R;
    if (!ctx->coarse_on_device) {
      const Csr& Rt = ctx->Rt;
      if ((rc = upload(ctx, &ctx->d_Rptr, R.ptr)) || (rc = upload(ctx, &ctx->d_Rcol, R.col)) ||
          (rc = upload(ctx, &ctx->d_Rval, R.val)) || (rc = upload(ctx, &ctx->d_Rtptr, Rt.ptr)) ||
          (rc = upload(ctx, &ctx->d_Rtcol, Rt.col)) || (rc = upload(ctx, &ctx->d_Rtval, Rt.val)))
        return rc;
    }
    std::vector<int32_t> rbR = row_blocks(R);
    ctx->nrbR = (int)rbR.size() - 1;
    if ((rc = upload(ctx, &ctx->d_Rrowblk, rbR))) return rc;
    double* d_acval;
    if ((rc = upload(ctx, &d_acval, ctx->AC.val))) return rc;
    // the factor holds A_c in reverse Cuthill-McKee order: its Cholesky
    // factor then stays inside a narrow envelope, and the factorisation and
    // the inverse skip the tiles outside it
    const int nc = ctx->nc;
    std::vector<int> cst{0, nc}, ident((size_t)nc), cinds;
    std::iota(ident.begin(), ident.end(), 0);
    block_rcm(ctx->AC, cst, ident, cinds);
    std::vector<BlockDesc> cdesc;
    std::vector<int> coffcount;
    int64_t celems = 0;
    // dense: the explicit inverse shares the factor's tile layout
    if ((rc = build_blocks(ctx, ctx->AC, d_acval, cst, cinds, false, false, cdesc, &ctx->d_cdesc,
                           &ctx->d_ctiles, &celems, coffcount)))
      return rc;
    // right-looking: serial depth of one tile product per step, the
    // trailing updates of a step run side by side
    ctx->cnt = cdesc[0].nt;
    const int cnt = ctx->cnt;
    // envelope: first tile column of each tile row's lower part, made
    // non-decreasing (no fill left of it), and the last tile row reaching
    // each column
    std::vector<int> pos((size_t)nc), fc((size_t)cnt), last((size_t)cnt);
    for (int k = 0; k < nc; k++) pos[cinds[k]] = k;
    for (int i = 0; i < cnt; i++) fc[i] = i;
    for (int k = 0; k < nc; k++) {
      const int g = cinds[k];
      for (int q = ctx->AC.ptr[g]; q < ctx->AC.ptr[g + 1]; q++) {
        const int l = pos[ctx->AC.col[q]];
        if (l < k) fc[k / kTile] = std::min(fc[k / kTile], l / kTile);
      }
    }
    for (int i = cnt - 2; i >= 0; i--) fc[i] = std::min(fc[i], fc[i + 1]);
    for (int j = 0, i = 0; j < cnt; j++) {
      while (i + 1 < cnt && fc[i + 1] <= j) i++;
      last[j] = std::max(i, j);
    }
    {
      std::vector<int32_t> perm((size_t)cnt * kTile), iperm((size_t)nc);
      for (int k = 0; k < cnt * kTile; k++) perm[k] = k < nc ? cinds[k] : k;
      for (int k = 0; k < nc; k++) iperm[cinds[k]] = k;
      int32_t* d_iperm;
      if ((rc = upload(ctx, &ctx->d_cperm, perm)) || (rc = upload(ctx, &d_iperm, iperm))) return rc;
      // the prolongation reads the coarse solution in the factor's order
      launch_map_index(ctx->d_Rtcol, doug_restriction_nnz(ctx), d_iperm, ctx->stream);
      HIPCHK(hipStreamSynchronize(ctx->stream));  // perm and iperm go out of scope
    }
    // the explicit inverse's step m follows the factorisation's step m on
    // the third stream, beside the factorisation's later steps (each a short
    // chain of small launches that leaves most of the GPU idle)
    while ((int)ctx->cevents.size() < cnt + 2) {
      hipEvent_t e;
      HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
      ctx->cevents.push_back(e);
    }
    if ((rc = dalloc(ctx, &ctx->d_cW, (size_t)celems))) return rc;
    HIPCHK(hipEventRecord(ctx->cevents[cnt + 1], ctx->stream));  // the coarse tiles are in place
    HIPCHK(hipStreamWaitEvent(ctx->stream3, ctx->cevents[cnt + 1], 0));
    launch_inv_init(ctx->d_cW, ctx->d_cdesc, cnt, ctx->stream3);
    for (int j = 0; j < cnt; j++) {
      launch_chol_diag(ctx->d_ctiles, ctx->d_cdesc, 1, j, ctx->d_flag + 1, ctx->stream, false);
      launch_chol_offdiag(ctx->d_ctiles, ctx->d_cdesc, 1, j, last[j] - j, ctx->stream, false);
      HIPCHK(hipEventRecord(ctx->cevents[j], ctx->stream));
      HIPCHK(hipStreamWaitEvent(ctx->stream3, ctx->cevents[j], 0));
      launch_inv_step(ctx->d_ctiles, ctx->d_cW, ctx->d_cdesc, cnt, j, last[j], ctx->stream3);
      launch_rl_update(ctx->d_ctiles, ctx->d_cdesc, cnt, j, last[j], ctx->stream);
    }
    HIPCHK(hipEventRecord(ctx->cevents[cnt], ctx->stream3));
    HIPCHK(hipStreamWaitEvent(ctx->stream, ctx->cevents[cnt], 0));
    if (ctx->precision == DOUG_PREC_FP32_FACTORS) {
      // the FP32-factor form applies the coarse inverse in FP32 too (half
      // the bytes per apply; its products and sums stay in FP64)
      if ((rc = dalloc(ctx, &ctx->d_cW32, (size_t)celems))) return rc;
      launch_to_float(ctx->d_cW, ctx->d_cW32, celems, ctx->stream);
    }
    HIPCHK(hipGetLastError());
    size_t cpad = (size_t)ctx->cnt * kTile;
    if ((rc = dalloc(ctx, &ctx->d_cpart, cpad * gemv_chunks(ctx->cnt)))) return rc;
    for (double** v : {&ctx->d_crhs, &ctx->d_cu, &ctx->d_csol}) {
      if ((rc = dalloc(ctx, v, cpad))) return rc;
      HIPCHK(hipMemsetAsync(*v, 0, cpad * sizeof(double), ctx->stream));
    }
    ctx->info.coarse_factor_bytes = 2 * celems * (int64_t)sizeof(double) +
                                    (ctx->d_cW32 ? celems * (int64_t)sizeof(float) : 0);
  }
  HIPCHK(hipEventRecord(ev3, ctx->stream));
  return DOUG_OK;
}

int doug_setup_host(doug_ctx* ctx) {
  if (!ctx) return DOUG_E_ARG;
  return setup_host(ctx);
}

int doug_setup(doug_ctx* ctx, int32_t precision) {
  if (!ctx) return DOUG_E_ARG;
  if (ctx->device < 0) return fail(ctx, DOUG_E_STATE, "host-only context: no GPU work");
  if (precision != DOUG_PREC_FP64 && precision != DOUG_PREC_FP32_FACTORS)
    return fail(ctx, DOUG_E_ARG, "precision must be DOUG_PREC_FP64 or DOUG_PREC_FP32_FACTORS");
  int rc = check_params(ctx);
  if (rc) return rc;
  ctx->precision = precision;
  HIPCHK(hipSetDevice(ctx->device));
  free_device(ctx, true);
  double t0 = now();
  // [0, 1] subdomain factorisation, [2, 3] coarse factorisation, [4] the
  // input of the subdomain factorisation, [5] its end
  hipEvent_t ev[6];
  for (auto& e : ev) HIPCHK(hipEventCreate(&e));
  std::vector<double>& scaled = ctx->scaled;
  std::vector<uint8_t>& strong = ctx->strong;
  // non-positive pivot flags: [0] subdomains, [1] coarse matrix
  if ((rc = dalloc(ctx, &ctx->d_flag, 2))) return rc;
  HIPCHK(hipMemsetAsync(ctx->d_flag, 0, 2 * sizeof(int), ctx->stream));
  // several processes aggregate their own rows of the values as read (the
  // pattern of A and A_input_val, which the scale round trip leaves alone):
  // they start before it, the round trip and the upload take half of the
  // hardware threads beside them
  std::thread agg;
  const bool early_agg = ctx->nranks > 0 && ctx->inj_nagr == 0 && !ctx->have_partition;
  // several processes with the device coarse build: the ranks aggregate
  // their own values as read, and the global round trip and strength go to
  // the device
  const bool dev_values = early_agg && ctx->ol_dist == 1 && ctx->smoothers == 1 &&
                          ctx->use_device_coarse && ctx->has_coarse;
  ctx->host_values = true;
  if (early_agg) {
    // the host threads are split between the two: the aggregation (the
    // longer) keeps all but `side` of them, this thread's round trip and
    // upload take `side` (one thread when the device forms the values)
    const int all = setup_threads(), side = dev_values ? 1 : std::max(1, all / 8);
    agg = std::thread([&ctx, &scaled, &strong, all, side]() {
      set_setup_thread_cap(std::max(1, all - side));
      phase_aggregate(ctx, scaled, strong);
    });
    set_setup_thread_cap(side);
  }
  if (!dev_values) phase_values(ctx, scaled, strong);
  // waits for the host thread that enqueues the subdomain work, if any
  std::function<int()> join_fine;
  // R = T S^T and A_c on the device for the smoothed prolongator
  auto coarse_space = [&]() -> int {
    if (rank_coarse(ctx)) {
      if (ctx->ol_dist == 1 && ctx->smoothers == 1 && ctx->use_device_coarse && ctx->d_valin) {
        int r = coarse_build_device(ctx, strong, true);
        if (r != DOUG_E_UNSUPPORTED) {
          for (RankLocal& L : ctx->rsetup.ranks) {
            L.A = Csr();
            L.agg = Aggregates();
          }
          return r;
        }
        (void)hipStreamSynchronize(ctx->stream);
      }
      if (join_fine)
        if (int r = join_fine()) return r;
      if (int r = host_values_now(ctx)) return r;
      if (!ctx->ranks_kept_local) {
        // the same decomposition again, now keeping each rank's local
        // matrix (the aggregation is deterministic: the same results)
        RankSetup again;
        run_virtual_ranks(ctx, again, nullptr, true);
        ctx->rsetup.ranks = std::move(again.ranks);
        ctx->ranks_kept_local = true;
      }
      phase_coarse_ranks(ctx, strong);
      return DOUG_OK;
    }
    if (ctx->have_inj_R) return phase_coarse_given(ctx);
    if (ctx->smoothers == 1 && ctx->use_device_coarse) return coarse_build_device(ctx, strong);
    phase_coarse_host(ctx, strong);
    return DOUG_OK;
  };
  bool coarse_done = false;
  if (ctx->have_partition) {
    // neither the matrix upload nor the subdomains depend on the
    // aggregation: a second host thread aggregates while this one uploads
    // the matrix, prepares the subdomains and the GPU factorises them.  The
    // two touch disjoint context members and read A, scaled and strong only.
    // The aggregation is sequential and the longest setup step, so this
    // thread's parallel loops leave it half of the hardware threads.
    agg = std::thread([&]() { phase_aggregate(ctx, scaled, strong); });
    set_setup_thread_cap(std::max(1, setup_threads() / 2 - 1));
    rc = upload_matrix(ctx);
    if (!rc) rc = upload_strong(ctx, strong);
    if (!rc) {
      phase_subdomains(ctx);
      rc = device_fine(ctx, ev[0], ev[1]);
    }
    set_setup_thread_cap(0);
    agg.join();
    if (rc) return rc;
    if ((rc = coarse_space())) return rc;
  } else {
    if (!early_agg) agg = std::thread([&]() { phase_aggregate(ctx, scaled, strong); });
    rc = upload_matrix(ctx, !dev_values);
    if (!rc) rc = dev_values ? device_values(ctx) : upload_strong(ctx, strong);
    set_setup_thread_cap(0);
    agg.join();
    if (rc) return rc;
    // the subdomain factorisation goes to the second stream, after the
    // matrix upload; the first stream waits for its end before anything
    // else.  The several-process subdomains are known once the ranks have
    // aggregated, so their factorisation runs beside the coarse space's
    // construction and factorisation; the one-process subdomains come from
    // the coarse aggregation, so they follow the coarse space, and their
    // factorisation runs beside the coarse one (a chain of small steps that
    // leaves most of the GPU idle).
    HIPCHK(hipEventRecord(ev[4], ctx->stream));
    auto fine = [&]() -> int {
      phase_subdomains(ctx);
      HIPCHK(hipStreamWaitEvent(ctx->stream2, ev[4], 0));
      int r;
      {
        StreamOverride on2(ctx->stream2);
        r = device_fine(ctx, ev[0], ev[1]);
      }
      if (r) return r;
      HIPCHK(hipEventRecord(ev[5], ctx->stream2));
      return DOUG_OK;
    };
    if (rank_coarse(ctx)) {
      // With exact subdomain solves and the device coarse build, a second
      // host thread enqueues the subdomain work (index sets, occurrence
      // lists, envelopes, factorisation) while this one builds the coarse
      // space: both spend most of their time waiting on the GPU.  They touch
      // disjoint context members; allocations and errors go through ctx->mu.
      // The host coarse build (the fallback) reads the host values, so it
      // waits for that thread first.
      const bool beside = ctx->p.fine_method == 1 && ctx->ol_dist == 1 && ctx->smoothers == 1 &&
                          ctx->use_device_coarse && ctx->d_valin;
      int rc_fine = DOUG_OK;
      std::thread fth;
      if (beside)
        fth = std::thread([&]() {
          rc_fine = hipSetDevice(ctx->device) == hipSuccess ? fine() : DOUG_E_HIP;
        });
      else if ((rc = fine()))
        return rc;
      join_fine = [&]() {
        if (fth.joinable()) fth.join();
        return rc_fine;
      };
      rc = coarse_space();
      if (!rc) rc = device_coarse(ctx, ev[2], ev[3]);
      const int rf = join_fine();
      join_fine = nullptr;
      if (rc || (rc = rf)) return rc;
    } else {
      if ((rc = coarse_space()) || (rc = device_coarse(ctx, ev[2], ev[3])) || (rc = fine())) return rc;
    }
    coarse_done = true;
    HIPCHK(hipStreamWaitEvent(ctx->stream, ev[5], 0));
  }
  fill_info(ctx);
  ctx->host_ready = true;
  if (!coarse_done && (rc = device_coarse(ctx, ev[2], ev[3]))) return rc;
  int flag[2] = {0, 0};
  HIPCHK(hipMemcpyAsync(flag, ctx->d_flag, 2 * sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  join_reaper(ctx);
  float ms_f = 0.f, ms_c = 0.f;
  HIPCHK(hipEventElapsedTime(&ms_f, ev[0], ev[1]));
  HIPCHK(hipEventElapsedTime(&ms_c, ev[2], ev[3]));
  for (auto& e : ev) (void)hipEventDestroy(e);
  ctx->info.t_factor = ms_f * 1e-3;
  ctx->info.t_coarse_factor = ms_c * 1e-3;
  if (!ctx->keep_setup_temps) free_scratch_slots(ctx);
  ctx->info.temp_bytes = scratch_slot_bytes(ctx);
  ctx->info.t_total = now() - t0;
  if (flag[0])
    return fail(ctx, DOUG_E_NUMERIC, "subdomain " + std::to_string(flag[0] - 1) + " is not positive definite");
  if (flag[1]) return fail(ctx, DOUG_E_NUMERIC, "the coarse matrix is not positive definite");
  ctx->ready = true;
  return DOUG_OK;
}

int doug_get_info(const doug_ctx* ctx, doug_info* info) {
  if (!ctx || !info) return DOUG_E_ARG;
  *info = ctx->info;
  return DOUG_OK;
}

static void put_aggr(const Aggregates& a, int32_t* num, int32_t* starts, int32_t* nodes) {
  if (num)
    for (size_t i = 0; i < a.num.size(); i++) num[i] = a.num[i];
  if (starts)
    for (size_t i = 0; i < a.starts.size(); i++) starts[i] = a.starts[i] + 1;
  if (nodes)
    for (size_t i = 0; i < a.nodes.size(); i++) nodes[i] = a.nodes[i] + 1;
}

int doug_get_fine_aggregates(const doug_ctx* ctx, int32_t* num, int32_t* starts, int32_t* nodes) {
  if (!ctx || !ctx->host_ready) return DOUG_E_STATE;
  put_aggr(ctx->fin, num, starts, nodes);
  return DOUG_OK;
}

int doug_get_coarse_aggregates(const doug_ctx* ctx, int32_t* num, int32_t* starts, int32_t* nodes) {
  if (!ctx || !ctx->host_ready) return DOUG_E_STATE;
  put_aggr(ctx->cfull, num, starts, nodes);
  return DOUG_OK;
}

int doug_get_subdomains(const doug_ctx* ctx, int32_t* starts, int32_t* inds) {
  if (!ctx || !ctx->host_ready) return DOUG_E_STATE;
  for (size_t i = 0; i < ctx->sstarts.size(); i++) starts[i] = ctx->sstarts[i] + 1;
  for (size_t i = 0; i < ctx->sinds.size(); i++) inds[i] = ctx->sinds[i] + 1;
  return DOUG_OK;
}

static void put_csr(const Csr& M, int32_t* rowptr, int32_t* cols, double* vals) {
  for (size_t i = 0; i < M.ptr.size(); i++) rowptr[i] = M.ptr[i];
  for (size_t i = 0; i < M.col.size(); i++) {
    cols[i] = M.col[i];
    vals[i] = M.val[i];
  }
}

int64_t doug_restriction_nnz(const doug_ctx* ctx) {
  if (!ctx) return -1;
  return ctx->coarse_on_device ? ctx->nnz_R : ctx->R.nnz();
}
int64_t doug_coarse_matrix_nnz(const doug_ctx* ctx) { return ctx ? ctx->AC.nnz() : -1; }

int doug_get_restriction(const doug_ctx* ctx, int32_t* rowptr, int32_t* cols, double* vals) {
  if (!ctx || !ctx->host_ready) return DOUG_E_STATE;
  if (!ctx->coarse_on_device) {
    put_csr(ctx->R, rowptr, cols, vals);
    return DOUG_OK;
  }
  for (size_t i = 0; i < ctx->R.ptr.size(); i++) rowptr[i] = ctx->R.ptr[i];
  if (hipSetDevice(ctx->device) != hipSuccess ||
      hipStreamSynchronize(ctx->stream) != hipSuccess ||
      hipMemcpy(cols, ctx->d_Rcol, (size_t)ctx->nnz_R * sizeof(int32_t), hipMemcpyDeviceToHost) != hipSuccess ||
      hipMemcpy(vals, ctx->d_Rval, (size_t)ctx->nnz_R * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    return DOUG_E_HIP;
  return DOUG_OK;
}

int doug_get_coarse_matrix(const doug_ctx* ctx, int32_t* rowptr, int32_t* cols, double* vals) {
  if (!ctx || !ctx->host_ready) return DOUG_E_STATE;
  put_csr(ctx->AC, rowptr, cols, vals);
  return DOUG_OK;
}

// ---------------------------------------------------------------------------
static void enqueue_spmv(doug_ctx* ctx, const double* x, double* y) {
  launch_spmv(ctx->nrb_a, ctx->d_rowblk_a, ctx->d_ptr, ctx->d_col, ctx->d_val, x, y, ctx->stream,
              ctx->spmv_shape);
}

// Work vectors of one preconditioner apply: the subdomain vectors and the
// coarse vectors with the chunked GEMVs' partial sums.
struct PrecWork {
  double *subsol, *crhs, *cu, *csol, *cpart;
};

static PrecWork ctx_work(doug_ctx* ctx) {
  return PrecWork{ctx->d_subsol, ctx->d_crhs, ctx->d_cu, ctx->d_csol, ctx->d_cpart};
}

static void enqueue_subdomain_solves(doug_ctx* ctx, const double* r, double* subsol,
                                     const double* skip = nullptr) {
  if (ctx->p.fine_method == 2) {  // FinePreconditioner_sgs_Apply
    launch_gather_positions((int64_t)ctx->sinds.size(), ctx->d_inds, r, subsol, ctx->stream);
    launch_sgs(ctx->d_sgs, ctx->nsub, ctx->sgs_mmax, ctx->d_sgs_levptr, ctx->d_sgs_lrows,
               ctx->d_sgs_rbeg, ctx->d_sgs_rlen, ctx->d_sgs_col, ctx->d_sgs_val, ctx->d_sgs_diag,
               subsol, ctx->sgs_iters, ctx->stream);
    return;
  }
  if (ctx->precision == DOUG_PREC_FP32_FACTORS)
    launch_trsv(ctx->d_tiles32, ctx->d_sub_trsv, ctx->nsub, ctx->maxnt, ctx->d_inds, r, subsol,
                ctx->stream, skip);
  else
    launch_trsv(ctx->d_tiles, ctx->d_sub_trsv, ctx->nsub, ctx->maxnt, ctx->d_inds, r, subsol,
                ctx->stream, skip);
}

static void enqueue_subdomain_solves(doug_ctx* ctx, const double* r) {
  enqueue_subdomain_solves(ctx, r, ctx->d_subsol);
}

// preconditioner (pcg.F90:216-270): fine solves, then coarse correction added.
// With `dot`, PCG's r.z (rho and beta in d_scal) is formed with z; with
// `skip`, the subdomain solves do nothing once *skip is set.
static void enqueue_prec(doug_ctx* ctx, const double* r, double* z, const PrecWork& w,
                         bool dot = false, const double* skip = nullptr) {
  const Params& p = ctx->p;
  if (p.method == 0) {
    launch_copy(z, r, ctx->n, ctx->stream);
    if (dot) launch_dot(r, z, ctx->n, ctx->d_partial, nullptr, 1, ctx->d_scal, ctx->stream);
    return;
  }
  // prec2Level_solve (pcg.F90:191-208) with several processes, where the
  // restriction has taken r: method 4 adds no coarse correction, method 2
  // (overlap > 0) keeps the coarse correction alone, 1, 3 and 5 add both
  const bool coarse = ctx->has_coarse && p.method != 4;
  const bool fine = !(ctx->has_coarse && p.method == 2);
  if (coarse) {
    launch_spmv_wave_rows(ctx->nc, ctx->d_Rptr, ctx->d_Rcol, ctx->d_Rval, r, w.crhs, ctx->stream);
    if (ctx->d_cW32) {
      launch_gemv_w(ctx->d_cW32, ctx->d_cdesc, ctx->cnt, w.crhs, w.cu, w.cpart, ctx->stream, ctx->d_cperm);
      launch_gemv_wt(ctx->d_cW32, ctx->d_cdesc, ctx->cnt, w.cu, w.csol, w.cpart, ctx->stream);
    } else {
      launch_gemv_w(ctx->d_cW, ctx->d_cdesc, ctx->cnt, w.crhs, w.cu, w.cpart, ctx->stream, ctx->d_cperm);
      launch_gemv_wt(ctx->d_cW, ctx->d_cdesc, ctx->cnt, w.cu, w.csol, w.cpart, ctx->stream);
    }
  }
  const double* fine_src = r;
  if (fine && p.fine_method >= 1) {
    enqueue_subdomain_solves(ctx, r, w.subsol, skip);
    fine_src = w.subsol;
  }
  if (dot)
    launch_overlap_sum_dot(ctx->n, fine ? ctx->d_occ_ptr : nullptr, ctx->d_occ_pos, fine_src,
                           coarse ? ctx->d_Rtptr : nullptr, ctx->d_Rtcol, ctx->d_Rtval, w.csol, z, r,
                           ctx->d_partial, ctx->d_scal, ctx->stream);
  else
    launch_overlap_sum(ctx->n, fine ? ctx->d_occ_ptr : nullptr, ctx->d_occ_pos, fine_src,
                       coarse ? ctx->d_Rtptr : nullptr, ctx->d_Rtcol, ctx->d_Rtval, w.csol, z,
                       ctx->stream);
}

static void enqueue_prec(doug_ctx* ctx, const double* r, double* z, bool dot = false,
                         const double* skip = nullptr) {
  enqueue_prec(ctx, r, z, ctx_work(ctx), dot, skip);
}

int doug_spmv_device(doug_ctx* ctx, const double* dx, double* dy) {
  if (!ctx) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_spmv before doug_setup");
  enqueue_spmv(ctx, dx, dy);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

int doug_spmv(doug_ctx* ctx, const double* x, double* y) {
  if (!ctx || !x || !y) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_spmv before doug_setup");
  size_t bytes = (size_t)ctx->n * sizeof(double);
  HIPCHK(hipMemcpyAsync(ctx->d_p, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  enqueue_spmv(ctx, ctx->d_p, ctx->d_q);
  HIPCHK(hipMemcpyAsync(y, ctx->d_q, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

int doug_prec_apply_device(doug_ctx* ctx, const double* dr, double* dz) {
  if (!ctx) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_prec_apply before doug_setup");
  enqueue_prec(ctx, dr, dz);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Allocation-free preconditioner apply: the work vectors of an apply live in
// one file-scope device array per GPU, of kNoAllocDoubles doubles (32 MiB),
// which the loaded code object holds until the process ends.  Host code
// never passes the array's name to a kernel: its device address comes from
// hipGetSymbolAddress, once per device, with that device current.
constexpr int64_t kNoAllocDoubles = 4194304;
static __device__ __align__(256) double scratch[kNoAllocDoubles];
static std::mutex g_scratch_mu;
static std::vector<double*> g_scratch_ptr;  // per device ordinal

static int scratch_address(doug_ctx* ctx, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= ctx->device) g_scratch_ptr.resize((size_t)ctx->device + 1, nullptr);
  if (!g_scratch_ptr[ctx->device]) {
    double* p = nullptr;
    hipError_t e = hipGetSymbolAddress((void**)&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
      (void)hipGetLastError();
      return fail(ctx, DOUG_E_HIP, std::string("hipGetSymbolAddress of the scratch array: ") +
                                       hipGetErrorString(e));
    }
    g_scratch_ptr[ctx->device] = p;
  }
  *out = g_scratch_ptr[ctx->device];
  return DOUG_OK;
}

extern "C" {

int doug_prec_apply_device_noalloc(doug_ctx* ctx, const double* dr, double* dz) {
  if (!ctx || !dr || !dz) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_prec_apply before doug_setup");
  // layout: subdomain vectors, then the three coarse vectors and the partial
  // sums, each part starting on a 256-byte boundary
  auto up = [](int64_t v) { return (v + 31) / 32 * 32; };
  const int64_t nsubv = up((int64_t)ctx->sinds.size());
  const bool coarse = ctx->has_coarse && ctx->p.method != 4 && ctx->p.method != 0;
  const int64_t cpad = coarse ? up((int64_t)ctx->cnt * kTile) : 0;
  const int64_t npart = coarse ? up((int64_t)ctx->cnt * kTile * gemv_chunks(ctx->cnt)) : 0;
  const int64_t need = nsubv + 3 * cpad + npart;
  if (need > kNoAllocDoubles)
    return fail(ctx, DOUG_E_ARG, "the allocation-free apply needs " + std::to_string(need) +
                                     " doubles of work space; its scratch array holds " +
                                     std::to_string(kNoAllocDoubles));
  HIPCHK(hipSetDevice(ctx->device));
  double* base = nullptr;
  if (int rc = scratch_address(ctx, &base)) return rc;
  PrecWork w{base, base + nsubv, base + nsubv + cpad, base + nsubv + 2 * cpad, base + nsubv + 3 * cpad};
  // the coarse vectors' padding beyond n_c enters the tile GEMVs: start at zero
  if (coarse) launch_zero(w.crhs, 3 * cpad, ctx->stream);
  enqueue_prec(ctx, dr, dz, w);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

int doug_prec_apply(doug_ctx* ctx, const double* r, double* z) {
  if (!ctx || !r || !z) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_prec_apply before doug_setup");
  size_t bytes = (size_t)ctx->n * sizeof(double);
  HIPCHK(hipMemcpyAsync(ctx->d_r, r, bytes, hipMemcpyHostToDevice, ctx->stream));
  enqueue_prec(ctx, ctx->d_r, ctx->d_z);
  HIPCHK(hipMemcpyAsync(z, ctx->d_z, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

// ---------------------------------------------------------------------------
// Lanczos condition estimate (pcg.F90:465-497) with EISPACK tql1
// (Bowdler, Martin, Reinsch, Wilkinson, Num. Math. 11 (1968) 293-306),
// 1-based arrays.
static double pythag(double a, double b) {
  double p = std::max(std::fabs(a), std::fabs(b));
  if (p == 0.0) return p;
  double r = std::min(std::fabs(a), std::fabs(b)) / p;
  r = r * r;
  for (;;) {
    double t = 4.0 + r;
    if (t == 4.0) break;
    double s = r / t, u = 1.0 + 2.0 * s;
    p = u * p;
    r = (s / u) * (s / u) * r;
  }
  return p;
}

static int tql1(int n, std::vector<double>& d, std::vector<double>& e) {
  if (n == 1) return 0;
  for (int i = 2; i <= n; i++) e[i - 1] = e[i];
  double f = 0.0, tst1 = 0.0;
  e[n] = 0.0;
  for (int l = 1; l <= n; l++) {
    int j = 0, m = l;
    double h = std::fabs(d[l]) + std::fabs(e[l]);
    tst1 = std::max(tst1, h);
    for (m = l; m <= n; m++)
      if (tst1 + std::fabs(e[m]) == tst1) break;
    if (m != l) {
      for (;;) {
        if (j == 30) return l;
        j++;
        int l1 = l + 1, l2 = l1 + 1;
        double g = d[l];
        double pp = (d[l1] - g) / (2.0 * e[l]);
        double r = pythag(pp, 1.0);
        double sr = pp >= 0 ? std::fabs(r) : -std::fabs(r);
        d[l] = e[l] / (pp + sr);
        d[l1] = e[l] * (pp + sr);
        double dl1 = d[l1];
        h = g - d[l];
        for (int i = l2; i <= n; i++) d[i] = d[i] - h;
        f = f + h;
        pp = d[m];
        double c = 1.0, c2 = c, c3 = c, el1 = e[l1], s = 0.0, s2 = 0.0;
        for (int ii = 1; ii <= m - l; ii++) {
          c3 = c2;
          c2 = c;
          s2 = s;
          int i = m - ii;
          g = c * e[i];
          h = c * pp;
          r = pythag(pp, e[i]);
          e[i + 1] = s * r;
          s = e[i] / r;
          c = pp / r;
          pp = c * d[i] - s * g;
          d[i + 1] = h + s * (c * g + s * d[i]);
        }
        pp = -s * s2 * c3 * el1 * e[l] / dl1;
        e[l] = s * pp;
        d[l] = c * pp;
        if (!(tst1 + std::fabs(e[l]) > tst1)) break;
      }
    }
    double pp = d[l] + f;
    int i = 1;
    for (int ii = 2; ii <= l; ii++) {
      int k = l + 2 - ii;
      if (pp >= d[k - 1]) {
        i = k;
        break;
      }
      d[k] = d[k - 1];
      i = 1;
    }
    d[i] = pp;
  }
  return 0;
}

static double lanczos_cond(int it, const std::vector<double>& alpha, const std::vector<double>& beta) {
  if (it < 1) return 0.0;
  std::vector<double> d((size_t)it + 1, 0.0), e((size_t)it + 1, 0.0);
  d[1] = 1.0 / alpha[1];
  for (int i = 2; i <= it; i++) {
    if (alpha[i] == 0 || beta[i] < 0) return 0.0;
    d[i] = 1.0 / alpha[i] + beta[i] / alpha[i - 1];
    e[i] = -std::sqrt(beta[i]) / alpha[i - 1];
  }
  tql1(it, d, e);
  return d[it] / d[1];
}

static int pcg_core(doug_ctx* ctx, double tol, int maxit, int32_t* it_out, double* cond,
                    double* res_hist) {
  if (maxit == -1) maxit = 100000;
  const int n = ctx->n;
  hipStream_t st = ctx->stream;
  double* scal = ctx->h_scal;
  // r = b - A x
  enqueue_spmv(ctx, ctx->d_x, ctx->d_tmp);
  launch_residual(n, ctx->d_b, ctx->d_tmp, ctx->d_r, st);
  launch_dot(ctx->d_r, ctx->d_r, n, ctx->d_partial, nullptr, 3, ctx->d_scal, st);
  HIPCHK(hipMemcpyAsync(scal, ctx->d_scal, S_N * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  double init_norm = scal[S_RES];
  if (init_norm == 0.0) init_norm = 1.0;
  std::vector<double> alpha((size_t)std::min(maxit, 1 << 20) + 2, 0.0), beta(alpha.size(), 0.0);
  // The convergence test runs on the device (k_reduce_finish): the host
  // keeps the next iteration queued while it waits for the current one, so
  // the GPU does not idle between iterations.  An iteration queued after
  // the test passed leaves x and r unchanged and skips the subdomain solves.
  const int64_t cap = (int64_t)alpha.size() - 2;
  if (cap > ctx->hist_cap) {
    int rc = dalloc(ctx, &ctx->d_hist, (size_t)cap * 3);
    if (rc) return rc;
    ctx->hist_cap = cap;
  }
  scal[S_DONE] = 0.0;
  scal[S_ITER] = 0.0;
  scal[S_THR] = tol * tol;
  scal[S_INIT] = init_norm;
  scal[S_MAXIT] = (double)maxit;
  scal[S_HCAP] = (double)cap;
  HIPCHK(hipMemcpyAsync(ctx->d_scal + S_DONE, scal + S_DONE, (S_HCAP - S_DONE + 1) * sizeof(double),
                        hipMemcpyHostToDevice, st));
  const double* done_flag = ctx->d_scal + S_DONE;
  // iteration body (pcg.F90:387-435); from the second iteration on it is
  // replayed from a captured HIP graph
  auto enqueue_iteration = [&](bool first) {
    enqueue_prec(ctx, ctx->d_r, ctx->d_z, true, done_flag);  // with rho = r.z and beta
    launch_p_update(n, ctx->d_z, ctx->d_p, ctx->d_scal, first ? 1 : 0, st);
    enqueue_spmv(ctx, ctx->d_p, ctx->d_q);
    launch_dot(ctx->d_p, ctx->d_q, n, ctx->d_partial, nullptr, 2, ctx->d_scal, st);
    launch_xr_update(n, ctx->d_x, ctx->d_r, ctx->d_p, ctx->d_q, ctx->d_partial, ctx->d_scal, st, ctx->d_hist);
  };
  hipEvent_t ev[2] = {nullptr, nullptr};
  for (hipEvent_t& e : ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  struct EventGuard {
    hipEvent_t* e;
    ~EventGuard() {
      for (int k = 0; k < 2; k++)
        if (e[k]) (void)hipEventDestroy(e[k]);
    }
  } guard{ev};
  int issued = 0;
  // queue iteration issued + 1 and a copy of the scalar state behind it
  auto issue = [&]() -> int {
    issued++;
    if (issued >= 2 && ctx->use_graph) {
      if (!ctx->iter_graph) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(false);
        HIPCHK(hipStreamEndCapture(st, &g));
        hipError_t ie = hipGraphInstantiate(&ctx->iter_graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
      }
      HIPCHK(hipGraphLaunch(ctx->iter_graph, st));
    } else {
      enqueue_iteration(issued == 1);
    }
    double* hs = scal + (issued & 1) * S_N;
    HIPCHK(hipMemcpyAsync(hs, ctx->d_scal, S_N * sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipEventRecord(ev[issued & 1], st));
    HIPCHK(hipGetLastError());
    return DOUG_OK;
  };
  // no iteration at all where the loop test fails at the start (ratio 1)
  bool done = maxit <= 0 || !(1.0 > tol * tol);
  int rc = done ? DOUG_OK : issue();
  if (rc) return rc;
  for (int checked = 0; !done;) {
    if (issued < maxit && (rc = issue())) return rc;
    checked++;
    HIPCHK(hipEventSynchronize(ev[checked & 1]));
    done = scal[(checked & 1) * S_N + S_DONE] != 0.0;
  }
  HIPCHK(hipStreamSynchronize(st));
  int it = 0;
  if (issued > 0) {
    HIPCHK(hipMemcpy(scal, ctx->d_scal, S_N * sizeof(double), hipMemcpyDeviceToHost));
    it = (int)scal[S_ITER];
    const int nh = (int)std::min<int64_t>(it, cap);
    std::vector<double> h((size_t)nh * 3);
    if (nh > 0) HIPCHK(hipMemcpy(h.data(), ctx->d_hist, h.size() * sizeof(double), hipMemcpyDeviceToHost));
    for (int k = 1; k <= nh; k++) {
      alpha[k] = h[3 * (k - 1)];
      beta[k] = h[3 * (k - 1) + 1];
      if (res_hist) res_hist[k - 1] = std::sqrt(h[3 * (k - 1) + 2]);
    }
  }
  if (it_out) *it_out = it;
  if (cond) *cond = lanczos_cond(std::min<int>(it, (int)alpha.size() - 2), alpha, beta);
  return DOUG_OK;
}

int doug_pcg_device(doug_ctx* ctx, const double* db, double* dx, double tol, int32_t maxit,
                    int32_t* it, double* cond, double* res_hist) {
  if (!ctx || !db || !dx) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_pcg before doug_setup");
  size_t bytes = (size_t)ctx->n * sizeof(double);
  HIPCHK(hipMemcpyAsync(ctx->d_b, db, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_x, dx, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  int rc = pcg_core(ctx, tol, maxit, it, cond, res_hist);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(dx, ctx->d_x, bytes, hipMemcpyDeviceToDevice, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

int doug_pcg(doug_ctx* ctx, const double* b, double* x, double tol, int32_t maxit, int32_t* it,
             double* cond, double* res_hist) {
  if (!ctx || !b || !x) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_pcg before doug_setup");
  size_t bytes = (size_t)ctx->n * sizeof(double);
  HIPCHK(hipMemcpyAsync(ctx->d_b, b, bytes, hipMemcpyHostToDevice, ctx->stream));
  HIPCHK(hipMemcpyAsync(ctx->d_x, x, bytes, hipMemcpyHostToDevice, ctx->stream));
  int rc = pcg_core(ctx, tol, maxit, it, cond, res_hist);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(x, ctx->d_x, bytes, hipMemcpyDeviceToHost, ctx->stream));
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}


// ---------------------------------------------------------------------------
// Several right-hand sides
static int next_pow2(int v) {
  int p = 1;
  while (p < v) p <<= 1;
  return p;
}

int doug_prepare_batch(doug_ctx* ctx) {
  if (!ctx) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_prepare_batch before doug_setup");
  if (ctx->batch_ready) return DOUG_OK;
  if (ctx->p.fine_method == 2)
    return fail(ctx, DOUG_E_UNSUPPORTED, "the batched solve has no Gauss-Seidel fine level (fine_method 2)");
  if (ctx->p.method >= 2 && ctx->p.method <= 4)
    return fail(ctx, DOUG_E_UNSUPPORTED, "the batched solve covers methods 0, 1 and 5");
  const int n = ctx->n, NR = kMaxRhs;
  hipStream_t st = ctx->stream;
  int rc;
  if (ctx->p.fine_method == 1) {
    // panels of the envelope solves; the factors themselves are shared with
    // the single right-hand-side path
    if (ctx->precision == DOUG_PREC_FP32_FACTORS) {
      float* u;
      if ((rc = dalloc(ctx, &u, (size_t)ctx->work_rows * NR))) return rc;
      ctx->d_U = u;
    } else {
      double* u;
      if ((rc = dalloc(ctx, &u, (size_t)ctx->work_rows * NR))) return rc;
      ctx->d_U = u;
    }
  }
  if (ctx->has_coarse) {
    size_t cpad = (size_t)ctx->cnt * kTile * NR;
    if ((rc = dalloc(ctx, &ctx->d_cPm, cpad * gemv_chunks(ctx->cnt)))) return rc;
    for (double** v : {&ctx->d_cG, &ctx->d_cU, &ctx->d_cS}) {
      if ((rc = dalloc(ctx, v, cpad))) return rc;
      HIPCHK(hipMemsetAsync(*v, 0, cpad * sizeof(double), st));
    }
  }
  if ((rc = dalloc(ctx, &ctx->d_subsolm, ctx->sinds.size() * (size_t)NR))) return rc;
  for (double** v : {&ctx->d_Xm, &ctx->d_Rm, &ctx->d_Zm, &ctx->d_Pm, &ctx->d_Pm2, &ctx->d_Qm, &ctx->d_Bm,
                     &ctx->d_Tm})
    if ((rc = dalloc(ctx, v, (size_t)n * NR))) return rc;
  // block partials: the column-pair reductions take kRedBlocksWide slots,
  // the fused SpMV one per row block
  if ((rc = dalloc(ctx, &ctx->d_scalm, (size_t)kMaxRhs * 8)) ||
      (rc = dalloc(ctx, &ctx->d_partialm, (size_t)std::max(kRedBlocksWide, ctx->nrb) * kMaxRhs)) ||
      (rc = dalloc(ctx, &ctx->d_counterm, 1)))
    return rc;
  HIPCHK(hipMemsetAsync(ctx->d_counterm, 0, sizeof(unsigned), st));
  HIPCHK(hipStreamSynchronize(st));
  ctx->batch_ready = true;
  return DOUG_OK;
}

// z = M^{-1} r for nrp columns (node-major, leading dimension nrp); with
// rz_dot (nrp even) the overlap sum also forms rho = r.z per column
static void enqueue_prec_multi(doug_ctx* ctx, int nrp, const double* r, double* z, bool rz_dot = false) {
  const Params& p = ctx->p;
  hipStream_t st = ctx->stream;
  const int n = ctx->n, NR = nrp <= 16 ? 16 : 32;
  if (p.method == 0) {
    launch_copy(z, r, n * nrp, st);
    return;
  }
  const bool coarse = ctx->has_coarse;
  if (coarse) {
    if (nrp % 2 == 0)  // restriction rows hold hundreds of entries
      launch_spmv_multi_rows(ctx->nc, ctx->d_Rptr, ctx->d_Rcol, ctx->d_Rval, r, nrp, ctx->d_cG, NR, nrp, st);
    else
      launch_spmv_multi(ctx->nrbR, ctx->d_Rrowblk, ctx->d_Rptr, ctx->d_Rcol, ctx->d_Rval, r, nrp,
                        ctx->d_cG, NR, nrp, st);
    launch_wgemm_fwd(ctx->d_cW, ctx->d_cdesc, ctx->cnt, ctx->d_cG, ctx->d_cU, ctx->d_cPm, NR, st,
                     ctx->d_cperm);
    launch_wgemm_bwd(ctx->d_cW, ctx->d_cdesc, ctx->cnt, ctx->cnt * kTile, ctx->d_cU, ctx->d_cS, NR, nrp,
                     ctx->d_cPm, NR, st);
  }
  const double* fine_src = r;
  if (p.fine_method == 1) {
    if (ctx->precision == DOUG_PREC_FP32_FACTORS)
      launch_trsm_multi(ctx->d_tiles32, ctx->d_sub_trsv, ctx->nsub, ctx->d_inds, r, nrp,
                        (float*)ctx->d_U, ctx->d_subsolm, nrp, nrp, NR, ctx->trsm_kb, ctx->trsm_cg, st);
    else
      launch_trsm_multi(ctx->d_tiles, ctx->d_sub_trsv, ctx->nsub, ctx->d_inds, r, nrp,
                        (double*)ctx->d_U, ctx->d_subsolm, nrp, nrp, NR, ctx->trsm_kb, ctx->trsm_cg, st);
    fine_src = ctx->d_subsolm;
  }
  if (rz_dot)
    launch_overlap_sum_dot_multi2(n, nrp, ctx->d_occ_ptr, ctx->d_occ_pos, fine_src,
                                  coarse ? ctx->d_Rtptr : nullptr, ctx->d_Rtcol, ctx->d_Rtval, ctx->d_cS, NR,
                                  z, r, ctx->d_partialm, ctx->d_counterm, ctx->d_scalm, st);
  else
    launch_overlap_sum_multi(n, nrp, ctx->d_occ_ptr, ctx->d_occ_pos, fine_src, coarse ? ctx->d_Rtptr : nullptr,
                             ctx->d_Rtcol, ctx->d_Rtval, ctx->d_cS, NR, z, st);
}

int doug_prec_apply_batch_device(doug_ctx* ctx, int32_t nrhs, const double* dr, double* dz) {
  if (!ctx || !dr || !dz || nrhs < 1 || nrhs > kMaxRhs) return DOUG_E_ARG;
  if (!ctx->batch_ready) return fail(ctx, DOUG_E_STATE, "doug_prec_apply_batch before doug_prepare_batch");
  const int nrp = next_pow2(nrhs), n = ctx->n;
  hipStream_t st = ctx->stream;
  HIPCHK(hipMemsetAsync(ctx->d_Rm, 0, (size_t)n * nrp * sizeof(double), st));
  HIPCHK(hipMemcpy2DAsync(ctx->d_Rm, nrp * sizeof(double), dr, nrhs * sizeof(double), nrhs * sizeof(double), n,
                          hipMemcpyDeviceToDevice, st));
  enqueue_prec_multi(ctx, nrp, ctx->d_Rm, ctx->d_Zm);
  HIPCHK(hipMemcpy2DAsync(dz, nrhs * sizeof(double), ctx->d_Zm, nrp * sizeof(double), nrhs * sizeof(double), n,
                          hipMemcpyDeviceToDevice, st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  return DOUG_OK;
}

// pcg_weigs run on every column at once; each column stops on its own test
static int pcg_batch_core(doug_ctx* ctx, int nrhs, double tol, int maxit, int32_t* its,
                          double* res_hist) {
  if (maxit == -1) maxit = 100000;
  const int n = ctx->n, nrp = next_pow2(nrhs);
  hipStream_t st = ctx->stream;
  double* hs = ctx->h_scalm;
  enum { SM_ACTIVE = 6 };
  // r = b - A x
  launch_spmv_multi(ctx->nrb, ctx->d_rowblk, ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_Xm, nrp,
                    ctx->d_Tm, nrp, nrp, st);
  launch_residual(n * nrp, ctx->d_Bm, ctx->d_Tm, ctx->d_Rm, st);
  HIPCHK(hipMemsetAsync(ctx->d_Pm, 0, (size_t)n * nrp * sizeof(double), st));
  HIPCHK(hipMemsetAsync(ctx->d_Pm2, 0, (size_t)n * nrp * sizeof(double), st));
  for (int k = 0; k < kMaxRhs * 8; k++) hs[k] = 0.0;
  for (int r = 0; r < nrhs; r++) hs[r * 8 + SM_ACTIVE] = 1.0;
  HIPCHK(hipMemcpyAsync(ctx->d_scalm, hs, kMaxRhs * 8 * sizeof(double), hipMemcpyHostToDevice, st));
  launch_dot_multi(ctx->d_Rm, ctx->d_Rm, n, nrp, ctx->d_partialm, ctx->d_counterm, 3, ctx->d_scalm, st);
  HIPCHK(hipMemcpyAsync(hs, ctx->d_scalm, kMaxRhs * 8 * sizeof(double), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  std::vector<double> init((size_t)nrhs);
  for (int r = 0; r < nrhs; r++) {
    init[r] = hs[r * 8 + S_RES] == 0.0 ? 1.0 : hs[r * 8 + S_RES];
    its[r] = 0;
  }
  int it = 0, nactive = nrhs;
  // With an even column count the iteration runs fused: the overlap sum
  // forms rho = r.z, and one SpMV pass forms p_new = z + beta p_old, q and
  // p.q, reading p from one buffer and writing the other (parity par of the
  // iteration); that pass also adds the previous iteration's alpha p to x,
  // and the last one is added after the loop.  Iterates are the same bits as
  // the unfused sequence below; only the summation order of p.q differs.
  const bool fused = nrp % 2 == 0;
  double* const P[2] = {ctx->d_Pm, ctx->d_Pm2};
  // iteration body; from the second iteration on it is replayed from a
  // captured HIP graph (one per padded column count and p parity)
  auto enqueue_iteration = [&](bool first, int par) {
    double* p = P[par ^ 1];
    if (fused) {
      enqueue_prec_multi(ctx, nrp, ctx->d_Rm, ctx->d_Zm, true);
      launch_spmv_p_multi2(ctx->nrb, ctx->d_rowblk, ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_Zm, P[par], p,
                           ctx->d_Qm, nrp, first ? 1 : 0, ctx->d_partialm, ctx->d_counterm, ctx->d_scalm,
                           ctx->d_Xm, st);
    } else {
      p = P[0];
      enqueue_prec_multi(ctx, nrp, ctx->d_Rm, ctx->d_Zm);
      launch_dot_multi(ctx->d_Rm, ctx->d_Zm, n, nrp, ctx->d_partialm, ctx->d_counterm, 1, ctx->d_scalm, st);
      launch_p_update_multi(n, nrp, ctx->d_Zm, p, ctx->d_scalm, first ? 1 : 0, st);
      launch_spmv_multi(ctx->nrb, ctx->d_rowblk, ctx->d_ptr, ctx->d_col, ctx->d_val, p, nrp, ctx->d_Qm, nrp,
                        nrp, st);
      launch_dot_multi(p, ctx->d_Qm, n, nrp, ctx->d_partialm, ctx->d_counterm, 2, ctx->d_scalm, st);
    }
    launch_xr_update_multi(n, nrp, ctx->d_Xm, ctx->d_Rm, p, ctx->d_Qm, ctx->d_partialm, ctx->d_counterm,
                           ctx->d_scalm, st);
    (void)hipMemcpyAsync(hs, ctx->d_scalm, kMaxRhs * 8 * sizeof(double), hipMemcpyDeviceToHost, st);
  };
  while (nactive > 0 && it < maxit) {
    it++;
    const int par = fused ? (it - 1) & 1 : 0;
    if (it >= 2 && ctx->use_graph) {
      if (ctx->batch_graph_nrp != nrp)
        for (hipGraphExec_t& g : ctx->batch_graph)
          if (g) {
            (void)hipGraphExecDestroy(g);
            g = nullptr;
          }
      ctx->batch_graph_nrp = nrp;
      if (!ctx->batch_graph[par]) {
        hipGraph_t g;
        HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
        enqueue_iteration(false, par);
        HIPCHK(hipStreamEndCapture(st, &g));
        hipError_t ie = hipGraphInstantiate(&ctx->batch_graph[par], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        HIPCHK(ie);
      }
      HIPCHK(hipGraphLaunch(ctx->batch_graph[par], st));
    } else {
      enqueue_iteration(it == 1, par);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(st));
    bool changed = false;
    for (int r = 0; r < nrhs; r++) {
      if (hs[r * 8 + SM_ACTIVE] == 0.0) continue;
      double res = hs[r * 8 + S_RES];
      if (res_hist) res_hist[(int64_t)(it - 1) * nrhs + r] = std::sqrt(res);
      its[r] = it;
      if (!(res / init[r] > tol * tol) || it >= maxit) {
        hs[r * 8 + SM_ACTIVE] = 0.0;
        nactive--;
        changed = true;
      }
    }
    if (changed)
      HIPCHK(hipMemcpyAsync(ctx->d_scalm, hs, kMaxRhs * 8 * sizeof(double), hipMemcpyHostToDevice, st));
  }
  // the last iteration's x += alpha p (alpha is 0 for columns frozen before it)
  if (fused && it > 0) {
    launch_x_final_multi2(n, nrp, ctx->d_Xm, P[((it - 1) & 1) ^ 1], ctx->d_scalm, st);
    HIPCHK(hipGetLastError());
  }
  return DOUG_OK;
}

int doug_pcg_batch_device(doug_ctx* ctx, int32_t nrhs, const double* dB, double* dX, double tol,
                          int32_t maxit, int32_t* its, double* res_hist) {
  if (!ctx || !dB || !dX || !its || nrhs < 1 || nrhs > kMaxRhs) return DOUG_E_ARG;
  if (!ctx->batch_ready) return fail(ctx, DOUG_E_STATE, "doug_pcg_batch before doug_prepare_batch");
  const int n = ctx->n, nrp = next_pow2(nrhs);
  hipStream_t st = ctx->stream;
  const size_t w = nrhs * sizeof(double), wp = nrp * sizeof(double);
  HIPCHK(hipMemsetAsync(ctx->d_Bm, 0, (size_t)n * wp, st));
  HIPCHK(hipMemsetAsync(ctx->d_Xm, 0, (size_t)n * wp, st));
  HIPCHK(hipMemcpy2DAsync(ctx->d_Bm, wp, dB, w, w, n, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpy2DAsync(ctx->d_Xm, wp, dX, w, w, n, hipMemcpyDeviceToDevice, st));
  int rc = pcg_batch_core(ctx, nrhs, tol, maxit, its, res_hist);
  if (rc) return rc;
  HIPCHK(hipMemcpy2DAsync(dX, w, ctx->d_Xm, wp, w, n, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  return DOUG_OK;
}

int doug_pcg_batch(doug_ctx* ctx, int32_t nrhs, const double* B, double* X, double tol,
                   int32_t maxit, int32_t* its, double* res_hist) {
  if (!ctx || !B || !X || !its || nrhs < 1 || nrhs > kMaxRhs) return DOUG_E_ARG;
  if (!ctx->batch_ready) return fail(ctx, DOUG_E_STATE, "doug_pcg_batch before doug_prepare_batch");
  const int n = ctx->n, nrp = next_pow2(nrhs);
  hipStream_t st = ctx->stream;
  // host rows of nrhs columns travel as one block through the workspace
  // d_Tm and are spread to (or gathered from) rows of nrp columns on the
  // device: a strided copy from pageable memory would go row by row
  const size_t bytes = (size_t)n * nrhs * sizeof(double);
  HIPCHK(hipMemcpyAsync(ctx->d_Tm, B, bytes, hipMemcpyHostToDevice, st));
  launch_repack(ctx->d_Bm, nrp, ctx->d_Tm, nrhs, n, nrhs, st);
  HIPCHK(hipMemcpyAsync(ctx->d_Tm, X, bytes, hipMemcpyHostToDevice, st));
  launch_repack(ctx->d_Xm, nrp, ctx->d_Tm, nrhs, n, nrhs, st);
  int rc = pcg_batch_core(ctx, nrhs, tol, maxit, its, res_hist);
  if (rc) return rc;
  launch_repack(ctx->d_Tm, nrhs, ctx->d_Xm, nrp, n, nrhs, st);
  HIPCHK(hipMemcpyAsync(X, ctx->d_Tm, bytes, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  return DOUG_OK;
}

int doug_set_option(doug_ctx* ctx, int32_t option, int64_t value) {
  if (!ctx) return DOUG_E_ARG;
  switch (option) {
    case DOUG_OPT_ITERATION_GRAPH:
      ctx->use_graph = value != 0;
      return DOUG_OK;
    case DOUG_OPT_SPMV_SHAPE: {
      if (value < 0 || value >= kSpmvShapes) return fail(ctx, DOUG_E_ARG, "SpMV shape out of range");
      ctx->spmv_shape = (int)value;
      if (!ctx->ready) return DOUG_OK;  // applied at the next doug_setup
      drop_iteration_graph(ctx);
      std::vector<int32_t> rb = row_blocks(ctx->A, spmv_shape_rows(ctx->spmv_shape));
      ctx->nrb_a = (int)rb.size() - 1;
      return upload(ctx, &ctx->d_rowblk_a, rb);
    }
    case DOUG_OPT_LOCAL_ORDER:
      if (value != 0 && value != 1) return fail(ctx, DOUG_E_ARG, "local order must be 0 or 1");
      ctx->local_order = (int)value;
      ctx->ready = false;  // takes effect at the next doug_setup
      return DOUG_OK;
    case DOUG_OPT_BATCH_TILES:
      if (value != 1 && value != 2) return fail(ctx, DOUG_E_ARG, "tiles per step must be 1 or 2");
      drop_iteration_graph(ctx);
      ctx->trsm_kb = (int)value;
      return DOUG_OK;
    case DOUG_OPT_BATCH_COLUMN_GROUPS:
      if (value < 1 || value > 3) return fail(ctx, DOUG_E_ARG, "column groups must be 1, 2 or 3");
      drop_iteration_graph(ctx);
      ctx->trsm_cg = (int)value;
      return DOUG_OK;
    case DOUG_OPT_DEVICE_BLOCKS:
      ctx->use_device_blocks = value != 0;
      ctx->ready = false;  // takes effect at the next doug_setup
      return DOUG_OK;
    case DOUG_OPT_DEVICE_COARSE:
      ctx->use_device_coarse = value != 0;
      ctx->ready = false;  // takes effect at the next doug_setup
      return DOUG_OK;
    case DOUG_OPT_KEEP_SETUP_TEMPS:
      ctx->keep_setup_temps = value != 0;
      if (!ctx->keep_setup_temps && ctx->device >= 0) {
        (void)hipSetDevice(ctx->device);
        free_scratch_slots(ctx);
        ctx->info.temp_bytes = 0;
      }
      return DOUG_OK;
    case DOUG_OPT_TRSV_ORDER:
      if (value != 0 && value != 1) return fail(ctx, DOUG_E_ARG, "subdomain order must be 0 or 1");
      if (!ctx->ready) {
        ctx->trsv_order = (int)value;
        return DOUG_OK;
      }
      drop_iteration_graph(ctx);
      return set_trsv_order(ctx, (int)value);
    default:
      return fail(ctx, DOUG_E_ARG, "unknown option " + std::to_string(option));
  }
}

int doug_value_grad_device(doug_ctx* ctx, const double* dlam, const double* dx, double* dgrad) {
  if (!ctx || !dlam || !dx || !dgrad) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_value_grad before doug_setup");
  if (!ctx->d_src) {  // each entry's input position, needed by the gradient alone
    if (int rc = upload(ctx, &ctx->d_src, ctx->A.src)) return rc;
    HIPCHK(hipStreamSynchronize(ctx->stream));
  }
  launch_value_grad(ctx->n, ctx->d_ptr, ctx->d_col, ctx->d_src, dlam, dx, dgrad, ctx->stream);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(ctx->stream));
  return DOUG_OK;
}

int doug_time_operator(doug_ctx* ctx, int32_t which, int32_t reps, double* ms) {
  if (!ctx || !ms || reps <= 0) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_time_operator before doug_setup");
  hipStream_t st = ctx->stream;
  auto run = [&]() {
    switch (which) {
      case 0: enqueue_spmv(ctx, ctx->d_p, ctx->d_q); break;
      case 1:
        if (ctx->p.fine_method >= 1) enqueue_subdomain_solves(ctx, ctx->d_r);
        break;
      case 2: enqueue_prec(ctx, ctx->d_r, ctx->d_z); break;
      case 3: launch_copy(ctx->d_q, ctx->d_p, ctx->n, st); break;
      case 4:
        if (ctx->batch_ready) enqueue_prec_multi(ctx, kMaxRhs, ctx->d_Rm, ctx->d_Zm);
        break;
      default: break;
    }
  };
  if (which < 0 || which > 4) return fail(ctx, DOUG_E_ARG, "unknown operator");
  if (which == 4 && !ctx->batch_ready) return fail(ctx, DOUG_E_STATE, "operator 4 needs doug_prepare_batch");
  run();  // warm-up
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++) run();
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *ms = (double)t / reps;
  return DOUG_OK;
}

}  // extern "C"

// Average milliseconds of `reps` launches of f after one warm-up launch.
template <typename F>
static int time_launches(doug_ctx* ctx, int reps, F&& f, double* ms) {
  hipStream_t st = ctx->stream;
  f();
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; r++) f();
  HIPCHK(hipEventRecord(e1, st));
  HIPCHK(hipEventSynchronize(e1));
  float t = 0.f;
  HIPCHK(hipEventElapsedTime(&t, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  HIPCHK(hipGetLastError());
  *ms = (double)t / reps;
  return DOUG_OK;
}

extern "C" {

int doug_autotune(doug_ctx* ctx, int32_t reps, doug_tuning* out) {
  if (!ctx) return DOUG_E_ARG;
  if (!ctx->ready) return fail(ctx, DOUG_E_STATE, "doug_autotune before doug_setup");
  if (reps <= 0) reps = 20;
  int rc;
  doug_tuning tu{};
  // SpMV: every shape over row blocks built for its row count
  int32_t* rbs[3] = {nullptr, nullptr, nullptr};
  int nrbs[3] = {0, 0, 0};
  for (int k = 0; k < 3; k++) {
    std::vector<int32_t> rb = row_blocks(ctx->A, spmv_shape_rows(2 * k));
    nrbs[k] = (int)rb.size() - 1;
    if ((rc = upload(ctx, &rbs[k], rb))) return rc;
  }
  for (int s = 0; s < kSpmvShapes; s++) {
    const int k = s >> 1;
    rc = time_launches(ctx, reps, [&]() {
      launch_spmv(nrbs[k], rbs[k], ctx->d_ptr, ctx->d_col, ctx->d_val, ctx->d_p, ctx->d_q, ctx->stream, s);
    }, &tu.spmv_ms[s]);
    if (rc) return rc;
  }
  int best = 0;
  for (int s = 1; s < kSpmvShapes; s++)
    if (tu.spmv_ms[s] < tu.spmv_ms[best]) best = s;
  ctx->spmv_shape = best;
  ctx->d_rowblk_a = rbs[best >> 1];
  ctx->nrb_a = nrbs[best >> 1];
  // subdomain solves: as partitioned, then largest first when that differs
  int order = 0;
  if (ctx->p.fine_method == 1 && ctx->nsub > 0) {
    const int rt = std::max(3, reps / 4);
    if ((rc = set_trsv_order(ctx, 0)) ||
        (rc = time_launches(ctx, rt, [&]() { enqueue_subdomain_solves(ctx, ctx->d_r); }, &tu.trsv_ms[0])))
      return rc;
    bool sorted = true;
    for (int i = 1; i < ctx->nsub; i++) sorted = sorted && ctx->h_sub[i - 1].nt >= ctx->h_sub[i].nt;
    if (!sorted) {
      if ((rc = set_trsv_order(ctx, 1)) ||
          (rc = time_launches(ctx, rt, [&]() { enqueue_subdomain_solves(ctx, ctx->d_r); }, &tu.trsv_ms[1])))
        return rc;
      order = tu.trsv_ms[1] < tu.trsv_ms[0] ? 1 : 0;
    }
    if ((rc = set_trsv_order(ctx, order))) return rc;
  }
  // batched envelope solve: off-diagonal tiles staged per step
  int kb = ctx->trsm_kb, cg = ctx->trsm_cg;
  if (ctx->batch_ready && ctx->p.fine_method == 1 && ctx->nsub > 0) {
    const int rt = std::max(2, reps / 8);
    double best_ms = 0.0, ms[3][2];
    for (int g = 1; g <= 3; g++)
      for (int q = 1; q <= 2; q++) {
        ctx->trsm_kb = q;
        ctx->trsm_cg = g;
        if ((rc = time_launches(ctx, rt, [&]() { enqueue_prec_multi(ctx, kMaxRhs, ctx->d_Rm, ctx->d_Zm); },
                                &ms[g - 1][q - 1])))
          return rc;
        if (best_ms == 0.0 || ms[g - 1][q - 1] < best_ms) {
          best_ms = ms[g - 1][q - 1];
          kb = q;
          cg = g;
        }
      }
    tu.trsm_ms[0] = ms[cg - 1][0];
    tu.trsm_ms[1] = ms[cg - 1][1];
  }
  ctx->trsm_kb = kb;
  ctx->trsm_cg = cg;
  tu.trsm_col_groups = cg;
  HIPCHK(hipStreamSynchronize(ctx->stream));
  tu.trsm_tiles_per_step = kb;
  tu.spmv_shape = best;
  tu.spmv_rows = spmv_shape_rows(best);
  tu.spmv_xcd_order = best & 1;
  tu.trsv_order = order;
  ctx->tuning = tu;
  drop_iteration_graph(ctx);  // its launches were recorded with the old choices
  if (out) *out = tu;
  return DOUG_OK;
}

int doug_get_tuning(const doug_ctx* ctx, doug_tuning* out) {
  if (!ctx || !out) return DOUG_E_ARG;
  *out = ctx->tuning;
  out->spmv_shape = ctx->spmv_shape;
  out->spmv_rows = spmv_shape_rows(ctx->spmv_shape);
  out->spmv_xcd_order = ctx->spmv_shape & 1;
  out->trsv_order = ctx->trsv_order;
  out->trsm_tiles_per_step = ctx->trsm_kb;
  out->trsm_col_groups = ctx->trsm_cg;
  return DOUG_OK;
}

}  // extern "C"
