// gfx950 kernels that build the coarse space on the device during setup:
// the smoothed restriction R = T S^T (CoarseMtx.F90:152-233, 422-425) and the
// Galerkin product A_c = R (A R^T) (CoarseMtx.F90:476-516 through
// SpMtx_op_AB.f90:39-200).  Every value is summed in the order the
// reference's sequential loops sum it and every output row lists its
// columns in the reference's order, so the results are bit-identical to the
// host build.  Integer bookkeeping (prefix sums, the stable sort that
// transposes R) uses hipcub.
#include <hipcub/hipcub.hpp>

#include "doug_kernels.hpp"

#pragma clang fp contract(off)

namespace doug {

namespace {

int cdiv64(int64_t a, int b) { return (int)((a + b - 1) / b); }

// Row j of R^T: the entries S(j, kk) of row j of A's pattern, grouped by the
// aggregate a of kk and visited in ascending (a, kk, position) order; each
// group is summed in that order.  S_jj = 1 - omega, S_jk = -omega/d_j * a_jk
// for strong off-diagonal entries, 0 for weak ones, with d_j = a_jj + the
// weak entries of row j.  Written to the slot starting at ptr[j].
__global__ void k_rt_rows(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                          const double* __restrict__ val, const uint8_t* __restrict__ strong,
                          const int32_t* __restrict__ num, double omega,
                          int32_t* __restrict__ slot_col, double* __restrict__ slot_val,
                          int32_t* __restrict__ cnt) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int p0 = ptr[j], p1 = ptr[j + 1];
  double diag = 0.0;
  for (int p = p0; p < p1; p++)
    if (col[p] == j || !strong[p]) diag = diag + val[p];
  const double one_minus = 1.0 - omega, negw = -omega;
  int m = 0;
  int64_t prev_key = -1;
  int prev_p = -1;
  double acc = 0.0;
  int acc_a = -1;
  while (true) {
    int64_t best_key = INT64_MAX;
    int bp = -1;
    for (int p = p0; p < p1; p++) {
      const int kk = col[p], a = num[kk] - 1;
      if (a < 0) continue;
      const int64_t key = ((int64_t)a << 32) | (uint32_t)kk;
      const bool after = key > prev_key || (key == prev_key && p > prev_p);
      if (after && (key < best_key || (key == best_key && p < bp))) {
        best_key = key;
        bp = p;
      }
    }
    if (bp < 0) break;
    const double s = col[bp] == j ? one_minus : (strong[bp] ? negw / diag * val[bp] : 0.0);
    const int a = (int)(best_key >> 32);
    if (a == acc_a) {
      acc = acc + 1.0 * s;
    } else {
      if (acc_a >= 0) {
        slot_col[p0 + m] = acc_a;
        slot_val[p0 + m] = acc;
        m++;
      }
      acc_a = a;
      acc = 1.0 * s;
    }
    prev_key = best_key;
    prev_p = bp;
  }
  if (acc_a >= 0) {
    slot_col[p0 + m] = acc_a;
    slot_val[p0 + m] = acc;
    m++;
  }
  cnt[j] = m;
}

// Bound on the entries of row i of X*Y: sum over X's row of Y's row lengths.
__global__ void k_product_bound(int nrows, const int32_t* __restrict__ xptr,
                                const int32_t* __restrict__ xcol, const int32_t* __restrict__ yptr,
                                int64_t* __restrict__ ub) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  int64_t s = 0;
  for (int p = xptr[i]; p < xptr[i + 1]; p++) s += yptr[xcol[p] + 1] - yptr[xcol[p]];
  ub[i] = s;
}

// Row i of C = X*Y in Gustavson order (SpMtx_op_AB.f90:125-176): X's row in
// stored order, for each entry Y's row in stored order; a column is appended
// when first met (value = the product) and accumulated afterwards.  One
// thread per row; the first kFirstTouchLds columns of the row are kept in
// LDS, later ones in the row's global slot, whose capacity is the bound.
constexpr int kFirstTouchLds = 64;
__global__ __launch_bounds__(64) void k_first_touch(
    int nrows, const int32_t* __restrict__ xptr, const int32_t* __restrict__ xcol,
    const double* __restrict__ xval, const int32_t* __restrict__ yptr,
    const int32_t* __restrict__ ycol, const double* __restrict__ yval,
    const int64_t* __restrict__ slot_off, int32_t* __restrict__ slot_col,
    double* __restrict__ slot_val, int32_t* __restrict__ cnt) {
  constexpr int C = kFirstTouchLds;
  __shared__ int32_t kc[C][64];
  __shared__ double kv[C][64];
  const int t = threadIdx.x;
  const int i = blockIdx.x * 64 + t;
  if (i >= nrows) return;
  const int64_t base = slot_off[i];
  int m = 0;
  for (int p = xptr[i]; p < xptr[i + 1]; p++) {
    const int k = xcol[p];
    const double xv = xval[p];
    for (int q = yptr[k]; q < yptr[k + 1]; q++) {
      const int jj = ycol[q];
      const double prod = xv * yval[q];
      const int mc = m < C ? m : C;
      int f = -1;
      for (int e = 0; e < mc; e++)
        if (kc[e][t] == jj) {
          f = e;
          break;
        }
      if (f >= 0) {
        kv[f][t] = kv[f][t] + prod;
        continue;
      }
      for (int e = C; e < m; e++)
        if (slot_col[base + e] == jj) {
          f = e;
          break;
        }
      if (f >= 0) {
        slot_val[base + f] = slot_val[base + f] + prod;
        continue;
      }
      if (m < C) {
        kc[m][t] = jj;
        kv[m][t] = prod;
      } else {
        slot_col[base + m] = jj;
        slot_val[base + m] = prod;
      }
      m++;
    }
  }
  const int mc = m < C ? m : C;
  for (int e = 0; e < mc; e++) {
    slot_col[base + e] = kc[e][t];
    slot_val[base + e] = kv[e][t];
  }
  cnt[i] = m;
}

// The same product with one wavefront per row, for rows that gather many
// terms (A_c = R T'): the row's distinct columns sit one per lane (up to 128,
// two per lane) and each product finds its column with one comparison
// across the lanes.  Products are taken in Gustavson order, one at a time,
// and added by the lane that holds their column, so every sum is the
// sequential one and the columns come out in first-touch order.  A row with
// more than 128 columns raises *overflow; the caller then uses
// k_first_touch.
__device__ __forceinline__ double lane_double(double v, int l) {
  const long long b = __double_as_longlong(v);
  const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), l);
  const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
  return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}

__global__ __launch_bounds__(256) void k_first_touch_wave(
    int nrows, const int32_t* __restrict__ xptr, const int32_t* __restrict__ xcol,
    const double* __restrict__ xval, const int32_t* __restrict__ yptr,
    const int32_t* __restrict__ ycol, const double* __restrict__ yval,
    const int64_t* __restrict__ slot_off, int32_t* __restrict__ slot_col,
    double* __restrict__ slot_val, int32_t* __restrict__ cnt, int* __restrict__ overflow) {
  const int lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (i >= nrows) return;  // the whole wavefront
  int c0 = -1, c1 = -1;
  double a0 = 0.0, a1 = 0.0;
  int m = 0;
  for (int p = xptr[i]; p < xptr[i + 1]; p++) {
    const int k = xcol[p];
    const double xv = xval[p];
    const int q1 = yptr[k + 1];
    for (int qb = yptr[k]; qb < q1; qb += 64) {
      const int q = qb + lane;
      const int jl = q < q1 ? ycol[q] : -1;
      const double pl = q < q1 ? xv * yval[q] : 0.0;
      const int L = q1 - qb < 64 ? q1 - qb : 64;
      for (int u = 0; u < L; u++) {
        const int jj = __builtin_amdgcn_readlane(jl, u);
        const double pr = lane_double(pl, u);
        const unsigned long long h0 = __ballot(c0 == jj);
        if (h0) {
          if (lane == __ffsll((long long)h0) - 1) a0 = a0 + pr;
          continue;
        }
        const unsigned long long h1 = __ballot(c1 == jj);
        if (h1) {
          if (lane == __ffsll((long long)h1) - 1) a1 = a1 + pr;
          continue;
        }
        if (m < 64) {
          if (lane == m) {
            c0 = jj;
            a0 = pr;
          }
        } else if (m < 128) {
          if (lane == m - 64) {
            c1 = jj;
            a1 = pr;
          }
        }
        m++;
      }
    }
  }
  if (m > 128) {
    if (lane == 0) atomicExch(overflow, 1);
    return;
  }
  const int64_t base = slot_off[i];
  if (lane < m) {
    slot_col[base + lane] = c0;
    slot_val[base + lane] = a0;
  }
  if (lane + 64 < m) {
    slot_col[base + 64 + lane] = c1;
    slot_val[base + 64 + lane] = a1;
  }
  if (lane == 0) cnt[i] = m;
}

// Copy each row's cnt[i] entries from its slot to the compacted CSR.
template <typename O>
__global__ void k_compact(int nrows, const O* __restrict__ slot_off, const int32_t* __restrict__ out_ptr,
                          const int32_t* __restrict__ slot_col, const double* __restrict__ slot_val,
                          int32_t* __restrict__ out_col, double* __restrict__ out_val) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  const int64_t s = slot_off[i];
  const int o = out_ptr[i], m = out_ptr[i + 1] - o;
  for (int e = 0; e < m; e++) {
    out_col[o + e] = slot_col[s + e];
    out_val[o + e] = slot_val[s + e];
  }
}

__global__ void k_row_ids(int nrows, const int32_t* __restrict__ ptr, int32_t* __restrict__ rowid) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nrows) return;
  for (int p = ptr[i]; p < ptr[i + 1]; p++) rowid[p] = i;
}

__global__ void k_iota(int64_t n, int32_t* __restrict__ v) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) v[e] = (int32_t)e;
}

// integer histogram (order-independent result)
__global__ void k_key_histogram(int64_t n, const int32_t* __restrict__ keys, int32_t* __restrict__ counts) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < n) atomicAdd(&counts[keys[e]], 1);
}

__global__ void k_gather_entries(int64_t n, const int32_t* __restrict__ order,
                                 const int32_t* __restrict__ rowid, const double* __restrict__ val,
                                 int32_t* __restrict__ out_col, double* __restrict__ out_val) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int32_t q = order[e];
  out_col[e] = rowid[q];
  out_val[e] = val[q];
}

// ---- the coarse space of a DOUG run on several processes (overlap 1) ----
//
// Each process smooths the rows of its own nodes with the filtered diagonal
// summed in its own entry order (columns of nodes it sends first, then the
// others, each ascending) and the values after the scaling round trip; the
// rows of the outside layer it receives (A_ghost) with their diagonal entry
// alone and the values as read.  An aggregate's row of R is built by its
// owner; A_c is the sum, in process order, of each process's part
// R(:, own) (A R^T)(own, :).

// The process layout from the process-major node order `order` (every
// rank's own nodes in its local order, rank after rank, rank r from
// off[r]): proc[v] = v's position in it; sent[v] = v is one of the first
// nsent[owner] own nodes of its owner (the nodes a rank sends lead its local
// order).  `order` is a permutation of [0, n).
__global__ void k_mp_layout(int n, const int32_t* __restrict__ order, const int32_t* __restrict__ owner,
                            const int32_t* __restrict__ off, const int32_t* __restrict__ nsent,
                            int32_t* __restrict__ proc, uint8_t* __restrict__ sent) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n) return;
  const int v = order[e], r = owner[v];
  proc[v] = e;
  sent[v] = (uint8_t)(e - off[r] < nsent[r]);
}

// Row j in its owner's entry order (group: the column is a node its owner
// sends; then ascending column, then stored position), with the values
// after the round trip; the filtered diagonal summed in that order; and the
// value as read of the row's last diagonal entry (the one A_ghost keeps).
__global__ void k_mp_own_rows(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                              const double* __restrict__ val, const double* __restrict__ valin,
                              const uint8_t* __restrict__ strong, const int32_t* __restrict__ owner,
                              const uint8_t* __restrict__ sent, int32_t* __restrict__ ocol,
                              double* __restrict__ oval, double* __restrict__ dfilt,
                              double* __restrict__ dghost) {
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= n) return;
  const int p0 = ptr[j], p1 = ptr[j + 1], oj = owner[j];
  double d = 0.0, dg = 0.0;
  int64_t prev = -1;
  for (int m = p0; m < p1; m++) {
    // next entry in (group, column, position) order
    int64_t best = INT64_MAX;
    int bp = -1;
    for (int p = p0; p < p1; p++) {
      const int c = col[p];
      const int g = (owner[c] == oj && sent[c]) ? 0 : 1;
      const int64_t key = ((int64_t)g << 62) | ((int64_t)c << 31) | (int64_t)(p - p0);
      if (key > prev && key < best) {
        best = key;
        bp = p;
      }
    }
    prev = best;
    ocol[m] = col[bp];
    oval[m] = val[bp];
    if (col[bp] == j || !strong[bp]) d = d + val[bp];
  }
  for (int p = p0; p < p1; p++)
    if (col[p] == j) dg = valin[p];
  dfilt[j] = d;
  dghost[j] = dg;
}

// Row j of R^T, rows taken in process order (node order[t]): the entries
// S(j, kk) grouped by the aggregate a of kk, visited in ascending (a, place
// of kk in its process, position), each group summed in that order; S as
// a's owner has row j (its own row, or A_ghost's).
__global__ void k_mp_rt_rows(int n, const int32_t* __restrict__ order, const int32_t* __restrict__ ptr,
                             const int32_t* __restrict__ col, const double* __restrict__ val,
                             const double* __restrict__ valin, const uint8_t* __restrict__ strong,
                             const int32_t* __restrict__ num, const int32_t* __restrict__ owner,
                             const int32_t* __restrict__ proc, const double* __restrict__ dfilt,
                             const double* __restrict__ dghost, double omega,
                             int32_t* __restrict__ slot_col, double* __restrict__ slot_val,
                             int32_t* __restrict__ cnt) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int j = order[t];
  const int p0 = ptr[j], p1 = ptr[j + 1], oj = owner[j];
  const double one_minus = 1.0 - omega, negw = -omega, dj = dfilt[j], gj = dghost[j];
  int m = 0;
  int64_t prev_key = -1;
  int prev_p = -1;
  double acc = 0.0;
  int acc_a = -1;
  while (true) {
    int64_t best_key = INT64_MAX;
    int bp = -1;
    for (int p = p0; p < p1; p++) {
      const int kk = col[p], a = num[kk] - 1;
      if (a < 0) continue;
      const int64_t key = ((int64_t)a << 32) | (uint32_t)proc[kk];
      const bool after = key > prev_key || (key == prev_key && p > prev_p);
      if (after && (key < best_key || (key == best_key && p < bp))) {
        best_key = key;
        bp = p;
      }
    }
    if (bp < 0) break;
    const int kk = col[bp];
    double s;
    if (owner[kk] == oj)
      s = kk == j ? one_minus : (strong[bp] ? negw / dj * val[bp] : 0.0);
    else
      s = kk == j ? 1.0 - omega / gj * valin[bp] : (strong[bp] ? negw / gj * valin[bp] : 0.0);
    const int a = (int)(best_key >> 32);
    if (a == acc_a) {
      acc = acc + 1.0 * s;
    } else {
      if (acc_a >= 0) {
        slot_col[p0 + m] = acc_a;
        slot_val[p0 + m] = acc;
        m++;
      }
      acc_a = a;
      acc = 1.0 * s;
    }
    prev_key = best_key;
    prev_p = bp;
  }
  if (acc_a >= 0) {
    slot_col[p0 + m] = acc_a;
    slot_val[p0 + m] = acc;
    m++;
  }
  cnt[j] = m;
}

// Entry ids of R^T listed row by row in process order, and their aggregates
// (the keys of the stable sort that gives R with each row in process order).
__global__ void k_mp_proc_entries(int n, const int32_t* __restrict__ order,
                                  const int32_t* __restrict__ rtptr, const int32_t* __restrict__ rtcol,
                                  const int32_t* __restrict__ off, int32_t* __restrict__ ids,
                                  int32_t* __restrict__ keys) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int j = order[t];
  const int q0 = rtptr[j], len = rtptr[j + 1] - q0, o = off[t];
  for (int q = 0; q < len; q++) {
    ids[o + q] = q0 + q;
    keys[o + q] = rtcol[q0 + q];
  }
}

__global__ void k_mp_row_lengths(int n, const int32_t* __restrict__ order,
                                 const int32_t* __restrict__ rtptr, int32_t* __restrict__ len) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const int j = order[t];
  len[t] = rtptr[j + 1] - rtptr[j];
}

// A_c row a, one wavefront: R's row a in process order; the products of one
// process's nodes go to partial sums (Gustavson order, first product
// assigned), added to the row's totals when the next process starts (in
// process order, the first part assigned); the totals are written in
// ascending column.  Up to 128 distinct columns (two per lane), else
// *overflow.
__global__ __launch_bounds__(256) void k_mp_galerkin_wave(
    int nrows, const int32_t* __restrict__ xptr, const int32_t* __restrict__ xcol,
    const double* __restrict__ xval, const int32_t* __restrict__ owner,
    const int32_t* __restrict__ yptr, const int32_t* __restrict__ ycol,
    const double* __restrict__ yval, const int64_t* __restrict__ slot_off,
    int32_t* __restrict__ slot_col, double* __restrict__ slot_val, int32_t* __restrict__ cnt,
    int* __restrict__ overflow) {
  const int lane = threadIdx.x & 63;
  const int a = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (a >= nrows) return;  // the whole wavefront
  int c0 = -1, c1 = -1;
  double p0 = 0.0, p1 = 0.0, s0 = 0.0, s1 = 0.0;
  bool ip0 = false, ip1 = false, it0 = false, it1 = false;
  int m = 0, cur = -1;
  auto fold = [&]() {
    if (ip0) {
      s0 = it0 ? s0 + p0 : p0;
      it0 = true;
      ip0 = false;
    }
    if (ip1) {
      s1 = it1 ? s1 + p1 : p1;
      it1 = true;
      ip1 = false;
    }
  };
  for (int p = xptr[a]; p < xptr[a + 1]; p++) {
    const int i = xcol[p];
    const double xv = xval[p];
    const int ow = owner[i];
    if (ow != cur) {
      fold();
      cur = ow;
    }
    const int q1 = yptr[i + 1];
    for (int qb = yptr[i]; qb < q1; qb += 64) {
      const int q = qb + lane;
      const int jl = q < q1 ? ycol[q] : -1;
      const double pl = q < q1 ? xv * yval[q] : 0.0;
      const int L = q1 - qb < 64 ? q1 - qb : 64;
      for (int u = 0; u < L; u++) {
        const int jj = __builtin_amdgcn_readlane(jl, u);
        const double pr = lane_double(pl, u);
        const unsigned long long h0 = __ballot(c0 == jj);
        if (h0) {
          if (lane == __ffsll((long long)h0) - 1) {
            p0 = ip0 ? p0 + pr : pr;
            ip0 = true;
          }
          continue;
        }
        const unsigned long long h1 = __ballot(c1 == jj);
        if (h1) {
          if (lane == __ffsll((long long)h1) - 1) {
            p1 = ip1 ? p1 + pr : pr;
            ip1 = true;
          }
          continue;
        }
        if (m < 64) {
          if (lane == m) {
            c0 = jj;
            p0 = pr;
            ip0 = true;
          }
        } else if (m < 128) {
          if (lane == m - 64) {
            c1 = jj;
            p1 = pr;
            ip1 = true;
          }
        }
        m++;
      }
    }
  }
  fold();
  if (m > 128) {
    if (lane == 0) atomicExch(overflow, 1);
    return;
  }
  // place of each column among the row's columns (all distinct)
  int r0 = 0, r1 = 0;
  for (int u = 0; u < 64 && u < m; u++) {
    const int cu = __builtin_amdgcn_readlane(c0, u);
    r0 += cu < c0;
    r1 += cu < c1;
  }
  for (int u = 0; u + 64 < m; u++) {
    const int cu = __builtin_amdgcn_readlane(c1, u);
    r0 += cu < c0;
    r1 += cu < c1;
  }
  const int64_t base = slot_off[a];
  if (lane < m) {
    slot_col[base + r0] = c0;
    slot_val[base + r0] = s0;
  }
  if (lane + 64 < m) {
    slot_col[base + r1] = c1;
    slot_val[base + r1] = s1;
  }
  if (lane == 0) cnt[a] = m;
}

}  // namespace

void launch_mp_own_rows(int n, const int32_t* ptr, const int32_t* col, const double* val,
                        const double* valin, const uint8_t* strong, const int32_t* owner,
                        const uint8_t* sent, int32_t* ocol, double* oval, double* dfilt,
                        double* dghost, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mp_own_rows, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, ptr, col, val, valin,
                     strong, owner, sent, ocol, oval, dfilt, dghost);
}

void launch_mp_layout(int n, const int32_t* order, const int32_t* owner, const int32_t* off,
                      const int32_t* nsent, int32_t* proc, uint8_t* sent, hipStream_t st) {
  hipLaunchKernelGGL(k_mp_layout, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, order, owner, off, nsent, proc,
                     sent);
}

void launch_mp_rt_rows(int n, const int32_t* order, const int32_t* ptr, const int32_t* col,
                       const double* val, const double* valin, const uint8_t* strong,
                       const int32_t* num, const int32_t* owner, const int32_t* proc,
                       const double* dfilt, const double* dghost, double omega, int32_t* slot_col,
                       double* slot_val, int32_t* cnt, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mp_rt_rows, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, order, ptr, col, val,
                     valin, strong, num, owner, proc, dfilt, dghost, omega, slot_col, slot_val, cnt);
}

void launch_mp_row_lengths(int n, const int32_t* order, const int32_t* rtptr, int32_t* len,
                           hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mp_row_lengths, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, order, rtptr, len);
}

void launch_mp_proc_entries(int n, const int32_t* order, const int32_t* rtptr, const int32_t* rtcol,
                            const int32_t* off, int32_t* ids, int32_t* keys, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_mp_proc_entries, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, order, rtptr,
                     rtcol, off, ids, keys);
}

void launch_mp_galerkin_wave(int nrows, const int32_t* xptr, const int32_t* xcol, const double* xval,
                             const int32_t* owner, const int32_t* yptr, const int32_t* ycol,
                             const double* yval, const int64_t* slot_off, int32_t* slot_col,
                             double* slot_val, int32_t* cnt, int* overflow, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_mp_galerkin_wave, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, st, nrows,
                     xptr, xcol, xval, owner, yptr, ycol, yval, slot_off, slot_col, slot_val, cnt,
                     overflow);
}

void launch_rt_rows(int n, const int32_t* ptr, const int32_t* col, const double* val,
                    const uint8_t* strong, const int32_t* num, double omega, int32_t* slot_col,
                    double* slot_val, int32_t* cnt, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_rt_rows, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, ptr, col, val, strong,
                     num, omega, slot_col, slot_val, cnt);
}

void launch_product_bound(int nrows, const int32_t* xptr, const int32_t* xcol, const int32_t* yptr,
                          int64_t* ub, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_product_bound, dim3(cdiv64(nrows, 256)), dim3(256), 0, st, nrows, xptr, xcol,
                     yptr, ub);
}

void launch_first_touch(int nrows, const int32_t* xptr, const int32_t* xcol, const double* xval,
                        const int32_t* yptr, const int32_t* ycol, const double* yval,
                        const int64_t* slot_off, int32_t* slot_col, double* slot_val, int32_t* cnt,
                        hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_first_touch, dim3(cdiv64(nrows, 64)), dim3(64), 0, st, nrows, xptr, xcol,
                     xval, yptr, ycol, yval, slot_off, slot_col, slot_val, cnt);
}

void launch_first_touch_wave(int nrows, const int32_t* xptr, const int32_t* xcol,
                             const double* xval, const int32_t* yptr, const int32_t* ycol,
                             const double* yval, const int64_t* slot_off, int32_t* slot_col,
                             double* slot_val, int32_t* cnt, int* overflow, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_first_touch_wave, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, st, nrows,
                     xptr, xcol, xval, yptr, ycol, yval, slot_off, slot_col, slot_val, cnt,
                     overflow);
}

void launch_compact(int nrows, const int32_t* slot_off, const int32_t* out_ptr,
                    const int32_t* slot_col, const double* slot_val, int32_t* out_col,
                    double* out_val, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_compact<int32_t>, dim3(cdiv64(nrows, 256)), dim3(256), 0, st, nrows,
                     slot_off, out_ptr, slot_col, slot_val, out_col, out_val);
}

void launch_compact(int nrows, const int64_t* slot_off, const int32_t* out_ptr,
                    const int32_t* slot_col, const double* slot_val, int32_t* out_col,
                    double* out_val, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_compact<int64_t>, dim3(cdiv64(nrows, 256)), dim3(256), 0, st, nrows,
                     slot_off, out_ptr, slot_col, slot_val, out_col, out_val);
}

void launch_row_ids(int nrows, const int32_t* ptr, int32_t* rowid, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_row_ids, dim3(cdiv64(nrows, 256)), dim3(256), 0, st, nrows, ptr, rowid);
}

void launch_iota(int64_t n, int32_t* v, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_iota, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, v);
}

void launch_key_histogram(int64_t n, const int32_t* keys, int32_t* counts, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_key_histogram, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, keys, counts);
}

void launch_gather_entries(int64_t n, const int32_t* order, const int32_t* rowid, const double* val,
                           int32_t* out_col, double* out_val, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_entries, dim3(cdiv64(n, 256)), dim3(256), 0, st, n, order, rowid, val,
                     out_col, out_val);
}

hipError_t exclusive_scan(void* tmp, size_t& bytes, const int32_t* in, int32_t* out, int n,
                          hipStream_t st) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st);
}

hipError_t exclusive_scan(void* tmp, size_t& bytes, const int64_t* in, int64_t* out, int n,
                          hipStream_t st) {
  return hipcub::DeviceScan::ExclusiveSum(tmp, bytes, in, out, n, st);
}

hipError_t stable_sort_pairs(void* tmp, size_t& bytes, const int32_t* keys_in, int32_t* keys_out,
                             const int32_t* vals_in, int32_t* vals_out, int n, int end_bit,
                             hipStream_t st) {
  return hipcub::DeviceRadixSort::SortPairs(tmp, bytes, keys_in, keys_out, vals_in, vals_out, n, 0,
                                            end_bit, st);
}

}  // namespace doug
