// gfx950 kernels of the two-level Schwarz PCG solve path.
//
// Dense work (subdomain and coarse factorisation) runs on the FP64 matrix
// cores (v_mfma_f64_16x16x4f64) over 64x64 tiles staged in LDS; every
// per-iteration operator is HBM-bound and written as streaming kernels with
// coalesced 512-byte row segments and wavefront reductions.  Kernels whose
// results are compared bit-for-bit with a sequential CPU loop (SpMV,
// restriction, overlap sum, vector updates) keep products and sums as
// separate roundings.
#include "doug_kernels.hpp"
#include <algorithm>

#pragma clang fp contract(off)

namespace doug {

namespace {

constexpr int LD = kTile + 1;  // padded LDS row stride (elements)

// Matrix-core tile product for each element type.  Both forms take A[l%16][l/16]
// and B[l/16][l%16] from lane l.  Result register r of lane l holds
//   f64 16x16x4: D[l/16 + 4r][l%16]   (single-row groups strided by 4)
//   f32 16x16x4: D[4(l/16) + r][l%16] (groups of four consecutive rows)
template <typename T>
struct Mma;
template <>
struct Mma<double> {
  typedef double v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(double a, double b, v4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return (lane >> 4) + 4 * r; }
};
template <>
struct Mma<float> {
  typedef float v4 __attribute__((ext_vector_type(4)));
  static __device__ __forceinline__ v4 op(float a, float b, v4 c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int r) { return 4 * (lane >> 4) + r; }
};

// element offset of tile (i, k) from the block's base; k inside row i's envelope
__device__ __forceinline__ int64_t toff(const BlockDesc& b, int i, int k) {
  const TileRow r = b.tr[i];
  return (int64_t)(r.off + k - r.fc) * kTileElems;
}

// 64x64 row-major global tile -> LDS (stride LD); 256 threads, 16 B loads.
__device__ __forceinline__ void tile_to_lds(double* __restrict__ S, const double* __restrict__ g) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 8; q++) {
    int e = (t + 256 * q) * 2;
    int r = e >> 6, c = e & 63;
    double2 v = *reinterpret_cast<const double2*>(g + e);
    S[r * LD + c] = v.x;
    S[r * LD + c + 1] = v.y;
  }
}
__device__ __forceinline__ void tile_to_lds(float* __restrict__ S, const float* __restrict__ g) {
  const int t = threadIdx.x;
#pragma unroll
  for (int q = 0; q < 4; q++) {
    int e = (t + 256 * q) * 4;
    int r = e >> 6, c = e & 63;
    float4 v = *reinterpret_cast<const float4*>(g + e);
    S[r * LD + c] = v.x;
    S[r * LD + c + 1] = v.y;
    S[r * LD + c + 2] = v.z;
    S[r * LD + c + 3] = v.w;
  }
}

// Wave w of 4 owns the 32x32 quadrant (wr, wc) = (w/2, w%2) of the 64x64
// output as 2x2 blocks of 16x16.
// acc += A * B^T  (both operands row-major tiles in LDS)
template <typename T>
__device__ __forceinline__ void mma_abt(typename Mma<T>::v4 (&acc)[2][2], const T* As, const T* Bs,
                                        int lane, int wr, int wc) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 16; kk++) {
    int k = 4 * kk + lk;
    T a0 = As[(wr * 32 + li) * LD + k];
    T a1 = As[(wr * 32 + 16 + li) * LD + k];
    T b0 = Bs[(wc * 32 + li) * LD + k];
    T b1 = Bs[(wc * 32 + 16 + li) * LD + k];
    acc[0][0] = Mma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mma<T>::op(a1, b1, acc[1][1]);
  }
}

// acc += A * B  (B row-major [k][c])
template <typename T>
__device__ __forceinline__ void mma_ab(typename Mma<T>::v4 (&acc)[2][2], const T* As, const T* Bs,
                                       int lane, int wr, int wc) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll 4
  for (int kk = 0; kk < 16; kk++) {
    int k = 4 * kk + lk;
    T a0 = As[(wr * 32 + li) * LD + k];
    T a1 = As[(wr * 32 + 16 + li) * LD + k];
    T b0 = Bs[k * LD + wc * 32 + li];
    T b1 = Bs[k * LD + wc * 32 + 16 + li];
    acc[0][0] = Mma<T>::op(a0, b0, acc[0][0]);
    acc[0][1] = Mma<T>::op(a0, b1, acc[0][1]);
    acc[1][0] = Mma<T>::op(a1, b0, acc[1][0]);
    acc[1][1] = Mma<T>::op(a1, b1, acc[1][1]);
  }
}

template <typename V>
__device__ __forceinline__ void zero_acc(V (&acc)[2][2]) {
#pragma unroll
  for (int a = 0; a < 2; a++)
#pragma unroll
    for (int b = 0; b < 2; b++) acc[a][b] = V{0, 0, 0, 0};
}

// element (row, col) held by accumulator (bi, bj, r) of this lane
template <typename T>
__device__ __forceinline__ int acc_row(int wr, int bi, int lane, int r) {
  return wr * 32 + bi * 16 + Mma<T>::row(lane, r);
}
__device__ __forceinline__ int acc_col(int wc, int bj, int lane) {
  return wc * 32 + bj * 16 + (lane & 15);
}

// ------------------------------------------------------------------------
// Setup kernels
// ------------------------------------------------------------------------
template <typename T>
__global__ void k_scatter(T* __restrict__ tiles, const int64_t* __restrict__ dst,
                          const int32_t* __restrict__ src, const double* __restrict__ val,
                          int64_t ne) {
  int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e < ne) tiles[dst[e]] = (T)val[src[e]];
}

// Entries that one (i, j) pair of M receives more than once: group g's
// values val[src[seg[g]..seg[g+1])] summed in stored order, in FP64, into
// tiles[dst[g]] (the first occurrence's plain store is overwritten).
template <typename T>
__global__ void k_scatter_sum(T* __restrict__ tiles, const int64_t* __restrict__ dst,
                              const int64_t* __restrict__ seg, const int32_t* __restrict__ src,
                              const double* __restrict__ val, int64_t ngroups) {
  int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= ngroups) return;
  double s = 0.0;
  for (int64_t k = seg[g]; k < seg[g + 1]; k++) s += val[src[k]];
  tiles[dst[g]] = (T)s;
}

template <typename T>
__global__ void k_pad_identity(T* __restrict__ tiles, const BlockDesc* __restrict__ d,
                               int maxpad) {
  const BlockDesc b = d[blockIdx.y];
  int p = blockIdx.x * blockDim.x + threadIdx.x;
  int npad = b.nt * kTile - b.s;
  if (p >= npad || p >= maxpad) return;
  int lr = kTile - npad + p;  // local row inside the last diagonal tile
  tiles[b.base + toff(b, b.nt - 1, b.nt - 1) + lr * kTile + lr] = (T)1;
}

// ---- Cholesky factor and inverse of one 64x64 tile, in 16x16 blocks ----
//
// A 16x16 block inside one wavefront: lane l holds row l%16, columns
// 4(l/16)..4(l/16)+3 in v[0..3].

// Lower Cholesky factor of the block in v, right-looking: with
// s_c = rsqrt(D_cc), L_cc = D_cc s_c, L_rc = D_rc s_c and
// D_rq -= L_rc L_qc for c < q <= r; zeros above the diagonal on return, and
// s_r (= 1 / L_rr) in rinv of every lane of row r.  A non-positive pivot is
// replaced by 1 and reported through *flag when `report` is set.
template <typename T>
__device__ __forceinline__ void potrf16(T (&v)[4], T& rinv, int lane, bool report, int* flag) {
  const int r = lane & 15, g = lane >> 4;
  rinv = (T)0;
#pragma unroll
  for (int c = 0; c < 16; c++) {
    const int cg = c >> 2, cq = c & 3;
    T piv = __shfl(v[cq], c + 16 * cg);
    if (!(piv > (T)0)) {
      if (report && lane == 0) atomicCAS(flag, 0, (int)blockIdx.x + 1);
      piv = (T)1;
    }
    const T sc = rsqrt(piv);
    if (r == c) rinv = sc;
    if (g == cg) v[cq] = r > c ? v[cq] * sc : (r == c ? piv * sc : (T)0);
    const T lrc = __shfl(v[cq], r + 16 * cg);  // L_rc
#pragma unroll
    for (int q = 0; q < 4; q++) {
      const int col = 4 * g + q;
      const T lqc = __shfl(v[cq], col + 16 * cg);  // L_col,c
      if (col > c && col <= r) v[q] = v[q] - lrc * lqc;
    }
  }
#pragma unroll
  for (int q = 0; q < 4; q++)
    if (4 * g + q > r) v[q] = (T)0;
}

// acc += A B^T and acc += A B for 16x16 blocks in LDS (row strides la, lb)
template <typename T>
__device__ __forceinline__ void mma16_abt(typename Mma<T>::v4& acc, const T* a, int la, const T* b, int lb,
                                          int lane) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 4; kk++) acc = Mma<T>::op(a[li * la + 4 * kk + lk], b[li * lb + 4 * kk + lk], acc);
}
template <typename T>
__device__ __forceinline__ void mma16_ab(typename Mma<T>::v4& acc, const T* a, int la, const T* b, int lb,
                                         int lane) {
  const int li = lane & 15, lk = lane >> 4;
#pragma unroll
  for (int kk = 0; kk < 4; kk++) acc = Mma<T>::op(a[li * la + 4 * kk + lk], b[(4 * kk + lk) * lb + li], acc);
}

// orders this wavefront's LDS writes before its later LDS reads of them
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// one wavefront's staging: a 16x16 block (stride 17), then 16 reciprocals
constexpr int kWs = 16 * 17 + 16;

// The symmetric 64x64 tile C in As (lower part read) becomes W = L^-1 in
// Cg (row-major, zeros above the diagonal), C = L L^T.  Blocked by 16: each
// diagonal block is factored by every wavefront in registers (the same
// values in each, so no barrier is needed before the panel); each
// wavefront solves one panel block below it, X L_kk^T = C_Ik, one row per
// lane with L_kk from its own staging copy; the trailing update runs on the
// matrix cores.  Then W: each wavefront inverts one diagonal block, one
// column per lane (forward substitution, no lane exchange), and the
// off-diagonal blocks follow by distance from the diagonal,
// W_IJ = -W_II sum_{J<=K<I} L_IK W_KJ.  As ends with L, Bs with W; Ws holds
// 4 x kWs elements and rv 64 reciprocals of L's diagonal.  256 threads.
template <typename T>
__device__ __forceinline__ void factor_invert_tile(T* As, T* Bs, T* Ws, T* rv, T* __restrict__ Cg, int* flag) {
  typedef typename Mma<T>::v4 V;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int r = lane & 15, g = lane >> 4;
  T* ws = Ws + w * kWs;
  for (int kb = 0; kb < 4; kb++) {
    T l[4], rinv;
#pragma unroll
    for (int q = 0; q < 4; q++) l[q] = As[(16 * kb + r) * LD + 16 * kb + 4 * g + q];
    potrf16(l, rinv, lane, w == 0, flag);
    const int I = kb + 1 + w;  // this wavefront's panel block
    if (I < 4) {
#pragma unroll
      for (int q = 0; q < 4; q++) ws[r * 17 + 4 * g + q] = l[q];
      if (g == 0) ws[16 * 17 + r] = rinv;
      wave_lds_sync();
      if (lane < 16) {  // row 16 I + lane of the panel: X L_kk^T = C, column by column
        T* row = As + (16 * I + lane) * LD + 16 * kb;
        T x[16];
#pragma unroll
        for (int c = 0; c < 16; c++) x[c] = row[c];
#pragma unroll
        for (int c = 0; c < 16; c++) {
          x[c] = x[c] * ws[16 * 17 + c];
#pragma unroll
          for (int c2 = c + 1; c2 < 16; c2++) x[c2] = x[c2] - x[c] * ws[c2 * 17 + c];
        }
#pragma unroll
        for (int c = 0; c < 16; c++) row[c] = x[c];
      }
    }
    __syncthreads();  // every wavefront has read C_kk: wavefront 3 may overwrite it
    if (w == 3) {
#pragma unroll
      for (int q = 0; q < 4; q++) As[(16 * kb + r) * LD + 16 * kb + 4 * g + q] = l[q];
      if (g == 0) rv[16 * kb + r] = rinv;
    }
    // trailing update of the blocks (I, J), kb < J <= I < 4
    const int m = 3 - kb;
    for (int idx = w; idx < m * (m + 1) / 2; idx += 4) {
      int ii = 0;
      while ((ii + 1) * (ii + 2) / 2 <= idx) ii++;
      const int I2 = kb + 1 + ii, J2 = kb + 1 + (idx - ii * (ii + 1) / 2);
      V acc = V{0, 0, 0, 0};
      mma16_abt<T>(acc, As + 16 * I2 * LD + 16 * kb, LD, As + 16 * J2 * LD + 16 * kb, LD, lane);
#pragma unroll
      for (int q = 0; q < 4; q++) {
        T* c = As + (16 * I2 + Mma<T>::row(lane, q)) * LD + 16 * J2 + (lane & 15);
        *c = *c - acc[q];
      }
    }
    __syncthreads();
  }
  // W_ww = L_ww^-1 by columns: lane j solves L_ww x = e_j
  if (lane < 16) {
    const T* L = As + 16 * w * LD + 16 * w;
    T x[16];
#pragma unroll
    for (int i = 0; i < 16; i++) x[i] = i == lane ? (T)1 : (T)0;
#pragma unroll
    for (int i = 0; i < 16; i++) {
      x[i] = x[i] * rv[16 * w + i];
#pragma unroll
      for (int i2 = i + 1; i2 < 16; i2++) x[i2] = x[i2] - L[i2 * LD + i] * x[i];
    }
#pragma unroll
    for (int i = 0; i < 16; i++) Bs[(16 * w + i) * LD + 16 * w + lane] = i >= lane ? x[i] : (T)0;
  }
  __syncthreads();
  // off-diagonal blocks of W by distance from the diagonal
  for (int dd = 1; dd < 4; dd++) {
    const int J = w, I = w + dd;
    if (I < 4) {
      V acc = V{0, 0, 0, 0};
      for (int K = J; K < I; K++) mma16_ab<T>(acc, As + 16 * I * LD + 16 * K, LD, Bs + 16 * K * LD + 16 * J, LD, lane);
#pragma unroll
      for (int q = 0; q < 4; q++) ws[Mma<T>::row(lane, q) * 17 + (lane & 15)] = acc[q];
      wave_lds_sync();
      V acc2 = V{0, 0, 0, 0};
      mma16_ab<T>(acc2, Bs + 16 * I * LD + 16 * I, LD, ws, 17, lane);
#pragma unroll
      for (int q = 0; q < 4; q++) Bs[(16 * I + Mma<T>::row(lane, q)) * LD + 16 * J + (lane & 15)] = -acc2[q];
    }
    __syncthreads();
  }
  // W to global memory, zeros above the diagonal blocks
#pragma unroll
  for (int q = 0; q < 16; q++) {
    const int e = t + 256 * q, row = e >> 6, col = e & 63;
    Cg[e] = (col >> 4) <= (row >> 4) ? Bs[row * LD + col] : (T)0;
  }
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_diag(T* __restrict__ tiles,
                                                   const BlockDesc* __restrict__ d, int j,
                                                   int* __restrict__ flag, int left) {
  __shared__ T As[kTile * LD];
  __shared__ T Bs[kTile * LD];
  __shared__ T Ws[4 * kWs];
  __shared__ T rv[kTile];
  typedef typename Mma<T>::v4 V;
  const BlockDesc b = d[blockIdx.x];
  if (j >= b.nt) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  T* Cg = tiles + b.base + toff(b, j, j);
  V acc[2][2];
  zero_acc(acc);
  for (int k = left ? b.tr[j].fc : j; k < j; k++) {
    tile_to_lds(As, tiles + b.base + toff(b, j, k));
    __syncthreads();
    mma_abt<T>(acc, As, As, lane, wr, wc);
    __syncthreads();
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = acc_row<T>(wr, bi, lane, r), col = acc_col(wc, bj, lane);
        As[row * LD + col] = Cg[row * kTile + col] - acc[bi][bj][r];
      }
  __syncthreads();
  factor_invert_tile<T>(As, Bs, Ws, rv, Cg, flag);
}

template <typename T>
__global__ __launch_bounds__(256) void k_chol_offdiag(T* __restrict__ tiles,
                                                      const BlockDesc* __restrict__ d, int j,
                                                      int left) {
  __shared__ T As[kTile * LD];
  __shared__ T Bs[kTile * LD];
  typedef typename Mma<T>::v4 V;
  const BlockDesc b = d[blockIdx.y];
  if (j >= b.nt) return;
  const int i = j + 1 + blockIdx.x;
  if (i > b.tr[j].lr) return;  // tile (i, j) outside the envelope
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  T* Cg = tiles + b.base + toff(b, i, j);
  V acc[2][2];
  zero_acc(acc);
  for (int k = left ? b.tr[i].fc : j; k < j; k++) {
    tile_to_lds(As, tiles + b.base + toff(b, i, k));
    tile_to_lds(Bs, tiles + b.base + toff(b, j, k));
    __syncthreads();
    mma_abt<T>(acc, As, Bs, lane, wr, wc);
    __syncthreads();
  }
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        int row = acc_row<T>(wr, bi, lane, r), col = acc_col(wc, bj, lane);
        As[row * LD + col] = Cg[row * kTile + col] - acc[bi][bj][r];
      }
  tile_to_lds(Bs, tiles + b.base + toff(b, j, j));  // inv(L_jj)
  __syncthreads();
  zero_acc(acc);
  mma_abt<T>(acc, As, Bs, lane, wr, wc);  // L_ij = C inv(L_jj)^T
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        Cg[acc_row<T>(wr, bi, lane, r) * kTile + acc_col(wc, bj, lane)] = acc[bi][bj][r];
}

// Right-looking trailing update of a dense block (d[0]) after step j:
// A_ik -= L_ij L_kj^T for every j < k <= i < nt, one tile per workgroup.
__device__ __forceinline__ void tri_pair(int idx, int& ii, int& kk) {
  int r = (int)((sqrt(8.0 * idx + 1.0) - 1.0) * 0.5);
  while ((r + 1) * (r + 2) / 2 <= idx) r++;
  while (r * (r + 1) / 2 > idx) r--;
  ii = r;
  kk = idx - r * (r + 1) / 2;
}

__global__ __launch_bounds__(256) void k_rl_update(double* __restrict__ tiles,
                                                   const BlockDesc* __restrict__ d, int j) {
  __shared__ double As[kTile * LD];
  __shared__ double Bs[kTile * LD];
  const BlockDesc b = d[0];
  int ii, kk;
  tri_pair(blockIdx.x, ii, kk);
  const int i = j + 1 + ii, k = j + 1 + kk;
  if (i >= b.nt) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  tile_to_lds(As, tiles + b.base + toff(b, i, j));
  tile_to_lds(Bs, tiles + b.base + toff(b, k, j));
  __syncthreads();
  Mma<double>::v4 acc[2][2];
  zero_acc(acc);
  mma_abt<double>(acc, As, Bs, lane, wr, wc);
  double* Cg = tiles + b.base + toff(b, i, k);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int e = acc_row<double>(wr, bi, lane, r) * kTile + acc_col(wc, bj, lane);
        Cg[e] = Cg[e] - acc[bi][bj][r];
      }
}

// Inverse W = L^{-1} of a dense factor by right-looking substitution on
// L W = I (W starts with zero tiles below the diagonal).  Step m reads only
// block column m of L, final once the factorisation's step m has solved it,
// so step m may run as soon as that step is done.  Step m: W_mm = inv(L_mm)
// (held by the factor's diagonal tile), and block row m is finished,
// W_mc = inv(L_mm) W_mc for c < m (grid: m + 1) ...
__global__ __launch_bounds__(256) void k_inv_finish(const double* __restrict__ L,
                                                    double* __restrict__ W,
                                                    const BlockDesc* __restrict__ d, int m) {
  __shared__ double As[kTile * LD];
  __shared__ double Bs[kTile * LD];
  const BlockDesc b = d[0];
  const int c = blockIdx.x;
  if (c == m) {
    const int64_t o = b.base + toff(b, m, m);
    for (int e = threadIdx.x; e < kTileElems; e += blockDim.x) W[o + e] = L[o + e];
    return;
  }
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  tile_to_lds(As, L + b.base + toff(b, m, m));  // inv(L_mm)
  tile_to_lds(Bs, W + b.base + toff(b, m, c));
  __syncthreads();
  Mma<double>::v4 acc[2][2];
  zero_acc(acc);
  mma_ab<double>(acc, As, Bs, lane, wr, wc);
  double* Wg = W + b.base + toff(b, m, c);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++)
        Wg[acc_row<double>(wr, bi, lane, r) * kTile + acc_col(wc, bj, lane)] = acc[bi][bj][r];
}

// ... then W_ic -= L_im W_mc for every i > m, c <= m (grid: c, i - m - 1).
__global__ __launch_bounds__(256) void k_inv_update(const double* __restrict__ L,
                                                    double* __restrict__ W,
                                                    const BlockDesc* __restrict__ d, int m) {
  __shared__ double As[kTile * LD];
  __shared__ double Bs[kTile * LD];
  const BlockDesc b = d[0];
  const int c = blockIdx.x, i = m + 1 + blockIdx.y;
  if (i >= b.nt) return;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, wr = w >> 1, wc = w & 1;
  tile_to_lds(As, L + b.base + toff(b, i, m));
  tile_to_lds(Bs, W + b.base + toff(b, m, c));
  __syncthreads();
  Mma<double>::v4 acc[2][2];
  zero_acc(acc);
  mma_ab<double>(acc, As, Bs, lane, wr, wc);
  double* Wg = W + b.base + toff(b, i, c);
#pragma unroll
  for (int bi = 0; bi < 2; bi++)
#pragma unroll
    for (int bj = 0; bj < 2; bj++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int e = acc_row<double>(wr, bi, lane, r) * kTile + acc_col(wc, bj, lane);
        Wg[e] = Wg[e] - acc[bi][bj][r];
      }
}

// W's tiles below the diagonal zero (its diagonal tiles come with each step)
__global__ void k_inv_init(double* __restrict__ W, const BlockDesc* __restrict__ d) {
  const BlockDesc b = d[0];
  const int i = blockIdx.y, k = blockIdx.x;
  if (k >= i) return;
  const int64_t o = b.base + toff(b, i, k);
  for (int e = threadIdx.x; e < kTileElems; e += blockDim.x) W[o + e] = 0.0;
}

// ------------------------------------------------------------------------
// Wavefront reductions
// ------------------------------------------------------------------------
// 16 row partial sums spread over the 64 lanes (lane = column) -> lane l
// returns the full sum of row (l >> 2); four reduce-scatter halvings then
// two butterflies.
template <typename T>
__device__ __forceinline__ T reduce_scatter16(T (&v)[16], int lane) {
#pragma unroll
  for (int q = 0; q < 8; q++) {
    bool hi = lane & 32;
    T keep = hi ? v[q + 8] : v[q], send = hi ? v[q] : v[q + 8];
    v[q] = keep + __shfl_xor(send, 32);
  }
#pragma unroll
  for (int q = 0; q < 4; q++) {
    bool hi = lane & 16;
    T keep = hi ? v[q + 4] : v[q], send = hi ? v[q] : v[q + 4];
    v[q] = keep + __shfl_xor(send, 16);
  }
#pragma unroll
  for (int q = 0; q < 2; q++) {
    bool hi = lane & 8;
    T keep = hi ? v[q + 2] : v[q], send = hi ? v[q] : v[q + 2];
    v[q] = keep + __shfl_xor(send, 8);
  }
  {
    bool hi = lane & 4;
    T keep = hi ? v[1] : v[0], send = hi ? v[0] : v[1];
    v[0] = keep + __shfl_xor(send, 4);
  }
  v[0] += __shfl_xor(v[0], 2);
  v[0] += __shfl_xor(v[0], 1);
  return v[0];
}

// A factor tile element, loaded non-temporally: the subdomain solve streams
// every tile once per pass, and the hint keeps the 6 GB stream from
// displacing what the iteration's other kernels read again (A, the vectors).
template <typename T>
__device__ __forceinline__ T ntl(const T* p) {
  return __builtin_nontemporal_load(p);
}

// ------------------------------------------------------------------------
// Per-iteration kernels
// ------------------------------------------------------------------------
// Forward then backward substitution for one subdomain per block, in the
// factor's precision T; the right-hand side is gathered from (and the
// solution written to) FP64 vectors.  Nothing when *skip is set (a PCG
// iteration queued after the device's convergence test passed).
template <typename T>
__global__ __launch_bounds__(256) void k_trsv(const T* __restrict__ tiles,
                                              const BlockDesc* __restrict__ d,
                                              const int32_t* __restrict__ inds,
                                              const double* __restrict__ rhs,
                                              double* __restrict__ sol, const double* skip) {
  if (skip && *skip != 0.0) return;
  // the subdomain vector: nt*64 entries in dynamic LDS sized for the largest block
  extern __shared__ __align__(16) unsigned char trsv_dyn[];
  T* y = reinterpret_cast<T*>(trsv_dyn);
  __shared__ T tv[kTile];
  __shared__ T red[4 * kTile];
  __shared__ int rbase[kMaxTrsvTiles], rfc[kMaxTrsvTiles], rlr[kMaxTrsvTiles];
  const BlockDesc b = d[blockIdx.x];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  const int nt = b.nt, npad = nt * kTile;
  const T* T0 = tiles + b.base + (16 * w) * kTile + lane;
  // the block's envelope table in LDS: tile (i, k) at T0 + (rbase[i] + k) * 4096
  for (int i = t; i < nt; i += 256) {
    const TileRow r = b.tr[i];
    rbase[i] = r.off - r.fc;
    rfc[i] = r.fc;
    rlr[i] = r.lr;
  }
  for (int p = t; p < npad; p += 256) y[p] = p < b.s ? (T)rhs[inds[b.vec_off + p]] : (T)0;
  __syncthreads();
  // forward: y_i = inv(L_ii) (r_i - sum_{k<i} L_ik y_k).  The diagonal tile
  // is fetched with the first tiles of the row and off-diagonal tiles are
  // taken two at a time, so that more loads are in flight per wait.
  for (int i = 0; i < nt; i++) {
    const T* Ri = T0 + (int64_t)rbase[i] * kTileElems;
    T dg[16], acc[16];
#pragma unroll
    for (int q = 0; q < 16; q++) {
      dg[q] = ntl(Ri + (int64_t)i * kTileElems + q * kTile);
      acc[q] = (T)0;
    }
    int k = rfc[i];
    for (; k + 2 <= i; k += 2) {
      const T* A0 = Ri + (int64_t)k * kTileElems;
      T a0[16], a1[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        a0[q] = ntl(A0 + q * kTile);
        a1[q] = ntl(A0 + kTileElems + q * kTile);
      }
      const T y0 = y[k * kTile + lane], y1 = y[(k + 1) * kTile + lane];
#pragma unroll
      for (int q = 0; q < 16; q++) acc[q] += a0[q] * y0 + a1[q] * y1;
    }
    if (k < i) {
      const T* A0 = Ri + (int64_t)k * kTileElems;
      const T y0 = y[k * kTile + lane];
#pragma unroll
      for (int q = 0; q < 16; q++) acc[q] += ntl(A0 + q * kTile) * y0;
    }
    T s = reduce_scatter16(acc, lane);
    if ((lane & 3) == 0) {
      int row = 16 * w + (lane >> 2);
      tv[row] = y[i * kTile + row] - s;
    }
    __syncthreads();
    const T tl = tv[lane];
#pragma unroll
    for (int q = 0; q < 16; q++) acc[q] = dg[q] * tl;
    s = reduce_scatter16(acc, lane);
    if ((lane & 3) == 0) y[i * kTile + 16 * w + (lane >> 2)] = s;
    __syncthreads();
  }
  // backward: x_j = inv(L_jj)^T (y_j - sum_{i>j} L_ij^T x_i)
  for (int j = nt - 1; j >= 0; j--) {
    T dg[16];
#pragma unroll
    for (int q = 0; q < 16; q++) dg[q] = ntl(T0 + ((int64_t)rbase[j] + j) * kTileElems + q * kTile);
    T acc = (T)0;
    const int lr = rlr[j];
    int i = j + 1;
    for (; i + 2 <= lr + 1; i += 2) {
      const T* A0 = T0 + ((int64_t)rbase[i] + j) * kTileElems;
      const T* A1 = T0 + ((int64_t)rbase[i + 1] + j) * kTileElems;
      T a0[16], a1[16];
#pragma unroll
      for (int q = 0; q < 16; q++) {
        a0[q] = ntl(A0 + q * kTile);
        a1[q] = ntl(A1 + q * kTile);
      }
      const T* x0 = y + i * kTile + 16 * w;
#pragma unroll
      for (int q = 0; q < 16; q++) acc += a0[q] * x0[q] + a1[q] * x0[kTile + q];
    }
    if (i <= lr) {
      const T* A0 = T0 + ((int64_t)rbase[i] + j) * kTileElems;
      const T* x0 = y + i * kTile + 16 * w;
#pragma unroll
      for (int q = 0; q < 16; q++) acc += ntl(A0 + q * kTile) * x0[q];
    }
    red[w * kTile + lane] = acc;
    __syncthreads();
    if (w == 0)
      tv[lane] = y[j * kTile + lane] -
                 ((red[lane] + red[kTile + lane]) + (red[2 * kTile + lane] + red[3 * kTile + lane]));
    __syncthreads();
    acc = (T)0;
#pragma unroll
    for (int q = 0; q < 16; q++) acc += dg[q] * tv[16 * w + q];
    red[w * kTile + lane] = acc;
    __syncthreads();
    if (w == 0)
      y[j * kTile + lane] =
          (red[lane] + red[kTile + lane]) + (red[2 * kTile + lane] + red[3 * kTile + lane]);
    __syncthreads();
  }
  for (int p = t; p < b.s; p += 256) sol[b.vec_off + p] = (double)y[p];
}

// u = W c and c' = W^T u for the dense coarse inverse, split into chunks of
// kGemvChunk tiles so that the whole chip streams W; each chunk writes a
// partial 64-vector, summed afterwards chunk by chunk in a fixed order.
// part_i(ch) = sum_{k in chunk ch, k <= i} W_ik c_k   (grid: chunk, tile row)
// Vector loads of W's tiles: VEC entries (16 bytes) per lane, LPR lanes per
// 64-entry tile row, RPI tile rows per wavefront load.
template <typename TW>
struct GemvVec {
  static constexpr int VEC = 16 / (int)sizeof(TW), LPR = kTile / VEC, RPI = 64 / LPR, NI = 16 / RPI;
};
template <typename TW>
struct Vec16Of;
template <>
struct Vec16Of<double> {
  typedef double2 type;
};
template <>
struct Vec16Of<float> {
  typedef float4 type;
};
template <typename V>
__device__ __forceinline__ double vget(const V& v, int e);
template <>
__device__ __forceinline__ double vget<double2>(const double2& v, int e) { return e == 0 ? v.x : v.y; }
template <>
__device__ __forceinline__ double vget<float4>(const float4& v, int e) {
  return (double)(e == 0 ? v.x : e == 1 ? v.y : e == 2 ? v.z : v.w);
}

// part_i(ch) = sum_{k in chunk ch, k <= i} W_ik c_k   (grid: chunk, tile row).
// Wavefront w takes tile rows 16w..16w+15; lane (rs, cg) reads columns
// cg*VEC.. of rows 16w + q*RPI + rs.  Each lane sums its VEC products in
// column order, then its tiles in order; a butterfly over the LPR lanes of a
// row finishes the row.
template <typename TW>
__global__ __launch_bounds__(256) void k_gemv_w_part(const TW* __restrict__ W,
                                                     const BlockDesc* __restrict__ d,
                                                     const double* __restrict__ c,
                                                     const int32_t* __restrict__ perm,
                                                     double* __restrict__ part, int nch) {
  typedef GemvVec<TW> G;
  typedef typename Vec16Of<TW>::type V;
  __shared__ double cs[kGemvChunk * kTile];
  const BlockDesc b = d[0];
  const int ch = blockIdx.x, i = blockIdx.y;
  const int k0 = ch * kGemvChunk;
  if (i >= b.nt || k0 > i) return;
  const int k1 = min(i + 1, k0 + kGemvChunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, cg = lane % G::LPR, rs = lane / G::LPR;
  // the chunk's slice of c, gathered through perm once
  for (int e = t; e < (k1 - k0) * kTile; e += 256) {
    const int ci = k0 * kTile + e;
    cs[e] = c[perm ? perm[ci] : ci];
  }
  __syncthreads();
  double acc[G::NI];
#pragma unroll
  for (int q = 0; q < G::NI; q++) acc[q] = 0.0;
  for (int k = k0; k < k1; k++) {
    const TW* Tb = W + b.base + toff(b, i, k) + (16 * w + rs) * kTile + cg * G::VEC;
    V wv[G::NI];
#pragma unroll
    for (int q = 0; q < G::NI; q++) wv[q] = *reinterpret_cast<const V*>(Tb + q * G::RPI * kTile);
    double cv[G::VEC];
#pragma unroll
    for (int e = 0; e < G::VEC; e++) cv[e] = cs[(k - k0) * kTile + cg * G::VEC + e];
#pragma unroll
    for (int q = 0; q < G::NI; q++) {
      double sdot = 0.0;
#pragma unroll
      for (int e = 0; e < G::VEC; e++) sdot = sdot + vget(wv[q], e) * cv[e];
      acc[q] = acc[q] + sdot;
    }
  }
#pragma unroll
  for (int q = 0; q < G::NI; q++)
#pragma unroll
    for (int o = G::LPR / 2; o > 0; o >>= 1) acc[q] += __shfl_xor(acc[q], o);
  if (cg == 0)
#pragma unroll
    for (int q = 0; q < G::NI; q++)
      part[((int64_t)i * nch + ch) * kTile + 16 * w + q * G::RPI + rs] = acc[q];
}

// part_j(ch) = sum_{i in chunk ch of [j, nt)} W_ij^T u_i   (grid: chunk, tile
// column): lane (rs, cg) accumulates its VEC columns over its rows in order,
// a butterfly over the RPI lanes of a column, then the four wavefronts in order.
template <typename TW>
__global__ __launch_bounds__(256) void k_gemv_wt_part(const TW* __restrict__ W,
                                                      const BlockDesc* __restrict__ d,
                                                      const double* __restrict__ u,
                                                      double* __restrict__ part, int nch) {
  typedef GemvVec<TW> G;
  typedef typename Vec16Of<TW>::type V;
  __shared__ double red[4 * kTile];
  const BlockDesc b = d[0];
  const int ch = blockIdx.x, j = blockIdx.y;
  const int i0 = j + ch * kGemvChunk;
  if (j >= b.nt || i0 >= b.nt) return;
  const int i1 = min(b.nt, i0 + kGemvChunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, cg = lane % G::LPR, rs = lane / G::LPR;
  double acc[G::VEC];
#pragma unroll
  for (int e = 0; e < G::VEC; e++) acc[e] = 0.0;
  for (int i = i0; i < i1; i++) {
    const TW* Tb = W + b.base + toff(b, i, j) + (16 * w + rs) * kTile + cg * G::VEC;
    V wv[G::NI];
#pragma unroll
    for (int q = 0; q < G::NI; q++) wv[q] = *reinterpret_cast<const V*>(Tb + q * G::RPI * kTile);
    const double* ui = u + i * kTile + 16 * w + rs;
#pragma unroll
    for (int q = 0; q < G::NI; q++) {
      const double uq = ui[q * G::RPI];
#pragma unroll
      for (int e = 0; e < G::VEC; e++) acc[e] = acc[e] + vget(wv[q], e) * uq;
    }
  }
#pragma unroll
  for (int e = 0; e < G::VEC; e++)
#pragma unroll
    for (int o = G::LPR; o < 64; o <<= 1) acc[e] += __shfl_xor(acc[e], o);
  if (rs == 0)
#pragma unroll
    for (int e = 0; e < G::VEC; e++) red[w * kTile + cg * G::VEC + e] = acc[e];
  __syncthreads();
  if (w == 0)
    part[((int64_t)j * nch + ch) * kTile + lane] =
        (red[lane] + red[kTile + lane]) + (red[2 * kTile + lane] + red[3 * kTile + lane]);
}

// out_r(tile i) = sum over the tile's chunks in order; rows 0 .. nt*64-1.
// transposed: chunks of tile j cover [j, nt), otherwise [0, j].
__global__ void k_gemv_sum(const double* __restrict__ part, int nt, int nch, int transposed,
                           double* __restrict__ out) {
  const int e = blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= nt * kTile) return;
  const int i = e / kTile, r = e - i * kTile;
  const int len = transposed ? nt - i : i + 1;
  const int nc = (len + kGemvChunk - 1) / kGemvChunk;
  double s = 0.0;
  for (int ch = 0; ch < nc; ch++) s = s + part[((int64_t)i * nch + ch) * kTile + r];
  out[e] = s;
}

// CSR-stream SpMV: the block's nonzeros are multiplied with coalesced loads
// into LDS, then each row is summed in stored order by one thread.
// Row blocks are dispatched round-robin over the XCDs (block b runs on XCD
// b % 8).  With SWZ the first 8*(nb/8) block ids are renumbered so that XCD k
// takes the k-th contiguous eighth of the rows: a row's stencil neighbours
// then mostly sit in the same XCD's L2.  Bijective on [0, nb).
__device__ __forceinline__ int xcd_swizzle(int b, int nb) {
  const int q = nb >> 3;
  return b < (q << 3) ? (b & 7) * q + (b >> 3) : b;
}

// CSR-stream SpMV: the block's products val*x[col] are formed with coalesced
// loads into LDS (8 per thread), then each thread sums one row in entry
// order, so every shape gives the same bits.  TH = rows per block.
template <int TH, bool SWZ>
__global__ __launch_bounds__(TH) void k_spmv(const int32_t* __restrict__ rowblk,
                                             const int32_t* __restrict__ ptr,
                                             const int32_t* __restrict__ col,
                                             const double* __restrict__ val,
                                             const double* __restrict__ x,
                                             double* __restrict__ y, int nb) {
  constexpr int CH = 8 * TH;
  __shared__ double prod[CH];
  const int b = SWZ ? xcd_swizzle(blockIdx.x, nb) : (int)blockIdx.x;
  const int r0 = rowblk[b], r1 = rowblk[b + 1];
  const int e0 = ptr[r0], e1 = ptr[r1];
  const int t = threadIdx.x;
  if (e1 - e0 <= CH) {
    for (int e = e0 + t; e < e1; e += TH) prod[e - e0] = val[e] * x[col[e]];
    __syncthreads();
    const int row = r0 + t;
    if (row < r1) {
      // in stored order; unrolled so that the LDS reads of long rows (the
      // restriction's) are issued ahead of the dependent additions
      double acc = 0.0;
      const int eb = ptr[row] - e0, ee = ptr[row + 1] - e0;
      int e = eb;
      for (; e + 8 <= ee; e += 8) {
        double v[8];
#pragma unroll
        for (int q = 0; q < 8; q++) v[q] = prod[e + q];
#pragma unroll
        for (int q = 0; q < 8; q++) acc = acc + v[q];
      }
      for (; e < ee; e++) acc = acc + prod[e];
      y[row] = acc;
    }
  } else {
    // a single long row, streamed through LDS chunk by chunk
    double acc = 0.0;
    for (int c0 = e0; c0 < e1; c0 += CH) {
      int c1 = min(e1, c0 + CH);
      for (int e = c0 + t; e < c1; e += TH) prod[e - c0] = val[e] * x[col[e]];
      __syncthreads();
      if (t == 0)
        for (int e = c0; e < c1; e++) acc = acc + prod[e - c0];
      __syncthreads();
    }
    if (t == 0) y[r0] = acc;
  }
}

__global__ void k_overlap_sum(int n, const int32_t* __restrict__ occ_ptr,
                              const int32_t* __restrict__ occ_pos, const double* __restrict__ sol,
                              const int32_t* __restrict__ rt_ptr, const int32_t* __restrict__ rt_col,
                              const double* __restrict__ rt_val, const double* __restrict__ csol,
                              double* __restrict__ z) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double acc = 0.0;
  if (occ_ptr)  // no fine sum: the coarse correction alone
    for (int p = occ_ptr[i]; p < occ_ptr[i + 1]; p++) acc = acc + sol[occ_pos[p]];
  if (rt_ptr) {
    double tmp = 0.0;
    for (int p = rt_ptr[i]; p < rt_ptr[i + 1]; p++) tmp = tmp + rt_val[p] * csol[rt_col[p]];
    acc = acc + tmp;
  }
  z[i] = acc;
}

// ------------------------------------------------------------------------
// Deterministic reductions: a fixed tree per workgroup, then the partials
// summed in a fixed order (k_reduce_finish, or the last workgroup of the
// several-column kernels) before the scalar recurrence.
// ------------------------------------------------------------------------
enum { S_RHO = 0, S_RHO_PREV = 1, S_BETA = 2, S_PQ = 3, S_ALPHA = 4, S_RES = 5 };
// the single-vector PCG's convergence test on the device (see k_reduce_finish)
enum { S_DONE = 8, S_ITER = 9, S_THR = 10, S_INIT = 11, S_MAXIT = 12, S_HCAP = 13 };

__device__ __forceinline__ double block_sum(double v, double* sh) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (t == 0) r = (sh[0] + sh[1]) + (sh[2] + sh[3]);
  return r;
}


// The second stage of the single-vector reductions: one workgroup sums the
// `count` workgroup partials (thread t takes t, t + 1024, ... in order, each
// wavefront's by a butterfly, the 16 wavefronts' in order) and applies the
// scalar recurrence of `mode`: 0 the sum to out,
// 1 rho = r.z with beta, 2 p.q with alpha, 3 the squared residual norm.
// Mode 3 with out (a PCG iteration's residual) also keeps the iteration on
// the device: unless S_DONE is set, it counts the iteration in S_ITER,
// writes (alpha, beta, residual) to out[3(it-1)..] for it <= S_HCAP, and
// sets S_DONE once !(res / S_INIT > S_THR) or it >= S_MAXIT, the host's
// test; after that the iteration's other kernels leave x and r alone.
__global__ __launch_bounds__(1024) void k_reduce_finish(const double* __restrict__ partial, int count,
                                                        double* out, int mode, double* scal) {
  __shared__ double sh[16];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  double v = 0.0;
  for (int q = t; q < count; q += 1024) v += partial[q];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o);
  if (lane == 0) sh[w] = v;
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int k = 0; k < 16; k++) s += sh[k];
    if (mode == 0) {
      out[0] = s;
    } else if (mode == 1) {  // rho = r.z
      scal[S_BETA] = s / scal[S_RHO];
      scal[S_RHO_PREV] = scal[S_RHO];
      scal[S_RHO] = s;
    } else if (mode == 2) {  // p.q
      scal[S_PQ] = s;
      scal[S_ALPHA] = scal[S_RHO] / s;
    } else if (!out) {
      scal[S_RES] = s;
    } else if (scal[S_DONE] == 0.0) {
      scal[S_RES] = s;
      const int it = (int)scal[S_ITER] + 1;
      scal[S_ITER] = it;
      if (it <= (int)scal[S_HCAP]) {
        out[3 * (it - 1)] = scal[S_ALPHA];
        out[3 * (it - 1) + 1] = scal[S_BETA];
        out[3 * (it - 1) + 2] = s;
      }
      if (!(s / scal[S_INIT] > scal[S_THR]) || it >= (int)scal[S_MAXIT]) scal[S_DONE] = 1.0;
    }
  }
}

// PCG vector kernels over chunks of kVecChunk consecutive entries, one
// chunk per 256-thread workgroup, each thread taking entries t, t+256, ...
// of its chunk (every load coalesced, four loads per array in flight per
// thread).  The reductions sum each thread's entries in ascending order,
// the workgroup by the fixed tree of block_sum, and k_reduce_finish the
// workgroups' partials: the bits depend on n alone.
constexpr int kVecPer = kVecChunk / 256;

__global__ __launch_bounds__(256) void k_dot_chunks(const double* __restrict__ a,
                                                    const double* __restrict__ b, int n,
                                                    double* partial) {
  __shared__ double sh[4];
  const int i0 = blockIdx.x * kVecChunk + threadIdx.x;
  double av[kVecPer], bv[kVecPer];
#pragma unroll
  for (int k = 0; k < kVecPer; k++) {
    const int i = i0 + 256 * k;
    av[k] = i < n ? a[i] : 0.0;
    bv[k] = i < n ? b[i] : 0.0;
  }
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < kVecPer; k++) v = v + av[k] * bv[k];
  double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ __launch_bounds__(256) void k_p_update_chunks(int n, const double* __restrict__ z,
                                                         double* __restrict__ p,
                                                         const double* __restrict__ scal, int first) {
  const int i0 = blockIdx.x * kVecChunk + threadIdx.x;
  const double beta = first ? 0.0 : scal[S_BETA];
  double zv[kVecPer], pv[kVecPer];
#pragma unroll
  for (int k = 0; k < kVecPer; k++) {
    const int i = i0 + 256 * k;
    zv[k] = i < n ? z[i] : 0.0;
    pv[k] = (i < n && !first) ? p[i] : 0.0;
  }
#pragma unroll
  for (int k = 0; k < kVecPer; k++) {
    const int i = i0 + 256 * k;
    if (i < n) p[i] = first ? zv[k] : zv[k] + beta * pv[k];
  }
}

// x += alpha p, r -= alpha q and r.r partials; nothing once the device's
// convergence test has passed (S_DONE).
__global__ __launch_bounds__(256) void k_xr_update_chunks(int n, double* __restrict__ x,
                                                          double* __restrict__ r,
                                                          const double* __restrict__ p,
                                                          const double* __restrict__ q,
                                                          const double* __restrict__ scal,
                                                          double* partial) {
  __shared__ double sh[4];
  if (scal[S_DONE] != 0.0) return;
  const double alpha = scal[S_ALPHA];
  const int i0 = blockIdx.x * kVecChunk + threadIdx.x;
  double xv[kVecPer], rv[kVecPer], pv[kVecPer], qv[kVecPer];
#pragma unroll
  for (int k = 0; k < kVecPer; k++) {
    const int i = i0 + 256 * k;
    const bool in = i < n;
    xv[k] = in ? x[i] : 0.0;
    rv[k] = in ? r[i] : 0.0;
    pv[k] = in ? p[i] : 0.0;
    qv[k] = in ? q[i] : 0.0;
  }
  double v = 0.0;
#pragma unroll
  for (int k = 0; k < kVecPer; k++) {
    const int i = i0 + 256 * k;
    const double ri = rv[k] - alpha * qv[k];
    if (i < n) {
      x[i] = xv[k] + alpha * pv[k];
      r[i] = ri;
    }
    v = v + ri * ri;
  }
  double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// k_overlap_sum with the preconditioned residual's dot r.z (PCG's rho and
// beta) folded in: one row per thread, the workgroup's products summed by
// block_sum into one partial per workgroup for k_reduce_finish.
__global__ __launch_bounds__(256) void k_overlap_sum_dot(
    int n, const int32_t* __restrict__ occ_ptr, const int32_t* __restrict__ occ_pos,
    const double* __restrict__ sol, const int32_t* __restrict__ rt_ptr,
    const int32_t* __restrict__ rt_col, const double* __restrict__ rt_val,
    const double* __restrict__ csol, double* __restrict__ z, const double* __restrict__ r,
    double* partial) {
  __shared__ double sh[4];
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  double v = 0.0;
  if (i < n) {
    double acc = 0.0;
    if (occ_ptr)
      for (int p = occ_ptr[i]; p < occ_ptr[i + 1]; p++) acc = acc + sol[occ_pos[p]];
    if (rt_ptr) {
      double tmp = 0.0;
      for (int p = rt_ptr[i]; p < rt_ptr[i + 1]; p++) tmp = tmp + rt_val[p] * csol[rt_col[p]];
      acc = acc + tmp;
    }
    z[i] = acc;
    v = r[i] * acc;
  }
  double s = block_sum(v, sh);
  if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

// SpMtx_scale / SpMtx_find_strong / SpMtx_unscale on the device, one row
// per thread, the same operations in the same order as the host's
// scale_round_trip and find_strong: sv_i = sqrt|a_ii| (symmetric scaling) or
// |a_ii| (the last diagonal entry of the row), a~_ij = (a_ij / sv_i) / sv_j,
// the round-trip value (a~_ij * sv_i) * sv_j, and the strength flag
// |a~_ij| > alpha * max_{k != i} |a~_ik| (the diagonal entry: a_ii > that bound).
__global__ void k_row_scale(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                            const double* __restrict__ vin, int symm, double* __restrict__ diag,
                            double* __restrict__ sv) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double d = 0.0;
  for (int p = ptr[i]; p < ptr[i + 1]; p++)
    if (col[p] == i) d = vin[p];
  diag[i] = d;
  sv[i] = symm ? sqrt(fabs(d)) : fabs(d);
}

__global__ void k_round_trip_strong(int n, const int32_t* __restrict__ ptr,
                                    const int32_t* __restrict__ col, const double* __restrict__ vin,
                                    const double* __restrict__ diag, const double* __restrict__ sv,
                                    int symm, double alpha, double* __restrict__ val,
                                    uint8_t* __restrict__ strong) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double si = sv[i];
  double maxnd = (double)(-1.0e15f);
  for (int p = ptr[i]; p < ptr[i + 1]; p++) {
    const int c = col[p];
    const double v = vin[p] / si;
    const double s = symm ? v / sv[c] : v;
    const double a = fabs(s);
    if (c != i && maxnd < a) maxnd = a;
  }
  maxnd = maxnd * alpha;
  for (int p = ptr[i]; p < ptr[i + 1]; p++) {
    const int c = col[p];
    const double sj = sv[c];
    double v = vin[p] / si;
    const double s = symm ? v / sj : v;
    v = s * si;
    val[p] = symm ? v * sj : v;
    strong[p] = c != i ? (fabs(s) > maxnd) : (diag[i] > maxnd);
  }
}


__global__ void k_to_float(const double* __restrict__ src, float* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = (float)src[i];
}

// y = A x for matrices with long rows (the restriction: hundreds of entries
// per row): one wavefront per row, lane l summing entries l, l+64, ... in
// order, then a fixed butterfly over the lanes.
__global__ __launch_bounds__(256) void k_spmv_wave_rows(int nrows, const int32_t* __restrict__ ptr,
                                                        const int32_t* __restrict__ col,
                                                        const double* __restrict__ val,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ y) {
  const int row = blockIdx.x * 4 + (threadIdx.x >> 6), lane = threadIdx.x & 63;
  if (row >= nrows) return;
  double acc = 0.0;
  int e = ptr[row] + lane;
  const int end = ptr[row + 1];
  // four strides loaded ahead, added in the same order as one at a time
  for (; e + 192 < end; e += 256) {
    const int c0 = col[e], c1 = col[e + 64], c2 = col[e + 128], c3 = col[e + 192];
    const double v0 = val[e], v1 = val[e + 64], v2 = val[e + 128], v3 = val[e + 192];
    const double x0 = x[c0], x1 = x[c1], x2 = x[c2], x3 = x[c3];
    acc = acc + v0 * x0;
    acc = acc + v1 * x1;
    acc = acc + v2 * x2;
    acc = acc + v3 * x3;
  }
  for (; e < end; e += 64) acc = acc + val[e] * x[col[e]];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o);
  if (lane == 0) y[row] = acc;
}

// dst(i, c) = src(i, c) for c < ncols and 0 for ncols <= c < ldd: a matrix
// of n rows between leading dimensions lds and ldd
__global__ void k_repack(double* __restrict__ dst, int ldd, const double* __restrict__ src, int lds,
                         int64_t n, int ncols) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * ldd) return;
  const int64_t i = e / ldd;
  const int c = (int)(e - i * ldd);
  dst[e] = c < ncols ? src[i * lds + c] : 0.0;
}

__global__ void k_copy(double* __restrict__ dst, const double* __restrict__ src, int n) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

__global__ void k_zero(double* __restrict__ dst, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = 0.0;
}

__global__ void k_residual(int n, const double* __restrict__ b, const double* __restrict__ ax,
                           double* __restrict__ r) {
  int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) r[i] = b[i] - ax[i];
}


// ------------------------------------------------------------------------
// Several right-hand sides: explicit inverse factors and batched GEMMs
// ------------------------------------------------------------------------
// 64 x NR block of a workspace -> LDS (stride NR+1)
template <typename T, int NR>
__device__ __forceinline__ void panel_to_lds(T* __restrict__ S, const T* __restrict__ g) {
  for (int e = threadIdx.x; e < kTile * NR; e += 256) S[(e / NR) * (NR + 1) + (e % NR)] = g[e];
}

// Register staging of 64x64 tiles and 64xNR panels: every thread loads its
// share with 16-byte loads, so that all loads of a step are in flight before
// the first LDS store.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float4 type;
  static constexpr int n = 4;
};
template <>
struct Vec16<double> {
  typedef double2 type;
  static constexpr int n = 2;
};

// Factor tiles are streamed once per pass: their loads are non-temporal, so
// that they do not evict the solved panels, which the next rows read again.
template <typename T, int NT = 256>
struct TileRegs {
  static constexpr int V = Vec16<T>::n, N = kTileElems / (NT * V);
  typename Vec16<T>::type r[N];
  typedef T native __attribute__((ext_vector_type(V)));
  __device__ __forceinline__ void load(const T* __restrict__ g) {
#pragma unroll
    for (int q = 0; q < N; q++) {
      const native v = __builtin_nontemporal_load(reinterpret_cast<const native*>(g) + threadIdx.x + NT * q);
      __builtin_memcpy(&r[q], &v, sizeof(v));
    }
  }
  __device__ __forceinline__ void store(T* __restrict__ S) const {
#pragma unroll
    for (int q = 0; q < N; q++) {
      const int e = (threadIdx.x + NT * q) * V, row = e >> 6, col = e & 63;
      const T* v = reinterpret_cast<const T*>(&r[q]);
#pragma unroll
      for (int c = 0; c < V; c++) S[row * LD + col + c] = v[c];
    }
  }
};

template <typename T, int NR, int NT = 256>
struct PanelRegs {
  static constexpr int V = Vec16<T>::n, N = kTile * NR / (NT * V);
  static_assert(N >= 1, "panel smaller than one load per thread");
  typename Vec16<T>::type r[N];
  __device__ __forceinline__ void load(const T* __restrict__ g) {
#pragma unroll
    for (int q = 0; q < N; q++)
      r[q] = reinterpret_cast<const typename Vec16<T>::type*>(g)[threadIdx.x + NT * q];
  }
  __device__ __forceinline__ void store(T* __restrict__ S) const {
#pragma unroll
    for (int q = 0; q < N; q++) {
      const int e = (threadIdx.x + NT * q) * V, row = e / NR, col = e % NR;
      const T* v = reinterpret_cast<const T*>(&r[q]);
#pragma unroll
      for (int c = 0; c < V; c++) S[row * (NR + 1) + col + c] = v[c];
    }
  }
};

// acc[c] += A(16w.., :) * B(:, cb + 16c..) from LDS, for the CPW columns
// from cb of an NR-column panel; TR: use A^T
template <typename T, int NR, bool TR, int CPW = NR>
__device__ __forceinline__ void panel_mma(typename Mma<T>::v4 (&acc)[CPW / 16], const T* As,
                                          const T* Bs, int w, int li, int lk, int cb = 0) {
#pragma unroll 4
  for (int ks = 0; ks < 16; ks++) {
    const int kk = 4 * ks + lk;
    const T a = TR ? As[kk * LD + 16 * w + li] : As[(16 * w + li) * LD + kk];
#pragma unroll
    for (int c = 0; c < CPW / 16; c++) acc[c] = Mma<T>::op(a, Bs[kk * (NR + 1) + cb + c * 16 + li], acc[c]);
  }
}

// Dense coarse inverse (block d[0]) times NR columns, in chunks of
// kGemvChunk tiles per workgroup so that the whole chip streams W; each chunk
// writes a partial 64 x NR panel, summed afterwards in a fixed order.
// Wave w owns rows 16w..16w+15 and all NR columns.
// forward, part (ch, i) = sum_{k in chunk ch, k <= i} W_ik G_k
template <int NR>
__global__ __launch_bounds__(256) void k_wgemm_fwd_part(const double* __restrict__ W,
                                                        const BlockDesc* __restrict__ d,
                                                        const double* __restrict__ G,
                                                        const int32_t* __restrict__ perm,
                                                        double* __restrict__ part, int nch) {
  __shared__ double As[kTile * LD];
  __shared__ double Bs[kTile * (NR + 1)];
  typedef Mma<double>::v4 V;
  const BlockDesc b = d[0];
  const int ch = blockIdx.x, i = blockIdx.y;
  const int k0 = ch * kGemvChunk;
  if (i >= b.nt || k0 > i) return;
  const int k1 = min(i + 1, k0 + kGemvChunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, lk = lane >> 4;
  V acc[NR / 16];
#pragma unroll
  for (int c = 0; c < NR / 16; c++) acc[c] = V{0, 0, 0, 0};
  for (int k = k0; k < k1; k++) {
    tile_to_lds(As, W + b.base + toff(b, i, k));
    if (perm)  // rows of G in the factor's order
      for (int e = threadIdx.x; e < kTile * NR; e += 256)
        Bs[(e / NR) * (NR + 1) + (e % NR)] = G[(int64_t)perm[k * kTile + e / NR] * NR + (e % NR)];
    else
      panel_to_lds<double, NR>(Bs, G + (int64_t)k * kTile * NR);
    __syncthreads();
    panel_mma<double, NR, false>(acc, As, Bs, w, li, lk);
    __syncthreads();
  }
  double* P = part + ((int64_t)i * nch + ch) * kTile * NR;
#pragma unroll
  for (int c = 0; c < NR / 16; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) P[(16 * w + Mma<double>::row(lane, r)) * NR + c * 16 + li] = acc[c][r];
}

// backward, part (ch, j) = sum_{i in chunk ch of [j, nt)} W_ij^T U_i
template <int NR>
__global__ __launch_bounds__(256) void k_wgemm_bwd_part(const double* __restrict__ W,
                                                        const BlockDesc* __restrict__ d,
                                                        const double* __restrict__ U,
                                                        double* __restrict__ part, int nch) {
  __shared__ double As[kTile * LD];
  __shared__ double Bs[kTile * (NR + 1)];
  typedef Mma<double>::v4 V;
  const BlockDesc b = d[0];
  const int ch = blockIdx.x, j = blockIdx.y;
  const int i0 = j + ch * kGemvChunk;
  if (j >= b.nt || i0 >= b.nt) return;
  const int i1 = min(b.nt, i0 + kGemvChunk);
  const int t = threadIdx.x, lane = t & 63, w = t >> 6, li = lane & 15, lk = lane >> 4;
  V acc[NR / 16];
#pragma unroll
  for (int c = 0; c < NR / 16; c++) acc[c] = V{0, 0, 0, 0};
  for (int i = i0; i < i1; i++) {
    tile_to_lds(As, W + b.base + toff(b, i, j));
    panel_to_lds<double, NR>(Bs, U + (int64_t)i * kTile * NR);
    __syncthreads();
    panel_mma<double, NR, true>(acc, As, Bs, w, li, lk);
    __syncthreads();
  }
  double* P = part + ((int64_t)j * nch + ch) * kTile * NR;
#pragma unroll
  for (int c = 0; c < NR / 16; c++)
#pragma unroll
    for (int r = 0; r < 4; r++) P[(16 * w + Mma<double>::row(lane, r)) * NR + c * 16 + li] = acc[c][r];
}

// out(p, c) = sum over the chunks of p's tile in order, for p < rows, c < ncols
// (leading dimension ldo); transposed: chunks of tile j cover [j, nt)
template <int NR>
__global__ void k_panel_sum(const double* __restrict__ part, int nt, int nch, int transposed,
                            double* __restrict__ out, int ldo, int rows, int ncols) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)nt * kTile * NR) return;
  const int p = (int)(e / NR), c = (int)(e - (int64_t)p * NR);
  if (p >= rows || c >= ncols) return;
  const int i = p / kTile, rr = p - i * kTile;
  const int len = transposed ? nt - i : i + 1;
  const int nc = (len + kGemvChunk - 1) / kGemvChunk;
  double s = 0.0;
  for (int ch = 0; ch < nc; ch++) s = s + part[(((int64_t)i * nch + ch) * kTile + rr) * NR + c];
  out[(int64_t)p * ldo + c] = s;
}

// A block's tile-row table packed into LDS, one word per tile row: off in
// the high 16 bits (a block has at most 256 * 257 / 2 tiles), fc and lr in
// a byte each (at most 256 tile rows).  Reading it from LDS keeps the
// solve's address arithmetic off the vector-memory counter, so that the
// waits for a step's tiles do not also wait for later loads.
struct PackedRows {
  const uint32_t* w;
  __device__ __forceinline__ int fc(int i) const { return __builtin_amdgcn_readfirstlane((int)((w[i] >> 8) & 255u)); }
  __device__ __forceinline__ int lr(int i) const { return __builtin_amdgcn_readfirstlane((int)(w[i] & 255u)); }
  // element offset of tile (i, k) from the block's base
  __device__ __forceinline__ int64_t toff(int i, int k) const {
    const uint32_t v = w[i];
    return (int64_t)(__builtin_amdgcn_readfirstlane((int)(v >> 16)) + k -
                     __builtin_amdgcn_readfirstlane((int)((v >> 8) & 255u))) *
           kTileElems;
  }
};

__device__ __forceinline__ void pack_rows(uint32_t* w, const BlockDesc& b, int nthreads) {
  for (int i = threadIdx.x; i < b.nt; i += nthreads) {
    const TileRow r = b.tr[i];
    w[i] = ((uint32_t)r.off << 16) | ((uint32_t)r.fc << 8) | (uint32_t)r.lr;
  }
}

// The tiles a tile row's solve starts with: its diagonal tile and the
// off-diagonal tiles of its first step (forward: row i, tiles (i, fc_i..);
// backward, TR: column j, tiles (j+1.., j)).
template <typename T, int KB, bool TR, int NT>
__device__ __forceinline__ void load_row_tiles(const T* Lb, const PackedRows& tr, int i, TileRegs<T, NT>& rd,
                                               TileRegs<T, NT> (&ra)[KB]) {
  rd.load(Lb + tr.toff(i, i));
  if (!TR) {
    const int fc = tr.fc(i);
#pragma unroll
    for (int u = 0; u < KB; u++)
      if (fc + u < i) ra[u].load(Lb + tr.toff(i, fc + u));
  } else {
    const int lr = tr.lr(i);
#pragma unroll
    for (int u = 0; u < KB; u++)
      if (i + 1 + u <= lr) ra[u].load(Lb + tr.toff(i + 1 + u, i));
  }
}

// Envelope forward/backward substitution for NR columns, one workgroup per
// block.  Wave w owns rows 16w..16w+15 of each 64 x NR panel; panels of
// finished tile rows live in U (global, re-read through L2); the workgroup's
// barriers order its own global writes before its later reads.  Each step
// stages KB off-diagonal tiles with their panels (a tile row's diagonal
// tile with its first step); the next step's loads are issued into
// registers before this step's products, so that their latency overlaps
// the matrix-core work.  The last step of a tile row loads the next row's
// first tiles in the same way, and a row's right-hand sides (forward) or
// Y_j (backward) are read at its start, ahead of the steps.
template <typename T, int NR, int KB, int CG, int WG = 4>
__global__ __launch_bounds__(64 * WG) void k_trsm_multi(const T* __restrict__ L,
                                                    const BlockDesc* __restrict__ d, int nblocks,
                                                    const int32_t* __restrict__ inds,
                                                    const double* __restrict__ rhs, int ldr,
                                                    T* __restrict__ U, double* __restrict__ out,
                                                    int ldo, int nrhs) {
  __shared__ T As[KB][kTile * LD];
  __shared__ T Ds[kTile * LD];
  __shared__ T Bs[KB][kTile * (NR + 1)];
  __shared__ uint32_t trw[kMaxTrsvTiles];
  typedef typename Mma<T>::v4 V;
  // CG column groups of NR columns each per subdomain; with two, the groups
  // of one subdomain are workgroups bid and bid + 8, so that they run on the
  // same XCD (workgroups go to the XCDs round-robin) and share its L2 copy of
  // the factor tiles
  int sub = blockIdx.x, grp = 0;
  if (CG == 2) {
    sub = (int)(blockIdx.x / 16) * 8 + (int)(blockIdx.x % 8);
    grp = (int)(blockIdx.x / 8) & 1;
    if (sub >= nblocks) return;
  }
  const BlockDesc b = d[sub];
  pack_rows(trw, b, 64 * WG);
  __syncthreads();
  const PackedRows tr{trw};
  // WG = 8: wavefronts w and w + 4 take the same rows, columns 0..NR/2-1 and
  // NR/2..NR-1, so that one staged tile serves the whole panel
  constexpr int NT = 64 * WG, CPW = NR * 4 / WG;
  static_assert(WG == 4 || (WG == 8 && CG == 1 && NR == 32), "eight waves take two 16-column halves");
  const int t = threadIdx.x, lane = t & 63, w = (t >> 6) & 3, li = lane & 15, lk = lane >> 4;
  const int cb = (t >> 8) * CPW;  // first column of this wavefront
  const T* Lb = L + b.base;
  T* Ub = U + b.woff * (NR * CG) + (int64_t)grp * b.nt * kTile * NR;
  const int c0 = grp * NR;  // first right-hand side of this group
  rhs += c0;
  out += c0;
  nrhs -= c0;
  V acc[CPW / 16];
  T rv[CPW / 16][4];  // right-hand side (forward) or Y_j (backward) of the row
  // tiles of the next tile row (its diagonal tile and first step), loaded
  // while the current row's last step runs
  TileRegs<T, NT> rdn, ran[KB];
  load_row_tiles<T, KB, false, NT>(Lb, tr, 0, rdn, ran);
  // forward: Y_i = inv(L_ii) (R_i - sum_{k=fc_i}^{i-1} L_ik Y_k)
  for (int i = 0; i < b.nt; i++) {
#pragma unroll
    for (int c = 0; c < CPW / 16; c++) acc[c] = V{0, 0, 0, 0};
    const int fc = tr.fc(i);
    {
      // steps of KB off-diagonal tiles; the next step's tiles and panels are
      // loaded into registers while this step's products run
      TileRegs<T, NT> ra[KB], rd = rdn;
      PanelRegs<T, NR, NT> rb[KB];
#pragma unroll
      for (int u = 0; u < KB; u++) {
        ra[u] = ran[u];
        if (fc + u < i) rb[u].load(Ub + (int64_t)(fc + u) * kTile * NR);
      }
      // this row's right-hand sides, gathered while the steps run
#pragma unroll
      for (int c = 0; c < CPW / 16; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
          const int row = 16 * w + Mma<T>::row(lane, r), col = cb + c * 16 + li;
          const int p = i * kTile + row;
          rv[c][r] = p < b.s && col < nrhs ? (T)rhs[(int64_t)inds[b.vec_off + p] * ldr + col] : (T)0;
        }
      rd.store(Ds);
      for (int k = fc;; k += KB) {
#pragma unroll
        for (int u = 0; u < KB; u++)
          if (k + u < i) {
            ra[u].store(As[u]);
            rb[u].store(Bs[u]);
          }
        __syncthreads();
        const int kn = k + KB;
        if (kn < i) {
#pragma unroll
          for (int u = 0; u < KB; u++)
            if (kn + u < i) {
              ra[u].load(Lb + tr.toff(i, kn + u));
              rb[u].load(Ub + (int64_t)(kn + u) * kTile * NR);
            }
        } else if (i + 1 < b.nt) {
          load_row_tiles<T, KB, false, NT>(Lb, tr, i + 1, rdn, ran);
        }
#pragma unroll
        for (int u = 0; u < KB; u++)
          if (k + u < i) panel_mma<T, NR, false, CPW>(acc, As[u], Bs[u], w, li, lk, cb);
        __syncthreads();
        if (kn >= i) break;
      }
    }
#pragma unroll
    for (int c = 0; c < CPW / 16; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * w + Mma<T>::row(lane, r), col = cb + c * 16 + li;
        Bs[0][row * (NR + 1) + col] = rv[c][r] - acc[c][r];
      }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CPW / 16; c++) acc[c] = V{0, 0, 0, 0};
    panel_mma<T, NR, false, CPW>(acc, Ds, Bs[0], w, li, lk, cb);
    T* Ui = Ub + (int64_t)i * kTile * NR;
#pragma unroll
    for (int c = 0; c < CPW / 16; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) Ui[(16 * w + Mma<T>::row(lane, r)) * NR + cb + c * 16 + li] = acc[c][r];
    __syncthreads();
  }
  // backward: X_j = inv(L_jj)^T (Y_j - sum_{i=j+1}^{lr_j} L_ij^T X_i), X_j over Y_j
  load_row_tiles<T, KB, true, NT>(Lb, tr, b.nt - 1, rdn, ran);
  for (int j = b.nt - 1; j >= 0; j--) {
#pragma unroll
    for (int c = 0; c < CPW / 16; c++) acc[c] = V{0, 0, 0, 0};
    const int lr = tr.lr(j);
    T* Uj = Ub + (int64_t)j * kTile * NR;
    {
      TileRegs<T, NT> ra[KB], rd = rdn;
      PanelRegs<T, NR, NT> rb[KB];
#pragma unroll
      for (int u = 0; u < KB; u++) {
        ra[u] = ran[u];
        if (j + 1 + u <= lr) rb[u].load(Ub + (int64_t)(j + 1 + u) * kTile * NR);
      }
      // Y_j, read while the steps run (the steps read X_i, i > j, only)
#pragma unroll
      for (int c = 0; c < CPW / 16; c++)
#pragma unroll
        for (int r = 0; r < 4; r++) rv[c][r] = Uj[(16 * w + Mma<T>::row(lane, r)) * NR + cb + c * 16 + li];
      rd.store(Ds);
      for (int i = j + 1;; i += KB) {
#pragma unroll
        for (int u = 0; u < KB; u++)
          if (i + u <= lr) {
            ra[u].store(As[u]);
            rb[u].store(Bs[u]);
          }
        __syncthreads();
        const int in = i + KB;
        if (in <= lr) {
#pragma unroll
          for (int u = 0; u < KB; u++)
            if (in + u <= lr) {
              ra[u].load(Lb + tr.toff(in + u, j));
              rb[u].load(Ub + (int64_t)(in + u) * kTile * NR);
            }
        } else if (j > 0) {
          load_row_tiles<T, KB, true, NT>(Lb, tr, j - 1, rdn, ran);
        }
#pragma unroll
        for (int u = 0; u < KB; u++)
          if (i + u <= lr) panel_mma<T, NR, true, CPW>(acc, As[u], Bs[u], w, li, lk, cb);
        __syncthreads();
        if (in > lr) break;
      }
    }
#pragma unroll
    for (int c = 0; c < CPW / 16; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * w + Mma<T>::row(lane, r), col = cb + c * 16 + li;
        Bs[0][row * (NR + 1) + col] = rv[c][r] - acc[c][r];
      }
    __syncthreads();
#pragma unroll
    for (int c = 0; c < CPW / 16; c++) acc[c] = V{0, 0, 0, 0};
    panel_mma<T, NR, true, CPW>(acc, Ds, Bs[0], w, li, lk, cb);  // inv(L_jj)^T
#pragma unroll
    for (int c = 0; c < CPW / 16; c++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = 16 * w + Mma<T>::row(lane, r), col = cb + c * 16 + li;
        Uj[row * NR + col] = acc[c][r];
        const int p = j * kTile + row;
        if (p < b.s && col < nrhs) out[(b.vec_off + p) * ldo + col] = (double)acc[c][r];
      }
    __syncthreads();
  }
}

__global__ __launch_bounds__(kSpmvThreads) void k_spmv_multi(
    const int32_t* __restrict__ rowblk, const int32_t* __restrict__ ptr,
    const int32_t* __restrict__ col, const double* __restrict__ val, const double* __restrict__ x,
    int ldx, double* __restrict__ y, int ldy, int nrhs) {
  __shared__ double sval[kSpmvChunk];
  __shared__ int32_t scol[kSpmvChunk];
  const int r0 = rowblk[blockIdx.x], r1 = rowblk[blockIdx.x + 1];
  const int e0 = ptr[r0], e1 = ptr[r1];
  const int t = threadIdx.x;
  const bool staged = e1 - e0 <= kSpmvChunk;
  if (staged) {
    for (int e = e0 + t; e < e1; e += kSpmvThreads) {
      sval[e - e0] = val[e];
      scol[e - e0] = col[e];
    }
    __syncthreads();
  }
  const int work = (r1 - r0) * nrhs;
  for (int idx = t; idx < work; idx += kSpmvThreads) {
    const int row = r0 + idx / nrhs, r = idx % nrhs;
    double acc = 0.0;
    if (staged) {
      for (int e = ptr[row]; e < ptr[row + 1]; e++)
        acc = acc + sval[e - e0] * x[(int64_t)scol[e - e0] * ldx + r];
    } else {
      for (int e = ptr[row]; e < ptr[row + 1]; e++) acc = acc + val[e] * x[(int64_t)col[e] * ldx + r];
    }
    y[(int64_t)row * ldy + r] = acc;
  }
}

__global__ void k_overlap_sum_multi(int n, int nrhs, const int32_t* __restrict__ occ_ptr,
                                    const int32_t* __restrict__ occ_pos,
                                    const double* __restrict__ sol,
                                    const int32_t* __restrict__ rt_ptr,
                                    const int32_t* __restrict__ rt_col,
                                    const double* __restrict__ rt_val,
                                    const double* __restrict__ csol, int ldc,
                                    double* __restrict__ z) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * nrhs) return;
  const int i = (int)(e / nrhs), r = (int)(e - (int64_t)i * nrhs);
  double acc = 0.0;
  for (int p = occ_ptr[i]; p < occ_ptr[i + 1]; p++) acc = acc + sol[(int64_t)occ_pos[p] * nrhs + r];
  if (rt_ptr) {
    double tmp = 0.0;
    for (int p = rt_ptr[i]; p < rt_ptr[i + 1]; p++)
      tmp = tmp + rt_val[p] * csol[(int64_t)rt_col[p] * ldc + r];
    acc = acc + tmp;
  }
  z[e] = acc;
}

// Column-pair forms (nrhs, ldx, ldy, ldc even): one thread per (row, pair of
// columns) with 16-byte loads; every column is summed in the same order as
// in the one-column forms above, so the results are the same bits.
__global__ __launch_bounds__(kSpmvThreads) void k_spmv_multi2(
    const int32_t* __restrict__ rowblk, const int32_t* __restrict__ ptr,
    const int32_t* __restrict__ col, const double* __restrict__ val, const double* __restrict__ x,
    int ldx, double* __restrict__ y, int ldy, int nrhs) {
  __shared__ double sval[kSpmvChunk];
  __shared__ int32_t scol[kSpmvChunk];
  const int r0 = rowblk[blockIdx.x], r1 = rowblk[blockIdx.x + 1];
  const int e0 = ptr[r0], e1 = ptr[r1];
  const int t = threadIdx.x, np = nrhs >> 1;
  const bool staged = e1 - e0 <= kSpmvChunk;
  if (staged) {
    for (int e = e0 + t; e < e1; e += kSpmvThreads) {
      sval[e - e0] = val[e];
      scol[e - e0] = col[e];
    }
    __syncthreads();
  }
  const int work = (r1 - r0) * np;
  for (int idx = t; idx < work; idx += kSpmvThreads) {
    const int row = r0 + idx / np, c = 2 * (idx % np);
    double2 acc = make_double2(0.0, 0.0);
    for (int e = ptr[row]; e < ptr[row + 1]; e++) {
      const double a = staged ? sval[e - e0] : val[e];
      const int k = staged ? scol[e - e0] : col[e];
      const double2 xv = *reinterpret_cast<const double2*>(x + (int64_t)k * ldx + c);
      acc.x = acc.x + a * xv.x;
      acc.y = acc.y + a * xv.y;
    }
    *reinterpret_cast<double2*>(y + (int64_t)row * ldy + c) = acc;
  }
}

// Long rows (the restriction R, hundreds of entries per row) for column
// pairs: one workgroup per row; thread t takes column pair t % np and every
// (256/np)-th entry from t / np, then the partials of a pair are added in a
// fixed tree order.
__global__ __launch_bounds__(256) void k_spmv_multi_rows(const int32_t* __restrict__ ptr,
                                                         const int32_t* __restrict__ col,
                                                         const double* __restrict__ val,
                                                         const double* __restrict__ x, int ldx,
                                                         double* __restrict__ y, int ldy,
                                                         int nrhs) {
  __shared__ double2 part[256];
  const int row = blockIdx.x, t = threadIdx.x;
  const int np = nrhs >> 1, lanes = 256 / np, pr = t % np, ln = t / np, c = 2 * pr;
  double2 acc = make_double2(0.0, 0.0);
  for (int e = ptr[row] + ln; e < ptr[row + 1]; e += lanes) {
    const double a = val[e];
    const double2 xv = *reinterpret_cast<const double2*>(x + (int64_t)col[e] * ldx + c);
    acc.x = acc.x + a * xv.x;
    acc.y = acc.y + a * xv.y;
  }
  part[t] = acc;
  __syncthreads();
  for (int o = lanes >> 1; o > 0; o >>= 1) {
    if (ln < o) {
      const double2 q = part[t + o * np];
      part[t].x = part[t].x + q.x;
      part[t].y = part[t].y + q.y;
    }
    __syncthreads();
  }
  if (ln == 0) *reinterpret_cast<double2*>(y + (int64_t)row * ldy + c) = part[t];
}

__global__ void k_overlap_sum_multi2(int n, int nrhs, const int32_t* __restrict__ occ_ptr,
                                     const int32_t* __restrict__ occ_pos,
                                     const double* __restrict__ sol,
                                     const int32_t* __restrict__ rt_ptr,
                                     const int32_t* __restrict__ rt_col,
                                     const double* __restrict__ rt_val,
                                     const double* __restrict__ csol, int ldc,
                                     double* __restrict__ z) {
  const int np = nrhs >> 1;
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (int64_t)n * np) return;
  const int i = (int)(e / np), c = 2 * (int)(e - (int64_t)i * np);
  double2 acc = make_double2(0.0, 0.0);
  for (int p = occ_ptr[i]; p < occ_ptr[i + 1]; p++) {
    const double2 s = *reinterpret_cast<const double2*>(sol + (int64_t)occ_pos[p] * nrhs + c);
    acc.x = acc.x + s.x;
    acc.y = acc.y + s.y;
  }
  if (rt_ptr) {
    double2 tmp = make_double2(0.0, 0.0);
    for (int p = rt_ptr[i]; p < rt_ptr[i + 1]; p++) {
      const double w = rt_val[p];
      const double2 cs = *reinterpret_cast<const double2*>(csol + (int64_t)rt_col[p] * ldc + c);
      tmp.x = tmp.x + w * cs.x;
      tmp.y = tmp.y + w * cs.y;
    }
    acc.x = acc.x + tmp.x;
    acc.y = acc.y + tmp.y;
  }
  *reinterpret_cast<double2*>(z + (int64_t)i * nrhs + c) = acc;
}

// Per-column reductions over an n x nrhs node-major block: thread t always
// handles column t % nrhs (the stride is a multiple of nrhs), block sums per
// column in a fixed order, and the last block finishes every column.
constexpr int kMultiMax = kMaxRhs;
enum { SM_ACTIVE = 6 };

// mode 1: v = r.z (beta, rho); mode 2: v = p.q (alpha; 0 for a frozen
// column); mode 3: v = r.r
__device__ __forceinline__ void apply_colsum(double* sc, double v, int mode) {
  const bool active = sc[SM_ACTIVE] != 0.0;
  if (mode == 1) {
    sc[S_BETA] = v / sc[S_RHO];
    sc[S_RHO_PREV] = sc[S_RHO];
    sc[S_RHO] = v;
  } else if (mode == 2) {
    sc[S_PQ] = v;
    sc[S_ALPHA] = active ? sc[S_RHO] / v : 0.0;
  } else {
    sc[S_RES] = v;
  }
}

__device__ void finish_multi(double* partial, unsigned* counter, int nrhs, int mode, double* scal,
                             double* sh) {
  __shared__ bool last;
  if (threadIdx.x == 0) {
    __threadfence();
    last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!last) return;
  __threadfence();
  // 256 / nrhs threads per column, each over a fixed stride of the block
  // partials, then a fixed-order tree (nrhs is a power of two <= 32)
  const int t = threadIdx.x, per = 256 / nrhs, r = t / per, sub = t - r * per;
  double v = 0.0;
  for (int b = sub; b < (int)gridDim.x; b += per)
    v += __hip_atomic_load(&partial[b * kMultiMax + r], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  sh[t] = v;
  __syncthreads();
  for (int o = per >> 1; o > 0; o >>= 1) {
    if (sub < o) sh[t] = sh[t] + sh[t + o];
    __syncthreads();
  }
  if (sub == 0) apply_colsum(scal + r * 8, sh[t], mode);
  if (threadIdx.x == 0) *counter = 0u;
}

// Column sums of per-block partials (partial[column * nb + b], nb blocks)
// by a kernel of their own, one workgroup per column: thread t adds blocks
// t, t + 256, ... in order, then a fixed tree.  Used after the
// wide column-pair kernels, whose nb is too large for finish_multi's single
// workgroup.
__global__ __launch_bounds__(256) void k_finish_cols(const double* __restrict__ partial, int nb, int mode,
                                                     double* scal) {
  __shared__ double sh[256];
  const int r = blockIdx.x, t = threadIdx.x;
  double v = 0.0;
  int b = t;
  for (; b + 3 * 256 < nb; b += 4 * 256) {
    double a[4];
#pragma unroll
    for (int u = 0; u < 4; u++) a[u] = partial[(int64_t)r * nb + b + u * 256];
#pragma unroll
    for (int u = 0; u < 4; u++) v = v + a[u];
  }
  for (; b < nb; b += 256) v = v + partial[(int64_t)r * nb + b];
  sh[t] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) sh[t] = sh[t] + sh[t + o];
    __syncthreads();
  }
  if (t == 0) apply_colsum(scal + r * 8, sh[0], mode);
}

__device__ __forceinline__ void block_colsums(double v, int nrhs, double* sh, double* partial) {
  const int t = threadIdx.x;
  sh[t] = v;
  __syncthreads();
  if (t < nrhs) {
    double s = 0.0;
    for (int q = t; q < 256; q += nrhs) s += sh[q];
    partial[blockIdx.x * kMultiMax + t] = s;
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_dot_multi(const double* __restrict__ a,
                                                   const double* __restrict__ b, int64_t total,
                                                   int nrhs, double* partial, unsigned* counter,
                                                   int mode, double* scal) {
  __shared__ double sh[256];
  const int64_t stride = ((int64_t)gridDim.x * 256 / nrhs) * nrhs;
  const int64_t start = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (start < stride)
    for (int64_t e = start; e < total; e += stride) v = v + a[e] * b[e];
  block_colsums(v, nrhs, sh, partial);
  finish_multi(partial, counter, nrhs, mode, scal, sh);
}

// Column pairs (nrhs even): 16-byte loads, thread t owns columns 2(t % (nrhs/2)) and +1.
__device__ __forceinline__ void block_colsums2(double v0, double v1, int nrhs, double* sh,
                                               double* partial) {
  const int t = threadIdx.x, nr2 = nrhs >> 1;
  sh[t] = v0;
  sh[256 + t] = v1;
  __syncthreads();
  if (t < nrhs) {
    const double* h = sh + (t & 1) * 256;
    double s = 0.0;
    for (int q = t >> 1; q < 256; q += nr2) s += h[q];
    partial[(int64_t)t * gridDim.x + blockIdx.x] = s;  // column-major, finished by k_finish_cols
  }
  __syncthreads();
}

__global__ __launch_bounds__(256) void k_dot_multi2(const double2* __restrict__ a,
                                                    const double2* __restrict__ b, int64_t total2,
                                                    int nrhs, double* partial) {
  __shared__ double sh[512];
  const int nr2 = nrhs >> 1;
  const int64_t stride = ((int64_t)gridDim.x * 256 / nr2) * nr2;
  const int64_t start = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v0 = 0.0, v1 = 0.0;
  if (start < stride)
    for (int64_t e = start; e < total2; e += stride) {
      const double2 x = a[e], y = b[e];
      v0 = v0 + x.x * y.x;
      v1 = v1 + x.y * y.y;
    }
  block_colsums2(v0, v1, nrhs, sh, partial);
}

__global__ void k_p_update_multi(int64_t total, int nrhs, const double* __restrict__ z,
                                 double* __restrict__ p, const double* __restrict__ scal,
                                 int first) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total) return;
  const int r = (int)(e % nrhs);
  if (scal[r * 8 + SM_ACTIVE] == 0.0) return;  // converged columns stay frozen
  if (first)
    p[e] = z[e];
  else
    p[e] = z[e] + scal[r * 8 + S_BETA] * p[e];
}

__global__ __launch_bounds__(256) void k_xr_update_multi(int64_t total, int nrhs,
                                                         double* __restrict__ x,
                                                         double* __restrict__ r,
                                                         const double* __restrict__ p,
                                                         const double* __restrict__ q,
                                                         double* partial, unsigned* counter,
                                                         double* scal) {
  __shared__ double sh[256];
  const int64_t stride = ((int64_t)gridDim.x * 256 / nrhs) * nrhs;
  const int64_t start = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v = 0.0;
  if (start < stride) {
    const double* sc = scal + (int)(start % nrhs) * 8;
    const double alpha = sc[S_ALPHA];
    if (sc[SM_ACTIVE] != 0.0) {
      for (int64_t e = start; e < total; e += stride) {
        x[e] = x[e] + alpha * p[e];
        const double ri = r[e] - alpha * q[e];
        r[e] = ri;
        v = v + ri * ri;
      }
    } else {  // converged columns stay frozen
      for (int64_t e = start; e < total; e += stride) v = v + r[e] * r[e];
    }
  }
  block_colsums(v, nrhs, sh, partial);
  finish_multi(partial, counter, nrhs, 3, scal, sh);
}

// r -= alpha q per active column, with the r.r partials; x's update waits
// for the next k_spmv_p_multi2 (or k_x_final_multi2), which reads p anyway.
__global__ __launch_bounds__(256) void k_xr_update_multi2(int64_t total2, int nrhs,
                                                          double2* __restrict__ r,
                                                          const double2* __restrict__ q,
                                                          double* partial, const double* scal) {
  __shared__ double sh[512];
  const int nr2 = nrhs >> 1;
  const int64_t stride = ((int64_t)gridDim.x * 256 / nr2) * nr2;
  const int64_t start = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v0 = 0.0, v1 = 0.0;
  if (start < stride) {
    const int c0 = 2 * (int)(start % nr2);
    const double* s0 = scal + c0 * 8;
    const double* s1 = s0 + 8;
    // a frozen column keeps r
    const bool f0 = s0[SM_ACTIVE] != 0.0, f1 = s1[SM_ACTIVE] != 0.0, any = f0 || f1;
    const double a0 = f0 ? s0[S_ALPHA] : 0.0, a1 = f1 ? s1[S_ALPHA] : 0.0;
    int64_t e = start;
    // four items per round with every load issued first; the same element
    // updates and the same order of the sums as one item at a time
    for (; e + 3 * stride < total2; e += 4 * stride) {
      double2 ri[4], qe[4];
#pragma unroll
      for (int u = 0; u < 4; u++) ri[u] = r[e + u * stride];
      if (any) {
#pragma unroll
        for (int u = 0; u < 4; u++) qe[u] = q[e + u * stride];
#pragma unroll
        for (int u = 0; u < 4; u++) {
          if (f0) ri[u].x = ri[u].x - a0 * qe[u].x;
          if (f1) ri[u].y = ri[u].y - a1 * qe[u].y;
          r[e + u * stride] = ri[u];
        }
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        v0 = v0 + ri[u].x * ri[u].x;
        v1 = v1 + ri[u].y * ri[u].y;
      }
    }
    for (; e < total2; e += stride) {
      double2 ri = r[e];
      if (any) {
        const double2 qe = q[e];
        if (f0) ri.x = ri.x - a0 * qe.x;
        if (f1) ri.y = ri.y - a1 * qe.y;
        r[e] = ri;
      }
      v0 = v0 + ri.x * ri.x;
      v1 = v1 + ri.y * ri.y;
    }
  }
  block_colsums2(v0, v1, nrhs, sh, partial);
}

// x += alpha p, the last iteration's deferred update (alpha is 0 for the
// columns frozen before it)
__global__ void k_x_final_multi2(int64_t total2, int nrhs, double2* __restrict__ x,
                                 const double2* __restrict__ p, const double* __restrict__ scal) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= total2) return;
  const int c0 = 2 * (int)(e % (nrhs >> 1));
  const double al0 = scal[c0 * 8 + S_ALPHA], al1 = scal[(c0 + 1) * 8 + S_ALPHA];
  double2 xe = x[e];
  const double2 pe = p[e];
  xe.x = xe.x + al0 * pe.x;
  xe.y = xe.y + al1 * pe.y;
  x[e] = xe;
}

// Fused forms of the batched iteration (column pairs, nrhs even).
//
// z = overlap sum of the subdomain solutions (+ coarse correction), and
// rho = r.z per column: the elements are visited in the partition of
// k_dot_multi2 (thread t of the grid takes e = t, t + stride, ...) and the
// partials finished by k_finish_cols, so z and rho are the same bits as
// k_overlap_sum_multi2 followed by k_dot_multi2 on the same grid.
__global__ __launch_bounds__(256) void k_overlap_sum_dot_multi2(
    int n, int nrhs, const int32_t* __restrict__ occ_ptr, const int32_t* __restrict__ occ_pos,
    const double* __restrict__ sol, const int32_t* __restrict__ rt_ptr,
    const int32_t* __restrict__ rt_col, const double* __restrict__ rt_val,
    const double* __restrict__ csol, int ldc, double* __restrict__ z, const double* __restrict__ r,
    double* partial) {
  __shared__ double sh[512];
  const int np = nrhs >> 1;
  const int64_t total2 = (int64_t)n * np;
  const int64_t stride = ((int64_t)gridDim.x * 256 / np) * np;
  const int64_t start = (int64_t)blockIdx.x * 256 + threadIdx.x;
  double v0 = 0.0, v1 = 0.0;
  if (start < stride) {
    const int c = 2 * (int)(start % np), di = (int)(stride / np);
    int i = (int)(start / np);
    // kU nodes per round in lockstep, so that the dependent loads (list
    // bounds, then positions, then values) of the kU nodes overlap; each
    // node's sums keep their order and the dot keeps the order of the items
    constexpr int kU = 2;
    for (int64_t e = start; e < total2; e += kU * stride, i += kU * di) {
      int p0[kU], p1[kU], q0[kU], q1[kU];
      bool in[kU];
      double2 acc[kU], tmp[kU], re[kU];
#pragma unroll
      for (int u = 0; u < kU; u++) {
        in[u] = e + u * stride < total2;
        const int iu = in[u] ? i + u * di : i;
        p0[u] = occ_ptr[iu];
        p1[u] = in[u] ? occ_ptr[iu + 1] : p0[u];
        q0[u] = rt_ptr ? rt_ptr[iu] : 0;
        q1[u] = rt_ptr && in[u] ? rt_ptr[iu + 1] : q0[u];
        re[u] = in[u] ? reinterpret_cast<const double2*>(r)[e + u * stride] : make_double2(0.0, 0.0);
        acc[u] = make_double2(0.0, 0.0);
        tmp[u] = make_double2(0.0, 0.0);
      }
      for (int j = 0;; j++) {
        bool g[kU], any = false;
#pragma unroll
        for (int u = 0; u < kU; u++) {
          g[u] = p0[u] + j < p1[u];
          any = any || g[u];
        }
        if (!any) break;
        int pos[kU];
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (g[u]) pos[u] = occ_pos[p0[u] + j];
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (g[u]) {
            const double2 sv = *reinterpret_cast<const double2*>(sol + (int64_t)pos[u] * nrhs + c);
            acc[u].x = acc[u].x + sv.x;
            acc[u].y = acc[u].y + sv.y;
          }
      }
      for (int j = 0;; j++) {
        bool g[kU], any = false;
#pragma unroll
        for (int u = 0; u < kU; u++) {
          g[u] = q0[u] + j < q1[u];
          any = any || g[u];
        }
        if (!any) break;
        int cc[kU];
        double w[kU];
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (g[u]) {
            cc[u] = rt_col[q0[u] + j];
            w[u] = rt_val[q0[u] + j];
          }
#pragma unroll
        for (int u = 0; u < kU; u++)
          if (g[u]) {
            const double2 cs = *reinterpret_cast<const double2*>(csol + (int64_t)cc[u] * ldc + c);
            tmp[u].x = tmp[u].x + w[u] * cs.x;
            tmp[u].y = tmp[u].y + w[u] * cs.y;
          }
      }
#pragma unroll
      for (int u = 0; u < kU; u++)
        if (in[u]) {
          if (rt_ptr) {
            acc[u].x = acc[u].x + tmp[u].x;
            acc[u].y = acc[u].y + tmp[u].y;
          }
          reinterpret_cast<double2*>(z)[e + u * stride] = acc[u];
          v0 = v0 + re[u].x * acc[u].x;
          v1 = v1 + re[u].y * acc[u].y;
        }
    }
  }
  block_colsums2(v0, v1, nrhs, sh, partial);
}

// p_new = z + beta p_old (p_new = z on the first iteration); a frozen column
// keeps p_old
__device__ __forceinline__ double2 p_next(double2 zv, double2 pv, double b0, double b1, bool a0,
                                          bool a1, bool first) {
  double2 o;
  o.x = a0 ? (first ? zv.x : zv.x + b0 * pv.x) : pv.x;
  o.y = a1 ? (first ? zv.y : zv.y + b1 * pv.y) : pv.y;
  return o;
}

// q = A p_new with p_new formed on the fly from z and p_old at every
// gathered node (the same expression as k_p_update_multi, so q is the same
// bits as the unfused sequence), p_new written for the block's own rows,
// and p.q per column from fixed-order block partials (one per row block).
// The own rows also take the previous iteration's x += alpha p_old, which
// k_xr_update_multi2 leaves to this pass (k_x_final_multi2 adds the last).  p_old and p_new are
// separate buffers: rows of other blocks still read p_old.  The block's
// entries are staged in LDS (as in k_spmv_multi2); a row's are taken four
// at a time, loads first, then added in entry order.
__global__ __launch_bounds__(kSpmvThreads) void k_spmv_p_multi2(
    const int32_t* __restrict__ rowblk, const int32_t* __restrict__ ptr,
    const int32_t* __restrict__ col, const double* __restrict__ val, const double* __restrict__ z,
    const double* __restrict__ pold, double* __restrict__ pnew, double* __restrict__ q, int nrhs,
    int first, double* partial, const double* scal, double* __restrict__ x) {
  __shared__ double sval[kSpmvChunk];
  __shared__ int32_t scol[kSpmvChunk];
  __shared__ double sh[512];
  const int r0 = rowblk[blockIdx.x], r1 = rowblk[blockIdx.x + 1];
  const int e0 = ptr[r0], e1 = ptr[r1];
  const int t = threadIdx.x, np = nrhs >> 1;
  const bool staged = e1 - e0 <= kSpmvChunk;
  if (staged) {
    for (int e = e0 + t; e < e1; e += kSpmvThreads) {
      sval[e - e0] = val[e];
      scol[e - e0] = col[e];
    }
    __syncthreads();
  }
  // np divides the block size, so the column pair of a thread is fixed
  const int pc = t % np;
  const double* s0 = scal + 2 * pc * 8;
  const double* s1 = s0 + 8;
  const bool a0 = s0[SM_ACTIVE] != 0.0, a1 = s1[SM_ACTIVE] != 0.0, fst = first != 0;
  const double b0 = s0[S_BETA], b1 = s1[S_BETA];
  // the previous iteration's x += alpha p, deferred to this pass (p_old is
  // that p; a column frozen since has alpha 0 from its next p.q)
  const double al0 = s0[S_ALPHA], al1 = s1[S_ALPHA];
  const double2* z2 = reinterpret_cast<const double2*>(z);
  const double2* po2 = reinterpret_cast<const double2*>(pold);
  const int work = (r1 - r0) * np;
  double v0 = 0.0, v1 = 0.0;
  for (int idx = t; idx < work; idx += kSpmvThreads) {
    const int row = r0 + idx / np;
    const int eb = ptr[row], ee = ptr[row + 1];
    double2 acc = make_double2(0.0, 0.0);
    int e = eb;
    for (; e + 4 <= ee; e += 4) {
      int k[4];
      double a[4];
      double2 zv[4], pv[4];
#pragma unroll
      for (int u = 0; u < 4; u++) {
        k[u] = staged ? scol[e + u - e0] : col[e + u];
        a[u] = staged ? sval[e + u - e0] : val[e + u];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        zv[u] = z2[(int64_t)k[u] * np + pc];
        pv[u] = po2[(int64_t)k[u] * np + pc];
      }
#pragma unroll
      for (int u = 0; u < 4; u++) {
        const double2 pk = p_next(zv[u], pv[u], b0, b1, a0, a1, fst);
        acc.x = acc.x + a[u] * pk.x;
        acc.y = acc.y + a[u] * pk.y;
      }
    }
    for (; e < ee; e++) {
      const int k = staged ? scol[e - e0] : col[e];
      const double a = staged ? sval[e - e0] : val[e];
      const double2 pk = p_next(z2[(int64_t)k * np + pc], po2[(int64_t)k * np + pc], b0, b1, a0, a1, fst);
      acc.x = acc.x + a * pk.x;
      acc.y = acc.y + a * pk.y;
    }
    const int64_t o = (int64_t)row * np + pc;
    const double2 pk = po2[o];
    const double2 pr = p_next(z2[o], pk, b0, b1, a0, a1, fst);
    reinterpret_cast<double2*>(pnew)[o] = pr;
    if (!fst) {
      double2 xe = reinterpret_cast<double2*>(x)[o];
      xe.x = xe.x + al0 * pk.x;
      xe.y = xe.y + al1 * pk.y;
      reinterpret_cast<double2*>(x)[o] = xe;
    }
    reinterpret_cast<double2*>(q)[o] = acc;
    v0 = v0 + pr.x * acc.x;
    v1 = v1 + pr.y * acc.y;
  }
  block_colsums2(v0, v1, nrhs, sh, partial);
}

// d loss / d a_ij = -lambda_i x_j for every stored entry, written to the
// entry's position in the caller's COO input
__global__ void k_value_grad(int n, const int32_t* __restrict__ ptr, const int32_t* __restrict__ col,
                             const int32_t* __restrict__ src, const double* __restrict__ lam,
                             const double* __restrict__ x, double* __restrict__ grad) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const double li = lam[i];
  for (int p = ptr[i]; p < ptr[i + 1]; p++) grad[src[p]] = -li * x[col[p]];
}

inline unsigned cdiv(int64_t a, int64_t b) { return (unsigned)((a + b - 1) / b); }

}  // namespace

template <typename T>
void launch_scatter_entries(T* tiles, const int64_t* dst, const int32_t* src, const double* val,
                            int64_t ne, hipStream_t st) {
  if (ne <= 0) return;
  hipLaunchKernelGGL(k_scatter<T>, dim3(cdiv(ne, 256)), dim3(256), 0, st, tiles, dst, src, val, ne);
}

template <typename T>
void launch_scatter_sums(T* tiles, const int64_t* dst, const int64_t* seg, const int32_t* src,
                         const double* val, int64_t ngroups, hipStream_t st) {
  if (ngroups <= 0) return;
  hipLaunchKernelGGL(k_scatter_sum<T>, dim3(cdiv(ngroups, 256)), dim3(256), 0, st, tiles, dst, seg,
                     src, val, ngroups);
}

template <typename T>
void launch_pad_identity(T* tiles, const BlockDesc* d, int nblocks, int maxpad, hipStream_t st) {
  if (maxpad <= 0 || nblocks <= 0) return;
  hipLaunchKernelGGL(k_pad_identity<T>, dim3(cdiv(maxpad, 64), nblocks), dim3(64), 0, st, tiles,
                     d, maxpad);
}

template <typename T>
void launch_chol_diag(T* tiles, const BlockDesc* d, int nblocks, int j, int* flag,
                      hipStream_t st, bool left) {
  hipLaunchKernelGGL(k_chol_diag<T>, dim3(nblocks), dim3(256), 0, st, tiles, d, j, flag, left ? 1 : 0);
}

template <typename T>
void launch_chol_offdiag(T* tiles, const BlockDesc* d, int nblocks, int j, int maxcount,
                         hipStream_t st, bool left) {
  if (maxcount <= 0) return;
  hipLaunchKernelGGL(k_chol_offdiag<T>, dim3(maxcount, nblocks), dim3(256), 0, st, tiles, d, j,
                     left ? 1 : 0);
}

void launch_rl_update(double* tiles, const BlockDesc* d, int nt, int j, int last, hipStream_t st) {
  const int m = std::min(last, nt - 1) - j;
  if (m <= 0) return;
  hipLaunchKernelGGL(k_rl_update, dim3(m * (m + 1) / 2), dim3(256), 0, st, tiles, d, j);
}

void launch_inv_init(double* W, const BlockDesc* d, int nt, hipStream_t st) {
  if (nt <= 0) return;
  hipLaunchKernelGGL(k_inv_init, dim3(nt, nt), dim3(256), 0, st, W, d);
}

void launch_inv_step(const double* L, double* W, const BlockDesc* d, int nt, int m, int last,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_inv_finish, dim3(m + 1), dim3(256), 0, st, L, W, d, m);
  const int rows = std::min(last, nt - 1) - m;
  if (rows > 0)
    hipLaunchKernelGGL(k_inv_update, dim3(m + 1, rows), dim3(256), 0, st, L, W, d, m);
}

__global__ void k_map_index(int32_t* __restrict__ idx, int64_t n, const int32_t* __restrict__ map) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) idx[p] = map[idx[p]];
}

void launch_values_strong(int n, const int32_t* ptr, const int32_t* col, const double* vin, int symm,
                          double alpha, double* diag, double* sv, double* val, uint8_t* strong,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_row_scale, dim3(cdiv(n, 256)), dim3(256), 0, st, n, ptr, col, vin, symm, diag, sv);
  hipLaunchKernelGGL(k_round_trip_strong, dim3(cdiv(n, 256)), dim3(256), 0, st, n, ptr, col, vin, diag,
                     sv, symm, alpha, val, strong);
}

void launch_map_index(int32_t* idx, int64_t n, const int32_t* map, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_map_index, dim3(cdiv(n, 256)), dim3(256), 0, st, idx, n, map);
}

template <typename T>
void prepare_trsv(int maxnt) {
  const int bytes = maxnt * kTile * (int)sizeof(T);
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_trsv<T>),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

template <typename T>
void launch_trsv(const T* tiles, const BlockDesc* d, int nblocks, int maxnt, const int32_t* inds,
                 const double* r, double* sol, hipStream_t st, const double* skip) {
  if (nblocks <= 0) return;
  const size_t bytes = (size_t)maxnt * kTile * sizeof(T);
  hipLaunchKernelGGL(k_trsv<T>, dim3(nblocks), dim3(256), bytes, st, tiles, d, inds, r, sol, skip);
}

#define DOUG_INSTANTIATE(T)                                                                    \
  template void launch_scatter_entries<T>(T*, const int64_t*, const int32_t*, const double*,   \
                                          int64_t, hipStream_t);                               \
  template void launch_scatter_sums<T>(T*, const int64_t*, const int64_t*, const int32_t*,     \
                                       const double*, int64_t, hipStream_t);                   \
  template void launch_pad_identity<T>(T*, const BlockDesc*, int, int, hipStream_t);           \
  template void launch_chol_diag<T>(T*, const BlockDesc*, int, int, int*, hipStream_t, bool);  \
  template void launch_chol_offdiag<T>(T*, const BlockDesc*, int, int, int, hipStream_t, bool); \
  template void launch_trsv<T>(const T*, const BlockDesc*, int, int, const int32_t*,          \
                               const double*, double*, hipStream_t, const double*);           \
  template void prepare_trsv<T>(int);
DOUG_INSTANTIATE(double)
DOUG_INSTANTIATE(float)
#undef DOUG_INSTANTIATE



// ------------------------------------------------------------------------
// Subdomain blocks from the device's copies of A, the blocks' unknown lists
// and each unknown's occurrence list (its positions q in the concatenated
// lists, ascending).  Row li of block b is unknown inds[s0 + li]; column c of
// A is local column q - s0 for the occurrence q of c inside [s0, s0 + s),
// found by scanning c's short list.  One workgroup per block.

__device__ __forceinline__ int local_col(const int32_t* __restrict__ occ_ptr,
                                         const int32_t* __restrict__ occ_pos, int c, int s0, int s) {
  for (int k = occ_ptr[c]; k < occ_ptr[c + 1]; k++) {
    const int q = occ_pos[k];
    if (q >= s0 + s) break;
    if (q >= s0) return q - s0;
  }
  return -1;
}

// First tile column of each tile row's envelope: min over the row's lower
// entries of their tile column; `fc` holds the tile row index on entry.
__global__ void k_block_envelope(const int32_t* __restrict__ starts, const int32_t* __restrict__ row0,
                                 const int32_t* __restrict__ inds, const int32_t* __restrict__ occ_ptr,
                                 const int32_t* __restrict__ occ_pos, const int32_t* __restrict__ a_ptr,
                                 const int32_t* __restrict__ a_col, int32_t* __restrict__ fc) {
  const int b = blockIdx.x;
  const int s0 = starts[b], s = starts[b + 1] - s0;
  int32_t* f = fc + row0[b];
  for (int li = threadIdx.x; li < s; li += blockDim.x) {
    const int g = inds[s0 + li];
    int m = li;
    for (int p = a_ptr[g]; p < a_ptr[g + 1]; p++) {
      const int lj = local_col(occ_ptr, occ_pos, a_col[p], s0, s);
      if (lj >= 0 && lj < m) m = lj;
    }
    atomicMin(&f[li / kTile], m / kTile);
  }
}

// The lower entries of each block into its tile storage (converted to T).
// Requires no duplicate (i, j) pair in A, which would race.
template <typename T>
__global__ void k_scatter_blocks(T* __restrict__ tiles, const BlockDesc* __restrict__ d,
                                 const int32_t* __restrict__ inds, const int32_t* __restrict__ occ_ptr,
                                 const int32_t* __restrict__ occ_pos, const int32_t* __restrict__ a_ptr,
                                 const int32_t* __restrict__ a_col, const double* __restrict__ a_val) {
  const BlockDesc b = d[blockIdx.x];
  const int s0 = (int)b.vec_off, s = b.s;
  for (int li = threadIdx.x; li < s; li += blockDim.x) {
    const int g = inds[s0 + li];
    const TileRow ri = b.tr[li / kTile];
    for (int p = a_ptr[g]; p < a_ptr[g + 1]; p++) {
      const int lj = local_col(occ_ptr, occ_pos, a_col[p], s0, s);
      if (lj < 0 || lj > li) continue;
      const int64_t pos = b.base + (int64_t)(ri.off + lj / kTile - ri.fc) * kTileElems +
                          (int64_t)(li % kTile) * kTile + (lj % kTile);
      tiles[pos] = (T)a_val[p];
    }
  }
}

void launch_block_envelope(const int32_t* starts, const int32_t* row0, int nblocks, const int32_t* inds,
                           const int32_t* occ_ptr, const int32_t* occ_pos, const int32_t* a_ptr,
                           const int32_t* a_col, int32_t* fc, hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_block_envelope, dim3(nblocks), dim3(256), 0, st, starts, row0, inds, occ_ptr,
                     occ_pos, a_ptr, a_col, fc);
}

template <typename T>
void launch_scatter_blocks(T* tiles, const BlockDesc* d, int nblocks, const int32_t* inds,
                           const int32_t* occ_ptr, const int32_t* occ_pos, const int32_t* a_ptr,
                           const int32_t* a_col, const double* a_val, hipStream_t st) {
  if (nblocks <= 0) return;
  hipLaunchKernelGGL(k_scatter_blocks<T>, dim3(nblocks), dim3(256), 0, st, tiles, d, inds, occ_ptr,
                     occ_pos, a_ptr, a_col, a_val);
}
template void launch_scatter_blocks<float>(float*, const BlockDesc*, int, const int32_t*, const int32_t*,
                                           const int32_t*, const int32_t*, const int32_t*, const double*,
                                           hipStream_t);
template void launch_scatter_blocks<double>(double*, const BlockDesc*, int, const int32_t*, const int32_t*,
                                            const int32_t*, const int32_t*, const int32_t*, const double*,
                                            hipStream_t);

// ------------------------------------------------------------------------
// Symmetric Gauss-Seidel sweeps (SymGaussSeidel, stationary.F90:33-88):
// x_i = (b_i - sum_{j != i} a_ij x_j) / a_ii in ascending column order, rows
// ascending then descending, rows without a diagonal entry left at zero.
// Level scheduling: row i of a forward sweep must follow every j < i it or
// its neighbours couple to and precede every such j > i, so rows at the same
// depth of that order are independent and one workgroup computes each
// level's rows side by side, a barrier between levels.  Each row still sums
// its own entries in column order, so the result is the sequential sweep's.
// ------------------------------------------------------------------------
namespace {

__global__ __launch_bounds__(256) void k_sgs(const SgsDesc* __restrict__ d,
                                             const int32_t* __restrict__ levptr,
                                             const int32_t* __restrict__ lrows,
                                             const int32_t* __restrict__ rbeg,
                                             const int32_t* __restrict__ rlen,
                                             const int32_t* __restrict__ ecol,
                                             const double* __restrict__ eval,
                                             const double* __restrict__ diag, double* sub,
                                             int iters) {
  extern __shared__ double xs[];
  const SgsDesc b = d[blockIdx.x];
  const int t = threadIdx.x;
  for (int r = t; r < b.m; r += 256) xs[r] = 0.0;
  __syncthreads();
  for (int it = 0; it < iters; it++) {
    for (int dir = 0; dir < 2; dir++) {
      const int nl = dir ? b.nlb : b.nlf;
      const int32_t* lp = levptr + (dir ? b.lb0 : b.lf0);
      const int32_t* rows = lrows + 2 * b.off + (dir ? b.m : 0);
      for (int lev = 0; lev < nl; lev++) {
        const int e = lp[lev + 1];
        for (int q = lp[lev] + t; q < e; q += 256) {
          const int r = rows[q];
          const int64_t pos = b.off + r;
          const double dg = diag[pos];
          if (dg != 0.0) {
            double xi = sub[pos];
            const int64_t e0 = b.e0 + rbeg[pos];
            const int len = rlen[pos];
            for (int k0 = 0; k0 < len; k0 += 8) {
              int c[8];
              double v[8];
#pragma unroll
              for (int k = 0; k < 8; k++) {
                c[k] = k0 + k < len ? ecol[e0 + k0 + k] : 0;
                v[k] = k0 + k < len ? eval[e0 + k0 + k] : 0.0;
              }
#pragma unroll
              for (int k = 0; k < 8; k++) {
                const double tk = xi - v[k] * xs[c[k]];
                xi = k0 + k < len ? tk : xi;
              }
            }
            xs[r] = xi / dg;
          }
        }
        __syncthreads();
      }
    }
  }
  for (int r = t; r < b.m; r += 256) sub[b.off + r] = xs[r];
}

__global__ void k_gather_positions(int64_t n, const int32_t* __restrict__ idx,
                                   const double* __restrict__ x, double* __restrict__ out) {
  const int64_t p = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (p < n) out[p] = x[idx[p]];
}

}  // namespace

void prepare_sgs(int mmax) {
  const int bytes = mmax * (int)sizeof(double);
  if (bytes > 64 * 1024)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(&k_sgs),
                              hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

void launch_gather_positions(int64_t n, const int32_t* idx, const double* x, double* out,
                             hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_gather_positions, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, n,
                     idx, x, out);
}

void launch_sgs(const SgsDesc* d, int nsub, int mmax, const int32_t* levptr, const int32_t* lrows,
                const int32_t* rbeg, const int32_t* rlen, const int32_t* ecol, const double* eval,
                const double* diag, double* sub, int iters, hipStream_t st) {
  if (nsub <= 0) return;
  const size_t bytes = (size_t)std::max(mmax, 1) * sizeof(double);
  hipLaunchKernelGGL(k_sgs, dim3(nsub), dim3(256), bytes, st, d, levptr, lrows, rbeg, rlen, ecol,
                     eval, diag, sub, iters);
}

void launch_spmv(int nrb, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                 const double* val, const double* x, double* y, hipStream_t st, int shape) {
  if (nrb <= 0) return;
  const dim3 g(nrb);
  switch (shape) {
    case 0: hipLaunchKernelGGL((k_spmv<128, false>), g, dim3(128), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
    case 1: hipLaunchKernelGGL((k_spmv<128, true>), g, dim3(128), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
    case 3: hipLaunchKernelGGL((k_spmv<256, true>), g, dim3(256), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
    case 4: hipLaunchKernelGGL((k_spmv<512, false>), g, dim3(512), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
    case 5: hipLaunchKernelGGL((k_spmv<512, true>), g, dim3(512), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
    default: hipLaunchKernelGGL((k_spmv<256, false>), g, dim3(256), 0, st, rowblk, ptr, col, val, x, y, nrb); break;
  }
}

template <typename TW>
void launch_gemv_w(const TW* W, const BlockDesc* d, int nt, const double* c, double* u,
                   double* part, hipStream_t st, const int32_t* perm) {
  const int nch = gemv_chunks(nt);
  hipLaunchKernelGGL(k_gemv_w_part<TW>, dim3(nch, nt), dim3(256), 0, st, W, d, c, perm, part, nch);
  hipLaunchKernelGGL(k_gemv_sum, dim3(cdiv(nt * kTile, 256)), dim3(256), 0, st, part, nt, nch, 0, u);
}

template <typename TW>
void launch_gemv_wt(const TW* W, const BlockDesc* d, int nt, const double* u, double* c,
                    double* part, hipStream_t st) {
  const int nch = gemv_chunks(nt);
  hipLaunchKernelGGL(k_gemv_wt_part<TW>, dim3(nch, nt), dim3(256), 0, st, W, d, u, part, nch);
  hipLaunchKernelGGL(k_gemv_sum, dim3(cdiv(nt * kTile, 256)), dim3(256), 0, st, part, nt, nch, 1, c);
}
template void launch_gemv_w<double>(const double*, const BlockDesc*, int, const double*, double*,
                                    double*, hipStream_t, const int32_t*);
template void launch_gemv_w<float>(const float*, const BlockDesc*, int, const double*, double*,
                                   double*, hipStream_t, const int32_t*);
template void launch_gemv_wt<double>(const double*, const BlockDesc*, int, const double*, double*,
                                     double*, hipStream_t);
template void launch_gemv_wt<float>(const float*, const BlockDesc*, int, const double*, double*,
                                    double*, hipStream_t);

void launch_to_float(const double* src, float* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_to_float, dim3(cdiv(n, 256)), dim3(256), 0, st, src, dst, n);
}

void launch_spmv_wave_rows(int nrows, const int32_t* ptr, const int32_t* col, const double* val,
                           const double* x, double* y, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_spmv_wave_rows, dim3(cdiv(nrows, 4)), dim3(256), 0, st, nrows, ptr, col, val, x, y);
}

void launch_overlap_sum(int n, const int32_t* occ_ptr, const int32_t* occ_pos, const double* sol,
                        const int32_t* rt_ptr, const int32_t* rt_col, const double* rt_val,
                        const double* csol, double* z, hipStream_t st) {
  hipLaunchKernelGGL(k_overlap_sum, dim3(cdiv(n, 256)), dim3(256), 0, st, n, occ_ptr, occ_pos,
                     sol, rt_ptr, rt_col, rt_val, csol, z);
}

void launch_dot(const double* a, const double* b, int n, double* partial, double* out, int mode,
                double* scal, hipStream_t st) {
  const int nb = (int)cdiv(n, kVecChunk);
  hipLaunchKernelGGL(k_dot_chunks, dim3(nb), dim3(256), 0, st, a, b, n, partial);
  hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(1024), 0, st, partial, nb, out, mode, scal);
}

void launch_p_update(int n, const double* z, double* p, const double* scal, int first,
                     hipStream_t st) {
  hipLaunchKernelGGL(k_p_update_chunks, dim3(cdiv(n, kVecChunk)), dim3(256), 0, st, n, z, p, scal,
                     first);
}

void launch_xr_update(int n, double* x, double* r, const double* p, const double* q,
                      double* partial, double* scal, hipStream_t st, double* hist) {
  const int nb = (int)cdiv(n, kVecChunk);
  hipLaunchKernelGGL(k_xr_update_chunks, dim3(nb), dim3(256), 0, st, n, x, r, p, q, scal, partial);
  hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(1024), 0, st, partial, nb, hist, 3, scal);
}

void launch_overlap_sum_dot(int n, const int32_t* occ_ptr, const int32_t* occ_pos, const double* sol,
                            const int32_t* rt_ptr, const int32_t* rt_col, const double* rt_val,
                            const double* csol, double* z, const double* r, double* partial,
                            double* scal, hipStream_t st) {
  const int nb = (int)cdiv(n, 256);
  hipLaunchKernelGGL(k_overlap_sum_dot, dim3(nb), dim3(256), 0, st, n, occ_ptr, occ_pos, sol, rt_ptr,
                     rt_col, rt_val, csol, z, r, partial);
  hipLaunchKernelGGL(k_reduce_finish, dim3(1), dim3(1024), 0, st, partial, nb, nullptr, 1, scal);
}

void launch_copy(double* dst, const double* src, int n, hipStream_t st) {
  hipLaunchKernelGGL(k_copy, dim3(cdiv(n, 256)), dim3(256), 0, st, dst, src, n);
}

void launch_repack(double* dst, int ldd, const double* src, int lds, int64_t n, int ncols,
                   hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_repack, dim3(cdiv(n * ldd, 256)), dim3(256), 0, st, dst, ldd, src, lds, n, ncols);
}

void launch_zero(double* dst, int64_t n, hipStream_t st) {
  if (n <= 0) return;
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, dst, n);
}

void launch_residual(int n, const double* b, const double* ax, double* r, hipStream_t st) {
  hipLaunchKernelGGL(k_residual, dim3(cdiv(n, 256)), dim3(256), 0, st, n, b, ax, r);
}

template <typename T>
void launch_trsm_multi(const T* L, const BlockDesc* d, int nblocks, const int32_t* inds,
                       const double* rhs, int ldr, T* U, double* out, int ldo, int nrhs, int NR,
                       int kb, int cg, hipStream_t st) {
  if (nblocks <= 0) return;
  const dim3 b(256);
  if (NR == 32 && cg == 3) {  // one eight-wave workgroup, two halves of 16 columns
    if (kb == 1)
      hipLaunchKernelGGL((k_trsm_multi<T, 32, 1, 1, 8>), dim3(nblocks), dim3(512), 0, st, L, d, nblocks, inds, rhs,
                         ldr, U, out, ldo, nrhs);
    else
      hipLaunchKernelGGL((k_trsm_multi<T, 32, 2, 1, 8>), dim3(nblocks), dim3(512), 0, st, L, d, nblocks, inds, rhs,
                         ldr, U, out, ldo, nrhs);
    return;
  }
  if (NR == 32 && cg == 2) {  // two groups of 16 columns
    const dim3 g(cdiv(nblocks, 8) * 16);
    if (kb == 1)
      hipLaunchKernelGGL((k_trsm_multi<T, 16, 1, 2>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
    else
      hipLaunchKernelGGL((k_trsm_multi<T, 16, 2, 2>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
    return;
  }
  const dim3 g(nblocks);
  if (NR == 16 && kb == 1)
    hipLaunchKernelGGL((k_trsm_multi<T, 16, 1, 1>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
  else if (NR == 16)
    hipLaunchKernelGGL((k_trsm_multi<T, 16, 2, 1>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
  else if (kb == 1)
    hipLaunchKernelGGL((k_trsm_multi<T, 32, 1, 1>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
  else
    hipLaunchKernelGGL((k_trsm_multi<T, 32, 2, 1>), g, b, 0, st, L, d, nblocks, inds, rhs, ldr, U, out, ldo, nrhs);
}

void launch_wgemm_fwd(const double* W, const BlockDesc* d, int nt, const double* G, double* U,
                      double* part, int NR, hipStream_t st, const int32_t* perm) {
  const int nch = gemv_chunks(nt);
  const dim3 g(nch, nt), gs(cdiv((int64_t)nt * kTile * NR, 256));
  if (NR == 16) {
    hipLaunchKernelGGL(k_wgemm_fwd_part<16>, g, dim3(256), 0, st, W, d, G, perm, part, nch);
    hipLaunchKernelGGL(k_panel_sum<16>, gs, dim3(256), 0, st, part, nt, nch, 0, U, NR, nt * kTile, NR);
  } else {
    hipLaunchKernelGGL(k_wgemm_fwd_part<32>, g, dim3(256), 0, st, W, d, G, perm, part, nch);
    hipLaunchKernelGGL(k_panel_sum<32>, gs, dim3(256), 0, st, part, nt, nch, 0, U, NR, nt * kTile, NR);
  }
}

void launch_wgemm_bwd(const double* W, const BlockDesc* d, int nt, int rows, const double* U,
                      double* out, int ldo, int nrhs, double* part, int NR, hipStream_t st) {
  const int nch = gemv_chunks(nt);
  const dim3 g(nch, nt), gs(cdiv((int64_t)nt * kTile * NR, 256));
  if (NR == 16) {
    hipLaunchKernelGGL(k_wgemm_bwd_part<16>, g, dim3(256), 0, st, W, d, U, part, nch);
    hipLaunchKernelGGL(k_panel_sum<16>, gs, dim3(256), 0, st, part, nt, nch, 1, out, ldo, rows, nrhs);
  } else {
    hipLaunchKernelGGL(k_wgemm_bwd_part<32>, g, dim3(256), 0, st, W, d, U, part, nch);
    hipLaunchKernelGGL(k_panel_sum<32>, gs, dim3(256), 0, st, part, nt, nch, 1, out, ldo, rows, nrhs);
  }
}

#define DOUG_INSTANTIATE_MULTI(T)                                                              \
  template void launch_trsm_multi<T>(const T*, const BlockDesc*, int, const int32_t*,          \
                                     const double*, int, T*, double*, int, int, int, int, int,  \
                                     hipStream_t);
DOUG_INSTANTIATE_MULTI(double)
DOUG_INSTANTIATE_MULTI(float)
#undef DOUG_INSTANTIATE_MULTI

void launch_spmv_multi(int nrb, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                       const double* val, const double* x, int ldx, double* y, int ldy, int nrhs,
                       hipStream_t st) {
  if (nrb <= 0) return;
  if (nrhs % 2 == 0 && ldx % 2 == 0 && ldy % 2 == 0)
    hipLaunchKernelGGL(k_spmv_multi2, dim3(nrb), dim3(kSpmvThreads), 0, st, rowblk, ptr, col, val, x,
                       ldx, y, ldy, nrhs);
  else
    hipLaunchKernelGGL(k_spmv_multi, dim3(nrb), dim3(kSpmvThreads), 0, st, rowblk, ptr, col, val, x,
                       ldx, y, ldy, nrhs);
}

void launch_spmv_multi_rows(int nrows, const int32_t* ptr, const int32_t* col, const double* val,
                            const double* x, int ldx, double* y, int ldy, int nrhs, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_spmv_multi_rows, dim3(nrows), dim3(256), 0, st, ptr, col, val, x, ldx, y, ldy,
                     nrhs);
}

void launch_overlap_sum_multi(int n, int nrhs, const int32_t* occ_ptr, const int32_t* occ_pos,
                              const double* sol, const int32_t* rt_ptr, const int32_t* rt_col,
                              const double* rt_val, const double* csol, int ldc, double* z,
                              hipStream_t st) {
  if (nrhs % 2 == 0 && ldc % 2 == 0)
    hipLaunchKernelGGL(k_overlap_sum_multi2, dim3(cdiv((int64_t)n * (nrhs / 2), 256)), dim3(256), 0, st,
                       n, nrhs, occ_ptr, occ_pos, sol, rt_ptr, rt_col, rt_val, csol, ldc, z);
  else
    hipLaunchKernelGGL(k_overlap_sum_multi, dim3(cdiv((int64_t)n * nrhs, 256)), dim3(256), 0, st, n,
                       nrhs, occ_ptr, occ_pos, sol, rt_ptr, rt_col, rt_val, csol, ldc, z);
}

void launch_dot_multi(const double* a, const double* b, int n, int nrhs, double* partial,
                      unsigned* counter, int mode, double* scal, hipStream_t st) {
  if (nrhs % 2 == 0) {
    hipLaunchKernelGGL(k_dot_multi2, dim3(kRedBlocksWide), dim3(256), 0, st,
                       reinterpret_cast<const double2*>(a), reinterpret_cast<const double2*>(b),
                       (int64_t)n * nrhs / 2, nrhs, partial);
    hipLaunchKernelGGL(k_finish_cols, dim3(nrhs), dim3(256), 0, st, partial, kRedBlocksWide, mode, scal);
  } else
    hipLaunchKernelGGL(k_dot_multi, dim3(kRedBlocks), dim3(256), 0, st, a, b, (int64_t)n * nrhs,
                       nrhs, partial, counter, mode, scal);
}

void launch_p_update_multi(int n, int nrhs, const double* z, double* p, const double* scal,
                           int first, hipStream_t st) {
  int64_t total = (int64_t)n * nrhs;
  hipLaunchKernelGGL(k_p_update_multi, dim3(cdiv(total, 256)), dim3(256), 0, st, total, nrhs, z, p,
                     scal, first);
}

void launch_xr_update_multi(int n, int nrhs, double* x, double* r, const double* p,
                            const double* q, double* partial, unsigned* counter, double* scal,
                            hipStream_t st) {
  if (nrhs % 2 == 0) {  // x's update is deferred to launch_spmv_p_multi2 / launch_x_final_multi2
    hipLaunchKernelGGL(k_xr_update_multi2, dim3(kRedBlocksWide), dim3(256), 0, st, (int64_t)n * nrhs / 2,
                       nrhs, reinterpret_cast<double2*>(r), reinterpret_cast<const double2*>(q), partial, scal);
    hipLaunchKernelGGL(k_finish_cols, dim3(nrhs), dim3(256), 0, st, partial, kRedBlocksWide, 3, scal);
  } else
    hipLaunchKernelGGL(k_xr_update_multi, dim3(kRedBlocks), dim3(256), 0, st, (int64_t)n * nrhs, nrhs,
                       x, r, p, q, partial, counter, scal);
}

void launch_overlap_sum_dot_multi2(int n, int nrhs, const int32_t* occ_ptr, const int32_t* occ_pos,
                                   const double* sol, const int32_t* rt_ptr, const int32_t* rt_col,
                                   const double* rt_val, const double* csol, int ldc, double* z,
                                   const double* r, double* partial, unsigned* counter, double* scal,
                                   hipStream_t st) {
  hipLaunchKernelGGL(k_overlap_sum_dot_multi2, dim3(kRedBlocksWide), dim3(256), 0, st, n, nrhs, occ_ptr,
                     occ_pos, sol, rt_ptr, rt_col, rt_val, csol, ldc, z, r, partial);
  hipLaunchKernelGGL(k_finish_cols, dim3(nrhs), dim3(256), 0, st, partial, kRedBlocksWide, 1, scal);
}

void launch_spmv_p_multi2(int nrb, const int32_t* rowblk, const int32_t* ptr, const int32_t* col,
                          const double* val, const double* z, const double* pold, double* pnew, double* q,
                          int nrhs, int first, double* partial, unsigned* counter, double* scal, double* x,
                          hipStream_t st) {
  hipLaunchKernelGGL(k_spmv_p_multi2, dim3(nrb), dim3(kSpmvThreads), 0, st, rowblk, ptr, col, val, z, pold,
                     pnew, q, nrhs, first, partial, scal, x);
  hipLaunchKernelGGL(k_finish_cols, dim3(nrhs), dim3(256), 0, st, partial, nrb, 2, scal);
}

void launch_x_final_multi2(int n, int nrhs, double* x, const double* p, const double* scal, hipStream_t st) {
  const int64_t total2 = (int64_t)n * nrhs / 2;
  hipLaunchKernelGGL(k_x_final_multi2, dim3(cdiv(total2, 256)), dim3(256), 0, st, total2, nrhs,
                     reinterpret_cast<double2*>(x), reinterpret_cast<const double2*>(p), scal);
}

void launch_value_grad(int n, const int32_t* ptr, const int32_t* col, const int32_t* src,
                       const double* lam, const double* x, double* grad, hipStream_t st) {
  hipLaunchKernelGGL(k_value_grad, dim3(cdiv(n, 256)), dim3(256), 0, st, n, ptr, col, src, lam, x, grad);
}

}  // namespace doug
