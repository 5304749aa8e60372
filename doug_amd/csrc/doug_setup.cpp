// Host setup of the two-level Schwarz preconditioner; see doug_setup.hpp.
// Reference citations are relative to the DOUG source tree.
#include "doug_setup.hpp"

#include <algorithm>
#include <atomic>
#include <cmath>
#include <numeric>
#include <queue>
#include <tuple>
#include <thread>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <stdexcept>

namespace doug {

static constexpr int kMaxInt = 2147483647;            // parameters.F90:99
static constexpr int kAggregated = kMaxInt;           // parameters.F90:101
static constexpr int kPending = kMaxInt / 2;          // parameters.F90:102

// Row-parallel loops of the setup: static contiguous chunks, one per thread,
// so that per-chunk outputs concatenate in row order.  Results do not
// depend on the thread count.
static thread_local int t_thread_cap = 0;

void set_setup_thread_cap(int cap) { t_thread_cap = cap; }

int setup_threads() {
  unsigned hw = std::thread::hardware_concurrency();
  const int n = (int)std::max(1u, std::min(hw ? hw : 1u, 16u));
  return t_thread_cap > 0 ? std::min(n, t_thread_cap) : n;
}

template <class F>
static void parallel_chunks(int n, int nthreads, F&& f) {
  if (nthreads <= 1 || n < 4096) {
    f(0, 0, n);
    return;
  }
  std::vector<std::thread> th;
  for (int t = 0; t < nthreads; t++) {
    int b = (int)((int64_t)n * t / nthreads), e = (int)((int64_t)n * (t + 1) / nthreads);
    th.emplace_back([&f, t, b, e]() { f(t, b, e); });
  }
  for (auto& x : th) x.join();
}

template <class F>
static void parallel_rows(int n, F&& f) {
  parallel_chunks(n, setup_threads(), [&f](int, int b, int e) {
    for (int i = b; i < e; i++) f(i);
  });
}

// Concatenate per-chunk CSR pieces (rows in chunk order).
struct CsrPiece {
  std::vector<int> rowlen, col;
  std::vector<double> val;
};
static void join_pieces(std::vector<CsrPiece>& pieces, int nrows, Csr& M) {
  M.ptr.assign((size_t)nrows + 1, 0);
  size_t total = 0;
  for (auto& p : pieces) total += p.col.size();
  M.col.resize(total);
  M.val.resize(total);
  int row = 0;
  size_t off = 0;
  for (auto& p : pieces) {
    for (int len : p.rowlen) {
      M.ptr[row + 1] = M.ptr[row] + len;
      row++;
    }
    std::copy(p.col.begin(), p.col.end(), M.col.begin() + off);
    std::copy(p.val.begin(), p.val.end(), M.val.begin() + off);
    off += p.col.size();
  }
}

Csr coo_to_csr(int n, int64_t nnz, const int32_t* indi, const int32_t* indj, const double* val,
               int base, std::string& err) {
  Csr A;
  A.nrows = A.ncols = n;
  A.ptr.assign((size_t)n + 1, 0);
  for (int64_t t = 0; t < nnz; t++) {
    int64_t i = (int64_t)indi[t] - base, j = (int64_t)indj[t] - base;
    if (i < 0 || i >= n || j < 0 || j >= n) {
      err = "matrix entry " + std::to_string(t) + " out of range";
      return Csr{};
    }
    A.ptr[(size_t)i + 1]++;
  }
  for (int i = 0; i < n; i++) A.ptr[i + 1] += A.ptr[i];
  A.col.resize((size_t)nnz);
  A.val.resize((size_t)nnz);
  A.src.resize((size_t)nnz);
  std::vector<int> fill(A.ptr.begin(), A.ptr.end() - 1);
  for (int64_t t = 0; t < nnz; t++) {
    int i = indi[t] - base;
    int p = fill[i]++;
    A.col[p] = indj[t] - base;
    A.val[p] = val[t];
    A.src[p] = (int)t;
  }
  return A;
}

void scale_values(const Csr& A, bool symm, std::vector<double>& scaled, std::vector<double>& diag) {
  int n = std::max(A.nrows, A.ncols);
  diag.assign((size_t)n, 0.0);
  parallel_rows(A.nrows, [&](int i) {
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++)
      if (A.col[p] == i) diag[i] = A.val[p];
  });
  scaled.resize(A.val.size());
  std::vector<double> sv((size_t)n);
  for (int i = 0; i < n; i++) sv[i] = symm ? std::sqrt(std::fabs(diag[i])) : std::fabs(diag[i]);
  parallel_rows(A.nrows, [&](int i) {
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      double v = A.val[p] / sv[i];
      scaled[p] = symm ? v / sv[A.col[p]] : v;
    }
  });
}

void scale_round_trip(const Csr& A, const std::vector<double>& input, bool symm,
                      std::vector<double>& scaled, std::vector<double>& diag,
                      std::vector<double>& round_trip) {
  const int n = std::max(A.nrows, A.ncols);
  diag.assign((size_t)n, 0.0);
  std::vector<double> sv((size_t)n, 0.0);
  parallel_rows(A.nrows, [&](int i) {
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++)
      if (A.col[p] == i) diag[i] = input[p];
    sv[i] = symm ? std::sqrt(std::fabs(diag[i])) : std::fabs(diag[i]);
  });
  scaled.resize(input.size());
  round_trip.resize(input.size());
  parallel_rows(A.nrows, [&](int i) {
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      const double si = sv[i], sj = sv[A.col[p]];
      double v = input[p] / si;
      const double s = symm ? v / sj : v;
      scaled[p] = s;
      v = s * si;
      round_trip[p] = symm ? v * sj : v;
    }
  });
}

void unscale_values(const Csr& A, bool symm, const std::vector<double>& diag,
                    std::vector<double>& vals) {
  int n = (int)diag.size();
  std::vector<double> sv((size_t)n);
  for (int i = 0; i < n; i++) sv[i] = symm ? std::sqrt(std::fabs(diag[i])) : std::fabs(diag[i]);
  parallel_rows(A.nrows, [&](int i) {
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      double v = vals[p] * sv[i];
      vals[p] = symm ? v * sv[A.col[p]] : v;
    }
  });
}

std::vector<uint8_t> find_strong(const Csr& A, const std::vector<double>& scaled,
                                 const std::vector<double>& diag, double alpha) {
  std::vector<uint8_t> strong((size_t)A.nnz(), 0);
  parallel_rows(A.nrows, [&](int i) {
    // single-precision literal in the reference (SpMtx_aggregation.F90:867)
    double maxndiag = (double)(-1.0e15f);
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++)
      if (A.col[p] != i) maxndiag = std::max(maxndiag, std::fabs(scaled[p]));
    maxndiag = maxndiag * alpha;
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      if (A.col[p] != i)
        strong[p] = std::fabs(scaled[p]) > maxndiag;
      else
        strong[p] = diag[i] > maxndiag;
    }
  });
  return strong;
}

// Compare-exchange of slots A < B: (key, pos) in descending key, ascending
// pos among equal keys; written with selects rather than branches, and
// with constant slot numbers, so that the eight slots stay in registers.
template <int A, int B>
static inline void cx_desc(double (&key)[8], int (&pos)[8]) {
  // the exchange as bit masks: a compiler may turn selects back into
  // branches, which mispredict on these data
  const bool sw = (key[B] > key[A]) | ((key[B] == key[A]) & (pos[B] < pos[A]));
  const uint64_t m = 0 - (uint64_t)sw;
  uint64_t ka, kb;
  std::memcpy(&ka, &key[A], 8);
  std::memcpy(&kb, &key[B], 8);
  const uint64_t na = (kb & m) | (ka & ~m), nb = (ka & m) | (kb & ~m);
  std::memcpy(&key[A], &na, 8);
  std::memcpy(&key[B], &nb, 8);
  const uint32_t m32 = (uint32_t)m, pa = (uint32_t)pos[A], pb = (uint32_t)pos[B];
  pos[A] = (int)((pb & m32) | (pa & ~m32));
  pos[B] = (int)((pa & m32) | (pb & ~m32));
}

template <int A, int B>
static inline void cx_asc(int64_t (&key)[8]) {
  const int64_t a = key[A], b = key[B];
  key[A] = std::min(a, b);
  key[B] = std::max(a, b);
}

// Eight slots through Batcher's odd-even merge network (19 exchanges).
static inline void sort8_desc(double (&k)[8], int (&p)[8]) {
  cx_desc<0, 1>(k, p), cx_desc<2, 3>(k, p), cx_desc<4, 5>(k, p), cx_desc<6, 7>(k, p);
  cx_desc<0, 2>(k, p), cx_desc<1, 3>(k, p), cx_desc<4, 6>(k, p), cx_desc<5, 7>(k, p);
  cx_desc<1, 2>(k, p), cx_desc<5, 6>(k, p);
  cx_desc<0, 4>(k, p), cx_desc<1, 5>(k, p), cx_desc<2, 6>(k, p), cx_desc<3, 7>(k, p);
  cx_desc<2, 4>(k, p), cx_desc<3, 5>(k, p);
  cx_desc<1, 2>(k, p), cx_desc<3, 4>(k, p), cx_desc<5, 6>(k, p);
}

// Eight distinct keys, ascending, by the same network.
static inline void sort8_asc(int64_t (&k)[8]) {
  cx_asc<0, 1>(k), cx_asc<2, 3>(k), cx_asc<4, 5>(k), cx_asc<6, 7>(k);
  cx_asc<0, 2>(k), cx_asc<1, 3>(k), cx_asc<4, 6>(k), cx_asc<5, 7>(k);
  cx_asc<1, 2>(k), cx_asc<5, 6>(k);
  cx_asc<0, 4>(k), cx_asc<1, 5>(k), cx_asc<2, 6>(k), cx_asc<3, 7>(k);
  cx_asc<2, 4>(k), cx_asc<3, 5>(k);
  cx_asc<1, 2>(k), cx_asc<3, 4>(k), cx_asc<5, 6>(k);
}

// SpMtx_build_refs_symm with sortdown (SpMtx_arrangement.F90:461-568):
// strong off-diagonal neighbours per row, |scaled| descending, ties keep
// their row order.
static void strong_refs(const Csr& A, const std::vector<double>& scaled,
                        const std::vector<uint8_t>& strong, std::vector<int>& srow,
                        std::vector<int>& scol) {
  int n = A.nrows;
  srow.assign((size_t)n + 1, 0);
  parallel_rows(n, [&](int i) {
    int c = 0;
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++)
      if (strong[p] && A.col[p] != i && A.col[p] < n) c++;
    srow[i + 1] = c;
  });
  for (int i = 0; i < n; i++) srow[i + 1] += srow[i];
  scol.resize((size_t)srow[n]);
  parallel_chunks(n, setup_threads(), [&](int, int b, int e) {
    std::vector<int> tmp;
    for (int i = b; i < e; i++) {
      const int cnt = srow[i + 1] - srow[i];
      const int p0 = A.ptr[i], len = A.ptr[i + 1] - p0;
      if (len <= 8) {
        // |scaled| descending, ties in row order: a strict order on
        // (|scaled|, position), so a sorting network over the row's (at most
        // eight) entries gives the stable order; entries that are not strong
        // off-diagonal neighbours get the key -1 and sort after all others.
        // Slots are filled by constant index (no unpredictable branches and
        // no stores at computed offsets).
        double key[8];
        int pos[8];
#define DOUG_SLOT(e)                                                              \
  {                                                                               \
    const int q = e < len ? p0 + e : p0;                                          \
    const int c = A.col[q];                                                       \
    const bool ok = (e < len) & (strong[q] != 0) & (c != i) & (c < n);            \
    key[e] = ok ? std::fabs(scaled[q]) : -1.0;                                    \
    pos[e] = q;                                                                   \
  }
        DOUG_SLOT(0) DOUG_SLOT(1) DOUG_SLOT(2) DOUG_SLOT(3)
        DOUG_SLOT(4) DOUG_SLOT(5) DOUG_SLOT(6) DOUG_SLOT(7)
#undef DOUG_SLOT
        sort8_desc(key, pos);
        // read back by constant index, so that key and pos stay in registers
        const int res[8] = {pos[0], pos[1], pos[2], pos[3], pos[4], pos[5], pos[6], pos[7]};
        int* out = scol.data() + srow[i];
        const int* col = A.col.data();
        for (int q = 0; q < cnt; q++) out[q] = col[res[q]];
        continue;
      }
      tmp.clear();
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++)
        if (strong[p] && A.col[p] != i && A.col[p] < n) tmp.push_back(p);
      // stable insertion sort, |scaled| descending
      for (size_t q = 1; q < tmp.size(); q++) {
        const int e = tmp[q];
        const double ve = std::fabs(scaled[e]);
        size_t r = q;
        for (; r > 0 && ve > std::fabs(scaled[tmp[r - 1]]); r--) tmp[r] = tmp[r - 1];
        tmp[r] = e;
      }
      for (size_t k = 0; k < tmp.size(); k++) scol[srow[i] + k] = A.col[tmp[k]];
    }
  });
}

namespace {

// Growing layers around a seed: lets_colour / can_add_roundlayer
// (Aggregate_mod.F90:287-497).  List positions are 0-based here; node ids are
// 0-based; `stat` keeps the reference's encoding.
struct Grower {
  const std::vector<int>& rs;
  const std::vector<int>& cn;
  std::vector<int>& stat;
  std::vector<int>& distance;
  std::vector<int>& nodes;
  int& allmind;
  int neighood, maxasize;
  std::vector<int> inlayer, newlayer, newdist, rounder;
  int nneigs = 0;

  Grower(const std::vector<int>& rs_, const std::vector<int>& cn_, std::vector<int>& st,
         std::vector<int>& di, std::vector<int>& nd, int& am, int nh, int mx)
      : rs(rs_), cn(cn_), stat(st), distance(di), nodes(nd), allmind(am), neighood(nh),
        maxasize(mx) {}

  void push(std::vector<int>& v, int& cnt, int x) {
    if ((size_t)cnt >= v.size()) v.resize(v.size() * 2 + 16);
    v[cnt++] = x;
  }

  int round_layer(int layer, int& ninnodes) {
    // An empty layer changes nothing (no hits, allmind untouched, the size
    // check holds since every earlier inner layer passed it) and leaves the
    // next layer empty: the recursion would end in 3.
    if (ninnodes == 0) return 3;
    const int nxtlayer = layer + 1;
    bool layerfits = true;
    int ntoadd = 0, nrounders = 0;
    if (layer <= 2 * neighood) {
      // each node not yet placed is counted once per hit; a fresh one (stat
      // 0) joins the new layer at its first hit, with the distance of its
      // discoverer + 1.  The new layer holds distinct nodes here, so n + 1
      // slots always suffice.  Written without branches (the hit pattern is
      // not predictable); the distances go to a side list first and into
      // `distance` after the scan (nothing in the scan reads them).
      int* nl = newlayer.data();
      int* nd = newdist.data();
      int* st = stat.data();
      const int* di = distance.data();
      const int* rsp = rs.data();
      const int* cnp = cn.data();
      for (int k = 0; k < ninnodes; k++) {
        const int ln = inlayer[k];
        const int dn = di[ln] + 1;
        const int e = rsp[ln + 1];
        for (int i = rsp[ln]; i < e; i++) {
          const int c = cnp[i];
          const int s = st[c];
          nl[ntoadd] = c;
          nd[ntoadd] = dn;
          ntoadd += s == 0;
          st[c] = s - (s <= 0);
        }
      }
      for (int k = 0; k < ntoadd; k++) distance[nl[k]] = nd[k];
    }
    if (layer <= neighood) {
      int mind = kMaxInt, maxd = 0;
      for (int k = 0; k < ntoadd; k++) {
        int d = -stat[newlayer[k]];
        mind = std::min(mind, d);
        maxd = std::max(maxd, d);
      }
      if (mind < allmind) allmind = mind;
      if (mind < maxd || mind > allmind) {
        for (int k = 0; k < ntoadd; k++) {
          if (stat[newlayer[k]] == -maxd) {
            nrounders++;
            if (nneigs + ninnodes + 1 > maxasize) {
              layerfits = false;
              break;
            }
            push(inlayer, ninnodes, newlayer[k]);
            if ((int)rounder.size() < nrounders) rounder.resize(nrounders);
            rounder[nrounders - 1] = newlayer[k];
            newlayer[k] = -1;
          }
        }
      }
      if (!layerfits) nrounders = 0;
      int j = 0;
      for (int i = 0; i < ntoadd; i++)
        if (newlayer[i] != -1) newlayer[j++] = newlayer[i];
      ntoadd = j;
      for (int k = 0; k < nrounders; k++) {
        int ln = rounder[k];
        for (int i = rs[ln]; i < rs[ln + 1]; i++) {
          int c = cn[i];
          if (stat[c] == 0) {
            push(newlayer, ntoadd, c);
            distance[c] = distance[ln] + 1;
          }
        }
      }
    }
    if (layer <= neighood && nneigs + ninnodes > maxasize) layerfits = false;
    // a rounder's fresh neighbours are appended without being marked, so a
    // node can enter a layer twice and the explored list can outgrow n when
    // the layers cover the whole graph (the reference's nodes(n) would
    // overflow there): grow the list instead
    if (nodes.size() < (size_t)nneigs + (size_t)ninnodes) nodes.resize(2 * ((size_t)nneigs + ninnodes));
    if (layerfits) {
      for (int k = 0; k < ninnodes; k++) {
        stat[inlayer[k]] = layer + 1;
        nodes[nneigs++] = inlayer[k];
      }
      if (layer <= 2 * neighood) {
        for (int k = 0; k < ntoadd; k++) stat[newlayer[k]] = kPending - 1;
        ninnodes = ntoadd;
        if ((int)inlayer.size() < ntoadd) inlayer.resize(ntoadd);
        for (int k = 0; k < ntoadd; k++) inlayer[k] = newlayer[k];
        return round_layer(nxtlayer, ninnodes);
      }
      return 3;
    }
    for (int k = 0; k < ninnodes; k++) distance[inlayer[k]] = 0;
    for (int k = 0; k < ntoadd; k++) distance[newlayer[k]] = 0;
    for (int k = 0; k < nneigs; k++) distance[nodes[k]] = 0;
    for (int k = 0; k < ninnodes; k++) nodes[nneigs++] = inlayer[k];
    if (nneigs > 0) return layer > neighood ? 2 : 1;
    return 0;
  }

  int colour(int seed) {
    nneigs = 0;
    stat[seed] = 1;
    distance[seed] = 0;
    // A seed whose strong neighbours are all taken grows no layer: every
    // round_layer level is then empty, allmind is untouched (its candidate
    // minimum stays kMaxInt) and the result is the seed alone with status 3.
    bool free_neighbour = false;
    for (int i = rs[seed]; i < rs[seed + 1] && !free_neighbour; i++) free_neighbour = stat[cn[i]] <= 0;
    if (!free_neighbour) {
      nodes[nneigs++] = seed;
      return 3;
    }
    int ninnodes = 1;
    if (inlayer.empty()) inlayer.resize(64);
    if (newlayer.size() < stat.size() + 1) newlayer.resize(stat.size() + 1);
    if (newdist.size() < stat.size() + 1) newdist.resize(stat.size() + 1);
    inlayer[0] = seed;
    return round_layer(0, ninnodes);
  }
};

// The reference's `nextgoodstart` list (SpMtx_aggregation.F90:185-223,
// 291-314): after each aggregate its best outer-layer nodes are appended;
// before each seed choice the nodes aggregated meanwhile are dropped (order
// kept); the seed is the first entry that repeats an earlier entry, else the
// first entry.  An aggregated node never returns to the list, so the same
// choice follows from append positions alone: the first repeat is the node
// whose second occurrence came earliest (a heap), the first entry is the
// oldest live one (a FIFO).  O(log n) per step instead of a scan of the list.
struct SeedList {
  std::vector<int> cnt;
  std::vector<std::pair<int64_t, int>> fifo;
  size_t head = 0;
  std::priority_queue<std::pair<int64_t, int>, std::vector<std::pair<int64_t, int>>,
                      std::greater<std::pair<int64_t, int>>>
      rep;
  int64_t pos = 0, live = 0;
  explicit SeedList(int n) : cnt((size_t)n, 0) {}
  void append(int v) {
    ++pos;
    ++live;
    fifo.emplace_back(pos, v);
    if (++cnt[v] == 2) rep.emplace(pos, v);
  }
  void aggregated(int v) {
    live -= cnt[v];
    cnt[v] = 0;
  }
  int choose(const std::vector<int>& stat) {
    while (!rep.empty() && stat[rep.top().second] >= kPending) rep.pop();
    if (!rep.empty()) return rep.top().second;
    while (stat[fifo[head].second] >= kPending) head++;
    return fifo[head].second;
  }
};

void form_aggr(Aggregates& ag, int nagrs, int n, const std::vector<int>& num) {
  ag.nagr = nagrs;
  ag.num = num;
  ag.starts.assign((size_t)nagrs + 1, 0);
  for (int i = 0; i < n; i++)
    if (num[i] > 0) ag.starts[num[i]]++;
  for (int a = 0; a < nagrs; a++) ag.starts[a + 1] += ag.starts[a];
  ag.nodes.resize((size_t)ag.starts[nagrs]);
  std::vector<int> fill(ag.starts.begin(), ag.starts.end() - 1);
  for (int i = 0; i < n; i++)
    if (num[i] > 0) ag.nodes[fill[num[i] - 1]++] = i;
}

}  // namespace

void aggregate(const Csr& A, const std::vector<double>& scaled, const std::vector<uint8_t>& strong,
               int neighood, int minasize, int maxasize, double alpha, bool coarse_level,
               ProcessState& ps, Aggregates& inner, Aggregates& full) {
  const int n = A.nrows;
  std::vector<int> aggrnum((size_t)n, 0);
  int nagrs = 0;
  bool toosmall = false;
  int maxasizelargest = 0;
  int sn = (int)std::sqrt((float)n);  // single-precision sqrt (SpMtx_aggregation.F90:122)
  if (alpha >= 0 || sn * sn != n) {
    std::vector<int> srow, scol;
    strong_refs(A, scaled, strong, srow, scol);
    std::vector<int> stat((size_t)n, 0), nodes((size_t)n + 1, 0), distance((size_t)n, 0);
    maxasizelargest = coarse_level ? 3 * maxasize : maxasize + 4 * (2 * neighood);
    if (!coarse_level) ps.beta = alpha / 4.0;
    SeedList seeds(n);
    std::vector<int> outer;  // outer-layer candidates of the current aggregate
    std::vector<int> layerlen((size_t)(2 * neighood + 3), 0);
    int firstwocol = 0;
    Grower g(srow, scol, stat, distance, nodes, ps.allmind, neighood, maxasize);
    while (firstwocol < n) {
      int startnode = -1;
      if (seeds.live == 0) {
        bool done = false;
        while (stat[firstwocol] >= kPending) {
          firstwocol++;
          if (firstwocol >= n) {
            done = true;
            break;
          }
        }
        if (done) break;
        startnode = firstwocol;
      } else {
        startnode = seeds.choose(stat);
      }
      int ok = g.colour(startnode);
      int nneigs = g.nneigs;
      int mindistance = kMaxInt;
      if (ok > 0) {
        nagrs++;
        int nouters = 0, next_start_layer;
        if (ok == 3) {
          std::fill(layerlen.begin(), layerlen.end(), 0);
          int maxlayer = 0;
          if (!toosmall && nneigs < minasize) toosmall = true;
          for (int j = nneigs - 1; j >= 0; j--) {
            int k = stat[nodes[j]];
            if (k > maxlayer) maxlayer = k;
            if (k <= neighood + 1) break;
            layerlen[k]++;
          }
          next_start_layer = neighood + 2;
          int k = layerlen[neighood + 2];
          for (int j = neighood + 3; j <= maxlayer; j++)
            if (k < layerlen[j]) {
              k = layerlen[j];
              next_start_layer = j;
            }
        } else if (ok == 2) {
          next_start_layer = neighood + 2;
          if (!toosmall && nneigs < minasize) toosmall = true;
        } else {
          next_start_layer = 0;
        }
        outer.clear();
        for (int j = 0; j < nneigs; j++) {
          int node = nodes[j];
          if (stat[node] <= neighood + 1 && j + 1 <= maxasize) {
            stat[node] = kAggregated;
            aggrnum[node] = nagrs;
            seeds.aggregated(node);
          } else if (next_start_layer > neighood + 1) {
            if (stat[node] == next_start_layer) {
              mindistance = std::min(mindistance, distance[node]);
              nouters++;
              outer.push_back(node);
            }
          }
        }
        // the reference parks them at the tail of nextgoodstart (first one
        // at the very end) and reads them back from the end: discovery order
        for (int j = 0; j < nouters; j++)
          if (distance[outer[j]] == mindistance) seeds.append(outer[j]);
        for (int j = 0; j < nneigs; j++) {
          int node = nodes[j];
          if (stat[node] != kAggregated) {
            distance[node] = 0;
            stat[node] = 0;
          }
        }
      }
    }
  } else {
    // Cartesian aggregates for a perfect-square n (SpMtx_aggregation.F90:333-346)
    int nsa = (sn + neighood - 1) / neighood;
    int k = 0;
    for (int i = 1; i <= sn; i++)
      for (int j = 1; j <= sn; j++)
        aggrnum[k++] = ((i + neighood - 1) / neighood - 1) * nsa + (j + neighood - 1) / neighood;
    nagrs = *std::max_element(aggrnum.begin(), aggrnum.end());
  }

  std::vector<int> fullnum((size_t)n, 0);
  int full_nagrs = 0;
  if (!toosmall) {
    fullnum = aggrnum;
    for (int v : fullnum) full_nagrs = std::max(full_nagrs, v);
    form_aggr(inner, nagrs, n, aggrnum);
  } else {
    // merge too-small aggregates (SpMtx_aggregation.F90:365-591).  Members
    // are kept as lists in insertion order (aggregate 0 = unaggregated), so
    // that a whole aggregate moves to its eater in O(1).
    std::vector<int> head((size_t)nagrs + 1, -1), tail((size_t)nagrs + 1, -1), next((size_t)n, -1),
        msize((size_t)nagrs + 1, 0);
    auto append = [&](int a, int node) {
      next[node] = -1;
      if (tail[a] < 0)
        head[a] = node;
      else
        next[tail[a]] = node;
      tail[a] = node;
      msize[a]++;
    };
    for (int i = 0; i < n; i++) append(aggrnum[i], i);
    std::vector<int> colsaround, mem;
    // position of each neighbouring aggregate in colsaround for the current
    // aggregate (stamped with its number): the same first-seen order and the
    // same sums as a search of colsaround
    std::vector<int> cstamp((size_t)nagrs + 1, 0), cslot((size_t)nagrs + 1, 0);
    std::vector<double> connw;
    int nagrs_new = nagrs;
    const double beta = ps.beta;
    for (int a = 1; a <= nagrs; a++) {
      if (msize[a] >= minasize) continue;
      // The small aggregates lie scattered over the graph and the loop is
      // bound by the latency of their rows: prefetch, a few aggregates
      // ahead, the row bounds (32 ahead), then the rows (16 ahead), then
      // the aggregate numbers of their columns (6 ahead).  Lists of later
      // aggregates can still grow; the prefetches are only hints.
      for (int s = 0; s < 3; s++) {
        const int pa = a + (s == 0 ? 32 : s == 1 ? 16 : 6);
        if (pa > nagrs || msize[pa] >= minasize) continue;
        for (int node = head[pa]; node >= 0; node = next[node]) {
          if (s == 0) {
            __builtin_prefetch(&A.ptr[node]);
          } else if (s == 1) {
            __builtin_prefetch(&A.col[A.ptr[node]]);
            __builtin_prefetch(&scaled[A.ptr[node]]);
          } else {
            for (int p = A.ptr[node]; p < A.ptr[node + 1]; p++)
              if (A.col[p] < n) __builtin_prefetch(&aggrnum[A.col[p]]);
          }
        }
      }
      colsaround.clear();
      connw.clear();
      for (int node = head[a]; node >= 0; node = next[node]) {
        for (int p = A.ptr[node]; p < A.ptr[node + 1]; p++) {
          int c = A.col[p];
          if (c >= n) continue;
          int colr = aggrnum[c];
          if (colr > 0 && colr != a) {
            if (cstamp[colr] == a) {
              connw[cslot[colr]] = connw[cslot[colr]] + std::fabs(scaled[p]);
            } else {
              cstamp[colr] = a;
              cslot[colr] = (int)colsaround.size();
              colsaround.push_back(colr);
              connw.push_back(std::fabs(scaled[p]));
            }
          }
        }
      }
      double maxcw = 0.0;
      int eater = 0;
      for (size_t j = 0; j < colsaround.size(); j++)
        if (maxcw < connw[j] && msize[colsaround[j]] + msize[a] <= maxasize) {
          maxcw = connw[j];
          eater = colsaround[j];
        }
      if (maxcw >= alpha || (coarse_level && maxcw >= beta)) {
        for (int node = head[a]; node >= 0; node = next[node]) aggrnum[node] = eater;
        if (head[a] >= 0) {
          if (tail[eater] < 0)
            head[eater] = head[a];
          else
            next[tail[eater]] = head[a];
          tail[eater] = tail[a];
          msize[eater] += msize[a];
        }
        head[a] = tail[a] = -1;
        msize[a] = 0;
        nagrs_new--;
      } else {
        // give the aggregate away node by node, from its last node back
        mem.clear();
        for (int node = head[a]; node >= 0; node = next[node]) mem.push_back(node);
        int nleft = (int)mem.size();
        int sz = nleft;
        bool reduced = true;
        while (nleft > 0 && reduced) {
          reduced = false;
          for (int k = sz - 1; k >= 0; k--) {
            int node = mem[k];
            if (node < 0) continue;
            for (int p = A.ptr[node]; p < A.ptr[node + 1]; p++) {
              int c = A.col[p];
              if (c >= n) continue;
              int colr = aggrnum[c];
              if (colr != a && msize[colr] < maxasizelargest && std::fabs(scaled[p]) >= beta) {
                append(colr, node);
                aggrnum[node] = colr;
                mem[k] = -1;
                nleft--;
                reduced = true;
                break;
              }
            }
          }
        }
        if (nleft < sz) {  // the nodes left stay, in their order
          head[a] = tail[a] = -1;
          msize[a] = 0;
          for (int j = 0; j < sz; j++)
            if (mem[j] >= 0) append(a, mem[j]);
          if (nleft == 0) nagrs_new--;
        }
      }
    }
    int64_t naggregated = 0;
    for (int a = 1; a <= nagrs; a++) naggregated += (int64_t)msize[a];
    full_nagrs = nagrs_new;
    int cnt = *std::max_element(aggrnum.begin(), aggrnum.end());
    if (cnt > nagrs_new) {
      std::vector<int> st((size_t)cnt + 1, 0);
      int col = 0;
      for (int i = 0; i < n; i++)
        if (aggrnum[i] > 0) {
          if (st[aggrnum[i]] == 0) st[aggrnum[i]] = ++col;
          aggrnum[i] = st[aggrnum[i]];
        }
    }
    if (naggregated == n) {
      fullnum = aggrnum;
      full_nagrs = *std::max_element(fullnum.begin(), fullnum.end());
    } else {
      for (int i = 0; i < n; i++) fullnum[i] = aggrnum[i] > 0 ? aggrnum[i] : ++full_nagrs;
    }
    form_aggr(inner, nagrs_new, n, aggrnum);
  }
  form_aggr(full, full_nagrs, n, fullnum);
}

Csr transpose(const Csr& A) {
  Csr T;
  T.nrows = A.ncols;
  T.ncols = A.nrows;
  const int nth = A.nrows >= 4096 ? setup_threads() : 1;
  // per-chunk column counts give every chunk its write offsets, so each
  // output row lists its entries in ascending source-row order
  std::vector<std::vector<int>> cnt((size_t)nth);
  parallel_chunks(A.nrows, nth, [&](int t, int b, int e) {
    cnt[t].assign((size_t)T.nrows, 0);
    for (int p = A.ptr[b]; p < A.ptr[e]; p++) cnt[t][A.col[p]]++;
  });
  T.ptr.assign((size_t)T.nrows + 1, 0);
  for (int c = 0; c < T.nrows; c++) {
    int tot = 0;
    for (int t = 0; t < nth; t++) {
      int v = cnt[t][c];
      cnt[t][c] = tot;
      tot += v;
    }
    T.ptr[c + 1] = T.ptr[c] + tot;
  }
  T.col.resize(A.col.size());
  T.val.resize(A.val.size());
  parallel_chunks(A.nrows, nth, [&](int t, int b, int e) {
    std::vector<int>& off = cnt[t];
    for (int i = b; i < e; i++)
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
        int c = A.col[p];
        int q = T.ptr[c] + off[c]++;
        T.col[q] = i;
        T.val[q] = A.val[p];
      }
  });
  return T;
}

// IntRestBuild with smoothers >= 2 (CoarseMtx.F90:236-417).  The second
// smoother S2 has the same entries as S but its own diagonal: the loop that
// forms it assigns a_ii when it meets the diagonal entry and adds the weak
// entries met afterwards (CoarseMtx.F90:238-245), so weak entries stored
// before the diagonal drop out (the reference's array is not cleared first;
// a row without a diagonal entry starts from zero here).  SS = S2 S, and the
// restriction is T (SS)^T for 2 (and above 8), T (SS S2) for 3, T (SS SS) for
// 4, T (SS SS S2) for 5, T (SS SS SS) for 6, T (SS SS SS S2) for 7 and
// T ((SS SS)(SS SS)) for 8, every product in first-touch order.  Rows of the
// result are returned sorted by column, as the later products use them.
static Csr restriction_powers(const Csr& A, const std::vector<uint8_t>& strong,
                              const Aggregates& ag, int smoothers, double omega,
                              const std::vector<uint8_t>* ghost) {
  const int n = A.nrows;
  std::vector<double> d1((size_t)n, 0.0), d2((size_t)n, 0.0);
  parallel_rows(n, [&](int i) {
    const bool gh = ghost && (*ghost)[i];
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      if (A.col[p] == i) {
        d1[i] = gh ? A.val[p] : d1[i] + A.val[p];
        d2[i] = A.val[p];
      } else if (!strong[p] && !gh) {
        d1[i] = d1[i] + A.val[p];
        d2[i] = d2[i] + A.val[p];
      }
    }
  });
  auto smoother = [&](const std::vector<double>& diag) {
    Csr S;
    S.nrows = S.ncols = n;
    S.ptr = A.ptr;
    S.col = A.col;
    S.val.resize(A.val.size());
    const double one_minus = 1.0 - omega;
    parallel_rows(n, [&](int i) {
      const bool gh = ghost && (*ghost)[i];
      for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
        if (A.col[p] == i)
          S.val[p] = gh ? 1.0 - omega / diag[i] * A.val[p] : one_minus;
        else if (strong[p])
          S.val[p] = -omega / diag[i] * A.val[p];
        else
          S.val[p] = 0.0;
      }
    });
    return S;
  };
  const Csr S = smoother(d1), S2 = smoother(d2);
  const Csr SS = first_touch_product(S2, S);
  Csr X;
  switch (smoothers) {
    case 3: X = first_touch_product(SS, S2); break;
    case 4: X = first_touch_product(SS, SS); break;
    case 5: X = first_touch_product(first_touch_product(SS, SS), S2); break;
    case 6: X = first_touch_product(first_touch_product(SS, SS), SS); break;
    case 7: X = first_touch_product(first_touch_product(first_touch_product(SS, SS), SS), S2); break;
    case 8: {
      const Csr S4 = first_touch_product(SS, SS);
      X = first_touch_product(S4, S4);
      break;
    }
    default: X = transpose(SS); break;  // 2, and the reference's fallback above 8
  }
  Csr T;
  T.nrows = ag.nagr;
  T.ncols = n;
  T.ptr = ag.starts;
  T.col = ag.nodes;
  T.val.assign(T.col.size(), 1.0);
  // the values do not depend on where a column first appears; sort each row
  return transpose(transpose(first_touch_product(T, X)));
}

Csr restriction(const Csr& A, const std::vector<uint8_t>& strong, const Aggregates& ag,
                int smoothers, double omega, const std::vector<uint8_t>* ghost) {
  const int n = A.nrows, nagr = ag.nagr;
  Csr R;
  R.nrows = nagr;
  R.ncols = n;
  if (smoothers == 0) {  // R = T, the aggregate indicator (CoarseMtx.F90:80-108)
    R.ptr = ag.starts;
    R.col = ag.nodes;
    R.val.assign(R.col.size(), 1.0);
    return R;
  }
  if (smoothers >= 2) return restriction_powers(A, strong, ag, smoothers, omega, ghost);
  // filtered diagonal and the Jacobi smoother S (CoarseMtx.F90:152-233); an
  // A_ghost row takes its diagonal entry (:164-170) and keeps a_ii/a_ii in
  // its diagonal smoother entry (:200-205)
  std::vector<double> diag((size_t)n, 0.0);
  parallel_rows(n, [&](int i) {
    const bool gh = ghost && (*ghost)[i];
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      if (gh) {
        if (A.col[p] == i) diag[i] = A.val[p];
      } else if (A.col[p] == i || !strong[p]) {
        diag[i] = diag[i] + A.val[p];
      }
    }
  });
  Csr S;
  S.nrows = S.ncols = n;
  S.ptr = A.ptr;
  S.col = A.col;
  S.val.resize(A.val.size());
  const double one_minus = 1.0 - omega;
  parallel_rows(n, [&](int i) {
    const bool gh = ghost && (*ghost)[i];
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      if (A.col[p] == i)
        S.val[p] = gh ? 1.0 - omega / diag[i] * A.val[p] : one_minus;
      else if (strong[p])
        S.val[p] = -omega / diag[i] * A.val[p];
      else
        S.val[p] = 0.0;  // weak entries stay as explicit zeros (CoarseMtx.F90:187-191)
    }
  });
  // R = T S^T (SpMtx_op_AB.f90:125-176): the reference visits, for row a of
  // R, the aggregate's nodes kk ascending and for each the entries S(j,kk);
  // R(a,j) is the first such S(j,kk) followed by the others in ascending kk.
  // Built here column by column: for node j the entries S(j,kk) of row j are
  // ordered by (aggregate of kk, kk) and summed per aggregate in that order,
  // which gives the same values; R^T comes out row-sorted, R by transposing.
  Csr Rt;
  Rt.nrows = n;
  Rt.ncols = nagr;
  const int nth = setup_threads();
  std::vector<CsrPiece> pieces((size_t)nth);
  parallel_chunks(n, nth, [&](int t, int b, int e) {
    CsrPiece& out = pieces[t];
    std::vector<std::pair<std::pair<int, int>, int>> ent;  // ((aggregate, kk), position)
    for (int j = b; j < e; j++) {
      ent.clear();
      for (int p = S.ptr[j]; p < S.ptr[j + 1]; p++) {
        int a = ag.num[S.col[p]] - 1;
        if (a >= 0) ent.push_back({{a, S.col[p]}, p});
      }
      std::stable_sort(ent.begin(), ent.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      int len = 0;
      for (size_t q = 0; q < ent.size();) {
        int a = ent[q].first.first;
        double acc = 1.0 * S.val[ent[q].second];
        size_t r = q + 1;
        for (; r < ent.size() && ent[r].first.first == a; r++) acc = acc + 1.0 * S.val[ent[r].second];
        out.col.push_back(a);
        out.val.push_back(acc);
        len++;
        q = r;
      }
      out.rowlen.push_back(len);
    }
  });
  join_pieces(pieces, n, Rt);
  return transpose(Rt);
}

// SpMtx_AB (SpMtx_op_AB.f90:39-200) for row-arranged operands: row i
// visits X's row in stored order and, for each entry (i, k), Y's row k in
// stored order; the columns of a row come in first-touch order, each entry
// first assigned its product, then accumulated.
Csr first_touch_product(const Csr& X, const Csr& Y) {
  const int nr = X.nrows, nc = Y.ncols;
  const int nth = setup_threads();
  Csr C;
  C.nrows = nr;
  C.ncols = nc;
  std::vector<CsrPiece> pieces((size_t)nth);
  parallel_chunks(nr, nth, [&](int t, int b, int e) {
    CsrPiece& out = pieces[t];
    std::vector<int> mark((size_t)nc, -1), pos((size_t)nc, 0);
    for (int i = b; i < e; i++) {
      size_t start = out.col.size();
      for (int p = X.ptr[i]; p < X.ptr[i + 1]; p++) {
        int k = X.col[p];
        for (int q = Y.ptr[k]; q < Y.ptr[k + 1]; q++) {
          int jj = Y.col[q];
          double prod = X.val[p] * Y.val[q];
          if (mark[jj] != i) {
            mark[jj] = i;
            pos[jj] = (int)out.col.size();
            out.col.push_back(jj);
            out.val.push_back(prod);
          } else {
            double& v = out.val[pos[jj]];
            v = v + prod;
          }
        }
      }
      out.rowlen.push_back((int)(out.col.size() - start));
    }
  });
  join_pieces(pieces, nr, C);
  return C;
}

Csr galerkin(const Csr& A, const Csr& R, const Csr& Rt) {
  // T' = A R^T: R^T's rows list aggregates ascending (R arranged by columns);
  // A_c = R T' with R's rows in ascending column order.
  return first_touch_product(R, first_touch_product(A, Rt));
}

void coarse_partition(const Aggregates& cfull, const Aggregates& ffull, int n,
                      std::vector<int>& starts, std::vector<int>& nodes) {
  std::vector<char> seen((size_t)n, 0);
  starts.assign((size_t)cfull.nagr + 1, 0);
  nodes.clear();
  for (int c = 0; c < cfull.nagr; c++) {
    size_t first = nodes.size();
    for (int q = cfull.starts[c]; q < cfull.starts[c + 1]; q++) {
      int fa = cfull.nodes[q];  // fine aggregate (0-based)
      for (int t = ffull.starts[fa]; t < ffull.starts[fa + 1]; t++) {
        int node = ffull.nodes[t];
        if (!seen[node]) {
          seen[node] = 1;
          nodes.push_back(node);
        }
      }
    }
    for (size_t t = first; t < nodes.size(); t++) seen[nodes[t]] = 0;
    starts[c + 1] = (int)nodes.size();
  }
}

void add_layers_all(const Csr& A, const std::vector<int>& pstarts, const std::vector<int>& pnodes,
                    int ol, std::vector<int>& sstarts, std::vector<int>& sinds) {
  const int n = A.nrows;
  const int np = (int)pstarts.size() - 1;
  std::vector<int> onfront((size_t)n, 0);
  sstarts.assign((size_t)np + 1, 0);
  sinds.clear();
  std::vector<int> nodes;
  for (int ip = 0; ip < np; ip++) {
    nodes.assign(pnodes.begin() + pstarts[ip], pnodes.begin() + pstarts[ip + 1]);
    int nn = (int)nodes.size();
    for (int v : nodes) onfront[v] = -1;
    // the reference re-scans the last node of the previous front at each
    // layer (frontstart(layer)=nfront); that visit adds nothing, so a plain
    // BFS over [begin, end) of the previous front gives the same order.
    int fbeg = 0, fend = nn;
    for (int layer = 1; layer <= ol; layer++) {
      int nfront = (int)nodes.size();
      for (int i = fbeg; i < fend; i++) {
        int node = nodes[i];
        for (int p = A.ptr[node]; p < A.ptr[node + 1]; p++) {
          int nb = A.col[p];
          if (onfront[nb] == 0) {
            onfront[nb] = layer;
            nodes.push_back(nb);
          }
        }
      }
      fbeg = nfront;
      fend = (int)nodes.size();
    }
    for (int v : nodes) onfront[v] = 0;
    sinds.insert(sinds.end(), nodes.begin(), nodes.end());
    sstarts[ip + 1] = (int)sinds.size();
  }
}

// ---------------------------------------------------------------------------
// Reverse Cuthill-McKee order inside each index block (E. Cuthill and
// J. McKee, Proc. 24th ACM National Conference, 1969; reversal: A. George,
// 1971; starting node from the pseudo-peripheral search of N. Gibbs,
// W. Poole and P. Stockmeyer, SIAM J. Numer. Anal. 13 (1976) 236-250).
// Only the device copy of the subdomain index lists is reordered: it decides
// where each unknown sits in the subdomain's dense factor, so that the
// factor's nonzero envelope is a narrow band of tiles.
namespace {

struct LocalGraph {
  std::vector<int> ptr, adj, deg;
};

// BFS levels from `root` over unvisited nodes (mark != stamp); returns the
// number of levels and leaves the last level in [last_b, last_e) of `order`.
int bfs_levels(const LocalGraph& g, int root, std::vector<int>& mark, int stamp,
               std::vector<int>& order, int& last_b, int& last_e) {
  // `order` needs room for every node (each enters once); the scan appends
  // without branches: a node is written at the end every time and kept
  // (the end advanced) only when it was unmarked
  if (order.size() < g.deg.size() + 1) order.resize(g.deg.size() + 1);
  int* ord = order.data();
  int* mk = mark.data();
  const int* ptr = g.ptr.data();
  const int* adj = g.adj.data();
  ord[0] = root;
  mk[root] = stamp;
  int lb = 0, le = 1, levels = 1, top = 1;
  while (true) {
    for (int q = lb; q < le; q++) {
      const int v = ord[q];
      for (int p = ptr[v]; p < ptr[v + 1]; p++) {
        const int u = adj[p];
        const int fresh = mk[u] != stamp;
        ord[top] = u;
        top += fresh;
        mk[u] = stamp;
      }
    }
    if (top == le) break;
    lb = le;
    le = top;
    levels++;
  }
  order.resize((size_t)top);
  last_b = lb;
  last_e = le;
  return levels;
}

}  // namespace

namespace {

struct RcmBuf {
  std::vector<int> mark, order, bfs, cbfs, nbr;
  std::vector<char> placed;
  LocalGraph g;
};

// One block [s0, s0 + s) of `inds` into `out`; `floc` is -1 on entry and exit.
void rcm_block(const Csr& A, const std::vector<int>& inds, int s0, int s, std::vector<int>& out,
               std::vector<int>& floc, RcmBuf& buf) {
  std::vector<int>& mark = buf.mark;
  std::vector<int>& order = buf.order;
  std::vector<int>& bfs = buf.bfs;
  std::vector<int>& nbr = buf.nbr;
  std::vector<char>& placed = buf.placed;
  LocalGraph& g = buf.g;
  if (s <= 1) {
    if (s == 1) out[s0] = inds[s0];
    return;
  }
  for (int li = 0; li < s; li++) floc[inds[s0 + li]] = li;
  g.ptr.assign((size_t)s + 1, 0);
  g.deg.assign((size_t)s, 0);
  {
    // room for every entry of the block's rows; an entry is written at the
    // end every time and kept when its column lies in the block (no branch)
    size_t cap = 0;
    for (int li = 0; li < s; li++) cap += (size_t)(A.ptr[inds[s0 + li] + 1] - A.ptr[inds[s0 + li]]);
    if (g.adj.size() < cap + 1) g.adj.resize(cap + 1);
    int* adj = g.adj.data();
    const int* fl = floc.data();
    const int* acol = A.col.data();
    int top = 0;
    for (int li = 0; li < s; li++) {
      const int gi = inds[s0 + li];
      for (int p = A.ptr[gi]; p < A.ptr[gi + 1]; p++) {
        const int lj = fl[acol[p]];
        adj[top] = lj;
        top += (lj >= 0) & (lj != li);
      }
      g.ptr[li + 1] = top;
      g.deg[li] = g.ptr[li + 1] - g.ptr[li];
    }
  }
  for (int li = 0; li < s; li++) floc[inds[s0 + li]] = -1;
  mark.assign((size_t)s, 0);
  placed.assign((size_t)s, 0);
  order.clear();
  int stamp = 0, next = 0;
  while ((int)order.size() < s) {
    while (placed[next]) next++;
    // minimum-degree unplaced node of this component (ties: lowest index).
    // The first component is usually the whole block: then it is the
    // block's minimum-degree node, and the search from it shows that it
    // reached every node
    int root = next, lb, le, ecc = -1;
    if (order.empty()) {
      int gmin = 0;
      for (int v = 1; v < s; v++)
        if (g.deg[v] < g.deg[gmin]) gmin = v;
      ecc = bfs_levels(g, gmin, mark, ++stamp, bfs, lb, le);
      if ((int)bfs.size() == s)
        root = gmin;
      else
        ecc = -1;
    }
    if (ecc < 0) {
      bfs_levels(g, next, mark, ++stamp, bfs, lb, le);
      for (int v : bfs)
        if (g.deg[v] < g.deg[root] || (g.deg[v] == g.deg[root] && v < root)) root = v;
      ecc = bfs_levels(g, root, mark, ++stamp, bfs, lb, le);
    }
    // pseudo-peripheral node: move to a minimum-degree node of the last
    // level while the eccentricity grows
    for (int round = 0; round < 8; round++) {
      int cand = bfs[lb];
      for (int q = lb; q < le; q++) {
        int v = bfs[q];
        if (g.deg[v] < g.deg[cand] || (g.deg[v] == g.deg[cand] && v < cand)) cand = v;
      }
      int clb, cle;
      std::vector<int>& cbfs = buf.cbfs;
      int cecc = bfs_levels(g, cand, mark, ++stamp, cbfs, clb, cle);
      if (cecc <= ecc) break;
      root = cand;
      ecc = cecc;
      bfs.swap(cbfs);
      lb = clb;
      le = cle;
    }
    // Cuthill-McKee: neighbours in ascending degree (then index), appended
    // to `order` (sized for the block) through raw pointers
    int head = (int)order.size(), tail = head;
    order.resize((size_t)s);
    int* ord = order.data();
    char* pl = placed.data();
    const int* adj = g.adj.data();
    const int* gp = g.ptr.data();
    const int* dg = g.deg.data();
    ord[tail++] = root;
    pl[root] = 1;
    while (head < tail) {
      const int v = ord[head++], p0 = gp[v], d = gp[v + 1] - p0;
      if (d <= 8) {
        // (degree, index) packed into one ordered key, an entry already
        // placed (or a repeat of one in this row) keyed last; eight slots
        // through the sorting network
        int64_t key[8];
        int m = 0;
#pragma unroll
        for (int q = 0; q < 8; q++) {
          const int u = q < d ? adj[p0 + q] : 0;
          const bool fresh = q < d && !pl[u];
          key[q] = fresh ? ((int64_t)dg[u] << 32) | (uint32_t)u : INT64_MAX;
          if (q < d) pl[u] = 1;
          m += fresh;
        }
        sort8_asc(key);
        for (int q = 0; q < m; q++) ord[tail + q] = (int)(uint32_t)key[q];
        tail += m;
      } else {
        nbr.resize((size_t)d);
        int m = 0;
        for (int p = p0; p < p0 + d; p++) {
          const int u = adj[p];
          nbr[m] = u;
          m += !pl[u];
          pl[u] = 1;
        }
        std::sort(nbr.begin(), nbr.begin() + m, [&](int x, int y) {
          return dg[x] != dg[y] ? dg[x] < dg[y] : x < y;
        });
        for (int q = 0; q < m; q++) ord[tail + q] = nbr[q];
        tail += m;
      }
    }
    order.resize((size_t)tail);
  }
  for (int q = 0; q < s; q++) out[s0 + q] = inds[s0 + order[s - 1 - q]];
}

}  // namespace

// node-indexed scratch of one virtual_ranks worker (kept between set-ups)
struct RankNode {
  int layer = -1, label = 0, gl = -1, stamp = 0;
};

struct RankScratch {
  int n = 0;
  std::vector<RankNode> nd;
  std::vector<int> floc;  // -1 between uses (block_rcm_serial scratch)
  int tick = 0;
  explicit RankScratch(int n_) : n(n_), nd((size_t)n_), floc((size_t)n_, -1) {}
};

void block_rcm(const Csr& A, const std::vector<int>& starts, const std::vector<int>& inds,
               std::vector<int>& out, RankWorkspaces* ws) {
  const int nb = (int)starts.size() - 1;
  out.assign(inds.begin(), inds.end());
  int nth = std::max(1, std::min(setup_threads(), nb));
  // as many threads as lent scratch arrays (a fresh one costs more than a
  // thread gains)
  if (ws && !ws->w.empty() && ws->w[0] && ws->w[0]->n == A.nrows) nth = std::min(nth, (int)ws->w.size());
  std::vector<std::thread> th;
  for (int t = 0; t < nth; t++)
    th.emplace_back([&, t]() {
      // a node-indexed scratch of the whole matrix: the rank workers' one
      // when it fits (-1 between uses), else a fresh one
      RankScratch* rs = ws && t < (int)ws->w.size() && ws->w[t] && ws->w[t]->n == A.nrows ? ws->w[t].get() : nullptr;
      std::vector<int> own;
      if (!rs) own.assign((size_t)A.nrows, -1);
      std::vector<int>& floc = rs ? rs->floc : own;
      RcmBuf buf;
      for (int b = t; b < nb; b += nth) rcm_block(A, inds, starts[b], starts[b + 1] - starts[b], out, floc, buf);
    });
  for (auto& x : th) x.join();
}

void block_rcm_serial(const Csr& A, const std::vector<int>& starts, const std::vector<int>& inds,
                      std::vector<int>& out, std::vector<int>& floc) {
  out.assign(inds.begin(), inds.end());
  RcmBuf buf;
  for (int b = 0; b + 1 < (int)starts.size(); b++)
    rcm_block(A, inds, starts[b], starts[b + 1] - starts[b], out, floc, buf);
}

// ---------------------------------------------------------------------------
// Graph partitioner standing in for METIS 4.0 (absent here): recursive
// bisection.  Each bisection grows one side from a pseudo-peripheral vertex,
// greedily by edge gain (greedy graph growing, G. Karypis and V. Kumar, SIAM
// J. Sci. Comput. 20 (1998) 359-392), until it holds its share of the vertex
// weight (disconnected pieces are taken in vertex order), then moves boundary
// vertices whose move cuts fewer edges while both sides stay within the
// balance tolerance (greedy refinement in the manner of B. Kernighan and
// S. Lin, Bell Syst. Tech. J. 49 (1970) 291-307).  Deterministic.
namespace {

struct Bisector {
  const int* xadj;
  const int* adj;
  const int* vw;
  std::vector<int>& part;
  std::vector<int> loc, mark, side, queue, gain;
  std::vector<int64_t> disc;
  int stamp = 0;

  Bisector(int n, const int* x, const int* a, const int* w, std::vector<int>& p)
      : xadj(x), adj(a), vw(w), part(p), loc((size_t)n, -1), mark((size_t)n, 0),
        side((size_t)n, 0), gain((size_t)n, 0), disc((size_t)n, 0) {}

  int64_t weight(int v) const { return vw ? vw[v] : 1; }

  // BFS over `verts` (loc >= 0) from root; returns the last vertex reached
  // with the smallest degree in the last level
  int far_vertex(int root) {
    ++stamp;
    queue.clear();
    queue.push_back(root);
    mark[root] = stamp;
    size_t lb = 0, le = 1;
    while (true) {
      for (size_t q = lb; q < le; q++) {
        int v = queue[q];
        for (int p = xadj[v]; p < xadj[v + 1]; p++) {
          int u = adj[p];
          if (loc[u] >= 0 && mark[u] != stamp) {
            mark[u] = stamp;
            queue.push_back(u);
          }
        }
      }
      if (queue.size() == le) break;
      lb = le;
      le = queue.size();
    }
    int best = queue[lb];
    for (size_t q = lb; q < le; q++) {
      int v = queue[q];
      int dv = xadj[v + 1] - xadj[v], db = xadj[best + 1] - xadj[best];
      if (dv < db || (dv == db && v < best)) best = v;
    }
    return best;
  }

  void run(std::vector<int>& verts, int k, int base) {
    if (k <= 1 || verts.size() <= 1) {
      for (int v : verts) part[v] = base;
      return;
    }
    const int k1 = k / 2;
    int64_t total = 0;
    for (int v : verts) total += weight(v);
    const int64_t target = total * k1 / k;
    const int64_t tol = std::max<int64_t>(1, total / (50 * (int64_t)k));
    for (size_t q = 0; q < verts.size(); q++) loc[verts[q]] = (int)q;
    // side 0 grows from a pseudo-peripheral vertex of the first component,
    // always taking the frontier vertex with the largest gain (edges into
    // side 0 minus edges out of it; ties: first discovered)
    int start = verts[0];
    for (int r = 0; r < 3; r++) start = far_vertex(start);
    for (int v : verts) {
      side[v] = 1;
      gain[v] = 0;
    }
    int64_t w0 = 0;
    ++stamp;  // mark: discovered (in the heap or on side 0)
    size_t next = 0;
    int64_t order = 0;
    std::priority_queue<std::tuple<int, int64_t, int>> heap;  // (gain, -order, v)
    auto discover = [&](int v) {
      mark[v] = stamp;
      int g = 0;
      for (int p = xadj[v]; p < xadj[v + 1]; p++)
        if (loc[adj[p]] >= 0) g += side[adj[p]] == 0 ? 1 : -1;
      gain[v] = g;
      disc[v] = order++;
      heap.emplace(g, -disc[v], v);
    };
    discover(start);
    while (w0 < target) {
      if (heap.empty()) {  // component exhausted: next undiscovered vertex
        while (next < verts.size() && mark[verts[next]] == stamp) next++;
        if (next == verts.size()) break;
        discover(verts[next]);
      }
      auto [g, no, v] = heap.top();
      heap.pop();
      if (side[v] == 0 || g != gain[v]) continue;  // stale entry
      side[v] = 0;
      w0 += weight(v);
      for (int p = xadj[v]; p < xadj[v + 1]; p++) {
        int u = adj[p];
        if (loc[u] < 0 || side[u] == 0) continue;
        if (mark[u] != stamp) {
          discover(u);
        } else {
          gain[u] += 2;
          heap.emplace(gain[u], -disc[u], u);
        }
      }
    }
    // boundary refinement: move a vertex when it has more neighbours on the
    // other side and both sides stay within target +- tol
    for (int pass = 0; pass < 8; pass++) {
      bool moved = false;
      for (int v : verts) {
        int same = 0, other = 0;
        for (int p = xadj[v]; p < xadj[v + 1]; p++) {
          int u = adj[p];
          if (loc[u] < 0) continue;
          (side[u] == side[v] ? same : other)++;
        }
        if (other <= same) continue;
        const int64_t nw0 = side[v] == 0 ? w0 - weight(v) : w0 + weight(v);
        if (nw0 < target - tol || nw0 > target + tol) continue;
        side[v] ^= 1;
        w0 = nw0;
        moved = true;
      }
      if (!moved) break;
    }
    std::vector<int> a, b;
    for (int v : verts) (side[v] == 0 ? a : b).push_back(v);
    for (int v : verts) loc[v] = -1;
    if (a.empty() || b.empty()) {  // cannot split further: share what there is
      std::vector<int>& all = a.empty() ? b : a;
      for (size_t q = 0; q < all.size(); q++) part[all[q]] = base + (int)((int64_t)q * k / all.size());
      return;
    }
    verts.clear();
    verts.shrink_to_fit();
    run(a, k1, base);
    run(b, k - k1, base + k1);
  }
};

}  // namespace

void partition_graph(int n, const int* xadj, const int* adjncy, const int* vwgt, int nparts,
                     std::vector<int>& part) {
  part.assign((size_t)n, 0);
  if (n <= 0 || nparts <= 1) return;
  Bisector bis(n, xadj, adjncy, vwgt, part);
  std::vector<int> verts((size_t)n);
  for (int v = 0; v < n; v++) verts[v] = v;
  bis.run(verts, nparts, 0);
}

void aggregate_graph(const Csr& A, const Aggregates& inner, const Aggregates& full,
                     std::vector<int>& xadj, std::vector<int>& adjncy) {
  // SpMtx_buildAggrAdjncy (SpMtx_arrangement.F90:2558-2639): aggregates c1 !=
  // c2 joined by an entry (i, j), i < j, of the inner aggregation; each pair
  // once per direction, neighbours in order of discovery
  const int na = full.nagr, n = A.nrows;
  std::vector<std::vector<int>> nb((size_t)na);
  for (int i = 0; i < n; i++)
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) {
      const int j = A.col[p];
      if (j >= n || i >= j) continue;
      const int c1 = inner.num[i] - 1, c2 = inner.num[j] - 1;
      if (c1 < 0 || c2 < 0 || c1 == c2) continue;
      std::vector<int>& l = nb[c1];
      if (std::find(l.begin(), l.end(), c2) == l.end()) l.push_back(c2);
    }
  xadj.assign((size_t)na + 1, 0);
  for (int c1 = 0; c1 < na; c1++)
    for (int c2 : nb[c1]) {
      xadj[c1 + 1]++;
      xadj[c2 + 1]++;
    }
  for (int c = 0; c < na; c++) xadj[c + 1] += xadj[c];
  adjncy.assign((size_t)xadj[na], 0);
  std::vector<int> fill(xadj.begin(), xadj.end() - 1);
  for (int c1 = 0; c1 < na; c1++)
    for (int c2 : nb[c1]) {
      adjncy[fill[c1]++] = c2;
      adjncy[fill[c2]++] = c1;
    }
}

// ---------------------------------------------------------------------------
// Virtual ranks.  The reference's master sends each rank the rows within
// max(1, 2 ol) graph layers of its nodes (SpMtx_distributeWithOverlap,
// SpMtx_arrangement.F90:895-1021, rows arranged with sorted columns); the
// rank then labels its nodes by graph distance (SpMtx_build_ghost,
// :1028-1710) and numbers them by label (SpMtx_Build_lggl, :744-866).  Those
// loops are restated here as distances: the layer of an outside node is its
// distance from the rank's node set; an own node is in a neighbour's inner
// overlap when it lies within `ol` steps of that neighbour's nodes and is sent
// to it when exactly `ol` steps away.

Csr sort_rows_by_column(const Csr& A, const std::vector<double>& val) {
  Csr S;
  S.nrows = A.nrows;
  S.ncols = A.ncols;
  S.ptr = A.ptr;
  S.col = A.col;
  S.val = val;
  S.src.resize(A.col.size());
  // rows are short: a stable insertion sort in place (strict <), nothing to
  // do for rows already in column order; src follows each entry
  parallel_rows(A.nrows, [&](int i) {
    int* c = S.col.data();
    double* x = S.val.data();
    int* o = S.src.data();
    for (int p = A.ptr[i]; p < A.ptr[i + 1]; p++) o[p] = p;
    for (int p = A.ptr[i] + 1; p < A.ptr[i + 1]; p++) {
      if (c[p - 1] <= c[p]) continue;
      const int ck = c[p], ok = o[p];
      const double xk = x[p];
      int q = p;
      for (; q > A.ptr[i] && c[q - 1] > ck; q--) {
        c[q] = c[q - 1];
        x[q] = x[q - 1];
        o[q] = o[q - 1];
      }
      c[q] = ck;
      x[q] = xk;
      o[q] = ok;
    }
  });
  return S;
}

bool rows_sorted_by_column(const Csr& A) {
  std::vector<char> ok((size_t)std::max(1, setup_threads()), 1);
  parallel_chunks(A.nrows, setup_threads(), [&](int t, int b, int e) {
    for (int i = b; i < e && ok[t]; i++)
      for (int p = A.ptr[i] + 1; p < A.ptr[i + 1]; p++)
        if (A.col[p - 1] > A.col[p]) {
          ok[t] = 0;
          break;
        }
  });
  for (char c : ok)
    if (!c) return false;
  return true;
}

// Per-node state of one rank's pass, packed so that a lookup touches one
// cache line: layer (-1 off the rank's local matrix), labels (bit 0 inner
// overlap of some neighbour, bit 1 sent to some neighbour), local number,
// visit stamp.

RankWorkspaces::RankWorkspaces() = default;
RankWorkspaces::~RankWorkspaces() = default;
void RankWorkspaces::clear() { w.clear(); }

namespace {



// Labels of own nodes, per neighbour rank: within L steps of that rank's
// nodes on the front (inner overlap), exactly L steps away (sent).
void rank_labels(const Csr& As, const std::vector<int>& eptn, const std::vector<int>& front, int L,
                 RankScratch& S, std::vector<int>& frontier, std::vector<int>& next) {
  std::vector<int> nbr;
  for (int w : front) nbr.push_back(eptn[w]);
  std::sort(nbr.begin(), nbr.end());
  nbr.erase(std::unique(nbr.begin(), nbr.end()), nbr.end());
  std::vector<std::vector<int>> by_rank(nbr.size());
  for (int w : front) by_rank[std::lower_bound(nbr.begin(), nbr.end(), eptn[w]) - nbr.begin()].push_back(w);
  for (size_t k = 0; k < nbr.size(); k++) {
    const int t = ++S.tick;
    frontier = by_rank[k];
    for (int w : frontier) S.nd[w].stamp = t;
    for (int d = 1; d <= L && !frontier.empty(); d++) {
      next.clear();
      for (int v : frontier)
        for (int p = As.ptr[v]; p < As.ptr[v + 1]; p++) {
          const int w = As.col[p];
          if (S.nd[w].layer < 0 || S.nd[w].stamp == t) continue;  // off the rank's local matrix, or seen
          S.nd[w].stamp = t;
          next.push_back(w);
          if (S.nd[w].layer == 0) S.nd[w].label |= (d == L) ? 3 : 1;
        }
      frontier.swap(next);
    }
  }
}

// Labels: bit 0 inner overlap of some neighbour, bit 1 sent to some neighbour.
void rank_local(const Csr& As, const std::vector<double>& val, const std::vector<int>& eptn, int r,
                const std::vector<int>& own, int ol, RankScratch& S, RankLocal& out, bool ghost,
                const int* pos) {
  const int L = std::max(ol, 1), D = 2 * L;
  std::vector<int> touched(own.begin(), own.end());
  for (int v : own) S.nd[v].layer = 0;
  // layers 1..D of outside nodes (build_ghost's first front)
  std::vector<int> front, frontier(own.begin(), own.end()), next;
  for (int l = 1; l <= D && !frontier.empty(); l++) {
    next.clear();
    for (int v : frontier)
      for (int p = As.ptr[v]; p < As.ptr[v + 1]; p++) {
        const int w = As.col[p];
        if (S.nd[w].layer < 0) {
          S.nd[w].layer = l;
          next.push_back(w);
          touched.push_back(w);
        }
      }
    front.insert(front.end(), next.begin(), next.end());
    frontier.swap(next);
  }
  if (L == 1) {
    // one step: an own node is sent (and in the inner overlap) when some
    // front node's row holds it, whichever neighbour that node belongs to
    for (int w : front)
      for (int p = As.ptr[w]; p < As.ptr[w + 1]; p++) {
        const int u = As.col[p];
        if (S.nd[u].layer == 0) S.nd[u].label = 3;
      }
  } else {
    rank_labels(As, eptn, front, L, S, frontier, next);
  }
  (void)r;
  // numbering: [sent | inner overlap | inner | outer overlap | received]
  std::vector<int> grp[5];
  for (int v : own) grp[(S.nd[v].label & 2) ? 0 : (S.nd[v].label & 1) ? 1 : 2].push_back(v);
  for (int w : front)
    if (S.nd[w].layer <= L) grp[S.nd[w].layer == L ? 4 : 3].push_back(w);
  std::sort(grp[3].begin(), grp[3].end());
  std::sort(grp[4].begin(), grp[4].end());
  out.lg.clear();
  for (auto& g : grp) out.lg.insert(out.lg.end(), g.begin(), g.end());
  out.ninner = (int)own.size();
  const int nl = (int)out.lg.size();
  for (int k = 0; k < nl; k++) S.nd[out.lg[k]].gl = k;
  // local matrix: own rows whole; rows of outside nodes nearer than L keep
  // the columns of own nodes and of outside nodes up to layer L; the outer
  // layer's rows are the ghost matrix, not part of it
  Csr& A = out.A;
  A.nrows = A.ncols = nl;
  A.ptr.assign((size_t)nl + 1, 0);
  A.col.clear();
  A.val.clear();
  A.src.clear();
  {
    // room for the rows kept here and (with `ghost`) the ghost rows that
    // the caller appends, so that neither grows the arrays again
    size_t cap = 0;
    for (int k = 0; k < nl; k++) {
      const int v = out.lg[k];
      if (S.nd[v].layer < L || (ghost && S.nd[v].layer == L)) cap += (size_t)(As.ptr[v + 1] - As.ptr[v]);
    }
    A.col.reserve(cap);
    A.val.reserve(cap);
    A.src.reserve(cap);
    // the kept rows are written into [0, cap) below, one slot of slack
    A.col.resize(cap + 1);
    A.val.resize(cap + 1);
    A.src.resize(cap + 1);
  }
  {
    // each kept row: the entries to nodes the rank sends first, then the
    // others, each group in stored order.  A first scan counts the first
    // group; the second writes every entry at its group's cursor and
    // advances that cursor only for an entry that is kept (no branches)
    int* acol = A.col.data();
    double* aval = A.val.data();
    int* asrc = A.src.data();
    const RankNode* nd = S.nd.data();
    int top = 0;
    for (int k = 0; k < nl; k++) {
      const int v = out.lg[k];
      if (nd[v].layer < L) {
        const bool ownrow = nd[v].layer == 0;
        const int p0 = As.ptr[v], p1 = As.ptr[v + 1];
        int ns = 0;
        for (int p = p0; p < p1; p++) {
          const RankNode& w = nd[As.col[p]];
          const bool keep = (w.layer >= 0) & (ownrow | (w.layer <= L));
          ns += keep & (w.layer == 0) & ((w.label & 2) != 0);
        }
        int a = top, b = top + ns;
        for (int p = p0; p < p1; p++) {
          const RankNode& w = nd[As.col[p]];
          const bool keep = (w.layer >= 0) & (ownrow | (w.layer <= L));
          const bool sent = keep & (w.layer == 0) & ((w.label & 2) != 0);
          const int at = sent ? a : b;
          acol[at] = w.gl;
          aval[at] = val[p];
          asrc[at] = pos ? pos[p] : p;
          a += sent;
          b += keep & !sent;
        }
        top = b;
      }
      A.ptr[k + 1] = top;
    }
    A.col.resize((size_t)top);
    A.val.resize((size_t)top);
    A.src.resize((size_t)top);
  }
  // A_ghost (SpMtx_build_ghost, SpMtx_arrangement.F90:1590-1612): the
  // received nodes' rows, entries to nodes up to layer L, values as read;
  // the entry order inside a ghost row decides no result
  out.nsent = (int)grp[0].size();
  out.nrecv = (int)grp[4].size();
  out.G = Csr();
  if (ghost) {
    Csr& G = out.G;
    G.nrows = out.nrecv;
    G.ncols = nl;
    G.ptr.assign((size_t)out.nrecv + 1, 0);
    for (int k = 0; k < out.nrecv; k++) {
      const int v = grp[4][k];
      for (int p = As.ptr[v]; p < As.ptr[v + 1]; p++) {
        const int w = As.col[p];
        const int lw = S.nd[w].layer;
        if (lw < 0 || lw > L) continue;
        G.col.push_back(S.nd[w].gl);
        G.val.push_back(val[p]);
        G.src.push_back(pos ? pos[p] : p);
      }
      G.ptr[k + 1] = (int)G.col.size();
    }
  }
  for (int v : touched) {
    S.nd[v].layer = -1;
    S.nd[v].label = 0;
    S.nd[v].gl = -1;
  }
}


}  // namespace

void virtual_ranks(const Csr& As, const std::vector<double>& val, const std::vector<int>& eptn,
                   int nranks, int ol_dist, int ol_sub,
                   int nsub_per_rank, int radius1, int minasize1, int maxasize1, double strong1,
                   RankSetup& out, const Csr* rcm_graph, RankWorkspaces* ws, bool keep_local,
                   const std::vector<int>* pos) {
  const int n = As.nrows;
  std::vector<std::vector<int>> own((size_t)nranks);
  {
    std::vector<int> cnt((size_t)nranks, 0);
    for (int i = 0; i < n; i++) cnt[eptn[i]]++;
    for (int r = 0; r < nranks; r++) own[r].reserve((size_t)cnt[r]);
  }
  for (int i = 0; i < n; i++) own[eptn[i]].push_back(i);
  out.ranks.assign((size_t)nranks, RankLocal{});
  out.own_off.assign((size_t)nranks + 1, 0);
  for (int r = 0; r < nranks; r++) out.own_off[r + 1] = out.own_off[r] + (int)own[r].size();
  out.order.resize((size_t)n);
  std::vector<Aggregates> inner((size_t)nranks), full((size_t)nranks);
  std::vector<std::vector<int>> sst((size_t)nranks), sin((size_t)nranks), dsin((size_t)nranks);
  // own local indices in ascending global order, and the own nodes listed in
  // the inner and full aggregates of each rank
  std::vector<std::vector<int>> by_global((size_t)nranks);
  std::vector<int> nlisted[2] = {std::vector<int>((size_t)nranks, 0), std::vector<int>((size_t)nranks, 0)};
  const int nparts = std::max(nsub_per_rank, 1);
  // the ranks are independent processes in the reference: one worker each
  const int nw = std::max(1, std::min(nranks, setup_threads()));
  if (ws) {  // scratch of earlier calls is reused when it fits this matrix
    if ((int)ws->w.size() < nw) ws->w.resize((size_t)nw);
    std::vector<std::thread> mk;
    for (int t = 0; t < nw; t++)
      if (!ws->w[t] || ws->w[t]->n != n)
        mk.emplace_back([ws, t, n]() { ws->w[t] = std::make_unique<RankScratch>(n); });
    for (auto& x : mk) x.join();
  }
  // ranks differ in cost (boundary boxes are smaller): each worker takes the
  // next rank not yet taken; a rank's results do not depend on its worker
  std::atomic<int> next_rank{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nw; t++)
    th.emplace_back([&, t]() {
      set_setup_thread_cap(1);
      std::unique_ptr<RankScratch> mine;
      RankScratch* sp = ws ? ws->w[t].get() : (mine = std::make_unique<RankScratch>(n)).get();
      RankScratch& S = *sp;
      for (int r = next_rank++; r < nranks; r = next_rank++) {
        RankLocal& R = out.ranks[r];
        rank_local(As, val, eptn, r, own[r], ol_dist, S, R, keep_local, pos ? pos->data() : nullptr);
        std::copy(R.lg.begin(), R.lg.begin() + R.ninner, out.order.begin() + out.own_off[r]);
        // getLocal (Aggregate_utils.f95:33-51): the own rows, which lead the
        // local matrix (these steps read rows < nrows alone); a new matrix
        // carries no symmetry flags, so the strength test scales by rows
        Csr& LA = R.A;
        const int local_rows = LA.nrows;
        LA.nrows = R.ninner;
        std::vector<double> scaled, diag;
        scale_values(LA, false, scaled, diag);
        std::vector<uint8_t> strong = find_strong(LA, scaled, diag, strong1);
        ProcessState ps;
        aggregate(LA, scaled, strong, radius1, minasize1, maxasize1, strong1, false, ps, inner[r], full[r]);
        for (int i = 0; i < R.ninner; i++) {
          nlisted[0][r] += inner[r].num[i] > 0;
          nlisted[1][r] += full[r].num[i] > 0;
        }
        // the own part of the numbering is a few ascending runs (sent, inner
        // overlap, inner): merged run by run into global order
        std::vector<int>& ord = by_global[r];
        ord.resize((size_t)R.ninner);
        std::iota(ord.begin(), ord.end(), 0);
        const auto by_gl = [&R](int a, int b) { return R.lg[a] < R.lg[b]; };
        for (int i = 1, run = 0; i <= R.ninner; i++)
          if (i == R.ninner || R.lg[i] < R.lg[i - 1]) {
            if (run > 0) std::inplace_merge(ord.begin(), ord.begin() + run, ord.begin() + i, by_gl);
            run = i;
          }
        // Partitionings_metis_InitCoarse: inner nodes in local order, grouped
        // by the part of their aggregate (one part: all of them)
        std::vector<int> part;
        if (nparts > 1) {
          std::vector<int> xadj, adjncy;
          aggregate_graph(LA, inner[r], inner[r], xadj, adjncy);
          partition_graph(inner[r].nagr, xadj.data(), adjncy.data(), nullptr, nparts, part);
        }
        LA.nrows = local_rows;
        std::vector<int> pst((size_t)nparts + 1, 0), pnd;
        for (int i = 0; i < R.ninner; i++) {
          const int a = inner[r].num[i];
          if (a > 0) pst[(nparts > 1 ? part[a - 1] : 0) + 1]++;
        }
        for (int q = 0; q < nparts; q++) pst[q + 1] += pst[q];
        pnd.assign((size_t)pst[nparts], 0);
        std::vector<int> fill(pst.begin(), pst.end() - 1);
        for (int i = 0; i < R.ninner; i++) {
          const int a = inner[r].num[i];
          if (a > 0) pnd[fill[nparts > 1 ? part[a - 1] : 0]++] = i;
        }
        std::vector<int> ps2{0};
        for (int q = 0; q < nparts; q++)
          if (pst[q + 1] > pst[q]) ps2.push_back(pst[q + 1]);
        add_layers_all(R.A, ps2, pnd, ol_sub, sst[r], sin[r]);
        for (int& v : sin[r]) v = R.lg[v];
        if (rcm_graph) block_rcm_serial(*rcm_graph, sst[r], sin[r], dsin[r], S.floc);
        if (keep_local) {
          // the outer layer's rows are the last ones: A_ghost fills them
          const int first = R.A.nrows - R.nrecv;
          const int base = R.A.ptr[first];
          R.A.col.insert(R.A.col.end(), R.G.col.begin(), R.G.col.end());
          R.A.val.insert(R.A.val.end(), R.G.val.begin(), R.G.val.end());
          R.A.src.insert(R.A.src.end(), R.G.src.begin(), R.G.src.end());
          for (int k = 0; k < R.nrecv; k++) R.A.ptr[first + k + 1] = base + R.G.ptr[k + 1];
          R.G = Csr();
          R.agg = inner[r];
        } else {
          R.A = Csr();  // only the numbering is kept
        }
      }
      set_setup_thread_cap(0);
    });
  for (auto& x : th) x.join();
  // rank-major global numbering of aggregates and subdomains: offsets by a
  // scan over the ranks, then every rank fills its own slices in parallel
  std::vector<int> off_in((size_t)nranks + 1, 0), off_full((size_t)nranks + 1, 0);
  std::vector<int> off_sub((size_t)nranks + 1, 0), off_ent((size_t)nranks + 1, 0);
  for (int r = 0; r < nranks; r++) {
    off_in[r + 1] = off_in[r] + inner[r].nagr;
    off_full[r + 1] = off_full[r] + full[r].nagr;
    off_sub[r + 1] = off_sub[r] + (int)sst[r].size() - 1;
    off_ent[r + 1] = off_ent[r] + (int)sin[r].size();
  }
  Aggregates* lists[2] = {&out.inner, &out.full};
  const int nagr_tot[2] = {off_in[nranks], off_full[nranks]};
  for (int w = 0; w < 2; w++) {
    lists[w]->nagr = nagr_tot[w];
    lists[w]->num.assign((size_t)n, 0);
    lists[w]->starts.assign((size_t)nagr_tot[w] + 1, 0);
    lists[w]->nodes.assign((size_t)n, 0);
  }
  out.sstarts.assign((size_t)off_sub[nranks] + 1, 0);
  out.sinds.assign((size_t)off_ent[nranks], 0);
  out.dinds.assign(rcm_graph ? (size_t)off_ent[nranks] : 0, 0);
  // first listed node of each rank in either list
  std::vector<int> node_off[2] = {std::vector<int>((size_t)nranks + 1, 0),
                                  std::vector<int>((size_t)nranks + 1, 0)};
  for (int r = 0; r < nranks; r++)
    for (int w = 0; w < 2; w++) node_off[w][r + 1] = node_off[w][r] + nlisted[w][r];
  std::vector<std::thread> th2;
  for (int t = 0; t < nw; t++)
    th2.emplace_back([&, t]() {
      std::vector<int> cnt;
      for (int r = t; r < nranks; r += nw) {
        const RankLocal& R = out.ranks[r];
        // own nodes ascend in global number; listed per aggregate in that order
        // (Form_Aggr lists each aggregate's nodes in node order)
        for (int w = 0; w < 2; w++) {
          const Aggregates& a = w == 0 ? inner[r] : full[r];
          Aggregates& g = *lists[w];
          const int off = w == 0 ? off_in[r] : off_full[r];
          cnt.assign((size_t)a.nagr + 1, 0);
          for (int i = 0; i < R.ninner; i++)
            if (a.num[i] > 0) cnt[a.num[i]]++;
          int base = node_off[w][r];
          for (int c = 1; c <= a.nagr; c++) {
            g.starts[off + c - 1] = base;
            base += cnt[c];
            cnt[c] = g.starts[off + c - 1];
          }
          for (int i : by_global[r]) {
            const int c = a.num[i];
            if (c <= 0) continue;
            const int gi = R.lg[i];
            g.num[gi] = c + off;
            g.nodes[cnt[c]++] = gi;
          }
        }
        for (size_t q = 1; q < sst[r].size(); q++) out.sstarts[off_sub[r] + q] = off_ent[r] + sst[r][q];
        std::copy(sin[r].begin(), sin[r].end(), out.sinds.begin() + off_ent[r]);
        if (rcm_graph) std::copy(dsin[r].begin(), dsin[r].end(), out.dinds.begin() + off_ent[r]);
      }
    });
  for (auto& x : th2) x.join();
  for (int w = 0; w < 2; w++) {
    Aggregates& g = *lists[w];
    const int tot = node_off[w][nranks];
    g.nodes.resize((size_t)tot);
    g.starts[g.nagr] = tot;
  }
}

void ranks_coarse_space(const RankSetup& rs, int n, const std::vector<double>& val_rt,
                        const std::vector<double>& val_in, const std::vector<uint8_t>& strong,
                        int smoothers, double omega, Csr& R, Csr& Rt, Csr& AC) {
  const int nranks = (int)rs.ranks.size();
  std::vector<int> off((size_t)nranks + 1, 0);
  for (int r = 0; r < nranks; r++) off[r + 1] = off[r] + rs.ranks[r].agg.nagr;
  const int nagr = off[nranks];
  const int nw = std::max(1, std::min(nranks, setup_threads()));
  // (1) every rank's rows of R: IntRestBuild on its matrix with A_ghost's
  // rows, for its own aggregates, in its local numbering
  std::vector<Csr> Rp((size_t)nranks);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nw; t++)
      th.emplace_back([&, t]() {
        set_setup_thread_cap(1);
        for (int r = t; r < nranks; r += nw) {
          const RankLocal& L = rs.ranks[r];
          const int nl = L.A.nrows, first_ghost = nl - L.nrecv;
          Csr M;
          M.nrows = M.ncols = nl;
          M.ptr = L.A.ptr;
          M.col = L.A.col;
          M.val.resize(L.A.col.size());
          std::vector<uint8_t> st(L.A.col.size()), gh((size_t)nl, 0);
          for (int i = 0; i < nl; i++) {
            const bool g = i >= first_ghost;
            gh[i] = g;
            for (int p = M.ptr[i]; p < M.ptr[i + 1]; p++) {
              const int e = L.A.src[p];
              M.val[p] = g ? val_in[e] : val_rt[e];
              st[p] = strong[e];
            }
          }
          Aggregates ag = L.agg;
          ag.num.resize((size_t)nl, 0);
          Rp[r] = restriction(M, st, ag, smoothers, omega, &gh);
        }
        set_setup_thread_cap(0);
      });
    for (auto& x : th) x.join();
  }
  // (2) the global R: rank r's rows are aggregates off[r].., columns mapped
  // to global numbers and sorted
  R = Csr();
  R.nrows = nagr;
  R.ncols = n;
  R.ptr.assign((size_t)nagr + 1, 0);
  for (int r = 0; r < nranks; r++)
    for (int a = 0; a < Rp[r].nrows; a++) R.ptr[off[r] + a + 1] = Rp[r].ptr[a + 1] - Rp[r].ptr[a];
  for (int a = 0; a < nagr; a++) R.ptr[a + 1] += R.ptr[a];
  R.col.resize((size_t)R.ptr[nagr]);
  R.val.resize((size_t)R.ptr[nagr]);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nw; t++)
      th.emplace_back([&, t]() {
        std::vector<std::pair<int, double>> row;
        for (int r = t; r < nranks; r += nw) {
          const Csr& P = Rp[r];
          const std::vector<int>& lg = rs.ranks[r].lg;
          for (int a = 0; a < P.nrows; a++) {
            row.clear();
            for (int p = P.ptr[a]; p < P.ptr[a + 1]; p++) row.push_back({lg[P.col[p]], P.val[p]});
            std::sort(row.begin(), row.end(),
                      [](const auto& x, const auto& y) { return x.first < y.first; });
            int q = R.ptr[off[r] + a];
            for (const auto& e : row) {
              R.col[q] = e.first;
              R.val[q] = e.second;
              q++;
            }
          }
          Rp[r] = Csr();
        }
      });
    for (auto& x : th) x.join();
  }
  Rt = transpose(R);
  // (3) each rank's part of A_c: T = A R^T on its own rows (its entry
  // order), then R(:, own) T with the own nodes in local order
  struct Part {
    std::vector<int> row, col;
    std::vector<double> val;
  };
  std::vector<Part> parts((size_t)nranks);
  {
    std::vector<std::thread> th;
    for (int t = 0; t < nw; t++)
      th.emplace_back([&, t]() {
        std::vector<int> mark((size_t)nagr, -1), pos((size_t)nagr, 0);
        std::vector<int> tptr, tcol, cnt, rtp, rti;
        std::vector<double> tval, rtv;
        for (int r = t; r < nranks; r += nw) {
          const RankLocal& L = rs.ranks[r];
          const int ni = L.ninner;
          // T rows of the own nodes (local order); first touch assigns
          tptr.assign((size_t)ni + 1, 0);
          tcol.clear();
          tval.clear();
          for (int i = 0; i < ni; i++) {
            const int start = (int)tcol.size();
            for (int p = L.A.ptr[i]; p < L.A.ptr[i + 1]; p++) {
              const double aij = val_rt[L.A.src[p]];
              const int j = L.lg[L.A.col[p]];
              for (int q = Rt.ptr[j]; q < Rt.ptr[j + 1]; q++) {
                const int b = Rt.col[q];
                const double prod = aij * Rt.val[q];
                if (mark[b] != i || pos[b] < start) {
                  mark[b] = i;
                  pos[b] = (int)tcol.size();
                  tcol.push_back(b);
                  tval.push_back(prod);
                } else {
                  tval[pos[b]] = tval[pos[b]] + prod;
                }
              }
            }
            tptr[i + 1] = (int)tcol.size();
          }
          for (int i = 0; i < ni; i++)
            for (int q = tptr[i]; q < tptr[i + 1]; q++) mark[tcol[q]] = -1;
          // R(:, own): the coarse rows reaching own nodes, each row's nodes
          // in local order
          std::vector<int> rows;
          for (int i = 0; i < ni; i++) {
            const int g = L.lg[i];
            for (int q = Rt.ptr[g]; q < Rt.ptr[g + 1]; q++)
              if (mark[Rt.col[q]] == -1) {
                mark[Rt.col[q]] = -2;
                rows.push_back(Rt.col[q]);
              }
          }
          for (size_t k = 0; k < rows.size(); k++) pos[rows[k]] = (int)k;
          cnt.assign(rows.size() + 1, 0);
          for (int i = 0; i < ni; i++) {
            const int g = L.lg[i];
            for (int q = Rt.ptr[g]; q < Rt.ptr[g + 1]; q++) cnt[pos[Rt.col[q]] + 1]++;
          }
          for (size_t k = 0; k < rows.size(); k++) cnt[k + 1] += cnt[k];
          rtp.assign(cnt.begin(), cnt.end());
          rti.assign((size_t)cnt.back(), 0);
          rtv.assign((size_t)cnt.back(), 0.0);
          for (int i = 0; i < ni; i++) {
            const int g = L.lg[i];
            for (int q = Rt.ptr[g]; q < Rt.ptr[g + 1]; q++) {
              const int k = pos[Rt.col[q]];
              rti[cnt[k]] = i;
              rtv[cnt[k]] = Rt.val[q];
              cnt[k]++;
            }
          }
          for (int a : rows) mark[a] = -1;
          // LAC = R(:, own) T, rows in the order found, columns first touch
          Part& out = parts[r];
          for (size_t k = 0; k < rows.size(); k++) {
            const int a = rows[k];
            const int start = (int)out.col.size();
            for (int p = rtp[k]; p < rtp[k + 1]; p++) {
              const int i = rti[p];
              const double rai = rtv[p];
              for (int q = tptr[i]; q < tptr[i + 1]; q++) {
                const int b = tcol[q];
                const double prod = rai * tval[q];
                if (mark[b] != (int)k || pos[b] < start) {
                  mark[b] = (int)k;
                  pos[b] = (int)out.col.size();
                  out.row.push_back(a);
                  out.col.push_back(b);
                  out.val.push_back(prod);
                } else {
                  out.val[pos[b]] = out.val[pos[b]] + prod;
                }
              }
            }
            for (size_t q = start; q < out.col.size(); q++) mark[out.col[q]] = -1;
          }
        }
      });
    for (auto& x : th) x.join();
  }
  // (4) gathered in rank order and consolidated: entries of a row sorted by
  // column (stable), equal ones added in rank order
  AC = Csr();
  AC.nrows = AC.ncols = nagr;
  std::vector<int> rcnt((size_t)nagr + 1, 0);
  for (const Part& P : parts)
    for (int a : P.row) rcnt[a + 1]++;
  for (int a = 0; a < nagr; a++) rcnt[a + 1] += rcnt[a];
  std::vector<int> gcol((size_t)rcnt[nagr]);
  std::vector<double> gval((size_t)rcnt[nagr]);
  {
    std::vector<int> fill(rcnt.begin(), rcnt.end() - 1);
    for (const Part& P : parts)
      for (size_t k = 0; k < P.row.size(); k++) {
        const int q = fill[P.row[k]]++;
        gcol[q] = P.col[k];
        gval[q] = P.val[k];
      }
  }
  parts.clear();
  std::vector<CsrPiece> pieces((size_t)nw);
  parallel_chunks(nagr, nw, [&](int t, int b, int e) {
    CsrPiece& out = pieces[t];
    std::vector<std::pair<int, double>> row;
    for (int a = b; a < e; a++) {
      row.clear();
      for (int q = rcnt[a]; q < rcnt[a + 1]; q++) row.push_back({gcol[q], gval[q]});
      std::stable_sort(row.begin(), row.end(),
                       [](const auto& x, const auto& y) { return x.first < y.first; });
      int len = 0;
      for (size_t q = 0; q < row.size();) {
        double v = row[q].second;
        size_t u = q + 1;
        for (; u < row.size() && row[u].first == row[q].first; u++) v = v + row[u].second;
        out.col.push_back(row[q].first);
        out.val.push_back(v);
        len++;
        q = u;
      }
      out.rowlen.push_back(len);
    }
  });
  join_pieces(pieces, nagr, AC);
}

}  // namespace doug
