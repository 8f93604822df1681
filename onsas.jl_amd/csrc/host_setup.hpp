// host_setup.hpp -- host algorithms of the one-time setup that touch no HIP
// API: the 3x3 block pattern of the tangent, the element -> block slot maps,
// the gather assembly's element order, and the multigrid's symbolic levels
// (aggregation, tentative prolongator, the patterns of P, A P and P^T A P with
// their fixed-order sum lists, the sliced-ELL layout of a coarse operator).
// setup.hip and amg.hip call these and upload what they return; the host
// tests compile this header alone with g++ under AddressSanitizer and
// UndefinedBehaviorSanitizer.
#pragma once

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <string>
#include <vector>

namespace onsas {
namespace host {

// ---- tangent pattern, slot maps, gather order --------------------------------------

// One element group: element-major connectivity, nnode nodes per element.
struct ConnView {
  const int32_t* conn;
  int nnode;
  int64_t ne;
};

// 3x3 block-CSR pattern of the tangent: row i holds every node that shares an
// element with node i, ascending; diag[i] is the index of block (i, i), -1 for
// a node that no element touches.
struct Pattern {
  std::vector<int32_t> row_ptr, col, diag;
};

inline Pattern block_pattern(int64_t nn, const std::vector<ConnView>& groups) {
  Pattern P;
  std::vector<int64_t> cnt(nn + 1, 0);
  for (const ConnView& g : groups)
    for (int64_t e = 0; e < g.ne; ++e)
      for (int a = 0; a < g.nnode; ++a) cnt[g.conn[e * g.nnode + a] + 1] += g.nnode;
  for (int64_t i = 0; i < nn; ++i) cnt[i + 1] += cnt[i];
  std::vector<int32_t> cand(cnt[nn]);
  {
    std::vector<int64_t> pos(cnt.begin(), cnt.end() - 1);
    for (const ConnView& g : groups)
      for (int64_t e = 0; e < g.ne; ++e)
        for (int a = 0; a < g.nnode; ++a) {
          const int32_t r = g.conn[e * g.nnode + a];
          for (int b = 0; b < g.nnode; ++b) cand[pos[r]++] = g.conn[e * g.nnode + b];
        }
  }
  P.row_ptr.assign(nn + 1, 0);
  for (int64_t i = 0; i < nn; ++i) {
    auto b = cand.begin() + cnt[i], e = cand.begin() + cnt[i + 1];
    std::sort(b, e);
    auto u = std::unique(b, e);
    P.row_ptr[i + 1] = P.row_ptr[i] + (int32_t)(u - b);
  }
  P.col.resize(P.row_ptr[nn]);
  P.diag.assign(nn, -1);
  for (int64_t i = 0; i < nn; ++i)
    std::copy(cand.begin() + cnt[i], cand.begin() + cnt[i] + (P.row_ptr[i + 1] - P.row_ptr[i]),
              P.col.begin() + P.row_ptr[i]);
  for (int64_t i = 0; i < nn; ++i) {
    auto b = P.col.begin() + P.row_ptr[i], e = P.col.begin() + P.row_ptr[i + 1];
    auto it = std::lower_bound(b, e, (int32_t)i);
    if (it != e && *it == i) P.diag[i] = (int32_t)(it - P.col.begin());
  }
  return P;
}

// Index of block (r, c) in the pattern (c must be in row r).
inline int32_t slot_of(const Pattern& P, int32_t r, int32_t c) {
  auto b = P.col.begin() + P.row_ptr[r], e = P.col.begin() + P.row_ptr[r + 1];
  return (int32_t)(std::lower_bound(b, e, c) - P.col.begin());
}

// Element order of the gather assembly.  Lane l of a gather wave handles row
// r0 + l; for its q-th contribution to be next to lane l+1's in the state
// planes, elements that are translates of each other must be adjacent.  Key:
// the element's shape signature (node ids relative to its first node) and then
// its first node: on a structured mesh every translate class is one run in
// node order.  When the signatures do not repeat (unstructured meshes, more
// than 512 classes) the key falls back to the smallest node id, which keeps
// the node-numbering locality.
inline std::vector<int32_t> gather_order(const ConnView& g) {
  const int nn_e = g.nnode;
  const int64_t ne = g.ne;
  std::vector<int32_t> order(ne);
  std::iota(order.begin(), order.end(), 0);
  auto sig_less = [&](int32_t x, int32_t y) {
    const int32_t* cx = &g.conn[(size_t)x * nn_e];
    const int32_t* cy = &g.conn[(size_t)y * nn_e];
    for (int a = 1; a < nn_e; ++a) {
      const int64_t dx = (int64_t)cx[a] - cx[0], dy = (int64_t)cy[a] - cy[0];
      if (dx != dy) return dx < dy;
    }
    if (cx[0] != cy[0]) return cx[0] < cy[0];
    return x < y;
  };
  std::sort(order.begin(), order.end(), sig_less);
  int64_t classes = 1;
  for (int64_t k = 1; k < ne; ++k) {
    const int32_t* cx = &g.conn[(size_t)order[k - 1] * nn_e];
    const int32_t* cy = &g.conn[(size_t)order[k] * nn_e];
    for (int a = 1; a < nn_e; ++a)
      if (cx[a] - cx[0] != cy[a] - cy[0]) {
        ++classes;
        break;
      }
  }
  if (classes > 512) {
    std::vector<int32_t> mn(ne);
    for (int64_t e = 0; e < ne; ++e) {
      const int32_t* ce = &g.conn[(size_t)e * nn_e];
      mn[e] = *std::min_element(ce, ce + nn_e);
    }
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int32_t x, int32_t y) { return mn[x] < mn[y]; });
  }
  return order;
}

// ---- multigrid symbolic levels ---------------------------------------------------------

// greedy aggregation of a node graph (CSR, self included)
inline int64_t aggregate(const std::vector<int32_t>& rp, const std::vector<int32_t>& cl, int64_t n,
                  std::vector<int32_t>& agg) {
  agg.assign(n, -1);
  int32_t na = 0;
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    bool free = true;
    for (int32_t k = rp[i]; k < rp[i + 1] && free; ++k) free = agg[cl[k]] < 0;
    if (!free) continue;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) agg[cl[k]] = na;
    agg[i] = na++;
  }
  for (int64_t i = 0; i < n; ++i) {
    if (agg[i] >= 0) continue;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      if (agg[cl[k]] >= 0) {
        agg[i] = agg[cl[k]];
        break;
      }
  }
  for (int64_t i = 0; i < n; ++i)
    if (agg[i] < 0) agg[i] = na++;
  return na;
}

// Tentative prolongator of one level: per aggregate, modified Gram-Schmidt
// (twice) of the rows of its member nodes in the near-nullspace B (bs x 6 per
// node); T gets Q, Bc (6 x 6 per aggregate) gets R.
inline void tentative(const std::vector<int64_t>& aptr, const std::vector<int32_t>& anodes, int bs,
               const std::vector<double>& B, std::vector<double>& T, std::vector<double>& Bc) {
  const int64_t na = (int64_t)aptr.size() - 1;
  std::vector<double> V, Q, col;
  for (int64_t a = 0; a < na; ++a) {
    const int64_t m = aptr[a + 1] - aptr[a];
    const int64_t rows = m * bs;
    V.assign(rows * 6, 0.0);
    Q.assign(rows * 6, 0.0);
    for (int64_t k = 0; k < m; ++k)
      for (int r = 0; r < bs; ++r)
        for (int q = 0; q < 6; ++q) V[(k * bs + r) * 6 + q] = B[anodes[aptr[a] + k] * bs * 6 + r * 6 + q];
    double R[36] = {0};
    for (int q = 0; q < 6; ++q) {
      col.assign(rows, 0.0);
      double n0 = 0;
      for (int64_t t = 0; t < rows; ++t) {
        col[t] = V[t * 6 + q];
        n0 += col[t] * col[t];
      }
      n0 = std::sqrt(n0);
      for (int pass = 0; pass < 2; ++pass)
        for (int j = 0; j < q; ++j) {
          double d = 0;
          for (int64_t t = 0; t < rows; ++t) d += Q[t * 6 + j] * col[t];
          R[j * 6 + q] += d;
          for (int64_t t = 0; t < rows; ++t) col[t] -= d * Q[t * 6 + j];
        }
      double nv = 0;
      for (int64_t t = 0; t < rows; ++t) nv += col[t] * col[t];
      nv = std::sqrt(nv);
      if (nv > 1e-10 * n0 && nv > 0) {
        R[q * 6 + q] = nv;
        for (int64_t t = 0; t < rows; ++t) Q[t * 6 + q] = col[t] / nv;
      }
    }
    for (int64_t k = 0; k < m; ++k)
      for (int r = 0; r < bs; ++r)
        for (int q = 0; q < 6; ++q) T[anodes[aptr[a] + k] * bs * 6 + r * 6 + q] = Q[(k * bs + r) * 6 + q];
    for (int q = 0; q < 36; ++q) Bc[a * 36 + q] = R[q];
  }
}

struct Trip {  // (column key, two list entries)
  int32_t key, a, b;
};

// Group the entries of one row by key (stable: generation order within a key):
// appends the distinct keys to cols and each group's entries to la/lb, closing
// each group in ptr.
inline void group_row(std::vector<Trip>& trips, std::vector<int32_t>& cols, std::vector<int64_t>& ptr,
               std::vector<int32_t>& la, std::vector<int32_t>* lb) {
  std::stable_sort(trips.begin(), trips.end(), [](const Trip& x, const Trip& y) { return x.key < y.key; });
  for (size_t t = 0; t < trips.size(); ++t) {
    if (t > 0 && trips[t].key != trips[t - 1].key) ptr.push_back((int64_t)la.size());
    if (t == 0 || trips[t].key != trips[t - 1].key) cols.push_back(trips[t].key);
    la.push_back(trips[t].a);
    if (lb) lb->push_back(trips[t].b);
  }
  if (!trips.empty()) ptr.push_back((int64_t)la.size());
}

// The host work of one level of the hierarchy: aggregation of the level's node
// graph (rp, cl; n nodes of bs dofs), its tentative prolongator from the
// near-nullspace Bcur (bs x 6 per node), the patterns of P, W = A P and the
// next level's operator P^T W with their fixed-order sum lists, and the
// sliced-ELL layout (rpw rows per slice) of the next level when it has more
// than max_coarse_nodes and at least sell6_min_rows nodes.
struct SymbolicLevel {
  int64_t na = 0;                                 // nodes of the next level
  std::vector<int32_t> agg;                       // node -> aggregate
  std::vector<double> T, Bc;                      // tentative P (bs x 6 per node), next modes
  std::vector<int32_t> p_ptr, p_col, p_row, pk;   // P pattern and its operator-block lists
  std::vector<int64_t> pk_ptr;
  int64_t nnz_p = 0, nnz_w = 0;
  std::vector<int32_t> wl_k, wl_p;                // W blocks: (operator block, P block) pairs
  std::vector<int64_t> wl_ptr;
  std::vector<int32_t> rt_ptr, rt_row, rt_pos;    // P^T order
  std::vector<int32_t> row_ptr, col, diag;        // next level's operator pattern
  std::vector<int32_t> gl_p, gl_w;                // its (P block, W block) pairs
  std::vector<int64_t> gl_ptr;
  int rpw = 0;                                    // rows per slice of its sliced ELL (0: none)
  std::vector<int32_t> s6_ptr, s6_col, s6_src, s6_row;
};

// The column window of each group of `slices_per_group` consecutive slices of
// a sliced-ELL level (one workgroup of the product kernel): lo (even) and span
// (even unless it ends at node n, at most cap) such that [lo, lo + span) holds
// the group's columns when they fit in cap nodes, and lo + span <= n.  Padding
// slots (s6_src < 0) are ignored.  Returns lo and span interleaved per group.
inline std::vector<int32_t> sell6_windows(const std::vector<int32_t>& s6_ptr,
                                          const std::vector<int32_t>& s6_col,
                                          const std::vector<int32_t>& s6_src, int64_t n,
                                          int slices_per_group, int32_t cap) {
  const int64_t nsl = (int64_t)s6_ptr.size() - 1;
  const int64_t ng = (nsl + slices_per_group - 1) / slices_per_group;
  std::vector<int32_t> w(2 * ng, 0);
  for (int64_t g = 0; g < ng; ++g) {
    int32_t lo = INT32_MAX, hi = -1;
    const int64_t s0 = g * slices_per_group, s1 = std::min<int64_t>(nsl, s0 + slices_per_group);
    for (int64_t q = 64 * (int64_t)s6_ptr[s0]; q < 64 * (int64_t)s6_ptr[s1]; ++q)
      if (s6_src[q] >= 0) {
        lo = std::min(lo, s6_col[q]);
        hi = std::max(hi, s6_col[q]);
      }
    if (hi < 0) lo = hi = 0;
    lo &= ~1;
    int64_t span = std::min<int64_t>((int64_t)hi - lo + 1, cap);
    span = (span + 1) & ~int64_t(1);
    span = std::min<int64_t>(span, n - lo);
    w[2 * g] = lo;
    w[2 * g + 1] = (int32_t)span;
  }
  return w;
}

inline bool symbolic_level(const std::vector<int32_t>& rp, const std::vector<int32_t>& cl,
                           int64_t n, int bs, const std::vector<double>& Bcur,
                           int64_t max_coarse_nodes, int64_t sell6_min_rows, int sell6_window,
                           int rpw, SymbolicLevel& S, std::string* err) {
  std::vector<int32_t>& agg = S.agg;
  const int64_t na = aggregate(rp, cl, n, agg);
  S.na = na;
  std::vector<int64_t> aptr(na + 1, 0);
  for (int64_t i = 0; i < n; ++i) aptr[agg[i] + 1]++;
  for (int64_t a = 0; a < na; ++a) aptr[a + 1] += aptr[a];
  std::vector<int32_t> anodes(n);
  {
    std::vector<int64_t> pos(aptr.begin(), aptr.end() - 1);
    for (int64_t i = 0; i < n; ++i) anodes[pos[agg[i]]++] = (int32_t)i;
  }
  S.T.assign(n * bs * 6, 0.0);
  S.Bc.assign(na * 36, 0.0);
  tentative(aptr, anodes, bs, Bcur, S.T, S.Bc);
  std::vector<Trip> trips;

  // P pattern (row i: the aggregates of its neighbours) and its lists: the
  // operator blocks of row i per aggregate, in row order
  std::vector<int32_t>&p_ptr = S.p_ptr, &p_col = S.p_col, &p_row = S.p_row, &pk = S.pk;
  std::vector<int64_t>& pk_ptr = S.pk_ptr;
  p_ptr.assign(n + 1, 0);
  pk_ptr.assign(1, 0);
  for (int64_t i = 0; i < n; ++i) {
    trips.clear();
    bool own = false;
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
      trips.push_back({agg[cl[k]], k, 0});
      own |= agg[cl[k]] == agg[i];
    }
    // every row gets a block in its own aggregate (a node without a diagonal
    // block in the pattern -- an isolated, fixed node -- through an empty entry)
    if (!own) trips.push_back({agg[i], -1, 0});
    const size_t before = p_col.size();
    group_row(trips, p_col, pk_ptr, pk, nullptr);
    p_row.insert(p_row.end(), p_col.size() - before, (int32_t)i);
    p_ptr[i + 1] = (int32_t)p_col.size();
  }
  const int64_t nnz_p = (int64_t)p_col.size();
  S.nnz_p = nnz_p;
  // W = A P pattern (row i: the P columns of its neighbours) and its lists of
  // (operator block, P block) pairs, in row and P order
  std::vector<int32_t> w_ptr(n + 1, 0), w_col;
  std::vector<int32_t>&wl_k = S.wl_k, &wl_p = S.wl_p;
  std::vector<int64_t>& wl_ptr = S.wl_ptr;
  wl_ptr.assign(1, 0);
  for (int64_t i = 0; i < n; ++i) {
    trips.clear();
    for (int32_t k = rp[i]; k < rp[i + 1]; ++k)
      for (int32_t p = p_ptr[cl[k]]; p < p_ptr[cl[k] + 1]; ++p) trips.push_back({p_col[p], k, p});
    group_row(trips, w_col, wl_ptr, wl_k, &wl_p);
    w_ptr[i + 1] = (int32_t)w_col.size();
  }
  const int64_t nnz_w = (int64_t)w_col.size();
  S.nnz_w = nnz_w;
  // P^T: per aggregate its P blocks, in row order
  std::vector<int32_t>&rt_ptr = S.rt_ptr, &rt_row = S.rt_row, &rt_pos = S.rt_pos;
  std::vector<int32_t> rt(nnz_p);
  rt_ptr.assign(na + 1, 0);
  rt_row.assign(nnz_p, 0);
  rt_pos.assign(nnz_p, 0);
  for (int64_t p = 0; p < nnz_p; ++p) rt_ptr[p_col[p] + 1]++;
  for (int64_t a = 0; a < na; ++a) rt_ptr[a + 1] += rt_ptr[a];
  {
    std::vector<int32_t> pos(rt_ptr.begin(), rt_ptr.end() - 1);
    for (int64_t p = 0; p < nnz_p; ++p) {
      const int32_t e = pos[p_col[p]]++;
      rt[e] = (int32_t)p;
      rt_row[e] = p_row[p];
      rt_pos[p] = e;
    }
  }
  // next level: pattern of P^T W and its lists of (P block, W block) pairs
  std::vector<int32_t>&c_row_ptr = S.row_ptr, &c_col = S.col, &gl_p = S.gl_p, &gl_w = S.gl_w,
                     &diag = S.diag;
  std::vector<int64_t>& gl_ptr = S.gl_ptr;
  c_row_ptr.assign(na + 1, 0);
  diag.assign(na, -1);
  gl_ptr.assign(1, 0);
  for (int64_t I = 0; I < na; ++I) {
    trips.clear();
    bool has_diag = false;
    for (int32_t e = rt_ptr[I]; e < rt_ptr[I + 1]; ++e) {
      const int32_t p = rt[e], i = p_row[p];
      for (int32_t w = w_ptr[i]; w < w_ptr[i + 1]; ++w) {
        trips.push_back({w_col[w], p, w});
        has_diag |= w_col[w] == I;
      }
    }
    if (!has_diag) trips.push_back({(int32_t)I, -1, -1});  // zero block: unit diagonal later
    const size_t before = c_col.size();
    group_row(trips, c_col, gl_ptr, gl_p, &gl_w);
    for (size_t k = before; k < c_col.size(); ++k)
      if (c_col[k] == I) diag[I] = (int32_t)k;
    c_row_ptr[I + 1] = (int32_t)c_col.size();
    if (diag[I] < 0) {
      *err = "multigrid: coarse node without a diagonal block";
      return false;
    }
  }
  // sliced-ELL layout of a large coarse operator (not the coarsest level):
  // rows ordered by decreasing length within windows, rpw per slice
  std::vector<int32_t>&s6_ptr = S.s6_ptr, &s6_col = S.s6_col, &s6_src = S.s6_src,
                     &s6_row = S.s6_row;
  s6_ptr.assign(1, 0);
  S.rpw = 0;
  if (na > max_coarse_nodes && na >= sell6_min_rows) {
    std::vector<int32_t> order(na);
    for (int64_t I = 0; I < na; ++I) order[I] = (int32_t)I;
    auto len = [&](int32_t I) { return c_row_ptr[I + 1] - c_row_ptr[I]; };
    for (int64_t w0 = 0; w0 < na; w0 += sell6_window)
      std::stable_sort(order.begin() + w0, order.begin() + std::min<int64_t>(w0 + sell6_window, na),
                       [&](int32_t x, int32_t y) { return len(x) > len(y); });
    const int lpr = 64 / rpw;
    S.rpw = rpw;
    for (int64_t r0 = 0; r0 < na; r0 += rpw) {
      int32_t width = 0;
      for (int r = 0; r < rpw; ++r) {
        const int32_t I = r0 + r < na ? order[r0 + r] : -1;
        s6_row.push_back(I);
        if (I >= 0) width = std::max(width, (len(I) + lpr - 1) / lpr);
      }
      for (int32_t m = 0; m < width; ++m)
        for (int l = 0; l < 64; ++l) {
          const int32_t I = s6_row[r0 + (l & (rpw - 1))];
          const int32_t slot = lpr * m + l / rpw;
          const bool ok = I >= 0 && slot < len(I);
          const int32_t k = ok ? c_row_ptr[I] + slot : -1;
          s6_src.push_back(k);
          s6_col.push_back(ok ? c_col[k] : 0);
        }
      s6_ptr.push_back(s6_ptr.back() + width);
    }
  }
  return true;
}

// ---- staged gather assembly ------------------------------------------------------------

// Work of one workgroup of the staged gather: `rows` consecutive tangent rows
// of one 64-row slice; workgroup g holds rows g * rows + [0, rows).  Its
// elements (every element with a node among the rows) are listed ascending in
// stage[stage_off, stage_off + nstage); the workgroup copies their state
// planes into LDS, column i for list entry i.  Its steps are
// steps[step0, step0 + nsteps).  pairs[pair_off, pair_off + npairs) are the
// (element, local node) incidences of its rows, row after row, each row's in
// the order of its diagonal block's contributions (e | a << 28): their nodal
// forces go to LDS too, row r's from pair_row[r] on.
struct StageGroup {
  int32_t row0, nstage, step0, nsteps;
  int64_t stage_off;
  int64_t pair_off;
  int32_t npairs, pad;
};
// One step of a wave: lane l on row row0 + l % rows at round m0 + l / rows.
// Its contribution entries are the words cbase + 64 k + l, k < qwidth / 2,
// two 16-bit entries (j << 4 | a << 2 | b, j the LDS column) per word, entry
// 2k in the low half; kNoStage pads.  qwidth is even.
struct StageStep {
  int64_t cbase;
  int32_t qwidth, m0;
};
constexpr uint16_t kNoStage = 0xffff;
constexpr int kMaxStage = 4095;  // LDS columns per workgroup (12-bit entry field)

struct StagedGather {
  int rows = 0;
  int max_stage = 0;
  int max_pairs = 0;
  std::vector<StageGroup> groups;
  std::vector<StageStep> steps;
  std::vector<int32_t> stage;
  std::vector<uint32_t> words;
  std::vector<uint32_t> pairs;
  std::vector<int32_t> pair_row;  // per row: its first pair, relative to its workgroup's
};

// Builds the staged gather of one element group with nnode = 4 over the
// tangent pattern (row_ptr; block (r, k) is round k of row r in the sliced-ELL
// layout, slices of 64 rows with slice_rounds[s] rounds).  slot[(a*4+b)*ne + e]
// is the block that element e's (a, b) block adds to; diag[r] the block (r, r)
// (-1: none).  rows divides 64.  The contributions of a block are listed in
// ascending element order, as the list-driven gather sums them.  Returns false
// (G untouched but `max_stage`) when a workgroup would stage more than cap
// elements.
inline bool staged_gather(int64_t nn, const std::vector<int32_t>& row_ptr,
                          const std::vector<int32_t>& diag, const int32_t* slot, int64_t ne,
                          const std::vector<int32_t>& slice_rounds, int rows, int cap,
                          StagedGather& G) {
  const int64_t nnzb = row_ptr[nn];
  cap = std::min(cap, kMaxStage);
  // contributions per block, ascending element order
  std::vector<int64_t> cptr(nnzb + 1, 0);
  for (int64_t k = 0; k < 16 * ne; ++k) cptr[slot[k] + 1]++;
  for (int64_t b = 0; b < nnzb; ++b) cptr[b + 1] += cptr[b];
  std::vector<uint32_t> cent(cptr[nnzb]);
  {
    std::vector<int64_t> pos(cptr.begin(), cptr.end() - 1);
    for (int64_t e = 0; e < ne; ++e)
      for (int ab = 0; ab < 16; ++ab) cent[pos[slot[(int64_t)ab * ne + e]]++] = (uint32_t)((e << 4) | ab);
  }
  const int64_t nslices = (int64_t)slice_rounds.size();
  const int per = 64 / rows;
  std::vector<int32_t> mark(ne, -1), loc(ne, 0);
  std::vector<int32_t> list;
  // first pass: the element lists (and the largest)
  StagedGather S;
  S.rows = rows;
  for (int64_t s = 0; s < nslices; ++s)
    for (int h = 0; h < per; ++h) {
      const int64_t r0 = 64 * s + (int64_t)h * rows;
      if (r0 >= nn) break;
      const int32_t gid = (int32_t)S.groups.size();
      list.clear();
      for (int64_t r = r0; r < std::min<int64_t>(nn, r0 + rows); ++r)
        for (int64_t q = cptr[row_ptr[r]]; q < cptr[row_ptr[r + 1]]; ++q) {
          const int64_t e = cent[q] >> 4;
          if (mark[e] != gid) {
            mark[e] = gid;
            list.push_back((int32_t)e);
          }
        }
      std::sort(list.begin(), list.end());
      S.max_stage = std::max(S.max_stage, (int)list.size());
      StageGroup g;
      g.row0 = (int32_t)r0;
      g.nstage = (int32_t)list.size();
      g.stage_off = (int64_t)S.stage.size();
      g.step0 = g.nsteps = 0;
      g.pair_off = 0;
      g.npairs = g.pad = 0;
      S.groups.push_back(g);
      S.stage.insert(S.stage.end(), list.begin(), list.end());
    }
  G.max_stage = S.max_stage;
  if (S.max_stage > cap) return false;
  // second pass: steps and entries
  struct Lane {
    int64_t blk;
    int32_t cnt;
  };
  std::vector<StageStep> wsteps;
  std::vector<std::vector<Lane>> wlanes;
  S.pair_row.assign(nn, 0);
  for (StageGroup& g : S.groups) {
    for (int32_t i = 0; i < g.nstage; ++i) loc[S.stage[g.stage_off + i]] = i;
    // the rows' incidences: the contributions of their diagonal blocks
    g.pair_off = (int64_t)S.pairs.size();
    for (int64_t r = g.row0; r < std::min<int64_t>(nn, g.row0 + (int64_t)rows); ++r) {
      S.pair_row[r] = (int32_t)(S.pairs.size() - g.pair_off);
      if (diag[r] < 0) continue;
      for (int64_t q = cptr[diag[r]]; q < cptr[diag[r] + 1]; ++q)
        S.pairs.push_back((uint32_t)((cent[q] >> 4) | (uint64_t)((cent[q] >> 2) & 3u) << 28));
    }
    g.npairs = (int32_t)(S.pairs.size() - g.pair_off);
    S.max_pairs = std::max(S.max_pairs, g.npairs);
    const int64_t s = g.row0 / 64;
    const int nst = (slice_rounds[s] + per - 1) / per;
    wsteps.assign(nst, StageStep{0, 0, 0});
    wlanes.assign(nst, std::vector<Lane>(64, Lane{-1, 0}));
    for (int t = 0; t < nst; ++t) {
      wsteps[t].m0 = t * per;
      for (int l = 0; l < 64; ++l) {
        const int64_t r = g.row0 + l % rows;
        const int m = t * per + l / rows;
        if (r >= nn || m >= row_ptr[r + 1] - row_ptr[r]) continue;
        const int64_t blk = row_ptr[r] + m;
        wlanes[t][l] = Lane{blk, (int32_t)(cptr[blk + 1] - cptr[blk])};
        wsteps[t].qwidth = std::max(wsteps[t].qwidth, wlanes[t][l].cnt);
      }
      wsteps[t].qwidth += wsteps[t].qwidth & 1;
    }
    // heaviest steps first: wave w takes steps w, w + waves, ... in this order
    std::vector<int> ord(nst);
    std::iota(ord.begin(), ord.end(), 0);
    std::stable_sort(ord.begin(), ord.end(),
                     [&](int x, int y) { return wsteps[x].qwidth > wsteps[y].qwidth; });
    g.step0 = (int32_t)S.steps.size();
    g.nsteps = nst;
    for (int t : ord) {
      StageStep st = wsteps[t];
      st.cbase = (int64_t)S.words.size();
      S.steps.push_back(st);
      S.words.resize(S.words.size() + 64 * (size_t)(st.qwidth / 2), 0xffffffffu);
      for (int l = 0; l < 64; ++l) {
        const Lane& ln = wlanes[t][l];
        for (int32_t q = 0; q < ln.cnt; ++q) {
          const uint32_t en = cent[cptr[ln.blk] + q];
          const uint32_t v = ((uint32_t)loc[en >> 4] << 4) | (en & 15u);
          uint32_t& w = S.words[st.cbase + 64 * (q / 2) + l];
          w = (q & 1) ? ((w & 0xffffu) | (v << 16)) : ((w & 0xffff0000u) | v);
        }
      }
    }
  }
  G = std::move(S);
  return true;
}

}  // namespace host
}  // namespace onsas
