// Host-side setup algorithms (onsas.jl_amd/csrc/host_setup.hpp) on meshes built
// here, with invariants checked; compiled with g++ under AddressSanitizer and
// UndefinedBehaviorSanitizer by tests/test_host.py.  Exit status 0 and a line
// "ok <mesh>" per mesh when every check holds.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "host_setup.hpp"

using namespace onsas::host;

#define CHECK(cond)                                                          \
  do {                                                                       \
    if (!(cond)) {                                                           \
      std::fprintf(stderr, "check failed: %s (line %d)\n", #cond, __LINE__); \
      std::exit(1);                                                          \
    }                                                                        \
  } while (0)

struct Mesh {
  int64_t nn = 0;
  std::vector<double> X;  // 3 per node
  std::vector<int32_t> tets, trusses;
  std::vector<uint8_t> fixmask;
};

// n^3 hexahedra split in six tets (the reference's split), x0 / y0 / z0 rollers
Mesh cube(int n) {
  Mesh m;
  const int np = n + 1;
  m.nn = (int64_t)np * np * np;
  auto id = [&](int i, int j, int k) { return (int32_t)(i + np * (j + np * k)); };
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i) m.X.insert(m.X.end(), {(double)i / n, (double)j / n, (double)k / n});
  for (int k = 0; k < n; ++k)
    for (int j = 0; j < n; ++j)
      for (int i = 0; i < n; ++i) {
        const int32_t n1 = id(i, j, k), n2 = id(i, j, k + 1), n3 = id(i, j + 1, k + 1),
                      n4 = id(i, j + 1, k), n5 = id(i + 1, j, k), n6 = id(i + 1, j, k + 1),
                      n7 = id(i + 1, j + 1, k + 1), n8 = id(i + 1, j + 1, k);
        const int32_t t[6][4] = {{n1, n4, n2, n6}, {n6, n2, n3, n4}, {n4, n3, n6, n7},
                                 {n4, n1, n5, n6}, {n4, n6, n5, n8}, {n4, n7, n6, n8}};
        for (auto& e : t) m.tets.insert(m.tets.end(), e, e + 4);
      }
  m.fixmask.assign(m.nn, 0);
  for (int k = 0; k < np; ++k)
    for (int j = 0; j < np; ++j)
      for (int i = 0; i < np; ++i) {
        uint8_t f = 0;
        if (i == 0) f |= 1;
        if (j == 0) f |= 2;
        if (k == 0) f |= 4;
        m.fixmask[id(i, j, k)] = f;
      }
  return m;
}

// the cube's nodes with the tets shuffled and renumbered (an unstructured
// order) plus trusses between random node pairs and one isolated, fixed node
Mesh scrambled(int n, unsigned seed) {
  Mesh m = cube(n);
  std::mt19937 rng(seed);
  const int64_t ne = (int64_t)m.tets.size() / 4;
  std::vector<int32_t> perm(m.nn);
  for (int64_t i = 0; i < m.nn; ++i) perm[i] = (int32_t)i;
  std::shuffle(perm.begin(), perm.end(), rng);
  std::vector<int32_t> order(ne);
  for (int64_t e = 0; e < ne; ++e) order[e] = (int32_t)e;
  std::shuffle(order.begin(), order.end(), rng);
  std::vector<int32_t> t2;
  for (int32_t e : order)
    for (int a = 0; a < 4; ++a) t2.push_back(perm[m.tets[4 * e + a]]);
  m.tets.swap(t2);
  std::vector<uint8_t> f2(m.nn + 1, 7);
  for (int64_t i = 0; i < m.nn; ++i) f2[perm[i]] = m.fixmask[i];
  m.fixmask.swap(f2);
  std::uniform_int_distribution<int32_t> pick(0, (int32_t)m.nn - 1);
  for (int b = 0; b < 40; ++b) {
    const int32_t a = pick(rng), c = pick(rng);
    if (a != c) m.trusses.insert(m.trusses.end(), {a, c});
  }
  m.nn += 1;  // the isolated node (fixed)
  m.X.insert(m.X.end(), {2.0, 2.0, 2.0});
  return m;
}

// staged gather of the tets: every (row, round) of the sliced-ELL layout is one
// lane of one step of its row's workgroup, and that lane's entries decode to
// the block's contributions in ascending element order
void check_staged(const Mesh& m, const Pattern& P) {
  const int64_t ne = (int64_t)m.tets.size() / 4, nn = m.nn;
  std::vector<int32_t> slot(16 * ne);
  for (int64_t e = 0; e < ne; ++e)
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b)
        slot[(a * 4 + b) * ne + e] = slot_of(P, m.tets[4 * e + a], m.tets[4 * e + b]);
  const int64_t nslices = (nn + 63) / 64;
  std::vector<int32_t> rounds(nslices, 0);
  for (int64_t r = 0; r < nn; ++r)
    rounds[r / 64] = std::max(rounds[r / 64], P.row_ptr[r + 1] - P.row_ptr[r]);
  // expected contributions per block, ascending element
  std::vector<std::vector<uint32_t>> want(P.col.size());
  for (int64_t e = 0; e < ne; ++e)
    for (int ab = 0; ab < 16; ++ab) want[slot[ab * ne + e]].push_back((uint32_t)((e << 4) | ab));
  for (int rows : {64, 32, 16, 8}) {
    StagedGather G;
    CHECK(!staged_gather(nn, P.row_ptr, P.diag, slot.data(), ne, rounds, rows, 0, G) || ne == 0);
    CHECK(staged_gather(nn, P.row_ptr, P.diag, slot.data(), ne, rounds, rows, kMaxStage, G));
    CHECK(G.rows == rows);
    std::vector<int> seen(P.col.size(), 0);
    for (const StageGroup& g : G.groups) {
      CHECK(g.row0 % rows == 0 && g.row0 < nn && g.nstage <= G.max_stage);
      const int32_t* sl = &G.stage[g.stage_off];
      for (int32_t i = 1; i < g.nstage; ++i) CHECK(sl[i - 1] < sl[i]);
      // the listed elements are exactly those with a node among the rows
      std::vector<int32_t> need;
      for (int64_t e = 0; e < ne; ++e)
        for (int a = 0; a < 4; ++a) {
          const int32_t r = m.tets[4 * e + a];
          if (r >= g.row0 && r < g.row0 + rows) {
            need.push_back((int32_t)e);
            break;
          }
        }
      CHECK((int32_t)need.size() == g.nstage && std::equal(need.begin(), need.end(), sl));
      for (int32_t t = 0; t < g.nsteps; ++t) {
        const StageStep& st = G.steps[g.step0 + t];
        CHECK(st.qwidth % 2 == 0 && st.m0 % (64 / rows) == 0);
        CHECK(st.cbase + 32 * (int64_t)st.qwidth <= (int64_t)G.words.size());
        for (int l = 0; l < 64; ++l) {
          const int64_t r = g.row0 + l % rows;
          const int mm = st.m0 + l / rows;
          const bool has = r < nn && mm < P.row_ptr[r + 1] - P.row_ptr[r];
          const std::vector<uint32_t>* w = has ? &want[P.row_ptr[r] + mm] : nullptr;
          if (has) seen[P.row_ptr[r] + mm]++;
          for (int q = 0; q < st.qwidth; ++q) {
            const uint32_t word = G.words[st.cbase + 64 * (q / 2) + l];
            const uint32_t en = (q & 1) ? word >> 16 : word & 0xffffu;
            if (!has || q >= (int)w->size()) {
              CHECK(en == kNoStage);
              continue;
            }
            const uint32_t j = en >> 4;
            CHECK((int32_t)j < g.nstage);
            CHECK((((uint32_t)sl[j]) << 4 | (en & 15u)) == (*w)[q]);
          }
        }
      }
    }
    for (size_t b = 0; b < P.col.size(); ++b) CHECK(seen[b] == 1);
    // workgroup g holds rows g * rows on; its incidences are its rows' diagonal
    // contributions, row after row, from pair_row[r] on
    CHECK((int64_t)G.pair_row.size() == nn);
    for (size_t gi = 0; gi < G.groups.size(); ++gi) {
      const StageGroup& g = G.groups[gi];
      CHECK(g.row0 == (int64_t)gi * rows && g.npairs <= G.max_pairs);
      int32_t off = 0;
      for (int64_t r = g.row0; r < std::min<int64_t>(nn, g.row0 + rows); ++r) {
        CHECK(G.pair_row[r] == off);
        if (P.diag[r] < 0) continue;
        for (uint32_t w : want[P.diag[r]]) {
          const uint32_t pr = G.pairs[g.pair_off + off++];
          CHECK((pr & 0x0fffffffu) == (w >> 4) && (pr >> 28) == ((w >> 2) & 3u));
          CHECK(m.tets[4 * (w >> 4) + ((w >> 2) & 3u)] == r);
        }
      }
      CHECK(off == g.npairs);
    }
    CHECK(G.pairs.size() == (size_t)(4 * ne));
  }
}

void check_mesh(const Mesh& m, const char* name) {
  std::vector<ConnView> views = {{m.tets.data(), 4, (int64_t)m.tets.size() / 4}};
  if (!m.trusses.empty()) views.push_back({m.trusses.data(), 2, (int64_t)m.trusses.size() / 2});
  const Pattern P = block_pattern(m.nn, views);
  CHECK((int64_t)P.row_ptr.size() == m.nn + 1 && P.row_ptr[0] == 0);
  CHECK((size_t)P.row_ptr[m.nn] == P.col.size());
  for (int64_t i = 0; i < m.nn; ++i) {
    for (int32_t k = P.row_ptr[i] + 1; k < P.row_ptr[i + 1]; ++k) CHECK(P.col[k - 1] < P.col[k]);
    const bool touched = P.row_ptr[i + 1] > P.row_ptr[i];
    CHECK(touched == (P.diag[i] >= 0));
    if (touched) CHECK(P.col[P.diag[i]] == i);
  }
  // every element pair has a slot, and the slot holds that column (symmetric pattern)
  for (const ConnView& g : views)
    for (int64_t e = 0; e < g.ne; ++e)
      for (int a = 0; a < g.nnode; ++a)
        for (int b = 0; b < g.nnode; ++b) {
          const int32_t r = g.conn[e * g.nnode + a], c = g.conn[e * g.nnode + b];
          const int32_t s = slot_of(P, r, c);
          CHECK(s >= P.row_ptr[r] && s < P.row_ptr[r + 1] && P.col[s] == c);
          const int32_t t = slot_of(P, c, r);
          CHECK(t >= P.row_ptr[c] && t < P.row_ptr[c + 1] && P.col[t] == r);
        }
  // gather order: a permutation
  for (const ConnView& g : views) {
    std::vector<int32_t> o = gather_order(g);
    std::vector<char> seen(g.ne, 0);
    for (int32_t e : o) {
      CHECK(e >= 0 && e < g.ne && !seen[e]);
      seen[e] = 1;
    }
  }
  // multigrid symbolic levels down to the coarsest, on the touched nodes' graph
  // (an untouched node gets a diagonal-only row, as the engine's fixed nodes do)
  std::vector<int32_t> rp(m.nn + 1, 0), cl;
  for (int64_t i = 0; i < m.nn; ++i) {
    if (P.row_ptr[i + 1] > P.row_ptr[i])
      cl.insert(cl.end(), P.col.begin() + P.row_ptr[i], P.col.begin() + P.row_ptr[i + 1]);
    else
      cl.push_back((int32_t)i);
    rp[i + 1] = (int32_t)cl.size();
  }
  std::vector<double> B(m.nn * 18, 0.0);
  for (int64_t i = 0; i < m.nn; ++i) {
    const double x = m.X[3 * i] - 0.5, y = m.X[3 * i + 1] - 0.5, z = m.X[3 * i + 2] - 0.5;
    double* b = &B[i * 18];
    for (int d = 0; d < 3; ++d) b[d * 6 + d] = 1.0;
    b[1 * 6 + 3] = -z, b[2 * 6 + 3] = y, b[0 * 6 + 4] = z, b[2 * 6 + 4] = -x;
    b[0 * 6 + 5] = -y, b[1 * 6 + 5] = x;
    for (int d = 0; d < 3; ++d)
      if ((m.fixmask[i] >> d) & 1u)
        for (int q = 0; q < 6; ++q) b[d * 6 + q] = 0;
  }
  int64_t n = m.nn;
  int bs = 3;
  for (int lev = 0; lev < 8; ++lev) {
    SymbolicLevel S;
    std::string err;
    CHECK(symbolic_level(rp, cl, n, bs, B, 20, 24, 16, 4, S, &err));
    CHECK(S.na >= 1 && S.na <= n);
    for (int32_t a : S.agg) CHECK(a >= 0 && a < S.na);
    CHECK((int64_t)S.p_ptr.size() == n + 1 && (int64_t)S.p_col.size() == S.nnz_p);
    for (int64_t i = 0; i < n; ++i) {
      bool own = false;
      for (int32_t p = S.p_ptr[i]; p < S.p_ptr[i + 1]; ++p) {
        CHECK(S.p_col[p] >= 0 && S.p_col[p] < S.na && S.p_row[p] == i);
        own |= S.p_col[p] == S.agg[i];
      }
      CHECK(own);
    }
    CHECK((int64_t)S.pk_ptr.size() == S.nnz_p + 1);
    for (int32_t k : S.pk) CHECK(k >= -1 && k < rp[n]);
    for (size_t q = 0; q < S.wl_k.size(); ++q)
      CHECK(S.wl_k[q] >= 0 && S.wl_k[q] < rp[n] && S.wl_p[q] >= 0 && S.wl_p[q] < S.nnz_p);
    CHECK((int64_t)S.rt_ptr.size() == S.na + 1 && S.rt_ptr[S.na] == S.nnz_p);
    for (int64_t p = 0; p < S.nnz_p; ++p) CHECK(S.rt_pos[p] >= 0 && S.rt_pos[p] < S.nnz_p);
    CHECK((int64_t)S.row_ptr.size() == S.na + 1 && (size_t)S.row_ptr[S.na] == S.col.size());
    for (int64_t I = 0; I < S.na; ++I) CHECK(S.diag[I] >= S.row_ptr[I] && S.col[S.diag[I]] == I);
    for (size_t q = 0; q < S.gl_p.size(); ++q) {
      CHECK(S.gl_p[q] >= -1 && S.gl_p[q] < S.nnz_p);
      CHECK(S.gl_w[q] >= -1 && S.gl_w[q] < S.nnz_w);
    }
    if (S.rpw > 0) {
      const int64_t nsl = (int64_t)S.s6_ptr.size() - 1;
      CHECK((int64_t)S.s6_row.size() == nsl * S.rpw);
      CHECK((int64_t)S.s6_col.size() == 64 * (int64_t)S.s6_ptr[nsl]);
      for (size_t q = 0; q < S.s6_src.size(); ++q) {
        CHECK(S.s6_src[q] >= -1 && S.s6_src[q] < (int32_t)S.col.size());
        CHECK(S.s6_col[q] >= 0 && S.s6_col[q] < S.na);
      }
    }
    // the tentative prolongator's columns are orthonormal per aggregate
    std::vector<double> G(S.na * 36, 0.0);
    for (int64_t i = 0; i < n; ++i)
      for (int r = 0; r < bs; ++r)
        for (int u = 0; u < 6; ++u)
          for (int v = 0; v < 6; ++v)
            G[S.agg[i] * 36 + u * 6 + v] += S.T[(i * bs + r) * 6 + u] * S.T[(i * bs + r) * 6 + v];
    for (int64_t a = 0; a < S.na; ++a)
      for (int u = 0; u < 6; ++u)
        for (int v = 0; v < 6; ++v) {
          const double g = G[a * 36 + u * 6 + v];
          if (u != v) CHECK(std::abs(g) < 1e-12);
          else CHECK(std::abs(g - 1.0) < 1e-12 || std::abs(g) < 1e-12);
        }
    if (S.na <= 20) break;
    rp = S.row_ptr;
    cl = S.col;
    n = S.na;
    bs = 6;
    B.swap(S.Bc);
  }
  check_staged(m, P);
  std::printf("ok %s (nodes %lld, blocks %zu)\n", name, (long long)m.nn, P.col.size());
}

int main() {
  check_mesh(cube(1), "cube1");
  check_mesh(cube(4), "cube4");
  check_mesh(cube(9), "cube9");
  check_mesh(scrambled(6, 7), "scrambled6");
  return 0;
}
