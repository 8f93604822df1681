// setup.hip -- one-time host work of onsas_finalize: reference-configuration
// shape gradients, element colouring, the 3x3 block-CSR pattern and the
// element->block slot maps, then the upload of everything to HBM.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <numeric>

#include "engine.hpp"
#include "host_setup.hpp"

namespace onsas {

namespace {

template <typename T>
int dev_upload(Ctx* c, T** dst, const T* src, size_t count) {
  size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
  ONSAS_HIP_TRY(c, hipMalloc((void**)dst, bytes));
  if (count) ONSAS_HIP_TRY(c, hipMemcpy(*dst, src, count * sizeof(T), hipMemcpyHostToDevice));
  return ONSAS_OK;
}

int dev_alloc(Ctx* c, void** dst, size_t bytes) {
  ONSAS_HIP_TRY(c, hipMalloc(dst, std::max<size_t>(bytes, 16)));
  ONSAS_HIP_TRY(c, hipMemset(*dst, 0, std::max<size_t>(bytes, 16)));
  return ONSAS_OK;
}

// Shape-function gradients of a tetrahedron in the reference configuration:
// J = X dXdz^T, V = det J / 6, dN = inv(J)^T dXdz (src/Entities/Tetrahedrons.jl:
// 129-138, 193-197, 269-271).  dN is node-major (n*3 + j).
bool tet_gradients(const double X[12], double dN[12], double* V) {
  static const double D[3][4] = {{1, -1, 0, 0}, {0, -1, 0, 1}, {0, -1, 1, 0}};
  double J[9];
  for (int i = 0; i < 3; ++i)
    for (int k = 0; k < 3; ++k) {
      double s = 0;
      for (int n = 0; n < 4; ++n) s += X[n * 3 + i] * D[k][n];
      J[i * 3 + k] = s;
    }
  double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
               J[2] * (J[3] * J[7] - J[4] * J[6]);
  *V = det / 6.0;
  if (!(*V > 0)) return false;
  double Ji[9] = {(J[4] * J[8] - J[5] * J[7]) / det, (J[2] * J[7] - J[1] * J[8]) / det,
                  (J[1] * J[5] - J[2] * J[4]) / det, (J[5] * J[6] - J[3] * J[8]) / det,
                  (J[0] * J[8] - J[2] * J[6]) / det, (J[2] * J[3] - J[0] * J[5]) / det,
                  (J[3] * J[7] - J[4] * J[6]) / det, (J[1] * J[6] - J[0] * J[7]) / det,
                  (J[0] * J[4] - J[1] * J[3]) / det};
  for (int j = 0; j < 3; ++j)
    for (int n = 0; n < 4; ++n) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += Ji[k * 3 + j] * D[k][n];
      dN[n * 3 + j] = s;
    }
  return true;
}

// Greedy first-fit colouring of the element conflict graph (two elements
// conflict when they share a node).  Elements keep their user order inside a
// colour, so structured meshes stay spatially coherent within each launch.
int color_group(Ctx* c, ElemGroup& g) {
  const int nw = 4;  // 256 colours max
  std::vector<uint64_t> used((size_t)c->nn * nw, 0);
  std::vector<int16_t> color(g.ne);
  int ncol = 0;
  for (int64_t e = 0; e < g.ne; ++e) {
    uint64_t m[nw] = {0, 0, 0, 0};
    for (int a = 0; a < g.nnode; ++a) {
      const uint64_t* u = &used[(size_t)g.conn[e * g.nnode + a] * nw];
      for (int w = 0; w < nw; ++w) m[w] |= u[w];
    }
    int col = -1;
    for (int w = 0; w < nw && col < 0; ++w)
      if (~m[w]) col = w * 64 + __builtin_ctzll(~m[w]);
    if (col < 0) return set_err(c, ONSAS_ERR_UNSUPPORTED, "element colouring needs > 256 colours");
    color[e] = (int16_t)col;
    ncol = std::max(ncol, col + 1);
    for (int a = 0; a < g.nnode; ++a)
      used[(size_t)g.conn[e * g.nnode + a] * nw + col / 64] |= 1ull << (col % 64);
  }
  g.color_off.assign(ncol + 1, 0);
  for (int64_t e = 0; e < g.ne; ++e) g.color_off[color[e] + 1]++;
  for (int k = 0; k < ncol; ++k) g.color_off[k + 1] += g.color_off[k];
  std::vector<int64_t> pos(g.color_off.begin(), g.color_off.end() - 1);
  g.perm.resize(g.ne);
  for (int64_t e = 0; e < g.ne; ++e) g.perm[pos[color[e]]++] = (int32_t)e;
  return ONSAS_OK;
}

std::vector<int32_t> gather_order(const ElemGroup& g) {
  return host::gather_order({g.conn.data(), g.nnode, g.ne});
}

// Staged gather, chosen by ONSAS_GATHER=staged for a model of one tet group:
// the largest row count per workgroup (64, 32, 16 or 8) whose largest element
// list fits the LDS budget (ONSAS_GATHER_LDS_KB, default 64 KiB) and the
// registers of the staging, or the count ONSAS_GATHER_ROWS names.  The
// list-driven gather is the default: on the 64^3 cube it is faster (each
// staged workgroup waits on about three dependent HBM round trips, with
// 16 waves per CU at most).
int build_staged_gather(Ctx* c, const std::vector<std::vector<int32_t>>& slots) {
  c->stage_rows = 0;
  const char* gm = std::getenv("ONSAS_GATHER");
  if (kSellLanes != 1 || c->groups.size() != 1 || c->groups[0].type != kTet ||
      !(gm && std::string(gm) == "staged"))
    return ONSAS_OK;
  const ElemGroup& g = c->groups[0];
  const char* kb = std::getenv("ONSAS_GATHER_LDS_KB");
  int lds_max = 0;
  ONSAS_HIP_TRY(c, hipDeviceGetAttribute(&lds_max, hipDeviceAttributeMaxSharedMemoryPerBlock,
                                         c->device));
  const int64_t budget =
      std::min<int64_t>(lds_max, 1024 * (int64_t)(kb ? std::max(8, std::atoi(kb)) : 64));
  const char* fr = std::getenv("ONSAS_GATHER_ROWS");
  const int forced = fr ? std::atoi(fr) : 0;
  const int64_t per_col = (int64_t)c->kt * (g.nstate + 1);
  // list entries per thread: each holds a state column and 3 forces in registers
  const int ks_max = std::max<int>(1, std::min<int>(kStageMaxEntries,
                                                    128 / (g.nstate * (int)c->kt / 4 + 6)));
  host::StagedGather S;
  for (int rows : {64, 32, 16, 8}) {
    if (forced && rows != forced) continue;
    int steps_cap = 0;
    for (int32_t w : c->slice_rounds) steps_cap = std::max(steps_cap, (w + 64 / rows - 1) / (64 / rows));
    const int64_t cap = std::min<int64_t>(
        {(budget - 16 * (int64_t)steps_cap) / per_col, (int64_t)ks_max * kStageThreads,
         host::kMaxStage});
    if (cap < 1) continue;
    host::StagedGather T;
    if (!host::staged_gather(c->nn, c->row_ptr, c->diag, slots[0].data(), g.ne, c->slice_rounds,
                             rows, (int)cap, T))
      continue;
    const int64_t lds = 16 * (int64_t)steps_cap + 24 * (int64_t)T.max_pairs +
                        per_col * std::max(1, T.max_stage);
    const int ks = (std::max({T.max_stage, T.max_pairs, 1}) + kStageThreads - 1) /
                   kStageThreads;
    if (lds > budget || ks > ks_max) continue;
    c->stage_steps_cap = steps_cap;
    c->stage_ks = ks;
    S = std::move(T);
    break;
  }
  if (S.rows == 0) return ONSAS_OK;  // nothing fits: the list-driven gather
  c->stage_rows = S.rows;
  c->stage_cap = std::max(1, S.max_stage);
  c->stage_pairs_cap = std::max(1, S.max_pairs);
  c->nstage_groups = (int64_t)S.groups.size();
  c->stage_total = (int64_t)S.stage.size();
  c->stage_nsteps = (int64_t)S.steps.size();
  std::vector<uint8_t> cm(c->nnzb);
  for (int64_t k = 0; k < c->nnzb; ++k) cm[k] = c->fixmask[c->col[k]];
  ONSAS_TRY(dev_upload(c, &c->d_stage_groups, S.groups.data(), S.groups.size()));
  ONSAS_TRY(dev_upload(c, &c->d_stage_steps, S.steps.data(), S.steps.size()));
  ONSAS_TRY(dev_upload(c, &c->d_stage, S.stage.data(), S.stage.size()));
  ONSAS_TRY(dev_upload(c, &c->d_stage_words, S.words.data(), S.words.size()));
  ONSAS_TRY(dev_upload(c, &c->d_stage_pairs, S.pairs.data(), std::max<size_t>(S.pairs.size(), 1)));
  ONSAS_TRY(dev_upload(c, &c->d_pair_row, S.pair_row.data(), S.pair_row.size()));
  ONSAS_TRY(dev_upload(c, &c->d_colmask, cm.data(), cm.size()));
  c->ncontrib = 2 * (int64_t)S.words.size();
  return stage_gather_prepare(c);
}

// Contribution lists of the gather assembly.  Block (r, k) of the tangent
// receives the element blocks (e, a, b) whose slot is that block; they are
// listed by (global element, a, b), so every assembly sums them in the same
// order.  The lists are stored per sliced-ELL group, interleaved by lane
// (entry q of lane l at cbase + 64 q + l) so that each contribution round is
// one contiguous 256-byte load for the wave; short lists are padded with
// 0xffffffff.  Also allocates the per-element state planes.
int build_gather(Ctx* c, const std::vector<std::vector<int32_t>>& slots, int64_t rstride) {
  const int64_t nnzb = c->nnzb, nn = c->nn;
  for (ElemGroup& g : c->groups) {
    if (g.type == kTet) {
      g.hasM = (g.mat == ONSAS_MAT_SVK);
      g.nstate = g.hasM ? kTetStateM : kTetStateBase;
    } else {
      g.hasM = false;
      g.nstate = kTrussState;
    }
    ONSAS_TRY(dev_alloc(c, &g.d_state, c->kt * g.nstate * (size_t)g.ne));
    ONSAS_TRY(dev_alloc(c, (void**)&g.d_fe,
                        sizeof(double) * (g.type == kTet ? 12 : 3) * (size_t)g.ne));
  }
  c->sell_rstride = rstride;
  ONSAS_TRY(build_staged_gather(c, slots));
  if (c->stage_rows > 0) return ONSAS_OK;
  std::vector<int64_t> ptr(nnzb + 1, 0);
  for (size_t gi = 0; gi < c->groups.size(); ++gi)
    for (int32_t blk : slots[gi]) ptr[blk + 1]++;
  for (int64_t b = 0; b < nnzb; ++b) ptr[b + 1] += ptr[b];
  std::vector<uint32_t> ent(ptr[nnzb]);
  {
    std::vector<int64_t> pos(ptr.begin(), ptr.end() - 1);
    for (size_t gi = 0; gi < c->groups.size(); ++gi) {
      const ElemGroup& g = c->groups[gi];
      const int nn_e = g.nnode;
      for (int64_t e = 0; e < g.ne; ++e)
        for (int a = 0; a < nn_e; ++a)
          for (int b = 0; b < nn_e; ++b) {
            const int32_t blk = slots[gi][(size_t)(a * nn_e + b) * g.ne + e];
            ent[pos[blk]++] = (uint32_t)(((g.goff + e) << 4) | (a << 2) | b);
          }
    }
  }
  // sliced-ELL groups, slice-major: the rounds of one slice run side by side
  // and share the element states of their rows in L2
  c->sell_groups.clear();
  auto add_group = [&](int64_t s, int m) {
    SellGroup sg;
    sg.base = (int32_t)(c->sell_ptr[s] + rstride * m);
    sg.row0 = (int32_t)(kSellRows * s);
    sg.round = m;
    sg.qwidth = 0;
    sg.cbase = 0;
    for (int lane = 0; lane < 64; ++lane) {
      const int64_t r = kSellRows * s + lane % kSellRows;
      const int k = kSellLanes * m + lane / kSellRows;
      if (r < nn && k < c->row_ptr[r + 1] - c->row_ptr[r]) {
        const int64_t blk = c->row_ptr[r] + k;
        sg.qwidth = std::max<int32_t>(sg.qwidth, (int32_t)(ptr[blk + 1] - ptr[blk]));
      }
    }
    c->sell_groups.push_back(sg);
  };
  for (int64_t s = 0; s < c->nslices; ++s)
    for (int m = 0; m < c->slice_rounds[s]; ++m) add_group(s, m);
  int64_t total = 0;
  for (SellGroup& sg : c->sell_groups) {
    sg.cbase = total;
    total += 64 * (int64_t)sg.qwidth;
  }
  c->ncontrib = total;
  std::vector<uint32_t> lists(std::max<int64_t>(total, 1), 0xffffffffu);
  for (const SellGroup& sg : c->sell_groups) {
    const int64_t s = sg.row0 / kSellRows;
    for (int lane = 0; lane < 64; ++lane) {
      const int64_t r = kSellRows * s + lane % kSellRows;
      const int k = kSellLanes * sg.round + lane / kSellRows;
      if (r >= nn || k >= c->row_ptr[r + 1] - c->row_ptr[r]) continue;
      const int64_t blk = c->row_ptr[r] + k;
      for (int64_t q = ptr[blk]; q < ptr[blk + 1]; ++q)
        lists[sg.cbase + 64 * (int64_t)(q - ptr[blk]) + lane] = ent[q];
    }
  }
  ONSAS_TRY(dev_upload(c, &c->d_contrib, lists.data(), lists.size()));
  ONSAS_TRY(dev_upload(c, &c->d_sell_groups, c->sell_groups.data(), c->sell_groups.size()));
  return ONSAS_OK;
}

}  // namespace

int finalize_model(Ctx* c) {
  if (c->nn <= 0) return set_err(c, ONSAS_ERR_STATE, "onsas_set_nodes was not called");
  if (c->groups.empty()) return set_err(c, ONSAS_ERR_STATE, "no element group was added");
  const int64_t nn = c->nn, n = 3 * nn;
  if (nn >= (int64_t(1) << 31) / 3) return set_err(c, ONSAS_ERR_UNSUPPORTED, "too many nodes");

  // fixed-dof mask
  c->fixmask.assign(nn, 0);
  for (int64_t d : c->fixed) c->fixmask[d / 3] |= (uint8_t)(1u << (d % 3));
  int64_t nfixed = 0;
  for (uint8_t m : c->fixmask) nfixed += __builtin_popcount(m);
  c->nfree = n - nfixed;

  // assembly mode (gather unless asked otherwise or out of its limits)
  {
    const char* am = std::getenv("ONSAS_ASSEMBLY");
    int64_t ntot = 0;
    for (const ElemGroup& g : c->groups) ntot += g.ne;
    c->gather = !(am && std::string(am) == "scatter") &&
                (int)c->groups.size() <= kMaxGatherGroups && ntot < (int64_t(1) << 28);
  }

  // colouring; in gather mode the arrays take the gather order and the colour
  // launches index them through cidx
  for (auto& g : c->groups) {
    ONSAS_TRY(color_group(c, g));
    if (c->gather) {
      std::vector<int32_t> order = gather_order(g);
      std::vector<int32_t> pos(g.ne);
      for (int64_t k = 0; k < g.ne; ++k) pos[order[k]] = (int32_t)k;
      g.cidx.resize(g.ne);
      for (int64_t k = 0; k < g.ne; ++k) g.cidx[k] = pos[g.perm[k]];
      g.perm = std::move(order);
    }
  }

  // block pattern: every (node_a, node_b) pair of every element
  {
    std::vector<host::ConnView> views;
    for (const auto& g : c->groups) views.push_back({g.conn.data(), g.nnode, g.ne});
    host::Pattern P = host::block_pattern(nn, views);
    c->row_ptr = std::move(P.row_ptr);
    c->col = std::move(P.col);
    c->diag = std::move(P.diag);
  }
  c->nnzb = c->row_ptr[nn];
  if (c->nnzb >= (int64_t(1) << 31) / 9)
    return set_err(c, ONSAS_ERR_UNSUPPORTED, "tangent too large for 32-bit slot indices");
  // nodes that no element touches get no diagonal block: they must be fixed
  for (int64_t i = 0; i < nn; ++i)
    if (c->diag[i] < 0 && c->fixmask[i] != 7)
      return set_err(c, ONSAS_ERR_ARG,
                     "node " + std::to_string(i) + " belongs to no element and is not fixed");
  host::Pattern pat{c->row_ptr, c->col, {}};
  auto slot_of = [&](int32_t r, int32_t cc) { return host::slot_of(pat, r, cc); };

  std::vector<std::vector<int32_t>> slots_h(c->groups.size());

  // per-group device arrays in colour-sorted order
  int64_t goff = 0;
  for (size_t gi = 0; gi < c->groups.size(); ++gi) {
    ElemGroup& g = c->groups[gi];
    const int64_t ne = g.ne;
    const int nn_e = g.nnode;
    g.goff = goff;
    goff += ne;
    std::vector<int32_t> conn((size_t)nn_e * ne);
    std::vector<int32_t>& slot = slots_h[gi];
    slot.resize((size_t)nn_e * nn_e * ne);
    int ngeo = (g.type == kTet) ? 13 : 5;
    std::vector<double> geo((size_t)ngeo * ne);
    std::vector<double> Gs;
    if (g.per_elem_G) Gs.resize(ne);
    for (int64_t s = 0; s < ne; ++s) {
      int64_t e = g.perm[s];
      const int32_t* ce = &g.conn[e * nn_e];
      for (int a = 0; a < nn_e; ++a) conn[(size_t)a * ne + s] = ce[a];
      for (int a = 0; a < nn_e; ++a)
        for (int b = 0; b < nn_e; ++b) slot[(size_t)(a * nn_e + b) * ne + s] = slot_of(ce[a], ce[b]);
      if (g.type == kTet) {
        double X[12], dN[12], V;
        for (int a = 0; a < 4; ++a) {
          X[a * 3 + 0] = c->X[ce[a]];
          X[a * 3 + 1] = c->Y[ce[a]];
          X[a * 3 + 2] = c->Z[ce[a]];
        }
        if (!tet_gradients(X, dN, &V))
          return set_err(c, ONSAS_ERR_NEGATIVE_VOLUME,
                         "Element with negative volume, check connectivity (group " +
                             std::to_string(gi) + ", element " + std::to_string(e) + ")");
        for (int q = 0; q < 12; ++q) geo[(size_t)q * ne + s] = dN[q];
        geo[(size_t)12 * ne + s] = V;
      } else {
        double d0[3] = {c->X[ce[1]] - c->X[ce[0]], c->Y[ce[1]] - c->Y[ce[0]],
                        c->Z[ce[1]] - c->Z[ce[0]]};
        double l0 = std::sqrt(d0[0] * d0[0] + d0[1] * d0[1] + d0[2] * d0[2]);
        if (!(l0 > 0))
          return set_err(c, ONSAS_ERR_ARG, "truss element " + std::to_string(e) + " has zero length");
        for (int q = 0; q < 3; ++q) geo[(size_t)q * ne + s] = d0[q];
        geo[(size_t)3 * ne + s] = l0;
        geo[(size_t)4 * ne + s] = g.area[e];
      }
      if (g.per_elem_G) Gs[s] = g.Gel[e];
    }
    ONSAS_TRY(dev_upload(c, &g.d_conn, conn.data(), conn.size()));
    ONSAS_TRY(dev_upload(c, &g.d_slot, slot.data(), slot.size()));
    ONSAS_TRY(dev_upload(c, &g.d_geo, geo.data(), geo.size()));
    ONSAS_TRY(dev_upload(c, &g.d_perm, g.perm.data(), g.perm.size()));
    if (!g.cidx.empty()) ONSAS_TRY(dev_upload(c, &g.d_cidx, g.cidx.data(), g.cidx.size()));
    if (g.per_elem_G) ONSAS_TRY(dev_upload(c, &g.d_G, Gs.data(), Gs.size()));
    if (!c->gather) slot.clear();
  }

  // nodes, pattern, masks
  ONSAS_TRY(dev_upload(c, &c->d_x, c->X.data(), nn));
  ONSAS_TRY(dev_upload(c, &c->d_y, c->Y.data(), nn));
  ONSAS_TRY(dev_upload(c, &c->d_z, c->Z.data(), nn));
  ONSAS_TRY(dev_upload(c, &c->d_fixmask, c->fixmask.data(), nn));
  ONSAS_TRY(dev_upload(c, &c->d_row_ptr, c->row_ptr.data(), nn + 1));
  ONSAS_TRY(dev_upload(c, &c->d_col, c->col.data(), c->nnzb));
  ONSAS_TRY(dev_upload(c, &c->d_diag, c->diag.data(), nn));

  // pressure faces: node -> incident faces CSR (pressure-group order, then face order)
  {
    std::vector<int32_t> fconn, fgrp;
    for (size_t pg = 0; pg < c->pressure_faces.size(); ++pg) {
      const auto& f = c->pressure_faces[pg];
      fconn.insert(fconn.end(), f.begin(), f.end());
      fgrp.insert(fgrp.end(), f.size() / 3, (int32_t)pg);
    }
    c->nfaces = (int64_t)fgrp.size();
    std::vector<int32_t> ptr(nn + 1, 0);
    for (int64_t f = 0; f < c->nfaces; ++f)
      for (int a = 0; a < 3; ++a) ptr[fconn[f * 3 + a] + 1]++;
    for (int64_t i = 0; i < nn; ++i) ptr[i + 1] += ptr[i];
    std::vector<int32_t> idx(ptr[nn]);
    std::vector<int32_t> pos(ptr.begin(), ptr.end() - 1);
    for (int64_t f = 0; f < c->nfaces; ++f)
      for (int a = 0; a < 3; ++a) idx[pos[fconn[f * 3 + a]]++] = (int32_t)f;
    ONSAS_TRY(dev_upload(c, &c->d_face_conn, fconn.data(), fconn.size()));
    ONSAS_TRY(dev_upload(c, &c->d_face_grp, fgrp.data(), fgrp.size()));
    ONSAS_TRY(dev_upload(c, &c->d_nf_ptr, ptr.data(), ptr.size()));
    ONSAS_TRY(dev_upload(c, &c->d_nf_idx, idx.data(), idx.size()));
    ONSAS_TRY(dev_alloc(c, (void**)&c->d_pressures,
                        sizeof(double) * std::max<size_t>(c->pressure_faces.size(), 1)));
  }

  // state and solver vectors
  c->kt = (c->precision == ONSAS_PRECISION_FP64) ? sizeof(double) : sizeof(float);
  ONSAS_TRY(dev_alloc(c, &c->d_K, c->kt * 9 * (size_t)c->nnzb));
  ONSAS_TRY(dev_alloc(c, &c->d_Dinv, c->kt * 9 * (size_t)nn));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_U, sizeof(double) * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_Uasm, sizeof(double) * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_fint, sizeof(double) * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_fext, sizeof(double) * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_res, sizeof(double) * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_nodal, sizeof(double) * n));
  for (void** v : {&c->d_px, &c->d_pr, &c->d_pz, &c->d_pp0, &c->d_pp1, &c->d_pq})
    ONSAS_TRY(dev_alloc(c, v, c->kt * n));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_part, sizeof(double) * kPartArrays * kMaxPartials));
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_sc, sizeof(Scalars)));
  ONSAS_HIP_TRY(c, hipHostMalloc((void**)&c->h_sc, sizeof(Scalars), hipHostMallocDefault));
  ONSAS_HIP_TRY(c, hipHostMalloc((void**)&c->h_stage, sizeof(double) * n, hipHostMallocDefault));
  for (double*& h : c->h_snap)
    ONSAS_HIP_TRY(c, hipHostMalloc((void**)&h, sizeof(double) * n, hipHostMallocDefault));
  for (double*& d : c->d_snap) ONSAS_TRY(dev_alloc(c, (void**)&d, sizeof(double) * n));
  ONSAS_HIP_TRY(c, hipHostMalloc((void**)&c->h_press,
                                 sizeof(double) * std::max<size_t>(c->pressure_faces.size(), 1),
                                 hipHostMallocDefault));
  {
    const char* w = std::getenv("ONSAS_WARM_START");
    c->warm_start = !(w && w[0] == '0');
    const char* q = std::getenv("ONSAS_PCG_PREDICT");
    c->pcg_predict = !(q && q[0] == '0');
  }
  if (c->warm_start)
    for (auto& bank : c->d_warm)
      for (void*& p : bank) ONSAS_TRY(dev_alloc(c, &p, c->kt * n));

  // reduction grid: one row per thread when it fits, whole 64-row slices per
  // workgroup, a multiple of 8 workgroups (XCD mapping), at most kMaxPartials
  int64_t want = (nn + kRedThreads - 1) / kRedThreads;
  int64_t np = std::min<int64_t>(kMaxPartials, std::max<int64_t>(8, ((want + 7) / 8) * 8));
  c->np = (int)np;
  c->chunk = (((nn + np - 1) / np) + kChunkRows - 1) / kChunkRows * kChunkRows;

  // sliced-ELL layout of the tangent (kSellRows block rows per slice, padded
  // with zero blocks whose column is the row itself)
  c->nslices = (nn + kSellRows - 1) / kSellRows;
  ONSAS_TRY(dev_alloc(c, (void**)&c->d_ppq, sizeof(double) * (size_t)((c->nslices + 3) / 4)));
  c->sell_ptr.assign(c->nslices + 1, 0);
  c->slice_rounds.assign(c->nslices, 0);
  c->ell_width = 0;
  for (int64_t s = 0; s < c->nslices; ++s) {
    int32_t w = 0;
    for (int64_t r = kSellRows * s; r < std::min<int64_t>(nn, kSellRows * s + kSellRows); ++r)
      w = std::max(w, c->row_ptr[r + 1] - c->row_ptr[r]);
    w = (w + kSellLanes - 1) / kSellLanes;  // slot rounds
    c->slice_rounds[s] = w;
    c->ell_width = std::max(c->ell_width, w);
    c->sell_ptr[s + 1] = c->sell_ptr[s] + 64 * w;
  }
  // round-major when its padding costs at most 25 % over the per-slice layout
  const char* lay = std::getenv("ONSAS_SELL_LAYOUT");
  const int64_t nell = (int64_t)c->ell_width * c->nslices * 64;
  c->ell_rounds = (lay == nullptr || std::string(lay) != "slices") &&
                  nell <= c->sell_ptr[c->nslices] + c->sell_ptr[c->nslices] / 4;
  if (c->ell_rounds)
    for (int64_t s = 0; s <= c->nslices; ++s) c->sell_ptr[s] = (int32_t)(64 * s);
  c->nsell = c->ell_rounds ? nell : c->sell_ptr[c->nslices];
  const int64_t rstride = c->ell_rounds ? 64 * c->nslices : 64;  // blocks between rounds
  if (c->nsell >= (int64_t(1) << 31) / 9)
    return set_err(c, ONSAS_ERR_UNSUPPORTED, "tangent too large for 32-bit sliced-ELL offsets");
  {
    std::vector<int32_t> scol(c->nsell, 0);
    for (int64_t s = 0; s < c->nslices; ++s) {
      const int32_t rounds = c->slice_rounds[s];
      for (int m = 0; m < rounds; ++m)
        for (int lane = 0; lane < 64; ++lane) {
          const int64_t r = kSellRows * s + lane % kSellRows;
          const int k = kSellLanes * m + lane / kSellRows;
          int32_t cc = 0;
          if (r < nn) cc = (k < c->row_ptr[r + 1] - c->row_ptr[r]) ? c->col[c->row_ptr[r] + k] : (int32_t)r;
          scol[c->sell_ptr[s] + rstride * m + lane] = cc;
        }
    }
    ONSAS_TRY(dev_upload(c, &c->d_slice_rounds, c->slice_rounds.data(), c->slice_rounds.size()));
    ONSAS_TRY(dev_upload(c, &c->d_sell_ptr, c->sell_ptr.data(), c->sell_ptr.size()));
    ONSAS_TRY(dev_upload(c, &c->d_sell_col, scol.data(), scol.size()));
    ONSAS_TRY(dev_alloc(c, &c->d_Ksell, c->kt * 9 * (size_t)c->nsell));
  }
  if (c->gather) ONSAS_TRY(build_gather(c, slots_h, rstride));

  if (c->gather) ONSAS_TRY(launch_gather_init(c));
  for (auto& e : c->ev) ONSAS_HIP_TRY(c, hipEventCreate(&e));
  for (auto& e : c->ev_snap) ONSAS_HIP_TRY(c, hipEventCreate(&e));
  for (auto& e : c->ev_snap_dev) ONSAS_HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableTiming));
  ONSAS_HIP_TRY(c, hipEventCreateWithFlags(&c->ev_press, hipEventDisableTiming));
  // timing-only events (read after a stream synchronize): no system-scope fence,
  // so the stamped dispatch pays no cache writeback and invalidation
  for (auto& e : c->ev_spmv)
    ONSAS_HIP_TRY(c, hipEventCreateWithFlags(&e, hipEventDisableSystemFence));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  c->finalized = true;
  return ONSAS_OK;
}

}  // namespace onsas
