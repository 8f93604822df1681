// engine.hpp -- internal declarations of libonsas_hip (not part of the C ABI).
//
// Device data layout (all in HBM, one context per GPU):
//   nodes     : x, y, z FP64 SoA (reference coordinates; only setup and face loads read them)
//   U, f_int, f_ext, residual : FP64, dof-interleaved (dof 3i+d of node i)
//   tangent   : 3x3 block CSR over nodes; row_ptr/col int32; values KT[9] per block,
//               row-major inside the block (KT = float in mixed mode, double in FP64 mode)
//   elements  : per group, SoA arrays in colour-sorted order:
//               conn[k][ne] int32, geo[q][ne] FP64 (tet: dN 12 + V; truss: d0 3 + l0 + A),
//               slot[a*nnode+b][ne] int32 (block index of (a, b) in the tangent)
//   PCG       : x, r, z, p0, p1, q in KT, per-node block-Jacobi inverse KT[9]
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/onsas_hip.h"
#include "host_setup.hpp"

namespace onsas {

constexpr int kNodesTet = 4;
constexpr int kNodesTruss = 2;
constexpr int kRedThreads = 512;   // threads per workgroup of every reduction-producing kernel
constexpr int kRowGroup = 16;      // lanes per block row in the block-CSR SpMV (matvec API)
#ifndef ONSAS_SELL_LANES
#define ONSAS_SELL_LANES 1
#endif
#ifndef ONSAS_SPMV_THREADS
#define ONSAS_SPMV_THREADS 512
#endif
#ifndef ONSAS_SPMV_UNROLL
#define ONSAS_SPMV_UNROLL 1
#endif
constexpr int kSellLanes = ONSAS_SELL_LANES;     // lanes per block row in the sliced-ELL SpMV
constexpr int kSellRows = 64 / kSellLanes;       // block rows per slice (one wave per slice)
constexpr int kSpmvThreads = ONSAS_SPMV_THREADS; // a reduction workgroup's waves x 64
constexpr int kChunkRows = (kSpmvThreads / 64) * kSellRows;  // rows per workgroup chunk unit (whole slices)
constexpr int kMaxPartials = 1024; // upper bound of the reduction grid (multiple of 8)

enum GroupType { kTet = 0, kTruss = 1 };

// Uniform grid over a tet group's bounding box for point location: bin b
// lists (in user element order) the elements whose bounding box overlaps it.
struct LocateGrid {
  bool built = false;
  double lo[3] = {0, 0, 0}, inv_h[3] = {0, 0, 0};
  int nb[3] = {0, 0, 0};
  int64_t* d_ptr = nullptr;
  int32_t* d_elem = nullptr;  // array positions
};

struct ElemGroup {
  int type = kTet;
  int mat = 0;          // ONSAS_MAT_* for tets, strain model for trusses
  double p0 = 0, p1 = 0;  // tet: material parameters; truss: E
  int64_t ne = 0;
  int nnode = 4;
  bool per_elem_G = false;
  std::vector<int32_t> conn;   // user order, element-major
  std::vector<double> area;    // truss areas (user order)
  std::vector<double> Gel;     // per-element G (user order)
  // colouring (filled by finalize)
  std::vector<int64_t> color_off;  // ncolors + 1 offsets into the sorted order
  std::vector<int32_t> perm;       // sorted position -> user element id
  // device arrays (sorted order, SoA with stride ne)
  int32_t* d_conn = nullptr;
  double* d_geo = nullptr;
  int32_t* d_slot = nullptr;
  double* d_G = nullptr;
  int32_t* d_perm = nullptr;
  // gather mode: the arrays above are in gather order (see setup) and the
  // colour launches reach them through cidx (colour-sorted position -> position)
  std::vector<int32_t> cidx;
  int32_t* d_cidx = nullptr;
  // gather assembly: per-element tangent factors (KT planes) and nodal forces
  // (FP64 planes), both SoA with stride ne in the sorted order
  void* d_state = nullptr;
  double* d_fe = nullptr;
  int nstate = 0;
  bool hasM = false;     // SVK: the c2 s_ab M term needs M = F F' (M = I otherwise)
  int64_t goff = 0;      // first global element id of the group (contribution lists)
  LocateGrid locate;     // built on the first point location
};

// Tet state, logical planes: 0 V c1, 1 V c2, 2..13 v[a][i], 14..23 g_ab (the
// coefficient of I), SVK only: 24..33 h_ab (the coefficient of M), 34..39 M
// (00 01 02 11 12 22); where each lives in memory is tet_plane
// (element_kernels.hip: per-node 4-vectors (v_a, g_aa), the (c1, c2) pairs,
// then planes of stride ne).  Truss planes (stride ne): 0 cdd, 1 cI, 2..4 d.
// Pairs a <= b numbered 00 01 02 03 11 12 13 22 23 33.
constexpr int kTetStateBase = 24, kTetStateM = 40, kTrussState = 5;
constexpr int kMaxGatherGroups = 8;
// staged gather: threads per workgroup, and the most list entries a thread
// stages (the kernel is instantiated for 1 .. kStageMaxEntries)
constexpr int kStageThreads = 256;
constexpr int kStageMaxEntries = 4;

// One sliced-ELL group = the 64 lanes of one slice at one round: the unit of
// work (one wave) of the gather assembly.
struct SellGroup {
  int32_t base;    // block index of lane 0 in the sliced-ELL arrays
  int32_t row0;    // first row of the slice
  int32_t round;   // m
  int32_t qwidth;  // contribution rounds (max over the lanes)
  int64_t cbase;   // first contribution entry (entries of round q at cbase + 64 q + lane)
};

// Smoothed-aggregation multigrid (amg.hip).  Level 0 is the fine tangent (its
// values are the sliced-ELL copy, its block-CSR unpacking and the block-Jacobi
// inverse of the context); levels >= 1 have 6 dofs per node and 6x6 block-CSR
// values.  Each level but the coarsest owns the transfer to the next one: the
// tentative prolongator T (bs x 6 per node), the smoothed prolongator
// P = (I - w D^-1 A) T (bs x 6 blocks, block-CSR over fine rows), W = A P and
// the fixed-order lists that form P, W and the next level's P^T W.
struct AmgLevel {
  int bs = 6;                         // dofs per node
  int64_t n = 0;                      // nodes
  std::vector<int32_t> row_ptr, col;  // operator pattern (levels >= 1)
  int64_t nnzb = 0;
  int32_t *d_row_ptr = nullptr, *d_col = nullptr, *d_diag = nullptr;
  void* d_A = nullptr;                // nnzb x 36 KT (levels >= 1)
  void* d_Dinv = nullptr;             // n x 36 KT (levels >= 1)
  // sliced-ELL copy of A for the cycle (large levels >= 1 but the coarsest):
  // rows ordered by decreasing length within windows, slices of R = rpw6 of
  // them, one wave each; lane l = R ph + r serves row s6_row[R s + r] (-1:
  // none), slot (64 / R) m + ph of round m.  Round q of the level (s6_ptr[s] <= q < s6_ptr[s+1])
  // holds lane l's block column at s6_col[64 q + l] and its 36 values at
  // A6s[2304 q + ...] in lane-interleaved 16-byte chunks (Sell6Layout in
  // amg.hip), BF16 in mixed mode, FP64 in FP64 mode: every product of the
  // cycle on this level reads that copy.
  // s6_src is the source block in A (-1: padding, zero values).
  int64_t nslices6 = 0, nrounds6 = 0;
  int rpw6 = 8;                       // rows per slice: 8, 4 or 2
  int ur6 = 4;                        // rounds whose loads a lane issues together
  int waves6 = 4;                     // slices (waves) per workgroup of the product
  int32_t *d_s6_ptr = nullptr, *d_s6_col = nullptr, *d_s6_src = nullptr, *d_s6_row = nullptr;
  int32_t* d_s6_win = nullptr;        // per workgroup of 4 slices: its column window (lo, span)
  void* d_A6s = nullptr;              // the sliced-ELL values (BF16 mixed, FP64 FP64)
  double lam = 0;                     // lambda_max(D^-1 A), power iteration
  void* d_ev = nullptr;               // its last iterate (warm start, KT, bs n)
  // cycle vectors (KT, bs n): right-hand side, iterate, residual, Chebyshev steps
  void *d_b = nullptr, *d_x = nullptr, *d_r = nullptr, *d_d0 = nullptr, *d_d1 = nullptr;
  // transfer to the next coarser level
  int32_t* d_agg = nullptr;           // node -> aggregate
  void* d_T = nullptr;                // tentative prolongator, bs x 6 per node (KT)
  int64_t nnz_p = 0;                  // blocks of P
  bool restrict_wide = false;         // restriction with 2 waves per aggregate
  int32_t *d_p_ptr = nullptr, *d_p_col = nullptr, *d_p_row = nullptr;
  void* d_P = nullptr;                // nnz_p x bs x 6 (KT), row-major blocks
  int64_t* d_pk_ptr = nullptr;        // per P block: the operator blocks of its row
  int32_t* d_pk = nullptr;            //   whose column lies in its aggregate
  // P^T rows: per aggregate its P blocks in row order (rt_ptr), each entry's
  // fine row (rt_row) and a copy of P in that order (PT); rt_pos: P block -> entry
  int32_t *d_rt_ptr = nullptr, *d_rt_row = nullptr, *d_rt_pos = nullptr;
  void* d_PT = nullptr;
  int64_t nnz_w = 0;                  // blocks of W = A P
  void* d_W = nullptr;                // nnz_w x bs x 6 (KT)
  int64_t* d_wl_ptr = nullptr;        // per W block: (operator block, P block) pairs
  int32_t *d_wl_k = nullptr, *d_wl_p = nullptr;
  int64_t* d_gl_ptr = nullptr;        // per block of the next level: (P block, W block)
  int32_t *d_gl_p = nullptr, *d_gl_w = nullptr;  //   pairs of P^T W
};

struct Amg {
  bool built = false, lam_set = false;
  int64_t gen = 0;         // numerics generation: captured launches bake in its
                           // smoothing coefficients
  bool stale = true;       // the coarse levels belong to an earlier tangent
  bool step_stale = true;  // ... to an earlier load step
  int fine_degree = 1;     // Chebyshev degree of the finest level's smoothing (1 or 2)
  int coarse_degree = 2;   // and of the coarser levels'
  std::vector<AmgLevel> lev;
  int nc = 0;                         // dofs of the coarsest level (dense inverse)
  double* d_dense = nullptr;          // nc x nc FP64 work (the inverse in place)
  double* d_gjc = nullptr;            // nc x 32 FP64: a pivot block's column
  int ldc = 0;                        // row length of d_Ainv (nc rounded up to 4)
  void* d_Ainv = nullptr;             // nc x ldc KT, zero padded
  double* d_part = nullptr;           // power-iteration partial sums
  double* d_rzs = nullptr;            // 2 x nslices r.z partials of the PCG (degree-1 fine level)
};

struct PcgState {     // device-resident scalars of the PCG iteration
  double tol2;        // (rtol ||b||)^2
  double rr_last;     // ||r_k||^2 at the last convergence check
  double bb;          // ||b||^2
  double rzs[2];      // r.z as k_pcg_A summed it, by iteration parity (k & 1)
  int32_t k;          // iterations completed
  int32_t done;       // 1 once converged or k reached maxiter
  int32_t iters;      // iteration count when done
  int32_t breakdown;  // 1 if p'Ap <= 0 or non-finite
  int32_t breakdown_k;
  int32_t maxiter;
  int32_t converged;  // 1 if ||r_k|| <= rtol ||b||
  int32_t pad;
};

struct NewtonStats {  // device-resident FP64 norms of one Newton step
  double r2;          // ||f_ext[free] - f_int[free]||^2
  double fext2;       // ||f_ext||^2
  double dU2;         // ||dU||^2
  double U2;          // ||U||^2 before the update
  double prev_b2;     // ||b||^2 of the previous inner solve of this load step (0: none)
  double eta;         // inner tolerance used by the last solve
};

struct Scalars {      // one allocation so a single copy returns everything
  PcgState pcg;
  NewtonStats nwt;
};

struct Ctx {
  int precision = ONSAS_PRECISION_MIXED;
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int injected = 0;
  bool finalized = false;
  size_t kt = 4;  // bytes per tangent / PCG scalar

  // model (host)
  int64_t nn = 0;
  std::vector<double> X, Y, Z;
  std::vector<ElemGroup> groups;
  std::vector<std::vector<int32_t>> pressure_faces;  // per pressure group, 3*nf
  std::vector<int64_t> fixed;
  std::vector<uint8_t> fixmask;  // per node, bit d = dof 3i+d fixed
  int64_t nfree = 0;

  // pattern (host copy kept for export)
  std::vector<int32_t> row_ptr, col, diag;
  int64_t nnzb = 0;

  // sliced-ELL copy of the tangent for the PCG SpMV: slices of kSellRows block
  // rows, each slice padded to width_s slots (a multiple of kSellLanes).  Lane
  // l = kSellRows * ph + r of the wave owning slice s reads slot k = kSellLanes m + ph
  // of row kSellRows s + r: block index sell_ptr[s] + 64 m + l, its 9 values at
  // 9 (sell_ptr[s] + 64 m) + 64 v + l (one 256-byte plane per value).
  std::vector<int32_t> sell_ptr;
  int64_t nslices = 0, nsell = 0;
  // round-major variant (ELL): when the padding allows it, round m of every
  // slice is stored contiguously -- block index (m * nslices + s) * 64 + lane --
  // so the waves sweeping round m at the same time read one contiguous window.
  bool ell_rounds = false;
  int32_t ell_width = 0;              // rounds of the widest slice
  std::vector<int32_t> slice_rounds;  // rounds per slice
  int32_t* d_slice_rounds = nullptr;
  int32_t* d_sell_ptr = nullptr;
  int32_t* d_sell_col = nullptr;
  void* d_Ksell = nullptr;

  // assembly mode: gather (default) writes every sliced-ELL block once from the
  // per-element states through fixed-order contribution lists; scatter adds the
  // element blocks colour by colour into the block-CSR copy and packs it.
  bool gather = true;
  std::vector<SellGroup> sell_groups;
  SellGroup* d_sell_groups = nullptr;
  uint32_t* d_contrib = nullptr;  // (global element << 4) | (a << 2) | b, 0xffffffff = none
  int64_t ncontrib = 0;           // stored entries including padding
  // staged gather (one tet group): per workgroup, the states of the elements
  // of `stage_rows` consecutive rows are copied into LDS and summed from there
  // (host_setup.hpp::staged_gather); stage_rows = 0 selects the lists above
  int stage_rows = 0;
  int stage_cap = 0;        // LDS columns per workgroup (the largest element list)
  int stage_pairs_cap = 0;  // the most incidences of a workgroup's rows
  int stage_steps_cap = 0;  // the most steps of a workgroup
  int stage_ks = 0;         // list entries per thread in the staging
  int64_t sell_rstride = 0; // sliced-ELL blocks between the rounds of a slice
  int64_t nstage_groups = 0;
  int64_t stage_total = 0;   // staged elements over all workgroups
  int64_t stage_nsteps = 0;
  host::StageGroup* d_stage_groups = nullptr;
  host::StageStep* d_stage_steps = nullptr;
  int32_t* d_stage = nullptr;
  uint32_t* d_stage_words = nullptr;
  uint32_t* d_stage_pairs = nullptr;
  int32_t* d_pair_row = nullptr;
  uint8_t* d_colmask = nullptr;  // per block: the fixed-dof mask of its column
  bool bsr_stale = false;         // d_K does not hold the last assembly yet

  // device
  double *d_x = nullptr, *d_y = nullptr, *d_z = nullptr;
  uint8_t* d_fixmask = nullptr;
  int32_t *d_row_ptr = nullptr, *d_col = nullptr, *d_diag = nullptr;
  double* d_noalloc = nullptr;  // this device's `scratch` (capi.hip), not owned
  void* d_K = nullptr;     // nnzb*9 KT
  void* d_Dinv = nullptr;  // nn*9 KT
  double *d_U = nullptr, *d_Uasm = nullptr, *d_fint = nullptr, *d_fext = nullptr, *d_res = nullptr;
  double* d_nodal = nullptr;
  void *d_px = nullptr, *d_pr = nullptr, *d_pz = nullptr, *d_pp0 = nullptr, *d_pp1 = nullptr,
       *d_pq = nullptr;  // PCG vectors (KT)
  double* d_part = nullptr;  // partial-sum arrays, kPartArrays x kMaxPartials
  double* d_ppq = nullptr;   // p.q partials of k_pcg_A, one per workgroup of 4 slices
  Scalars* d_sc = nullptr;
  Scalars* h_sc = nullptr;   // pinned
  double* h_stage = nullptr;  // pinned, 3 nn values: staging of state vector copies
  double* h_snap[2] = {};     // pinned, 3 nn values each: U snapshots (onsas_snapshot_U)
  hipEvent_t ev_snap[2] = {};  // ... their copies' completion
  // A snapshot copies U on the stream into d_snap (device to device, a few
  // microseconds), then d_snap to h_snap on copy_stream, so the link transfer
  // overlaps the next load step instead of holding the stream.
  double* d_snap[2] = {};
  hipEvent_t ev_snap_dev[2] = {};       // the device copy is done
  bool snap_pending[2] = {false, false};  // a transfer from d_snap[slot] was queued
  hipStream_t copy_stream = nullptr;    // default priority
  double* h_press = nullptr;            // pinned staging of the pressure values
  hipEvent_t ev_press = nullptr;        // ... their last upload has read it
  bool press_pending = false;
  // pressure faces
  int32_t *d_face_conn = nullptr, *d_face_grp = nullptr;  // all faces, 3*nf and nf
  int32_t *d_nf_ptr = nullptr, *d_nf_idx = nullptr;       // node -> faces CSR
  double* d_pressures = nullptr;
  int64_t nfaces = 0;

  // reduction grid
  int np = 8;          // workgroups of reduction kernels (multiple of 8)
  int64_t chunk = 1;   // rows (nodes) per logical workgroup, a multiple of kChunkRows

  // linear solver settings
  onsas_linear_settings lin{};

  // PCG graphs: [0] 2 iterations (a solve's opening), [1] the repeated chunk,
  // [2] 4 iterations; each keyed by its length, preconditioner and multigrid
  // generation
  struct PcgGraph {
    hipGraphExec_t exec = nullptr;
    int iters = 0;
    int precond = -1;
    int64_t gen = -1;
  } pcg_graph[3];
  int64_t pcg_solves = 0;  // PCG solves so far (every kSampleEvery-th is timed)
  // Warm starts of the Newton solves: the PCG solution of Newton iteration j
  // (j < kWarmSlots) of the current load step goes to warm[cur][j]; iteration j
  // of the next step starts from it (scaled on the device).  set_loads flips
  // the banks, set_U empties both.
  static constexpr int kWarmSlots = 4;
  void* d_warm[2][kWarmSlots] = {};
  bool warm_ok[2][kWarmSlots] = {};
  int warm_cur = 0;
  int newton_it = 0;        // Newton iterations so far in this load step
  bool warm_start = true;   // ONSAS_WARM_START=0 turns it off (read at finalize)
  // PCG iterations of Newton iteration j (j < kPredSlots) of the current load
  // step (bank warm_cur) and of the previous one (-1: unknown).  A solve
  // launches the previous step's count before its first host synchronization;
  // banks flip and clear with the warm-start banks.  ONSAS_PCG_PREDICT=0 turns
  // it off (read at finalize).
  static constexpr int kPredSlots = 8;
  static_assert(kPredSlots == 8, "one -1 per slot below");
  int pcg_pred[2][kPredSlots] = {{-1, -1, -1, -1, -1, -1, -1, -1}, {-1, -1, -1, -1, -1, -1, -1, -1}};
  bool pcg_predict = true;

  // aggregation multigrid preconditioner (built on its first use)
  Amg amg;

  // GMRES workspace (allocated on first use)
  void* d_gm_V = nullptr;   // (m+1) * n KT
  void* d_gm_w = nullptr;   // n KT
  double* d_gm_h = nullptr; // partial dots
  int gm_m = 0;

  // timers
  hipEvent_t ev[10] = {};  // 0/3 assembly, 1/8/2 element+gather passes, 4..7 Newton step
  hipEvent_t ev_spmv[2 * 32] = {};  // brackets of the sampled SpMV launches
  onsas_timings tm{};
  double last_assemble_ms = 0;
  bool asm_timing_pending = false;  // an assembly's events not yet read (read after a sync)
  bool assembled = false;
};

enum PartArray {  // rows of d_part
  kPartRZ0 = 0, kPartRZ1, kPartRR, kPartPQ, kPartBB, kPartR2, kPartF2, kPartDU2, kPartU2,
  kPartArrays
};

inline double* part(Ctx* c, int a) { return c->d_part + (size_t)a * kMaxPartials; }

// ---- setup.cpp ----
int finalize_model(Ctx* c);

// ---- element_kernels.hip ----
int launch_assemble(Ctx* c);                                    // zero + scatter + Dinv
int launch_element_outputs(Ctx* c, int g, const double* U, double* f, double* K, double* stress,
                           double* strain);
int launch_fext(Ctx* c);                                        // nodal + pressure faces
int launch_adjoint_shear(Ctx* c, int g, const double* lambda_dev, double* dJdG_dev);
int stage_gather_prepare(Ctx* c);     // the staged gather's LDS attribute
size_t stage_gather_lds(const Ctx* c);  // its LDS bytes per workgroup
int launch_gather_init(Ctx* c);                                 // zero padding, identity Dinv
int ensure_bsr(Ctx* c);                                         // d_K from d_Ksell if stale
int launch_tangent_apply_f64(Ctx* c, const double* v_dev, double* y_dev);  // y = K(U) v, FP64
int launch_locate(Ctx* c, int g, const double* px, const double* py, const double* pz, int64_t np,
                  double tol, int64_t* elem, double* w);

// ---- amg.hip ----
int amg_build(Ctx* c);                                          // hierarchy (host) + arrays
int amg_numeric(Ctx* c);                                        // Galerkin, inverses, omega
// x = M^-1 b; with d0_ready the first fine-level smoother step is already in
// the buffer amg_fine_smoother names (fused into the caller's kernel)
// ev (optional): 4 events, start/stop pairs stamped by the dispatches of the
// cycle's two exact fine-level products (r = b - A x), in cycle order
// prz (optional, degree-1 fine smoothing only): one r.z partial per fine-level
// slice, b.x of the cycle's input and output, written by its last product
// chk_rr (optional, with d0_ready): the PCG's r.r partials; the cycle's first
// tangent product then runs the PCG's stop check in its prologue (and writes
// the stop into st), so a converged PCG skips the rest of the cycle
int launch_vcycle(Ctx* c, const void* b, void* x, const PcgState* st, bool d0_ready,
                  hipEvent_t* ev = nullptr, double* prz = nullptr,
                  const double* chk_rr = nullptr);
double amg_fine_smoother(Ctx* c, void** d0);  // that buffer and its scale 1/theta
void amg_free(Ctx* c);
int amg_nlevels(Ctx* c);  // hierarchy inspection (0 before it is built)
int amg_export_info(Ctx* c, int l, int64_t* n, int32_t* bs, int64_t* nnzb, int64_t* nnz_p,
                    double* lam);
int amg_export_operator(Ctx* c, int l, int32_t* row_ptr, int32_t* col, double* vals);
int amg_export_prolongator(Ctx* c, int l, int32_t* p_ptr, int32_t* p_col, double* vals);
int amg_time_level_products(Ctx* c, int level, int launches, double* ms);  // batch timing

// ---- solver_kernels.hip ----
int launch_residual(Ctx* c);                                    // d_res, partial R2/F2
// x0: optional warm start (n KT values, zero on fixed dofs), scaled on the
// device by the factor minimising ||b - a K x0||
int launch_pcg_solve(Ctx* c, const double* b_dev, bool forcing, int* iters, double* relres,
                     int* converged, int* breakdown, const void* x0 = nullptr, int expect = -1);
double linear_rtol(const Ctx* c);                               // configured / default rtol
int reset_forcing(Ctx* c);                                      // new load step
int launch_gmres_solve(Ctx* c, const double* b_dev, int* iters, double* relres, int* converged);
int launch_update(Ctx* c);                                      // U += x, partial DU2/U2
int launch_refine_residual(Ctx* c, const double* b_dev, const double* y_dev,
                           double* r_dev);                      // r = b - y, partial R2/F2
int launch_accumulate_x(Ctx* c, double* lam_dev);               // lam += PCG x (free dofs)
int launch_finalize_newton(Ctx* c);                             // partials -> d_sc->nwt
int launch_matvec(Ctx* c, const double* x_dev, double* y_dev);
int launch_pack_sell(Ctx* c);                                  // block-CSR -> sliced-ELL copy
int launch_x_to_double(Ctx* c, double* out_dev);               // PCG x (KT) -> FP64 (fixed = 0)
int launch_convert_K(Ctx* c, double* out_dev);

// ---- error helpers ----
int set_err(Ctx* c, int code, const std::string& msg);

}  // namespace onsas

#define ONSAS_HIP_TRY(c, expr)                                                            \
  do {                                                                                    \
    hipError_t e__ = (expr);                                                              \
    if (e__ != hipSuccess)                                                                \
      return ::onsas::set_err((c), ONSAS_ERR_HIP,                                         \
                              std::string(#expr " failed: ") + hipGetErrorString(e__));  \
  } while (0)

#define ONSAS_TRY(expr)        \
  do {                         \
    int s__ = (expr);          \
    if (s__ != ONSAS_OK) return s__; \
  } while (0)

struct onsas_ctx {
  onsas::Ctx impl;
};
