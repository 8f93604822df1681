// capi.hip -- the extern "C" boundary of libonsas_hip (declared in
// include/onsas_hip.h).  Argument checking, ownership and the orchestration of
// one Newton iteration; the numerics live in element_kernels.hip and
// solver_kernels.hip, the setup in setup.hip.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "engine.hpp"

// ---- device memory of onsas_tangent_matvec_noalloc -----------------------------
// x and y of that entry point, kNoAllocMaxDofs doubles each, in one file-scope
// array (32 MiB) that every GPU the library runs on holds until the process
// ends.  Host code never passes `scratch` itself to a kernel or a copy: the
// array's name is a host-side stand-in, not a device address.  onsas_create
// looks the address up once per device (hipGetSymbolAddress, with that device
// current) and every later use goes through that pointer.
static constexpr int64_t kNoAllocMaxDofs = int64_t(1) << 21;  // 2,097,152 dofs
static __device__ __align__(256) double scratch[2 * kNoAllocMaxDofs];

namespace onsas {

int set_err(Ctx* c, int code, const std::string& msg) {
  if (c) c->err = msg;
  return code;
}

}  // namespace onsas

using onsas::Ctx;
using onsas::ElemGroup;
using onsas::set_err;

namespace {

std::string g_create_err;

std::mutex g_scratch_mu;
std::vector<double*> g_scratch_ptr;  // per device id: the device address of `scratch`

// The device address of `scratch` on device `dev`, which must be current:
// looked up on the first call for that device, then reused.
hipError_t scratch_address(int dev, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if ((int)g_scratch_ptr.size() <= dev) g_scratch_ptr.resize(dev + 1, nullptr);
  if (!g_scratch_ptr[dev]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch_ptr[dev] = (double*)p;
  }
  *out = g_scratch_ptr[dev];
  return hipSuccess;
}

Ctx* C(onsas_ctx* ctx) { return ctx ? &ctx->impl : nullptr; }

int check_ready(Ctx* c) {
  if (!c) return ONSAS_ERR_ARG;
  if (c->injected) {
    int code = c->injected;
    c->injected = 0;
    return set_err(c, code, "injected error");
  }
  if (!c->finalized) return set_err(c, ONSAS_ERR_STATE, "onsas_finalize was not called");
  return ONSAS_OK;
}

// State vectors (3 nn doubles) cross PCIe through the pinned stage: one DMA at
// full link rate instead of the runtime's chunked pageable copy.
int download_state(Ctx* c, double* dst, const double* d_src) {
  const size_t bytes = 3 * (size_t)c->nn * sizeof(double);
  ONSAS_HIP_TRY(c, hipMemcpyAsync(c->h_stage, d_src, bytes, hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  std::memcpy(dst, c->h_stage, bytes);
  return ONSAS_OK;
}

int upload_state(Ctx* c, double* d_dst, const double* src) {
  const size_t bytes = 3 * (size_t)c->nn * sizeof(double);
  std::memcpy(c->h_stage, src, bytes);
  ONSAS_HIP_TRY(c, hipMemcpyAsync(d_dst, c->h_stage, bytes, hipMemcpyHostToDevice, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

int check_building(Ctx* c) {
  if (!c) return ONSAS_ERR_ARG;
  if (c->injected) {
    int code = c->injected;
    c->injected = 0;
    return set_err(c, code, "injected error");
  }
  if (c->finalized) return set_err(c, ONSAS_ERR_STATE, "model already finalized");
  return ONSAS_OK;
}

// scratch device buffer for host<->device transfers of n doubles
struct DevScratch {
  double* p = nullptr;
  ~DevScratch() {
    if (p) (void)hipFree(p);
  }
};

int to_device(Ctx* c, DevScratch& s, const double* h, size_t count) {
  ONSAS_HIP_TRY(c, hipMalloc((void**)&s.p, std::max<size_t>(count, 1) * sizeof(double)));
  if (h && count)
    ONSAS_HIP_TRY(c, hipMemcpyAsync(s.p, h, count * sizeof(double), hipMemcpyHostToDevice, c->stream));
  return ONSAS_OK;
}

float elapsed(hipEvent_t a, hipEvent_t b) {
  float ms = 0;
  if (hipEventElapsedTime(&ms, a, b) != hipSuccess) return 0.f;
  return ms;
}

// Solve K x = b_dev with the configured linear solver; x left in the PCG x buffer (KT).
int linear_solve(Ctx* c, const double* b_dev, bool forcing, int* iters, double* relres,
                 int* converged, int* solver, const void* x0 = nullptr, int expect = -1) {
  int it = 0, conv = 0, brk = 0;
  double rr = 1.0;
  *solver = 0;
  *iters = 0;
  const int kind = c->lin.kind;
  if (kind == ONSAS_LINSOLVE_PCG || kind == ONSAS_LINSOLVE_AUTO) {
    ONSAS_TRY(onsas::launch_pcg_solve(c, b_dev, forcing, &it, &rr, &conv, &brk, x0, expect));
    *solver = 1;
    *iters = it;
    c->tm.pcg_iters += it;
    if (brk && kind == ONSAS_LINSOLVE_PCG)
      return set_err(c, ONSAS_ERR_LINEAR_SOLVE,
                     "PCG breakdown (p'Kp <= 0): the tangent is not positive definite");
    if (!brk) {
      *relres = rr;
      *converged = conv;
      return ONSAS_OK;
    }
  }
  int git = 0;
  ONSAS_TRY(onsas::launch_gmres_solve(c, b_dev, &git, &rr, &conv));
  *solver = (*solver == 1) ? 3 : 2;
  *iters += git;
  c->tm.gmres_iters += git;
  *relres = rr;
  *converged = conv;
  return ONSAS_OK;
}

// Sums of the R2 / F2 partial rows written by launch_refine_residual (host,
// index order, so the result is deterministic).
int read_r2_b2(Ctx* c, double* r2, double* b2) {
  std::vector<double> h(2 * (size_t)c->np);
  ONSAS_HIP_TRY(c, hipMemcpyAsync(h.data(), onsas::part(c, onsas::kPartR2), c->np * sizeof(double),
                                  hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(h.data() + c->np, onsas::part(c, onsas::kPartF2),
                                  c->np * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  double a = 0, b = 0;
  for (int i = 0; i < c->np; ++i) {
    a += h[i];
    b += h[c->np + i];
  }
  *r2 = a;
  *b2 = b;
  return ONSAS_OK;
}

// Solve K(U) x = b on the free dofs to ||b - K x|| <= tol ||b|| by iterative
// refinement: the residual uses the FP64 element-wise tangent product (the
// blocks are re-evaluated in FP64, never stored), each correction is a solve
// with the assembled tangent (FP32 in mixed mode) by the configured solver.
// x, y, r: n doubles of device scratch.
int refined_solve(Ctx* c, const double* b, double* x, double* y, double* r, double tol,
                  int max_pass, int* iters, double* relres) {
  const size_t nb = 3 * (size_t)c->nn * sizeof(double);
  ONSAS_HIP_TRY(c, hipMemsetAsync(x, 0, nb, c->stream));
  ONSAS_HIP_TRY(c, hipMemsetAsync(y, 0, nb, c->stream));
  ONSAS_TRY(onsas::launch_refine_residual(c, b, y, r));
  double r2 = 0, b2 = 0;
  ONSAS_TRY(read_r2_b2(c, &r2, &b2));
  *iters = 0;
  *relres = 0;
  if (!(b2 > 0)) return ONSAS_OK;
  const double bn = std::sqrt(b2);
  double rel = std::sqrt(r2) / bn;
  for (int p = 0; p < max_pass && rel > tol; ++p) {
    int it = 0, conv = 0, solver = 0;
    double rr = 0;
    ONSAS_TRY(linear_solve(c, r, false, &it, &rr, &conv, &solver));
    *iters += it;
    ONSAS_TRY(onsas::launch_accumulate_x(c, x));
    ONSAS_TRY(onsas::launch_tangent_apply_f64(c, x, y));
    ONSAS_TRY(onsas::launch_refine_residual(c, b, y, r));
    ONSAS_TRY(read_r2_b2(c, &r2, &b2));
    const double next = std::sqrt(r2) / bn;
    if (!std::isfinite(next)) return set_err(c, ONSAS_ERR_NOT_FINITE, "non-finite refinement residual");
    if (next > 0.5 * rel) {  // the correction no longer contracts
      rel = next;
      break;
    }
    rel = next;
  }
  *relres = rel;
  if (rel > tol)
    return set_err(c, ONSAS_ERR_LINEAR_SOLVE,
                   "refined solve stalled at relative residual " + std::to_string(rel));
  return ONSAS_OK;
}

// f_int and K at the current U (the body of onsas_assemble).
int assemble_now(Ctx* c) {
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[0], c->stream));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(c->d_Uasm, c->d_U, 3 * c->nn * sizeof(double),
                                  hipMemcpyDeviceToDevice, c->stream));
  ONSAS_TRY(onsas::launch_assemble(c));
  if (!c->gather) ONSAS_TRY(onsas::launch_pack_sell(c));
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[3], c->stream));
  c->amg.stale = true;  // its Galerkin products belong to the previous tangent
  return ONSAS_OK;
}

// Read the last assembly's event timings (the stream has been synchronized).
void account_assembly(Ctx* c) {
  if (!c->asm_timing_pending) return;
  c->asm_timing_pending = false;
  c->last_assemble_ms = elapsed(c->ev[0], c->ev[3]);
  c->tm.assemble_ms += c->last_assemble_ms;
  c->tm.scatter_ms += elapsed(c->ev[1], c->ev[2]);
  if (c->gather) c->tm.elem_ms += elapsed(c->ev[1], c->ev[8]);
  c->tm.assembly_mode = c->gather ? (c->stage_rows > 0 ? 2 : 1) : 0;
  c->tm.n_assemble++;
}

}  // namespace

extern "C" {

// target relative residual of the adjoint solve (FP64 refinement)
static constexpr double kAdjointTol = 1e-11;

const char* onsas_version(void) { return "onsas_hip 0.1.0 (gfx950)"; }

int onsas_create(int32_t precision, onsas_ctx** out) {
  if (!out) return ONSAS_ERR_ARG;
  *out = nullptr;
  if (precision != ONSAS_PRECISION_MIXED && precision != ONSAS_PRECISION_FP64) {
    g_create_err = "precision must be ONSAS_PRECISION_MIXED or ONSAS_PRECISION_FP64";
    return ONSAS_ERR_ARG;
  }
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev <= 0) {
    g_create_err = std::string("no HIP device available: ") + hipGetErrorString(e);
    return ONSAS_ERR_HIP;
  }
  onsas_ctx* ctx = new (std::nothrow) onsas_ctx();
  if (!ctx) return ONSAS_ERR_ALLOC;
  Ctx* c = &ctx->impl;
  c->precision = precision;
  (void)hipGetDevice(&c->device);
  e = scratch_address(c->device, &c->d_noalloc);
  if (e != hipSuccess) {
    g_create_err = std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e);
    delete ctx;
    return ONSAS_ERR_HIP;
  }
  e = hipStreamCreate(&c->stream);  // default priority
  if (e == hipSuccess) e = hipStreamCreate(&c->copy_stream);  // default priority
  if (e != hipSuccess) {
    g_create_err = std::string("hipStreamCreate: ") + hipGetErrorString(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete ctx;
    return ONSAS_ERR_HIP;
  }
  c->lin.kind = ONSAS_LINSOLVE_AUTO;
  c->lin.maxiter = 0;
  c->lin.rtol = 0;
  c->lin.gmres_restart = 40;
  c->lin.check_every = 16;
  *out = ctx;
  return ONSAS_OK;
}

void onsas_destroy(onsas_ctx* ctx) {
  if (!ctx) return;
  Ctx* c = &ctx->impl;
  (void)hipStreamSynchronize(c->stream);
  if (c->copy_stream) (void)hipStreamSynchronize(c->copy_stream);
  for (auto& g : c->pcg_graph)
    if (g.exec) (void)hipGraphExecDestroy(g.exec);
  onsas::amg_free(c);
  for (auto& g : c->groups)
    for (void* p : {(void*)g.d_conn, (void*)g.d_geo, (void*)g.d_slot, (void*)g.d_G, (void*)g.d_perm,
                    g.d_state, (void*)g.d_fe, (void*)g.d_cidx, (void*)g.locate.d_ptr,
                    (void*)g.locate.d_elem})
      if (p) (void)hipFree(p);
  void* ptrs[] = {c->d_x, c->d_y, c->d_z, c->d_fixmask, c->d_row_ptr, c->d_col, c->d_diag, c->d_K,
                  c->d_Dinv, c->d_U, c->d_Uasm, c->d_fint, c->d_fext, c->d_res, c->d_nodal, c->d_px, c->d_pr,
                  c->d_pz, c->d_pp0, c->d_pp1, c->d_pq, c->d_part, c->d_sc, c->d_face_conn,
                  c->d_face_grp, c->d_nf_ptr, c->d_nf_idx, c->d_pressures, c->d_gm_V, c->d_gm_w,
                  c->d_gm_h, c->d_sell_ptr, c->d_sell_col, c->d_Ksell, c->d_slice_rounds,
                  c->d_contrib, c->d_sell_groups, (void*)c->d_stage_groups,
                  (void*)c->d_stage_steps, c->d_stage, c->d_stage_words, c->d_stage_pairs,
                  c->d_pair_row, c->d_colmask, c->d_snap[0], c->d_snap[1], c->d_ppq};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  if (c->h_sc) (void)hipHostFree(c->h_sc);
  if (c->h_press) (void)hipHostFree(c->h_press);
  if (c->ev_press) (void)hipEventDestroy(c->ev_press);
  for (auto& e : c->ev_snap_dev)
    if (e) (void)hipEventDestroy(e);
  if (c->h_stage) (void)hipHostFree(c->h_stage);
  for (double* h : c->h_snap)
    if (h) (void)hipHostFree(h);
  for (auto& bank : c->d_warm)
    for (void* p : bank)
      if (p) (void)hipFree(p);
  for (auto& e : c->ev_snap)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->ev)
    if (e) (void)hipEventDestroy(e);
  for (auto& e : c->ev_spmv)
    if (e) (void)hipEventDestroy(e);
  (void)hipStreamDestroy(c->stream);
  if (c->copy_stream) (void)hipStreamDestroy(c->copy_stream);
  delete ctx;
}

const char* onsas_last_error(const onsas_ctx* ctx) {
  return ctx ? ctx->impl.err.c_str() : g_create_err.c_str();
}

int onsas_set_nodes(onsas_ctx* ctx, int64_t nn, const double* x, const double* y, const double* z) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  if (nn <= 0 || !x || !y || !z) return set_err(c, ONSAS_ERR_ARG, "onsas_set_nodes: bad arguments");
  c->nn = nn;
  c->X.assign(x, x + nn);
  c->Y.assign(y, y + nn);
  c->Z.assign(z, z + nn);
  return ONSAS_OK;
}

int onsas_add_tet_group(onsas_ctx* ctx, int64_t ne, const int32_t* conn, int32_t material,
                        const double* params, const double* shear_per_elem, int32_t* group) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  if (c->nn <= 0) return set_err(c, ONSAS_ERR_STATE, "call onsas_set_nodes first");
  if (ne <= 0 || !conn || !params)
    return set_err(c, ONSAS_ERR_ARG, "onsas_add_tet_group: bad arguments");
  if (material != ONSAS_MAT_SVK && material != ONSAS_MAT_NEOHOOKEAN && material != ONSAS_MAT_LINEAR)
    return set_err(c, ONSAS_ERR_ARG, "onsas_add_tet_group: unknown material");
  if (shear_per_elem && material == ONSAS_MAT_LINEAR)
    return set_err(c, ONSAS_ERR_UNSUPPORTED, "per-element shear modulus needs SVK or NeoHookean");
  for (int64_t k = 0; k < 4 * ne; ++k)
    if (conn[k] < 0 || conn[k] >= c->nn)
      return set_err(c, ONSAS_ERR_ARG, "onsas_add_tet_group: node id out of range");
  ElemGroup g;
  g.type = onsas::kTet;
  g.nnode = 4;
  g.mat = material;
  g.p0 = params[0];
  g.p1 = params[1];
  g.ne = ne;
  g.conn.assign(conn, conn + 4 * ne);
  if (shear_per_elem) {
    g.per_elem_G = true;
    g.Gel.assign(shear_per_elem, shear_per_elem + ne);
  }
  c->groups.push_back(std::move(g));
  if (group) *group = (int32_t)c->groups.size() - 1;
  return ONSAS_OK;
}

int onsas_add_truss_group(onsas_ctx* ctx, int64_t ne, const int32_t* conn, int32_t strain_model,
                          double E, const double* area, int32_t* group) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  if (c->nn <= 0) return set_err(c, ONSAS_ERR_STATE, "call onsas_set_nodes first");
  if (ne <= 0 || !conn || !area) return set_err(c, ONSAS_ERR_ARG, "onsas_add_truss_group: bad arguments");
  if (strain_model != ONSAS_STRAIN_ROTATED && strain_model != ONSAS_STRAIN_GREEN)
    return set_err(c, ONSAS_ERR_ARG, "onsas_add_truss_group: unknown strain model");
  for (int64_t k = 0; k < 2 * ne; ++k)
    if (conn[k] < 0 || conn[k] >= c->nn)
      return set_err(c, ONSAS_ERR_ARG, "onsas_add_truss_group: node id out of range");
  ElemGroup g;
  g.type = onsas::kTruss;
  g.nnode = 2;
  g.mat = strain_model;
  g.p0 = E;
  g.ne = ne;
  g.conn.assign(conn, conn + 2 * ne);
  g.area.assign(area, area + ne);
  c->groups.push_back(std::move(g));
  if (group) *group = (int32_t)c->groups.size() - 1;
  return ONSAS_OK;
}

int onsas_add_pressure_faces(onsas_ctx* ctx, int64_t nf, const int32_t* conn, int32_t* group) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  if (c->nn <= 0) return set_err(c, ONSAS_ERR_STATE, "call onsas_set_nodes first");
  if (nf <= 0 || !conn) return set_err(c, ONSAS_ERR_ARG, "onsas_add_pressure_faces: bad arguments");
  for (int64_t k = 0; k < 3 * nf; ++k)
    if (conn[k] < 0 || conn[k] >= c->nn)
      return set_err(c, ONSAS_ERR_ARG, "onsas_add_pressure_faces: node id out of range");
  c->pressure_faces.emplace_back(conn, conn + 3 * nf);
  if (group) *group = (int32_t)c->pressure_faces.size() - 1;
  return ONSAS_OK;
}

int onsas_set_fixed_dofs(onsas_ctx* ctx, int64_t nfixed, const int64_t* dofs) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  if (c->nn <= 0) return set_err(c, ONSAS_ERR_STATE, "call onsas_set_nodes first");
  if (nfixed < 0 || (nfixed > 0 && !dofs)) return set_err(c, ONSAS_ERR_ARG, "bad fixed dofs");
  for (int64_t k = 0; k < nfixed; ++k)
    if (dofs[k] < 0 || dofs[k] >= 3 * c->nn) return set_err(c, ONSAS_ERR_ARG, "fixed dof out of range");
  c->fixed.assign(dofs, dofs + nfixed);
  return ONSAS_OK;
}

int onsas_finalize(onsas_ctx* ctx) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_building(c));
  ONSAS_HIP_TRY(c, hipSetDevice(c->device));
  return onsas::finalize_model(c);
}

int onsas_set_linear_solver(onsas_ctx* ctx, const onsas_linear_settings* s) {
  Ctx* c = C(ctx);
  if (!c || !s) return ONSAS_ERR_ARG;
  if (s->kind < ONSAS_LINSOLVE_AUTO || s->kind > ONSAS_LINSOLVE_GMRES)
    return set_err(c, ONSAS_ERR_ARG, "unknown linear solver kind");
  if (s->precond != ONSAS_PRECOND_BLOCK_JACOBI && s->precond != ONSAS_PRECOND_AMG)
    return set_err(c, ONSAS_ERR_ARG, "unknown preconditioner");
  if (s->amg_refresh != ONSAS_AMG_REFRESH_STEP && s->amg_refresh != ONSAS_AMG_REFRESH_ASSEMBLY &&
      s->amg_refresh != ONSAS_AMG_REFRESH_ONCE)
    return set_err(c, ONSAS_ERR_ARG, "unknown multigrid refresh policy");
  c->lin = *s;
  if (c->lin.gmres_restart <= 0) c->lin.gmres_restart = 40;
  if (c->lin.check_every <= 0) c->lin.check_every = 16;
  // the multigrid hierarchy's symbolic part (host work): here rather than in the
  // first solve, when the model is complete; otherwise at the first solve
  if (c->finalized && c->lin.precond == ONSAS_PRECOND_AMG && !c->amg.built)
    ONSAS_TRY(onsas::amg_build(c));
  return ONSAS_OK;
}

int onsas_set_loads(onsas_ctx* ctx, const double* nodal, int32_t npressure, const double* pressures) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (npressure != (int32_t)c->pressure_faces.size() || (npressure > 0 && !pressures))
    return set_err(c, ONSAS_ERR_ARG, "onsas_set_loads: one pressure value per pressure group expected");
  const size_t n = 3 * (size_t)c->nn;
  // a new load step: last step's updates become the warm starts
  c->warm_cur ^= 1;
  for (bool& ok : c->warm_ok[c->warm_cur]) ok = false;
  for (int& n : c->pcg_pred[c->warm_cur]) n = -1;
  c->newton_it = 0;
  if (nodal) {
    std::memcpy(c->h_stage, nodal, n * sizeof(double));
    ONSAS_HIP_TRY(c, hipMemcpyAsync(c->d_nodal, c->h_stage, n * sizeof(double), hipMemcpyHostToDevice, c->stream));
  }
  else
    ONSAS_HIP_TRY(c, hipMemsetAsync(c->d_nodal, 0, n * sizeof(double), c->stream));
  if (npressure > 0) {
    // pinned staging, so the upload is asynchronous; the last one has read it
    if (c->press_pending) ONSAS_HIP_TRY(c, hipEventSynchronize(c->ev_press));
    std::memcpy(c->h_press, pressures, npressure * sizeof(double));
    ONSAS_HIP_TRY(c, hipMemcpyAsync(c->d_pressures, c->h_press, npressure * sizeof(double),
                                    hipMemcpyHostToDevice, c->stream));
    ONSAS_HIP_TRY(c, hipEventRecord(c->ev_press, c->stream));
    c->press_pending = true;
  }
  ONSAS_TRY(onsas::launch_fext(c));
  ONSAS_TRY(onsas::reset_forcing(c));  // a new load step: no previous inner residual
  c->amg.step_stale = true;            // ... and, by default, new multigrid coarse levels
  // the nodal values' upload reads h_stage, which the state copies reuse;
  // without them the step's work stays queued
  if (nodal) ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

int onsas_assemble(onsas_ctx* ctx, double* fint_out) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (c->asm_timing_pending) {  // the previous assembly's events are about to be reused
    ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
    account_assembly(c);
  }
  ONSAS_TRY(assemble_now(c));
  c->asm_timing_pending = true;
  c->assembled = true;
  if (fint_out) {
    ONSAS_HIP_TRY(c, hipMemcpyAsync(fint_out, c->d_fint, 3 * c->nn * sizeof(double),
                                    hipMemcpyDeviceToHost, c->stream));
    ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
    account_assembly(c);
  }
  // otherwise the work stays queued: the next call that synchronizes (the Newton
  // step) reads the timings, and the host does not wait on the GPU in between
  return ONSAS_OK;
}

int onsas_newton_step(onsas_ctx* ctx, onsas_step_stats* out) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!out) return set_err(c, ONSAS_ERR_ARG, "onsas_newton_step: NULL stats");
  if (!c->assembled) return set_err(c, ONSAS_ERR_STATE, "onsas_assemble must precede onsas_newton_step");
  std::memset(out, 0, sizeof *out);
  hipStream_t s = c->stream;
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[4], s));
  ONSAS_TRY(onsas::launch_residual(c));
  int iters = 0, conv = 0, solver = 0;
  double relres = 0;
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[5], s));
  // warm start from the previous load step's update of this Newton iteration
  const int j = c->newton_it++;
  const int prev = c->warm_cur ^ 1;
  const bool slot = c->warm_start && j < Ctx::kWarmSlots;
  const void* x0 = slot && c->warm_ok[prev][j] ? c->d_warm[prev][j] : nullptr;
  const bool pslot = c->pcg_predict && j < Ctx::kPredSlots;
  ONSAS_TRY(linear_solve(c, c->d_res, true, &iters, &relres, &conv, &solver, x0,
                         pslot ? c->pcg_pred[prev][j] : -1));
  if (pslot) c->pcg_pred[c->warm_cur][j] = solver == 1 ? iters : -1;
  if (slot) {  // this step's update, for the next step (PCG solutions only)
    c->warm_ok[c->warm_cur][j] = solver == 1;
    if (solver == 1)
      ONSAS_HIP_TRY(c, hipMemcpyAsync(c->d_warm[c->warm_cur][j], c->d_px, c->kt * 3 * c->nn,
                                      hipMemcpyDeviceToDevice, s));
  }
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[6], s));
  ONSAS_TRY(onsas::launch_update(c));
  ONSAS_TRY(onsas::launch_finalize_newton(c));
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[7], s));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(&c->h_sc->nwt, &c->d_sc->nwt, sizeof(onsas::NewtonStats),
                                  hipMemcpyDeviceToHost, s));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(s));
  account_assembly(c);
  const onsas::NewtonStats& nw = c->h_sc->nwt;
  out->r_norm = std::sqrt(nw.r2);
  out->fext_norm = std::sqrt(nw.fext2);
  out->dU_norm = std::sqrt(nw.dU2);
  out->U_norm = std::sqrt(nw.U2);
  out->r_rel = out->r_norm / out->fext_norm;
  out->dU_rel = out->dU_norm / out->U_norm;
  out->lin_iters = iters;
  out->lin_solver = solver;
  out->lin_converged = conv;
  out->lin_rel_residual = relres;
  out->lin_rtol = (c->lin.forcing && solver == 1) ? nw.eta : onsas::linear_rtol(c);
  out->assemble_ms = c->last_assemble_ms;
  out->solve_ms = elapsed(c->ev[4], c->ev[7]);
  c->tm.solve_ms += out->solve_ms;
  c->tm.pcg_ms += elapsed(c->ev[5], c->ev[6]);
  c->tm.n_steps++;
  c->assembled = false;  // the tangent belongs to the previous U
  if (!std::isfinite(out->r_norm) || !std::isfinite(out->dU_norm))
    return set_err(c, ONSAS_ERR_NOT_FINITE, "non-finite residual or update norm");
  return ONSAS_OK;
}

int onsas_get_U(onsas_ctx* ctx, double* U) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!U) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  ONSAS_TRY(download_state(c, U, c->d_U));
  return ONSAS_OK;
}

int onsas_snapshot_U(onsas_ctx* ctx, int32_t slot) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (slot < 0 || slot > 1) return set_err(c, ONSAS_ERR_ARG, "snapshot slot must be 0 or 1");
  const size_t bytes = 3 * c->nn * sizeof(double);
  // the slot's last transfer has left d_snap[slot] before it is overwritten
  if (c->snap_pending[slot]) ONSAS_HIP_TRY(c, hipStreamWaitEvent(c->stream, c->ev_snap[slot], 0));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(c->d_snap[slot], c->d_U, bytes, hipMemcpyDeviceToDevice, c->stream));
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev_snap_dev[slot], c->stream));
  // the transfer to the host runs on the copy stream, beside the next step
  ONSAS_HIP_TRY(c, hipStreamWaitEvent(c->copy_stream, c->ev_snap_dev[slot], 0));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(c->h_snap[slot], c->d_snap[slot], bytes, hipMemcpyDeviceToHost,
                                  c->copy_stream));
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev_snap[slot], c->copy_stream));
  c->snap_pending[slot] = true;
  return ONSAS_OK;
}

int onsas_read_snapshot(onsas_ctx* ctx, int32_t slot, double* U) {
  Ctx* c = C(ctx);
  if (!c || !c->finalized || slot < 0 || slot > 1 || !U) return ONSAS_ERR_ARG;
  // no stream work and no shared state but the slot: safe beside the
  // context's other calls on another thread
  if (hipEventSynchronize(c->ev_snap[slot]) != hipSuccess) return ONSAS_ERR_HIP;
  std::memcpy(U, c->h_snap[slot], 3 * c->nn * sizeof(double));
  return ONSAS_OK;
}

int onsas_set_U(onsas_ctx* ctx, const double* U) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!U) return set_err(c, ONSAS_ERR_ARG, "NULL input");
  ONSAS_TRY(upload_state(c, c->d_U, U));
  c->assembled = false;
  for (auto& bank : c->warm_ok)  // another state: no warm starts from the old one
    for (bool& ok : bank) ok = false;
  for (auto& bank : c->pcg_pred)
    for (int& n : bank) n = -1;
  c->newton_it = 0;
  return ONSAS_OK;
}

int onsas_get_fext(onsas_ctx* ctx, double* f) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!f) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  ONSAS_TRY(download_state(c, f, c->d_fext));
  return ONSAS_OK;
}

int onsas_get_fint(onsas_ctx* ctx, double* f) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!f) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  ONSAS_TRY(download_state(c, f, c->d_fint));
  return ONSAS_OK;
}

int onsas_gather_layout(onsas_ctx* ctx, int64_t* out) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!out) return set_err(c, ONSAS_ERR_ARG, "null output");
  out[0] = c->stage_rows;
  out[1] = c->stage_rows ? c->nstage_groups : 0;
  out[2] = c->stage_rows ? c->stage_total : 0;
  out[3] = c->gather ? c->ncontrib : 0;
  out[4] = c->stage_rows ? (int64_t)onsas::stage_gather_lds(c) : 0;
  out[5] = c->stage_rows ? c->stage_nsteps : 0;
  return ONSAS_OK;
}

int onsas_get_sizes(onsas_ctx* ctx, int64_t* nn, int64_t* nnzb, int64_t* ncolors) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (nn) *nn = c->nn;
  if (nnzb) *nnzb = c->nnzb;
  if (ncolors) {
    int64_t k = 0;
    for (auto& g : c->groups) k += (int64_t)g.color_off.size() - 1;
    *ncolors = k;
  }
  return ONSAS_OK;
}

int onsas_get_pattern(onsas_ctx* ctx, int32_t* row_ptr, int32_t* col) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (row_ptr) std::memcpy(row_ptr, c->row_ptr.data(), (c->nn + 1) * sizeof(int32_t));
  if (col) std::memcpy(col, c->col.data(), c->nnzb * sizeof(int32_t));
  return ONSAS_OK;
}

int onsas_get_tangent(onsas_ctx* ctx, double* vals) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!vals) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  DevScratch tmp;
  ONSAS_TRY(to_device(c, tmp, nullptr, 9 * c->nnzb));
  ONSAS_TRY(onsas::launch_convert_K(c, tmp.p));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(vals, tmp.p, 9 * c->nnzb * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

int onsas_time_cycle_products(onsas_ctx* ctx, int32_t launches, double* ms_per_launch) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!ms_per_launch) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  return onsas::amg_time_level_products(c, 0, launches, ms_per_launch);
}

int onsas_time_level_products(onsas_ctx* ctx, int32_t level, int32_t launches,
                              double* ms_per_launch) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!ms_per_launch) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  return onsas::amg_time_level_products(c, level, launches, ms_per_launch);
}

int onsas_amg_levels(onsas_ctx* ctx, int32_t* nlevels) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!nlevels) return set_err(c, ONSAS_ERR_ARG, "NULL output");
  *nlevels = onsas::amg_nlevels(c);
  return ONSAS_OK;
}

int onsas_amg_level_info(onsas_ctx* ctx, int32_t level, int64_t* n, int32_t* bs, int64_t* nnzb,
                         int64_t* nnz_p, double* lam) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  return onsas::amg_export_info(c, level, n, bs, nnzb, nnz_p, lam);
}

int onsas_amg_get_operator(onsas_ctx* ctx, int32_t level, int32_t* row_ptr, int32_t* col,
                           double* vals) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  return onsas::amg_export_operator(c, level, row_ptr, col, vals);
}

int onsas_amg_get_prolongator(onsas_ctx* ctx, int32_t level, int32_t* p_ptr, int32_t* p_col,
                              double* vals) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  return onsas::amg_export_prolongator(c, level, p_ptr, p_col, vals);
}

int onsas_tangent_matvec(onsas_ctx* ctx, const double* x, double* y) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!x || !y) return set_err(c, ONSAS_ERR_ARG, "NULL vector");
  const size_t n = 3 * (size_t)c->nn;
  DevScratch dx, dy;
  ONSAS_TRY(to_device(c, dx, x, n));
  ONSAS_TRY(to_device(c, dy, nullptr, n));
  ONSAS_TRY(onsas::launch_matvec(c, dx.p, dy.p));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(y, dy.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

// The same product with no allocation: x and y go through the device's
// `scratch` array (via the pointer onsas_create looked up) and the context's
// pinned stage, both allocated before this call.  Calls that overlap on one GPU
// share that array and race.
int onsas_tangent_matvec_noalloc(onsas_ctx* ctx, const double* x, double* y) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!x || !y) return set_err(c, ONSAS_ERR_ARG, "NULL vector");
  const int64_t n = 3 * c->nn;
  if (n > kNoAllocMaxDofs)
    return set_err(c, ONSAS_ERR_ARG,
                   "onsas_tangent_matvec_noalloc: " + std::to_string(n) +
                       " dofs exceed its limit of 2097152 (699050 nodes); use onsas_tangent_matvec");
  if (!c->d_noalloc) return set_err(c, ONSAS_ERR_STATE, "no scratch address for this device");
  double* dx = c->d_noalloc;
  double* dy = c->d_noalloc + kNoAllocMaxDofs;
  ONSAS_TRY(upload_state(c, dx, x));  // every x value is overwritten: nothing to clear
  ONSAS_TRY(onsas::launch_matvec(c, dx, dy));  // writes every y value
  return download_state(c, y, dy);
}

namespace {
int element_outputs(Ctx* c, int32_t group, const double* U, double* f, double* K, double* stress,
                    double* strain) {
  ONSAS_TRY(check_ready(c));
  if (group < 0 || group >= (int32_t)c->groups.size()) return set_err(c, ONSAS_ERR_ARG, "bad group");
  const ElemGroup& g = c->groups[group];
  const int nd = 3 * g.nnode;
  const size_t ne = g.ne;
  DevScratch df, dK, ds, de;
  if (f) ONSAS_TRY(to_device(c, df, nullptr, ne * nd));
  if (K) ONSAS_TRY(to_device(c, dK, nullptr, ne * nd * nd));
  if (stress) ONSAS_TRY(to_device(c, ds, nullptr, ne * 9));
  if (strain) ONSAS_TRY(to_device(c, de, nullptr, ne * 9));
  ONSAS_TRY(onsas::launch_element_outputs(c, group, U, df.p, dK.p, ds.p, de.p));
  if (f) ONSAS_HIP_TRY(c, hipMemcpyAsync(f, df.p, ne * nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (K) ONSAS_HIP_TRY(c, hipMemcpyAsync(K, dK.p, ne * nd * nd * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (stress) ONSAS_HIP_TRY(c, hipMemcpyAsync(stress, ds.p, ne * 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  if (strain) ONSAS_HIP_TRY(c, hipMemcpyAsync(strain, de.p, ne * 9 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

}  // namespace

int onsas_element_forces(onsas_ctx* ctx, int32_t group, double* f, double* K, double* stress,
                         double* strain) {
  Ctx* c = C(ctx);
  return element_outputs(c, group, c ? c->d_U : nullptr, f, K, stress, strain);
}

int onsas_get_stress_strain(onsas_ctx* ctx, int32_t group, int32_t at_last_assembly,
                            double* stress, double* strain) {
  Ctx* c = C(ctx);
  if (c && at_last_assembly && c->tm.n_assemble == 0)
    return set_err(c, ONSAS_ERR_STATE, "no assembly has been done yet");
  return element_outputs(c, group, c ? (at_last_assembly ? c->d_Uasm : c->d_U) : nullptr, nullptr,
                         nullptr, stress, strain);
}

int onsas_linear_solve(onsas_ctx* ctx, const double* b, double* x, int32_t* iters, double* rel_residual) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (!b || !x) return set_err(c, ONSAS_ERR_ARG, "NULL vector");
  if (!c->assembled && c->tm.n_assemble == 0) return set_err(c, ONSAS_ERR_STATE, "assemble first");
  const size_t n = 3 * (size_t)c->nn;
  DevScratch db, dx;
  ONSAS_TRY(to_device(c, db, b, n));
  ONSAS_TRY(to_device(c, dx, nullptr, n));
  int it = 0, conv = 0, solver = 0;
  double rr = 0;
  ONSAS_TRY(linear_solve(c, db.p, false, &it, &rr, &conv, &solver));
  ONSAS_TRY(onsas::launch_x_to_double(c, dx.p));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(x, dx.p, n * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (iters) *iters = it;
  if (rel_residual) *rel_residual = rr;
  if (!conv) return set_err(c, ONSAS_ERR_LINEAR_SOLVE, "linear solver did not reach its tolerance");
  return ONSAS_OK;
}

int onsas_adjoint_shear(onsas_ctx* ctx, int32_t group, const double* dJdU, double* dJdG, int32_t* iters) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (group < 0 || group >= (int32_t)c->groups.size()) return set_err(c, ONSAS_ERR_ARG, "bad group");
  const ElemGroup& g = c->groups[group];
  if (g.type != onsas::kTet || (g.mat != ONSAS_MAT_NEOHOOKEAN && g.mat != ONSAS_MAT_SVK))
    return set_err(c, ONSAS_ERR_UNSUPPORTED, "adjoint shear sensitivity needs an SVK or NeoHookean tet group");
  if (!dJdU || !dJdG) return set_err(c, ONSAS_ERR_ARG, "NULL vector");
  const size_t n = 3 * (size_t)c->nn;
  // the tangent at the current U (after a converged Newton loop U has moved by
  // the last correction since the last assembly)
  ONSAS_TRY(onsas_assemble(ctx, nullptr));
  DevScratch db, dlam, dy, dr, dout;
  ONSAS_TRY(to_device(c, db, dJdU, n));
  ONSAS_TRY(to_device(c, dlam, nullptr, n));
  ONSAS_TRY(to_device(c, dy, nullptr, n));
  ONSAS_TRY(to_device(c, dr, nullptr, n));
  ONSAS_TRY(to_device(c, dout, nullptr, g.ne));
  int it = 0;
  double rr = 0;
  ONSAS_TRY(refined_solve(c, db.p, dlam.p, dy.p, dr.p, kAdjointTol, 8, &it, &rr));
  ONSAS_TRY(onsas::launch_adjoint_shear(c, group, dlam.p, dout.p));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(dJdG, dout.p, g.ne * sizeof(double), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  if (iters) *iters = it;
  return ONSAS_OK;
}

int onsas_set_shear_moduli(onsas_ctx* ctx, int32_t group, const double* G) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (group < 0 || group >= (int32_t)c->groups.size()) return set_err(c, ONSAS_ERR_ARG, "bad group");
  ElemGroup& g = c->groups[group];
  if (g.type != onsas::kTet || (g.mat != ONSAS_MAT_NEOHOOKEAN && g.mat != ONSAS_MAT_SVK))
    return set_err(c, ONSAS_ERR_UNSUPPORTED, "per-element shear modulus needs an SVK or NeoHookean tet group");
  if (!G) return set_err(c, ONSAS_ERR_ARG, "NULL vector");
  std::vector<double> Gs(g.ne);
  for (int64_t s = 0; s < g.ne; ++s) {
    const double v = G[g.perm[s]];
    if (!std::isfinite(v)) return set_err(c, ONSAS_ERR_ARG, "non-finite shear modulus");
    Gs[s] = v;
  }
  if (!g.d_G) ONSAS_HIP_TRY(c, hipMalloc((void**)&g.d_G, g.ne * sizeof(double)));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(g.d_G, Gs.data(), g.ne * sizeof(double), hipMemcpyHostToDevice,
                                  c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  g.per_elem_G = true;
  g.Gel.assign(G, G + g.ne);
  c->assembled = false;  // K and f_int belong to the old moduli
  return ONSAS_OK;
}

namespace {
// Uniform bin grid of a tet group (reference configuration): about two
// elements per bin; each bin lists, in user element order, the elements whose
// bounding box (widened by a relative 1e-9) overlaps it.
int build_locate_grid(Ctx* c, ElemGroup& g) {
  onsas::LocateGrid& lg = g.locate;
  const int64_t ne = g.ne;
  double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
  const std::vector<double>* P[3] = {&c->X, &c->Y, &c->Z};
  for (int64_t k = 0; k < 4 * ne; ++k)
    for (int d = 0; d < 3; ++d) {
      const double v = (*P[d])[g.conn[k]];
      lo[d] = std::min(lo[d], v);
      hi[d] = std::max(hi[d], v);
    }
  double ext[3], vol = 1;
  const double diag = std::sqrt((hi[0] - lo[0]) * (hi[0] - lo[0]) + (hi[1] - lo[1]) * (hi[1] - lo[1]) +
                                (hi[2] - lo[2]) * (hi[2] - lo[2]));
  const double pad = 1e-9 * diag;
  for (int d = 0; d < 3; ++d) {
    lo[d] -= pad;
    hi[d] += pad;
    ext[d] = hi[d] - lo[d];
    vol *= ext[d];
  }
  const double h = std::cbrt(vol / std::max<double>(1.0, ne / 2.0));
  int64_t nbin = 1;
  for (int d = 0; d < 3; ++d) {
    lg.nb[d] = (int)std::min<double>(512.0, std::max(1.0, std::ceil(ext[d] / h)));
    lg.lo[d] = lo[d];
    lg.inv_h[d] = lg.nb[d] / ext[d];
    nbin *= lg.nb[d];
  }
  std::vector<int32_t> pos(ne);
  for (int64_t k = 0; k < ne; ++k) pos[g.perm[k]] = (int32_t)k;
  auto bin_range = [&](int64_t u, int b0[3], int b1[3]) {
    for (int d = 0; d < 3; ++d) {
      double mn = 1e300, mx = -1e300;
      for (int a = 0; a < 4; ++a) {
        const double v = (*P[d])[g.conn[4 * u + a]];
        mn = std::min(mn, v);
        mx = std::max(mx, v);
      }
      b0[d] = std::max(0, (int)std::floor((mn - pad - lg.lo[d]) * lg.inv_h[d]));
      b1[d] = std::min(lg.nb[d] - 1, (int)std::floor((mx + pad - lg.lo[d]) * lg.inv_h[d]));
    }
  };
  std::vector<int64_t> ptr(nbin + 1, 0);
  int b0[3], b1[3];
  for (int64_t u = 0; u < ne; ++u) {
    bin_range(u, b0, b1);
    for (int z = b0[2]; z <= b1[2]; ++z)
      for (int y = b0[1]; y <= b1[1]; ++y)
        for (int x = b0[0]; x <= b1[0]; ++x) ptr[((int64_t)z * lg.nb[1] + y) * lg.nb[0] + x + 1]++;
  }
  for (int64_t b = 0; b < nbin; ++b) ptr[b + 1] += ptr[b];
  std::vector<int32_t> list(std::max<int64_t>(ptr[nbin], 1));
  std::vector<int64_t> fill(ptr.begin(), ptr.end() - 1);
  for (int64_t u = 0; u < ne; ++u) {  // user order: every bin's list is sorted by it
    bin_range(u, b0, b1);
    for (int z = b0[2]; z <= b1[2]; ++z)
      for (int y = b0[1]; y <= b1[1]; ++y)
        for (int x = b0[0]; x <= b1[0]; ++x) list[fill[((int64_t)z * lg.nb[1] + y) * lg.nb[0] + x]++] = pos[u];
  }
  ONSAS_HIP_TRY(c, hipMalloc((void**)&lg.d_ptr, ptr.size() * sizeof(int64_t)));
  ONSAS_HIP_TRY(c, hipMemcpy(lg.d_ptr, ptr.data(), ptr.size() * sizeof(int64_t), hipMemcpyHostToDevice));
  ONSAS_HIP_TRY(c, hipMalloc((void**)&lg.d_elem, list.size() * sizeof(int32_t)));
  ONSAS_HIP_TRY(c, hipMemcpy(lg.d_elem, list.data(), list.size() * sizeof(int32_t), hipMemcpyHostToDevice));
  lg.built = true;
  return ONSAS_OK;
}
}  // namespace

int onsas_locate_points(onsas_ctx* ctx, int32_t group, int64_t npts, const double* x,
                        const double* y, const double* z, int64_t* elem, double* weights) {
  Ctx* c = C(ctx);
  ONSAS_TRY(check_ready(c));
  if (group < 0 || group >= (int32_t)c->groups.size()) return set_err(c, ONSAS_ERR_ARG, "bad group");
  ElemGroup& g = c->groups[group];
  if (g.type != onsas::kTet) return set_err(c, ONSAS_ERR_UNSUPPORTED, "point location needs a tet group");
  if (npts < 0 || (npts > 0 && (!x || !y || !z || !elem || !weights)))
    return set_err(c, ONSAS_ERR_ARG, "onsas_locate_points: bad arguments");
  if (npts == 0) return ONSAS_OK;
  if (!g.locate.built) ONSAS_TRY(build_locate_grid(c, g));
  DevScratch dx, dy, dz, dw, de;
  ONSAS_TRY(to_device(c, dx, x, npts));
  ONSAS_TRY(to_device(c, dy, y, npts));
  ONSAS_TRY(to_device(c, dz, z, npts));
  ONSAS_TRY(to_device(c, dw, nullptr, 4 * (size_t)npts));
  ONSAS_TRY(to_device(c, de, nullptr, npts));  // int64 results in a double-sized buffer
  ONSAS_TRY(onsas::launch_locate(c, group, dx.p, dy.p, dz.p, npts, 1e-12, (int64_t*)de.p, dw.p));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(elem, de.p, npts * sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipMemcpyAsync(weights, dw.p, 4 * npts * sizeof(double), hipMemcpyDeviceToHost,
                                  c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}

int onsas_get_timings(onsas_ctx* ctx, onsas_timings* t) {
  Ctx* c = C(ctx);
  if (!c || !t) return ONSAS_ERR_ARG;
  if (c->asm_timing_pending) {
    ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
    account_assembly(c);
  }
  *t = c->tm;
  return ONSAS_OK;
}

int onsas_reset_timings(onsas_ctx* ctx) {
  Ctx* c = C(ctx);
  if (!c) return ONSAS_ERR_ARG;
  c->tm = onsas_timings{};
  return ONSAS_OK;
}

int onsas_inject_error(onsas_ctx* ctx, int32_t code) {
  Ctx* c = C(ctx);
  if (!c || code >= 0) return ONSAS_ERR_ARG;
  c->injected = code;
  return ONSAS_OK;
}

}  // extern "C"
