// element_kernels.hip -- tet4 / truss internal force + tangent, coloured
// block-CSR scatter, block-Jacobi inverse, triangle-face pressure loads,
// per-element outputs and the adjoint contraction.  gfx950, wave64.
//
// Tangent of a tet in closed form.  With n_a the reference gradient of shape
// function a and V the reference volume, the reference builds
//   K = B' dSdE B V + (n' S n V) (x) I3          (src/Entities/Tetrahedrons.jl:207-215)
// Contracting the Voigt algebra by hand gives, per 3x3 node block (a, b):
//   K_ab = V [ c1 v_a v_b' + c2 v_b v_a' + c2 s_ab M + sigma_ab I ]
//   SVK  (SVKMaterial.jl:89-100):          v = F n, s_ab = n_a.n_b, M = F F', c1 = lambda, c2 = G
//   NeoHookean (NeoHookeanMaterial.jl:94-129): v = F^-T n, s_ab = v_a.v_b, M = I,
//        c1 = K J (2J - 1), c2 = G - K J (J - 1), J = sqrt(det C)
//   linear (Tetrahedrons.jl:240-266):      v = n, s_ab = n_a.n_b, M = I, c1 = lambda, c2 = G,
//        no geometric term
//   sigma_ab = n_a . S n_b   (geometric stiffness, hyperelastic only)
// and f_a = V F S n_a  (= B' voigt(S) V restricted to node a).  f and S are
// always FP64; the block is evaluated in the tangent precision KT.
//
// Truss (src/Entities/Trusses.jl:126-184), with d = X2 - X1 + u2 - u1, l = |d|,
// s_0 = -1, s_1 = +1:
//   K_ab = s_a s_b [ cI I + cdd d d' ],  f_a = s_a cf d
//   rotated: eps = (l^2 - l0^2) / (l0 (l0 + l)), S11 = E eps,
//            cI = S11 A / l, cdd = (E A / l0 - S11 A / l) / l^2, cf = A S11 / l
//   Green:   eps = (l^2 - l0^2) / (2 l0^2),  S11 = E eps,
//            cI = S11 A / l0, cdd = E A / l0^3, cf = A S11 / l0
#include "engine.hpp"

namespace onsas {

namespace {

constexpr int kScatterThreads = 256;

struct TetEval {
  double F[9];     // deformation gradient (I for linear)
  double S[9];     // second Piola-Kirchhoff (sigma for linear)
  double v[4][3];  // per-material gradient vectors (see header comment)
  double M[9];
  double c1, c2;
  bool geo;        // geometric stiffness present
  bool sv;         // s_ab from v (NeoHookean) instead of n
};

// p0, p1: SVK (lambda, G), NeoHookean (K, G), linear (E, nu)
__device__ inline void tet_eval(int mat, double p0, double p1, const double n[4][3],
                                const double u[4][3], TetEval& t) {
  double H[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      H[i * 3 + j] = u[0][i] * n[0][j] + u[1][i] * n[1][j] + u[2][i] * n[2][j] + u[3][i] * n[3][j];
#pragma unroll
  for (int q = 0; q < 9; ++q) {
    t.F[q] = H[q] + ((q % 4) == 0 ? 1.0 : 0.0);
    t.M[q] = (q % 4) == 0 ? 1.0 : 0.0;
  }
  if (mat == ONSAS_MAT_LINEAR) {
    // stress! (IsotropicLinearElasticMaterial.jl:80-92), lame_parameters (:70-76)
    const double lam = p0 * p1 / ((1 + p1) * (1 - 2 * p1));
    const double G = p0 / (2 * (1 + p1));
    double tr = H[0] + H[4] + H[8];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        t.S[i * 3 + j] = G * (H[i * 3 + j] + H[j * 3 + i]) + (i == j ? lam * tr : 0.0);
#pragma unroll
    for (int q = 0; q < 9; ++q) t.F[q] = (q % 4) == 0 ? 1.0 : 0.0;
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) t.v[b][d] = n[b][d];
    t.c1 = lam;
    t.c2 = G;
    t.geo = false;
    t.sv = false;
    return;
  }
  const double* F = t.F;
  if (mat == ONSAS_MAT_SVK) {
    // E = 1/2 (H + H' + H'H) (Tetrahedrons.jl:201); S = lambda tr(E) I + 2G E
    const double lam = p0, G = p1;
    double E[9];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        E[i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i] + H[i] * H[j] + H[3 + i] * H[3 + j] +
                              H[6 + i] * H[6 + j]);
    double tr = E[0] + E[4] + E[8];
#pragma unroll
    for (int q = 0; q < 9; ++q) t.S[q] = 2 * G * E[q] + ((q % 4) == 0 ? lam * tr : 0.0);
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int i = 0; i < 3; ++i)
        t.v[b][i] = F[i * 3 + 0] * n[b][0] + F[i * 3 + 1] * n[b][1] + F[i * 3 + 2] * n[b][2];
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        t.M[i * 3 + j] = F[i * 3] * F[j * 3] + F[i * 3 + 1] * F[j * 3 + 1] + F[i * 3 + 2] * F[j * 3 + 2];
    t.c1 = lam;
    t.c2 = G;
    t.geo = true;
    t.sv = false;
    return;
  }
  // NeoHookean: S = G (I - C^-1) + K J (J - 1) C^-1, C^-1 = F^-1 F^-T, J = sqrt(det C) = |det F|
  const double K = p0, G = p1;
  double cof[9];  // cofactor matrix of F: F^-T = cof / det F
  cof[0] = F[4] * F[8] - F[5] * F[7];
  cof[1] = F[5] * F[6] - F[3] * F[8];
  cof[2] = F[3] * F[7] - F[4] * F[6];
  cof[3] = F[2] * F[7] - F[1] * F[8];
  cof[4] = F[0] * F[8] - F[2] * F[6];
  cof[5] = F[1] * F[6] - F[0] * F[7];
  cof[6] = F[1] * F[5] - F[2] * F[4];
  cof[7] = F[2] * F[3] - F[0] * F[5];
  cof[8] = F[0] * F[4] - F[1] * F[3];
  const double detF = F[0] * cof[0] + F[1] * cof[1] + F[2] * cof[2];
  const double J = fabs(detF);
  double FiT[9];  // F^-T
#pragma unroll
  for (int q = 0; q < 9; ++q) FiT[q] = cof[q] / detF;
  double Ci[9];  // C^-1 = F^-1 F^-T = (F^-T)' (F^-T)
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      Ci[i * 3 + j] = FiT[i] * FiT[j] + FiT[3 + i] * FiT[3 + j] + FiT[6 + i] * FiT[6 + j];
  const double kjj = K * J * (J - 1);
#pragma unroll
  for (int q = 0; q < 9; ++q) t.S[q] = G * (((q % 4) == 0 ? 1.0 : 0.0) - Ci[q]) + kjj * Ci[q];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int i = 0; i < 3; ++i)
      t.v[b][i] = FiT[i * 3] * n[b][0] + FiT[i * 3 + 1] * n[b][1] + FiT[i * 3 + 2] * n[b][2];
  t.c1 = K * J * (2 * J - 1);
  t.c2 = G - kjj;
  t.geo = true;
  t.sv = true;
}

// f_a = V F S n_a
__device__ inline void tet_force_row(const TetEval& t, const double n[4][3], double V, int a,
                                     double fa[3]) {
  double Sn[3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    Sn[i] = t.S[i * 3] * n[a][0] + t.S[i * 3 + 1] * n[a][1] + t.S[i * 3 + 2] * n[a][2];
#pragma unroll
  for (int i = 0; i < 3; ++i)
    fa[i] = V * (t.F[i * 3] * Sn[0] + t.F[i * 3 + 1] * Sn[1] + t.F[i * 3 + 2] * Sn[2]);
}

// Element data converted once to the tangent precision.
template <typename T>
struct TetBlocks {
  T n[4][3], v[4][3], M[9], S[9], c1, c2, V;
  bool geo, sv;
  __device__ inline void load(const TetEval& t, const double nn[4][3], double Vol) {
#pragma unroll
    for (int b = 0; b < 4; ++b)
#pragma unroll
      for (int d = 0; d < 3; ++d) {
        n[b][d] = (T)nn[b][d];
        v[b][d] = (T)t.v[b][d];
      }
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      M[q] = (T)t.M[q];
      S[q] = (T)t.S[q];
    }
    c1 = (T)t.c1;
    c2 = (T)t.c2;
    V = (T)Vol;
    geo = t.geo;
    sv = t.sv;
  }
  __device__ inline void block(int a, int b, T out[9]) const {
    T s = sv ? (v[a][0] * v[b][0] + v[a][1] * v[b][1] + v[a][2] * v[b][2])
             : (n[a][0] * n[b][0] + n[a][1] * n[b][1] + n[a][2] * n[b][2]);
    T sig = 0;
    if (geo) {
      T Sn0 = S[0] * n[b][0] + S[1] * n[b][1] + S[2] * n[b][2];
      T Sn1 = S[3] * n[b][0] + S[4] * n[b][1] + S[5] * n[b][2];
      T Sn2 = S[6] * n[b][0] + S[7] * n[b][1] + S[8] * n[b][2];
      sig = n[a][0] * Sn0 + n[a][1] * Sn1 + n[a][2] * Sn2;
    }
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j)
        out[i * 3 + j] = V * (c1 * v[a][i] * v[b][j] + c2 * v[b][i] * v[a][j] + c2 * s * M[i * 3 + j] +
                              (i == j ? sig : T(0)));
  }
};

struct TrussEval {
  double d[3], l, S11, eps, cI, cdd, cf, sigma;
};

__device__ inline void truss_eval(int strain, double E, double A, const double d0[3], double l0,
                                  const double du[3], TrussEval& t) {
#pragma unroll
  for (int i = 0; i < 3; ++i) t.d[i] = d0[i] + du[i];
  t.l = sqrt(t.d[0] * t.d[0] + t.d[1] * t.d[1] + t.d[2] * t.d[2]);
  // l^2 - l0^2 = 2 d0.du + du.du (no cancellation)
  const double dl2 = 2 * (d0[0] * du[0] + d0[1] * du[1] + d0[2] * du[2]) +
                     (du[0] * du[0] + du[1] * du[1] + du[2] * du[2]);
  if (strain == ONSAS_STRAIN_ROTATED) {
    t.eps = dl2 / (l0 * (l0 + t.l));
    t.S11 = E * t.eps;
    t.cI = t.S11 * A / t.l;
    t.cdd = (E * A / l0 - t.S11 * A / t.l) / (t.l * t.l);
    t.cf = A * t.S11 / t.l;
  } else {
    t.eps = dl2 / (2 * l0 * l0);
    t.S11 = E * t.eps;
    t.cI = t.S11 * A / l0;
    t.cdd = E * A / (l0 * l0 * l0);
    t.cf = A * t.S11 / l0;
  }
  t.sigma = t.S11 * t.l / l0;
}

struct TetArgs {
  const int32_t* conn;
  const double* geo;
  const int32_t* slot;
  const double* G;
  const int32_t* idx;  // colour-sorted position -> array position (null: identity)
  int64_t ne, e0, e1;
  int mat;
  double p0, p1;
};

__device__ inline void tet_load(const TetArgs& ta, int64_t e, const double* __restrict__ U,
                                int32_t nd[4], double n[4][3], double& V, double u[4][3],
                                double& p1) {
  const int64_t ne = ta.ne;
#pragma unroll
  for (int b = 0; b < 4; ++b) nd[b] = ta.conn[b * ne + e];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int d = 0; d < 3; ++d) n[b][d] = ta.geo[(b * 3 + d) * ne + e];
  V = ta.geo[12 * ne + e];
#pragma unroll
  for (int b = 0; b < 4; ++b)
#pragma unroll
    for (int d = 0; d < 3; ++d) u[b][d] = U[3 * (int64_t)nd[b] + d];
  p1 = ta.G ? ta.G[e] : ta.p1;
}

// One lane per (element, row node a): f_a and the four blocks (a, b) of
// row a, read-modify-written without atomics (the colour makes every node of
// the launch exclusive to one element).  Entries in a fixed row or column are
// not added, so those rows/columns keep the identity written by k_fixed_identity.
template <typename KT>
__global__ __launch_bounds__(kScatterThreads) void k_tet_scatter(TetArgs ta,
                                                                  const double* __restrict__ U,
                                                                  double* __restrict__ fint,
                                                                  KT* __restrict__ K,
                                                                  const uint8_t* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * kScatterThreads + threadIdx.x;
  const int64_t k = ta.e0 + (t >> 2);
  const int a = (int)(t & 3);
  if (k >= ta.e1) return;
  const int64_t e = ta.idx ? (int64_t)ta.idx[k] : k;
  int32_t nd[4];
  double n[4][3], u[4][3], V, p1;
  tet_load(ta, e, U, nd, n, V, u, p1);
  TetEval ev;
  tet_eval(ta.mat, ta.p0, p1, n, u, ev);
  double fa[3];
  tet_force_row(ev, n, V, a, fa);
  double* fp = fint + 3 * (int64_t)nd[a];
  fp[0] += fa[0];
  fp[1] += fa[1];
  fp[2] += fa[2];
  TetBlocks<KT> tb;
  tb.load(ev, n, V);
  const uint32_t ma = mask[nd[a]];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint32_t mb = mask[nd[b]];
    const int64_t s = ta.slot[(a * 4 + b) * ta.ne + e];
    KT blk[9];
    tb.block(a, b, blk);
    KT* kp = K + 9 * s;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const bool off = ((ma >> i) & 1u) | ((mb >> j) & 1u);
        kp[i * 3 + j] += off ? KT(0) : blk[i * 3 + j];
      }
  }
}

struct TrussArgs {
  const int32_t* conn;
  const double* geo;  // d0 (3), l0, A
  const int32_t* slot;
  const int32_t* idx;  // colour-sorted position -> array position (null: identity)
  int64_t ne, e0, e1;
  int strain;
  double E;
};

template <typename KT>
__global__ __launch_bounds__(kScatterThreads) void k_truss_scatter(TrussArgs ta,
                                                                    const double* __restrict__ U,
                                                                    double* __restrict__ fint,
                                                                    KT* __restrict__ K,
                                                                    const uint8_t* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * kScatterThreads + threadIdx.x;
  const int64_t k = ta.e0 + (t >> 1);
  const int a = (int)(t & 1);
  if (k >= ta.e1) return;
  const int64_t e = ta.idx ? (int64_t)ta.idx[k] : k;
  const int64_t ne = ta.ne;
  const int32_t nd0 = ta.conn[e], nd1 = ta.conn[ne + e];
  double d0[3] = {ta.geo[e], ta.geo[ne + e], ta.geo[2 * ne + e]};
  const double l0 = ta.geo[3 * ne + e], A = ta.geo[4 * ne + e];
  double du[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) du[d] = U[3 * (int64_t)nd1 + d] - U[3 * (int64_t)nd0 + d];
  TrussEval tv;
  truss_eval(ta.strain, ta.E, A, d0, l0, du, tv);
  const double sa = a ? 1.0 : -1.0;
  const int32_t na = a ? nd1 : nd0;
  double* fp = fint + 3 * (int64_t)na;
#pragma unroll
  for (int d = 0; d < 3; ++d) fp[d] += sa * tv.cf * tv.d[d];
  const uint32_t ma = mask[na];
  const KT cI = (KT)tv.cI, cdd = (KT)tv.cdd;
  const KT dd[3] = {(KT)tv.d[0], (KT)tv.d[1], (KT)tv.d[2]};
#pragma unroll
  for (int b = 0; b < 2; ++b) {
    const int32_t nb = b ? nd1 : nd0;
    const uint32_t mb = mask[nb];
    const KT sab = (KT)(sa * (b ? 1.0 : -1.0));
    const int64_t s = ta.slot[(a * 2 + b) * ne + e];
    KT* kp = K + 9 * s;
#pragma unroll
    for (int i = 0; i < 3; ++i)
#pragma unroll
      for (int j = 0; j < 3; ++j) {
        const bool off = ((ma >> i) & 1u) | ((mb >> j) & 1u);
        const KT v = sab * ((i == j ? cI : KT(0)) + cdd * dd[i] * dd[j]);
        kp[i * 3 + j] += off ? KT(0) : v;
      }
  }
}

// ---- gather assembly ---------------------------------------------------------
// Pass 1 (one lane per element, coalesced SoA): the FP64 nodal forces f_a and
// the factors of the tangent blocks
//   K_ab = Vc1 v_a v_b' + Vc2 v_b v_a' + g_ab I (+ h_ab M for SVK),
//   g_ab = V (c2 s_ab + sigma_ab) (NeoHookean, linear; M = I there),
//   g_ab = V sigma_ab, h_ab = V c2 s_ab (SVK),
// rounded once to KT.  Pass 2 (one lane per stored tangent block) sums the
// blocks listed for it and writes it once.
__host__ __device__ constexpr int pair_index(int a, int b) {  // a <= b over 4 nodes -> 0..9
  return a * 4 - a * (a - 1) / 2 + (b - a);
}

// pair_index(min(a, b), max(a, b)) as nibble 4 a + b (a lookup with no branches)
constexpr uint64_t kPairTable = 0x9863875265413210ull;
// the off-diagonal pairs 01 02 03 12 13 23 -> 0..5, nibble 4 a + b (a != b)
constexpr uint64_t kOffPairTable = 0x0542503143002100ull;
constexpr bool pair_tables_ok() {
  for (int a = 0; a < 4; ++a)
    for (int b = 0; b < 4; ++b) {
      const int lo = a < b ? a : b, hi = a < b ? b : a;
      if ((int)((kPairTable >> (16 * a + 4 * b)) & 15u) != pair_index(lo, hi)) return false;
      const int j = lo == 0 ? hi - 1 : (lo == 1 ? hi + 1 : 5);
      if (a != b && (int)((kOffPairTable >> (16 * a + 4 * b)) & 15u) != j) return false;
    }
  return true;
}
static_assert(pair_tables_ok(), "pair tables");

// 4 and 2 consecutive values (16- and 8-byte aligned in FP32, 32 and 16 in FP64)
__device__ inline void st4(float* p, float a, float b, float c, float d) {
  *(float4*)p = make_float4(a, b, c, d);
}
__device__ inline void st4(double* p, double a, double b, double c, double d) {
  ((double2*)p)[0] = make_double2(a, b);
  ((double2*)p)[1] = make_double2(c, d);
}
__device__ inline void st2(float* p, float a, float b) { *(float2*)p = make_float2(a, b); }
__device__ inline void st2(double* p, double a, double b) { *(double2*)p = make_double2(a, b); }

// Where the tet state lives (stride ne, element e), by logical plane p as
// numbered below: node a's record (v_a0, v_a1, v_a2, g_aa) at 4 (a ne + e),
// (V c1, V c2) at 16 ne + 2 e, the off-diagonal g_ab at (18 + j) ne + e with
// j = 0..5 for ab = 01 02 03 12 13 23, and SVK's h_ab and M at p ne + e
// (p >= 24).  The lanes of a wave read consecutive elements, so one 16-byte
// load (FP32) brings a node's v_a with its g_aa and one 8-byte load c1 and c2.
__host__ __device__ constexpr int64_t tet_plane(int p, int64_t ne, int64_t e) {
  if (p < 2) return 16 * ne + 2 * e + p;
  if (p < 14) return 4 * (((p - 2) / 3) * ne + e) + (p - 2) % 3;
  if (p < 24) {
    const int q = p - 14;  // pair index
    const int a = q < 4 ? 0 : (q < 7 ? 1 : (q < 9 ? 2 : 3));
    const int b = q - pair_index(a, a) + a;
    if (a == b) return 4 * (a * ne + e) + 3;
    return (18 + (int)((kOffPairTable >> (16 * a + 4 * b)) & 15u)) * ne + e;
  }
  return p * ne + e;
}
// every logical plane of the base record lands on its own slot (a bijection
// onto [0, 24 ne) for ne = 5)
constexpr bool tet_plane_ok() {
  bool hit[120] = {};
  for (int p = 0; p < 24; ++p)
    for (int e = 0; e < 5; ++e) {
      const int64_t o = tet_plane(p, 5, e);
      if (o < 0 || o >= 120 || hit[o]) return false;
      hit[o] = true;
    }
  return tet_plane(3, 5, 2) == 4 * (0 * 5 + 2) + 1 && tet_plane(14 + 4, 5, 1) == 4 * (5 + 1) + 3 &&
         tet_plane(14 + 5, 5, 1) == (18 + 3) * 5 + 1 && tet_plane(1, 5, 4) == 16 * 5 + 9;
}
static_assert(tet_plane_ok(), "tet state layout");

// Shape-function gradients and volume of a tet from its reference coordinates
// X (node-major): the operations of the set-up's tet_gradients (setup.hip),
// which fills the stored geometry the other element kernels read, in the same
// order and without contraction: J = X^T D, V = det J / 6, dN = D^T J^-T.
__device__ inline void tet_gradients_dev(const double X[12], double n[4][3], double& V) {
#pragma clang fp contract(off)
  constexpr double D[3][4] = {{1, -1, 0, 0}, {0, -1, 0, 1}, {0, -1, 1, 0}};
  double J[9];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 3; ++k) {
      double s = 0;
#pragma unroll
      for (int b = 0; b < 4; ++b) s += X[b * 3 + i] * D[k][b];
      J[i * 3 + k] = s;
    }
  const double det = J[0] * (J[4] * J[8] - J[5] * J[7]) - J[1] * (J[3] * J[8] - J[5] * J[6]) +
                     J[2] * (J[3] * J[7] - J[4] * J[6]);
  V = det / 6.0;
  const double Ji[9] = {(J[4] * J[8] - J[5] * J[7]) / det, (J[2] * J[7] - J[1] * J[8]) / det,
                        (J[1] * J[5] - J[2] * J[4]) / det, (J[5] * J[6] - J[3] * J[8]) / det,
                        (J[0] * J[8] - J[2] * J[6]) / det, (J[2] * J[3] - J[0] * J[5]) / det,
                        (J[3] * J[7] - J[4] * J[6]) / det, (J[1] * J[6] - J[0] * J[7]) / det,
                        (J[0] * J[4] - J[1] * J[3]) / det};
#pragma unroll
  for (int j = 0; j < 3; ++j)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double s = 0;
#pragma unroll
      for (int k = 0; k < 3; ++k) s += Ji[k * 3 + j] * D[k][b];
      n[b][j] = s;
    }
}

// Element pass of the gather assembly: the tet's gradients are recomputed
// from the nodal reference coordinates (read with U at the same nodes, mostly
// from L2) instead of streaming the stored 104-byte geometry record.
template <typename KT>
__global__ __launch_bounds__(256) void k_tet_state(TetArgs ta, const double* __restrict__ U,
                                                   const double* __restrict__ Xr,
                                                   const double* __restrict__ Yr,
                                                   const double* __restrict__ Zr,
                                                   KT* __restrict__ st, double* __restrict__ fe,
                                                   int hasM) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ta.ne) return;
  const int64_t ne = ta.ne;
  int32_t nd[4];
  double n[4][3], u[4][3], V, X[12];
#pragma unroll
  for (int b = 0; b < 4; ++b) nd[b] = ta.conn[b * ne + e];
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    X[b * 3] = Xr[nd[b]];
    X[b * 3 + 1] = Yr[nd[b]];
    X[b * 3 + 2] = Zr[nd[b]];
#pragma unroll
    for (int d = 0; d < 3; ++d) u[b][d] = U[3 * (int64_t)nd[b] + d];
  }
  const double p1 = ta.G ? ta.G[e] : ta.p1;
  tet_gradients_dev(X, n, V);
  TetEval ev;
  tet_eval(ta.mat, ta.p0, p1, n, u, ev);
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double fa[3];
    tet_force_row(ev, n, V, a, fa);
#pragma unroll
    for (int d = 0; d < 3; ++d) fe[(a * 3 + d) * ne + e] = fa[d];
  }
  KT g[10];
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = a; b < 4; ++b) {
      const int p = pair_index(a, b);
      const double s = ev.sv ? ev.v[a][0] * ev.v[b][0] + ev.v[a][1] * ev.v[b][1] + ev.v[a][2] * ev.v[b][2]
                             : n[a][0] * n[b][0] + n[a][1] * n[b][1] + n[a][2] * n[b][2];
      double sig = 0;
      if (ev.geo) {
        double Sn[3];
#pragma unroll
        for (int i = 0; i < 3; ++i)
          Sn[i] = ev.S[i * 3] * n[b][0] + ev.S[i * 3 + 1] * n[b][1] + ev.S[i * 3 + 2] * n[b][2];
        sig = n[a][0] * Sn[0] + n[a][1] * Sn[1] + n[a][2] * Sn[2];
      }
      if (hasM) {
        g[p] = (KT)(V * sig);
        st[(24 + p) * ne + e] = (KT)(V * ev.c2 * s);
      } else {
        g[p] = (KT)(V * (ev.c2 * s + sig));
      }
    }
  // the layout of tet_plane, written with vector stores
#pragma unroll
  for (int a = 0; a < 4; ++a)
    st4(st + 4 * (a * ne + e), (KT)ev.v[a][0], (KT)ev.v[a][1], (KT)ev.v[a][2], g[pair_index(a, a)]);
  st2(st + 16 * ne + 2 * e, (KT)(V * ev.c1), (KT)(V * ev.c2));
  st[18 * ne + e] = g[pair_index(0, 1)];
  st[19 * ne + e] = g[pair_index(0, 2)];
  st[20 * ne + e] = g[pair_index(0, 3)];
  st[21 * ne + e] = g[pair_index(1, 2)];
  st[22 * ne + e] = g[pair_index(1, 3)];
  st[23 * ne + e] = g[pair_index(2, 3)];
  if (hasM) {
    st[34 * ne + e] = (KT)ev.M[0];
    st[35 * ne + e] = (KT)ev.M[1];
    st[36 * ne + e] = (KT)ev.M[2];
    st[37 * ne + e] = (KT)ev.M[4];
    st[38 * ne + e] = (KT)ev.M[5];
    st[39 * ne + e] = (KT)ev.M[8];
  }
}

template <typename KT>
__global__ __launch_bounds__(256) void k_truss_state(TrussArgs ta, const double* __restrict__ U,
                                                     KT* __restrict__ st, double* __restrict__ fe) {
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= ta.ne) return;
  const int64_t ne = ta.ne;
  const int32_t nd0 = ta.conn[e], nd1 = ta.conn[ne + e];
  double d0[3] = {ta.geo[e], ta.geo[ne + e], ta.geo[2 * ne + e]};
  const double l0 = ta.geo[3 * ne + e], A = ta.geo[4 * ne + e];
  double du[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) du[d] = U[3 * (int64_t)nd1 + d] - U[3 * (int64_t)nd0 + d];
  TrussEval tv;
  truss_eval(ta.strain, ta.E, A, d0, l0, du, tv);
  st[e] = (KT)tv.cdd;
  st[ne + e] = (KT)tv.cI;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    st[(2 + d) * ne + e] = (KT)tv.d[d];
    fe[d * ne + e] = tv.cf * tv.d[d];  // f_1 = +cf d, f_0 = -cf d
  }
}

struct GatherGroup {
  const void* state;
  const double* fe;
  int64_t ne, goff;
  int type, hasM;
};

struct GatherArgs {
  GatherGroup g[kMaxGatherGroups];
  int ng;
};

constexpr uint32_t kNoContrib = 0xffffffffu;

// One element block (a, b) in factored form, loaded from the state planes.
// w = 0 marks a padding entry (the loads then read element 0 of group 0 and
// the block adds nothing), so the loop body has no branches on the entry.
template <typename KT, bool HM>
struct Contrib {
  KT w, c1, c2, g, va[3], vb[3];
  KT h, M[HM ? 6 : 1];
  double f[3];
  int tet, hasM;
};

// HM: some group of the model carries the SVK M term (compiled out otherwise)
template <typename KT, bool HM>
__device__ inline void load_contrib(const GatherArgs& ga, uint32_t en, bool dg, Contrib<KT, HM>& c) {
  const bool valid = en != kNoContrib;
  const uint32_t ev = valid ? en : 0u;
  const int64_t ge = ev >> 4;
  const int a = (ev >> 2) & 3, b = ev & 3;
  const void* sp = ga.g[0].state;
  const double* fp = ga.g[0].fe;
  int64_t ne = ga.g[0].ne, goff = ga.g[0].goff;
  int type = ga.g[0].type, hasM = ga.g[0].hasM;
#pragma unroll
  for (int j = 1; j < kMaxGatherGroups; ++j)
    if (j < ga.ng && ge >= ga.g[j].goff) {
      sp = ga.g[j].state;
      fp = ga.g[j].fe;
      ne = ga.g[j].ne;
      goff = ga.g[j].goff;
      type = ga.g[j].type;
      hasM = ga.g[j].hasM;
    }
  const int64_t e = ge - goff;
  const KT* st = (const KT*)sp;
  c.w = valid ? KT(1) : KT(0);
  c.tet = type == kTet;
  c.hasM = hasM;
  c.h = KT(0);
  if (c.tet) {
    c.c1 = st[tet_plane(0, ne, e)];
    c.c2 = st[tet_plane(1, ne, e)];
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      c.va[i] = st[tet_plane(2 + 3 * a + i, ne, e)];
      c.vb[i] = st[tet_plane(2 + 3 * b + i, ne, e)];
    }
    const int p = a <= b ? pair_index(a, b) : pair_index(b, a);
    c.g = st[tet_plane(14 + p, ne, e)];
    if (HM && hasM) {
      c.h = st[(24 + p) * ne + e];
#pragma unroll
      for (int q = 0; q < 6; ++q) c.M[q] = st[(34 + q) * ne + e];
    }
    if (dg) {
#pragma unroll
      for (int d = 0; d < 3; ++d) c.f[d] = fp[(3 * a + d) * ne + e];
    }
  } else {
    c.c1 = st[e];
    c.c2 = st[ne + e];
    // truss: K_ab = s_a s_b [cI I + cdd d d'] in the same factored form
    const KT sab = (a == b) ? KT(1) : KT(-1);
#pragma unroll
    for (int i = 0; i < 3; ++i) {
      c.va[i] = st[(2 + i) * ne + e];
      c.vb[i] = sab * c.va[i];
    }
    c.g = sab * c.c2;  // c2 plane holds cI
    c.c2 = KT(0);
    if (dg) {
      const double sa = a ? 1.0 : -1.0;
#pragma unroll
      for (int d = 0; d < 3; ++d) c.f[d] = sa * fp[d * ne + e];
    }
  }
  if (!dg || !valid) c.f[0] = c.f[1] = c.f[2] = 0.0;
}

template <typename KT, bool HM>
__device__ inline void add_contrib(const Contrib<KT, HM>& c, KT acc[9], double& f0, double& f1,
                                   double& f2) {
  const KT c1 = c.w * c.c1, c2 = c.w * c.c2, g = c.w * c.g;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i * 3 + j] += c1 * c.va[i] * c.vb[j] + c2 * c.vb[i] * c.va[j];
  acc[0] += g;
  acc[4] += g;
  acc[8] += g;
  if (HM && c.hasM) {
    const KT h = c.w * c.h;
    acc[0] += h * c.M[0];
    acc[1] += h * c.M[1];
    acc[2] += h * c.M[2];
    acc[3] += h * c.M[1];
    acc[4] += h * c.M[3];
    acc[5] += h * c.M[4];
    acc[6] += h * c.M[2];
    acc[7] += h * c.M[4];
    acc[8] += h * c.M[5];
  }
  f0 += c.f[0];
  f1 += c.f[1];
  f2 += c.f[2];
}

// The end of a gathered block: fixed dofs zero their rows and columns (unit
// diagonal entry, as the scatter path does), the block is stored at o (value v
// at o[64 v]), and on the diagonal the FP64-computed inverse of the finished
// block (block-Jacobi) and the row's f_int are written.
template <typename KT>
__device__ inline void gather_finish(KT acc[9], uint32_t mr, uint32_t mc, bool dg, KT* __restrict__ o,
                                     KT* __restrict__ Dinv, double* __restrict__ fint, int64_t r,
                                     double f0, double f1, double f2) {
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
      if (((mr >> i) & 1u) | ((mc >> j) & 1u)) acc[i * 3 + j] = (dg && i == j) ? KT(1) : KT(0);
#pragma unroll
  for (int v = 0; v < 9; ++v) o[64 * v] = acc[v];
  if (dg) {
    double A[9];
#pragma unroll
    for (int q = 0; q < 9; ++q) A[q] = (double)acc[q];
    const double c0 = A[4] * A[8] - A[5] * A[7], cf1 = A[5] * A[6] - A[3] * A[8],
                 cf2 = A[3] * A[7] - A[4] * A[6];
    const double rdet = 1.0 / (A[0] * c0 + A[1] * cf1 + A[2] * cf2);
    KT* di = Dinv + 9 * r;
    di[0] = (KT)(c0 * rdet);
    di[1] = (KT)((A[2] * A[7] - A[1] * A[8]) * rdet);
    di[2] = (KT)((A[1] * A[5] - A[2] * A[4]) * rdet);
    di[3] = (KT)(cf1 * rdet);
    di[4] = (KT)((A[0] * A[8] - A[2] * A[6]) * rdet);
    di[5] = (KT)((A[2] * A[3] - A[0] * A[5]) * rdet);
    di[6] = (KT)(cf2 * rdet);
    di[7] = (KT)((A[1] * A[6] - A[0] * A[7]) * rdet);
    di[8] = (KT)((A[0] * A[4] - A[1] * A[3]) * rdet);
    fint[3 * r] = f0;
    fint[3 * r + 1] = f1;
    fint[3 * r + 2] = f2;
  }
}

// One wave per sliced-ELL group (64 rows at one round), one lane per block.
// The contribution rounds are uniform over the wave (padded lists), the
// element-group lookup is a predicated select (no dynamic indexing of the
// kernel arguments).  The lane holding a diagonal block also writes the row's
// f_int (its contributions are exactly the node's element incidences) and the
// FP64-computed inverse of the finished block (block-Jacobi).  Fixed dofs:
// their rows and columns are zeroed and the diagonal entry set to 1, as the
// scatter path does.
template <typename KT, bool HM>
__global__ __launch_bounds__(256) void k_gather(GatherArgs ga, const SellGroup* __restrict__ groups,
                                                int64_t ngroups, const uint32_t* __restrict__ contrib,
                                                const int32_t* __restrict__ row_ptr,
                                                const int32_t* __restrict__ col,
                                                const uint8_t* __restrict__ mask,
                                                KT* __restrict__ Ks, KT* __restrict__ Dinv,
                                                double* __restrict__ fint, int64_t nn) {
  // XCD-aware: hardware workgroup ids go round-robin over the 8 XCDs; logical
  // block (b % 8) * (nb / 8) + b / 8 gives each XCD a contiguous run of slices,
  // so the element states a slice's rounds share stay in one L2
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t L = (b & 7) * (nb >> 3) + (b >> 3);
  const int64_t w = L * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (w >= ngroups) return;
  const SellGroup sg = groups[w];
  const int64_t r = sg.row0 + lane % kSellRows;
  const int k = kSellLanes * sg.round + lane / kSellRows;
  if (r >= nn) return;
  const int32_t rs = row_ptr[r];
  if (k >= row_ptr[r + 1] - rs) return;
  const int32_t cc = col[rs + k];
  const bool dg = (cc == r);
  KT acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = KT(0);
  double f0 = 0, f1 = 0, f2 = 0;
  const uint32_t* cp = contrib + sg.cbase + lane;
  const int qw = sg.qwidth;
  // two contributions per step; the entries of the next step are loaded before
  // this step's element states, so each step waits on one memory latency
  uint32_t n0 = qw > 0 ? cp[0] : kNoContrib;
  uint32_t n1 = qw > 1 ? cp[64] : kNoContrib;
  for (int q = 0; q < qw; q += 2) {
    const uint32_t e0 = n0, e1 = n1;
    n0 = q + 2 < qw ? cp[64 * (int64_t)(q + 2)] : kNoContrib;
    n1 = q + 3 < qw ? cp[64 * (int64_t)(q + 3)] : kNoContrib;
    Contrib<KT, HM> c0, c1;
    load_contrib(ga, e0, dg, c0);
    load_contrib(ga, e1, dg, c1);
    add_contrib(c0, acc, f0, f1, f2);
    add_contrib(c1, acc, f0, f1, f2);
  }
  gather_finish(acc, mask[r], mask[cc], dg, Ks + 9 * (int64_t)sg.base + lane, Dinv, fint, r, f0, f1,
                f2);
}

// base[idx] through a 32-bit byte offset off a wave-uniform base (the host
// checks that every plane offset fits), which the compiler turns into a scalar
// base plus one 32-bit lane register per load
template <typename T>
__device__ inline T ld32(const T* __restrict__ base, uint32_t idx) {
  return *(const T*)((const char*)base + idx * (uint32_t)sizeof(T));
}

// 4 and 2 consecutive values at base[idx] (32-bit byte offsets, as ld32)
__device__ inline void ld4(const float* __restrict__ base, uint32_t idx, float v[4]) {
  const float4 t = *(const float4*)((const char*)base + idx * 4u);
  v[0] = t.x;
  v[1] = t.y;
  v[2] = t.z;
  v[3] = t.w;
}
__device__ inline void ld4(const double* __restrict__ base, uint32_t idx, double v[4]) {
  const double2* q = (const double2*)((const char*)base + idx * 8u);
  const double2 t0 = q[0], t1 = q[1];
  v[0] = t0.x;
  v[1] = t0.y;
  v[2] = t1.x;
  v[3] = t1.y;
}
__device__ inline void ld2(const float* __restrict__ base, uint32_t idx, float v[2]) {
  const float2 t = *(const float2*)((const char*)base + idx * 4u);
  v[0] = t.x;
  v[1] = t.y;
}
__device__ inline void ld2(const double* __restrict__ base, uint32_t idx, double v[2]) {
  const double2 t = *(const double2*)((const char*)base + idx * 8u);
  v[0] = t.x;
  v[1] = t.y;
}

// One contribution of a model with a single tet group (entries carry the
// element's position in the group: its offset is 0).
template <typename KT, bool HM>
__device__ inline void load_tet(const KT* __restrict__ st, const double* __restrict__ fe, uint32_t ne,
                                uint32_t en, bool dg, Contrib<KT, HM>& c) {
  const bool valid = en != kNoContrib;
  const uint32_t ev = valid ? en : 0u;
  const uint32_t e = ev >> 4;
  const int a = (ev >> 2) & 3, b = ev & 3;
  c.w = valid ? KT(1) : KT(0);
  c.tet = 1;
  c.hasM = HM;
  // tet_plane's records: node a's (v_a, g_aa) and (c1, c2) in one vector load
  // each; node b's record and the off-diagonal g_ab only when b != a
  KT A[4], C2[2];
  ld4(st, 4u * ((uint32_t)a * ne + e), A);
  ld2(st, 16u * ne + 2u * e, C2);
  c.c1 = C2[0];
  c.c2 = C2[1];
#pragma unroll
  for (int i = 0; i < 3; ++i) c.va[i] = c.vb[i] = A[i];
  c.g = A[3];
  if (a != b) {
    KT B[4];
    ld4(st, 4u * ((uint32_t)b * ne + e), B);
#pragma unroll
    for (int i = 0; i < 3; ++i) c.vb[i] = B[i];
    c.g = ld32(st, (18u + (uint32_t)((kOffPairTable >> (16 * a + 4 * b)) & 15u)) * ne + e);
  }
  const int p = (int)((kPairTable >> (16 * a + 4 * b)) & 15u);
  c.h = KT(0);
  if (HM) {
    c.h = ld32(st, (uint32_t)(24 + p) * ne + e);
#pragma unroll
    for (int q = 0; q < 6; ++q) c.M[q] = ld32(st, (uint32_t)(34 + q) * ne + e);
  }
  c.f[0] = c.f[1] = c.f[2] = 0.0;
  if (dg && valid) {
#pragma unroll
    for (int d = 0; d < 3; ++d) c.f[d] = ld32(fe, (uint32_t)(3 * a + d) * ne + e);
  }
}

// k_gather for a model of one tet group: the same contributions added in the
// same order, the group's state at a wave-uniform base with 32-bit offsets (no
// per-lane group lookup or 64-bit address arithmetic), read as tet_plane's
// vector records (3 to 4 loads per contribution instead of 9 scalar ones), NC
// contributions per step.
template <typename KT, bool HM, int NC>
__global__ __launch_bounds__(256) void k_gather_tet(
    const KT* __restrict__ st, const double* __restrict__ fe, uint32_t ne,
    const SellGroup* __restrict__ groups, int64_t ngroups, const uint32_t* __restrict__ contrib,
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
    const uint8_t* __restrict__ mask, KT* __restrict__ Ks, KT* __restrict__ Dinv,
    double* __restrict__ fint, int64_t nn) {
  const int64_t nb = gridDim.x, b = blockIdx.x;
  const int64_t L = (b & 7) * (nb >> 3) + (b >> 3);  // XCD-aware, as k_gather
  // wave-uniform (scalar loads of the group record, scalar loop control)
  const int32_t w = __builtin_amdgcn_readfirstlane((int32_t)(L * 4 + (threadIdx.x >> 6)));
  const int lane = threadIdx.x & 63;
  if (w >= ngroups) return;
  const SellGroup sg = groups[w];
  const int64_t r = sg.row0 + lane % kSellRows;
  const int k = kSellLanes * sg.round + lane / kSellRows;
  if (r >= nn) return;
  const int32_t rs = row_ptr[r];
  if (k >= row_ptr[r + 1] - rs) return;
  const int32_t cc = col[rs + k];
  const bool dg = (cc == r);
  KT acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = KT(0);
  double f0 = 0, f1 = 0, f2 = 0;
  const uint32_t* cp = contrib + sg.cbase + lane;
  const int qw = sg.qwidth;
  uint32_t nx[NC];
#pragma unroll
  for (int j = 0; j < NC; ++j) nx[j] = j < qw ? cp[64 * j] : kNoContrib;
  for (int q = 0; q < qw; q += NC) {
    uint32_t ex[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) {
      ex[j] = nx[j];
      nx[j] = q + NC + j < qw ? cp[64 * (int64_t)(q + NC + j)] : kNoContrib;
    }
    Contrib<KT, HM> cx[NC];
#pragma unroll
    for (int j = 0; j < NC; ++j) load_tet(st, fe, ne, ex[j], dg, cx[j]);
#pragma unroll
    for (int j = 0; j < NC; ++j) add_contrib(cx[j], acc, f0, f1, f2);
  }
  gather_finish(acc, mask[r], mask[cc], dg, Ks + 9 * (int64_t)sg.base + lane, Dinv, fint, r, f0, f1,
                f2);
}

// ---- staged gather ---------------------------------------------------------------
// One workgroup per `rows` consecutive rows of a 64-row slice (the lists of
// host_setup.hpp::staged_gather).  First the states of the workgroup's elements
// go to LDS, element i of its ascending list at column i (SP values apart);
// the loads run along the list, which on a structured mesh is a few runs of
// consecutive elements, so each plane's load is coalesced.  Then lane l of a
// step sums its block (row, round) from LDS through 16-bit entries, in the
// lists' element order -- the sums and order of k_gather -- and ends it as
// k_gather does; the diagonal's lane reads the nodal forces from HBM.  The
// next step's entries and column are loaded before this step's sums.
constexpr int kStageWords = 12;  // entry words of a step held in registers (24 contributions)

template <typename KT, bool HM, int SP>
__device__ inline void stage_add(const KT* __restrict__ ls, uint32_t en, KT acc[9]) {
  const bool valid = en != host::kNoStage;
  const KT* x = ls + (valid ? (int)(en >> 4) : 0) * SP;
  const int a = (en >> 2) & 3, b = en & 3;
  const KT c1 = valid ? x[0] : KT(0), c2 = valid ? x[1] : KT(0);
  KT va[3], vb[3];
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    va[i] = x[2 + 3 * a + i];
    vb[i] = x[2 + 3 * b + i];
  }
  const int p = a <= b ? pair_index(a, b) : pair_index(b, a);
  const KT g = valid ? x[14 + p] : KT(0);
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i * 3 + j] += c1 * va[i] * vb[j] + c2 * vb[i] * va[j];
  acc[0] += g;
  acc[4] += g;
  acc[8] += g;
  if (HM) {
    const KT h = valid ? x[24 + p] : KT(0);
    KT M[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) M[q] = x[34 + q];
    acc[0] += h * M[0];
    acc[1] += h * M[1];
    acc[2] += h * M[2];
    acc[3] += h * M[1];
    acc[4] += h * M[3];
    acc[5] += h * M[4];
    acc[6] += h * M[2];
    acc[7] += h * M[4];
    acc[8] += h * M[5];
  }
}

// KS: list entries per thread in the staging (elements and incidences alike),
// all loads issued before the first store
template <typename KT, int NST, int NT, int KS>
__global__ __launch_bounds__(NT) void k_gather_lds(
    const KT* __restrict__ st, const double* __restrict__ fe, int64_t ne,
    const host::StageGroup* __restrict__ groups, int64_t ngroups, const int32_t* __restrict__ stage,
    const host::StageStep* __restrict__ steps, const uint32_t* __restrict__ words,
    const uint32_t* __restrict__ pairs, const int32_t* __restrict__ pair_row, int rlog2,
    int cap, int pairs_cap, int steps_cap, const int32_t* __restrict__ row_ptr,
    const uint8_t* __restrict__ colmask, const int32_t* __restrict__ diag,
    const uint8_t* __restrict__ mask, const int32_t* __restrict__ sell_ptr, int64_t rstride,
    KT* __restrict__ Ks, KT* __restrict__ Dinv, double* __restrict__ fint, int64_t nn) {
  constexpr bool HM = NST == kTetStateM;
  constexpr int SP = NST + 1;  // odd stride: the staging stores hit distinct banks
  constexpr int NW = NT / 64;
  extern __shared__ __align__(16) unsigned char lds_raw[];
  host::StageStep* lst = reinterpret_cast<host::StageStep*>(lds_raw);
  double* lfe = reinterpret_cast<double*>(lst + steps_cap);  // 3 per incidence
  KT* ls = reinterpret_cast<KT*>(lfe + 3 * (size_t)pairs_cap);
  (void)cap;
  const int64_t nb = gridDim.x, bid = blockIdx.x;
  const int64_t L = (bid & 7) * (nb >> 3) + (bid >> 3);  // XCD-aware, as k_gather
  if (L >= ngroups) return;  // the whole workgroup, before its barrier
  const host::StageGroup g = groups[L];
  // workgroup L holds rows L << rlog2 on: the row loads need no group record
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int64_t row0 = L << rlog2;
  const int64_t r = row0 + (lane & ((1 << rlog2) - 1));
  const bool live = r < nn;
  const int32_t rs = live ? row_ptr[r] : 0;
  const int32_t len = live ? row_ptr[r + 1] - rs : 0;
  const int32_t dround = live ? diag[r] - rs : -1;
  const uint32_t mr = live ? mask[r] : 0u;
  const int32_t pr = live ? pair_row[r] : 0;
  const int64_t sbase = sell_ptr[row0 >> 6];
  KT* const orow = Ks + (r & 63);

  // this wave's first step (table entry, entry words, column mask), in flight
  // during the staging
  host::StageStep cur{0, 0, 0};
  uint32_t wc[kStageWords];
  uint32_t mc = 0;
  if (wave < g.nsteps) {
    cur = steps[g.step0 + wave];
    const int m = cur.m0 + (lane >> rlog2);
    if (m < len) mc = colmask[rs + m];
  }
#pragma unroll
  for (int k = 0; k < kStageWords; ++k)
    wc[k] = 2 * k < cur.qwidth ? words[cur.cbase + 64 * k + lane] : 0xffffffffu;

  // staging: the step table, the element states and the rows' nodal forces
  for (int i = threadIdx.x; i < g.nsteps; i += NT) lst[i] = steps[g.step0 + i];
  int32_t ei[KS];
  uint32_t pi[KS];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int i = (int)threadIdx.x + k * NT;
    ei[k] = i < g.nstage ? stage[g.stage_off + i] : -1;
    pi[k] = i < g.npairs ? pairs[g.pair_off + i] : 0xffffffffu;
  }
  KT v[KS][NST];
  double fv[KS][3];
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    if (ei[k] >= 0) {
#pragma unroll
      for (int p = 0; p < NST; ++p) v[k][p] = st[tet_plane(p, ne, ei[k])];
    }
    if (pi[k] != 0xffffffffu) {
      const int64_t e = pi[k] & 0x0fffffffu;
      const int a = (int)(pi[k] >> 28);
#pragma unroll
      for (int d = 0; d < 3; ++d) fv[k][d] = fe[(3 * a + d) * ne + e];
    }
  }
#pragma unroll
  for (int k = 0; k < KS; ++k) {
    const int i = (int)threadIdx.x + k * NT;
    if (ei[k] >= 0) {
#pragma unroll
      for (int p = 0; p < NST; ++p) ls[i * SP + p] = v[k][p];
    }
    if (pi[k] != 0xffffffffu) {
#pragma unroll
      for (int d = 0; d < 3; ++d) lfe[3 * i + d] = fv[k][d];
    }
  }
  __syncthreads();

  for (int t = wave; t < g.nsteps; t += NW) {
    // the next step's loads first
    const int tn = t + NW;
    host::StageStep nx{0, 0, 0};
    uint32_t wn[kStageWords];
    uint32_t mn = 0;
    if (tn < g.nsteps) {
      nx = lst[tn];
      const int m2 = nx.m0 + (lane >> rlog2);
      if (m2 < len) mn = colmask[rs + m2];
    }
#pragma unroll
    for (int k = 0; k < kStageWords; ++k)
      wn[k] = 2 * k < nx.qwidth ? words[nx.cbase + 64 * k + lane] : 0xffffffffu;

    const int m = cur.m0 + (lane >> rlog2);
    if (m < len) {
      const bool dg = m == dround;
      const int nw = cur.qwidth >> 1;
      KT acc[9];
#pragma unroll
      for (int q = 0; q < 9; ++q) acc[q] = KT(0);
#pragma unroll
      for (int k = 0; k < kStageWords; ++k)
        if (k < nw) {  // uniform over the wave
          stage_add<KT, HM, SP>(ls, wc[k] & 0xffffu, acc);
          stage_add<KT, HM, SP>(ls, wc[k] >> 16, acc);
        }
      for (int k = kStageWords; k < nw; ++k) {  // longer lists (unstructured meshes)
        const uint32_t w = words[cur.cbase + 64 * k + lane];
        stage_add<KT, HM, SP>(ls, w & 0xffffu, acc);
        stage_add<KT, HM, SP>(ls, w >> 16, acc);
      }
      // f_int: the row's incidences in its diagonal block's contribution order
      double f0 = 0, f1 = 0, f2 = 0;
      if (dg) {  // valid entries are a prefix of the list
#pragma unroll
        for (int k = 0; k < kStageWords; ++k)
#pragma unroll
          for (int h = 0; h < 2; ++h)
            if (k < nw && ((wc[k] >> (16 * h)) & 0xffffu) != host::kNoStage) {
              const double* x = lfe + 3 * (pr + 2 * k + h);
              f0 += x[0];
              f1 += x[1];
              f2 += x[2];
            }
        for (int k = kStageWords; k < nw; ++k) {
          const uint32_t w = words[cur.cbase + 64 * k + lane];
          for (int h = 0; h < 2; ++h)
            if (((w >> (16 * h)) & 0xffffu) != host::kNoStage) {
              const double* x = lfe + 3 * (pr + 2 * k + h);
              f0 += x[0];
              f1 += x[1];
              f2 += x[2];
            }
        }
      }
      gather_finish(acc, mr, mc, dg, orow + 9 * (sbase + rstride * m), Dinv, fint, r, f0, f1, f2);
    }
    cur = nx;
    mc = mn;
#pragma unroll
    for (int k = 0; k < kStageWords; ++k) wc[k] = wn[k];
  }
}

// Point location (PointEvalHandler, src/Meshes/Handlers.jl:67-87): one lane per
// query point, the candidates of its grid bin scanned in user element order;
// the first tetrahedron whose barycentric coordinates are all >= -tol holds the
// point (the reference's Serial search keeps the first containing element,
// src/Meshes/Searches.jl:64-82).  lambda_a = dN_a . (p - X_b) for any node b != a
// (linear shape functions vanish at the other nodes), i.e. the weights of
// Tetrahedrons.jl:311-313.
__global__ void k_locate(TetArgs ta, const double* __restrict__ X, const double* __restrict__ Y,
                         const double* __restrict__ Z, const double* __restrict__ px,
                         const double* __restrict__ py, const double* __restrict__ pz, int64_t np,
                         const int64_t* __restrict__ bin_ptr, const int32_t* __restrict__ bin_elem,
                         double3 lo, double3 inv_h, int3 nb, const int32_t* __restrict__ perm,
                         double tol, int64_t* __restrict__ out_elem, double* __restrict__ out_w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= np) return;
  const double p[3] = {px[i], py[i], pz[i]};
  int64_t found = -1;
  double w[4] = {0, 0, 0, 0};
  const int bx = (int)floor((p[0] - lo.x) * inv_h.x), by = (int)floor((p[1] - lo.y) * inv_h.y),
            bz = (int)floor((p[2] - lo.z) * inv_h.z);
  if (bx >= 0 && by >= 0 && bz >= 0 && bx < nb.x && by < nb.y && bz < nb.z) {
    const int64_t b = ((int64_t)bz * nb.y + by) * nb.x + bx;
    for (int64_t q = bin_ptr[b]; q < bin_ptr[b + 1] && found < 0; ++q) {
      const int64_t e = bin_elem[q];
      double lam[4];
      bool inside = true;
#pragma unroll
      for (int a = 0; a < 4; ++a) {
        const int32_t nbn = ta.conn[((a + 1) & 3) * ta.ne + e];
        const double d0 = p[0] - X[nbn], d1 = p[1] - Y[nbn], d2 = p[2] - Z[nbn];
        lam[a] = ta.geo[(a * 3) * ta.ne + e] * d0 + ta.geo[(a * 3 + 1) * ta.ne + e] * d1 +
                 ta.geo[(a * 3 + 2) * ta.ne + e] * d2;
        inside = inside && lam[a] >= -tol;
      }
      if (inside) {
        found = perm[e];
#pragma unroll
        for (int a = 0; a < 4; ++a) w[a] = lam[a];
      }
    }
  }
  out_elem[i] = found;
#pragma unroll
  for (int a = 0; a < 4; ++a) out_w[4 * i + a] = w[a];
}

// Block-Jacobi rows of nodes without a diagonal block (fully fixed, no element).
template <typename KT>
__global__ void k_dinv_identity(const int32_t* __restrict__ diag, KT* __restrict__ Dinv, int64_t nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn || diag[i] >= 0) return;
#pragma unroll
  for (int q = 0; q < 9; ++q) Dinv[9 * i + q] = (q % 4) == 0 ? KT(1) : KT(0);
}

// y += K(U) v with the element blocks evaluated in FP64 and never stored: the
// FP64 tangent product used to refine mixed-precision solves.  Lanes, colours
// and the fixed-dof masking are those of k_tet_scatter, so the operator is the
// assembled one (fixed rows and columns are completed by k_fixed_copy).
__global__ __launch_bounds__(kScatterThreads) void k_tet_apply(TetArgs ta,
                                                               const double* __restrict__ U,
                                                               const double* __restrict__ v,
                                                               double* __restrict__ y,
                                                               const uint8_t* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * kScatterThreads + threadIdx.x;
  const int64_t k = ta.e0 + (t >> 2);
  const int a = (int)(t & 3);
  if (k >= ta.e1) return;
  const int64_t e = ta.idx ? (int64_t)ta.idx[k] : k;
  int32_t nd[4];
  double n[4][3], u[4][3], V, p1;
  tet_load(ta, e, U, nd, n, V, u, p1);
  TetEval ev;
  tet_eval(ta.mat, ta.p0, p1, n, u, ev);
  TetBlocks<double> tb;
  tb.load(ev, n, V);
  double acc[3] = {0, 0, 0};
#pragma unroll
  for (int b = 0; b < 4; ++b) {
    const uint32_t mb = mask[nd[b]];
    double vb[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) vb[j] = ((mb >> j) & 1u) ? 0.0 : v[3 * (int64_t)nd[b] + j];
    double blk[9];
    tb.block(a, b, blk);
#pragma unroll
    for (int i = 0; i < 3; ++i) acc[i] += blk[i * 3] * vb[0] + blk[i * 3 + 1] * vb[1] + blk[i * 3 + 2] * vb[2];
  }
  const uint32_t ma = mask[nd[a]];
  double* yp = y + 3 * (int64_t)nd[a];
#pragma unroll
  for (int i = 0; i < 3; ++i) yp[i] += ((ma >> i) & 1u) ? 0.0 : acc[i];
}

__global__ __launch_bounds__(kScatterThreads) void k_truss_apply(TrussArgs ta,
                                                                 const double* __restrict__ U,
                                                                 const double* __restrict__ v,
                                                                 double* __restrict__ y,
                                                                 const uint8_t* __restrict__ mask) {
  const int64_t t = (int64_t)blockIdx.x * kScatterThreads + threadIdx.x;
  const int64_t k = ta.e0 + (t >> 1);
  const int a = (int)(t & 1);
  if (k >= ta.e1) return;
  const int64_t e = ta.idx ? (int64_t)ta.idx[k] : k;
  const int64_t ne = ta.ne;
  const int32_t nd0 = ta.conn[e], nd1 = ta.conn[ne + e];
  double d0[3] = {ta.geo[e], ta.geo[ne + e], ta.geo[2 * ne + e]};
  const double l0 = ta.geo[3 * ne + e], A = ta.geo[4 * ne + e];
  double du[3];
#pragma unroll
  for (int d = 0; d < 3; ++d) du[d] = U[3 * (int64_t)nd1 + d] - U[3 * (int64_t)nd0 + d];
  TrussEval tv;
  truss_eval(ta.strain, ta.E, A, d0, l0, du, tv);
  // K_ab v_b summed over b: s_a [cI w + cdd d (d.w)], w = v_1 - v_0 (masked)
  const uint32_t m0 = mask[nd0], m1 = mask[nd1];
  double w[3];
#pragma unroll
  for (int j = 0; j < 3; ++j)
    w[j] = (((m1 >> j) & 1u) ? 0.0 : v[3 * (int64_t)nd1 + j]) -
           (((m0 >> j) & 1u) ? 0.0 : v[3 * (int64_t)nd0 + j]);
  const double dw = tv.d[0] * w[0] + tv.d[1] * w[1] + tv.d[2] * w[2];
  const double sa = a ? 1.0 : -1.0;
  const int32_t na = a ? nd1 : nd0;
  const uint32_t ma = a ? m1 : m0;
  double* yp = y + 3 * (int64_t)na;
#pragma unroll
  for (int i = 0; i < 3; ++i) yp[i] += ((ma >> i) & 1u) ? 0.0 : sa * (tv.cI * w[i] + tv.cdd * tv.d[i] * dw);
}

// y[fixed] = v[fixed] (the identity rows of the assembled operator).
__global__ void k_fixed_copy(const double* __restrict__ v, double* __restrict__ y,
                             const uint8_t* __restrict__ mask, int64_t nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const uint32_t m = mask[i];
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if ((m >> d) & 1u) y[3 * i + d] = v[3 * i + d];
}

template <typename T>
__global__ void k_fill(T* __restrict__ p, int64_t n, T v) {
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * blockDim.x)
    p[i] = v;
}

template <typename KT>
__global__ void k_fixed_identity(KT* __restrict__ K, const int32_t* __restrict__ diag,
                                 const uint8_t* __restrict__ mask, int64_t nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const uint32_t m = mask[i];
  if (!m || diag[i] < 0) return;
#pragma unroll
  for (int d = 0; d < 3; ++d)
    if ((m >> d) & 1u) K[9 * (int64_t)diag[i] + 4 * d] = KT(1);
}

// Inverse of each 3x3 diagonal block (computed in FP64, stored in KT).
template <typename KT>
__global__ void k_block_jacobi(const KT* __restrict__ K, const int32_t* __restrict__ diag,
                               KT* __restrict__ Dinv, int64_t nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double A[9];
  KT* o = Dinv + 9 * i;
  if (diag[i] < 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) o[q] = (q % 4) == 0 ? KT(1) : KT(0);
    return;
  }
  const KT* k = K + 9 * (int64_t)diag[i];
#pragma unroll
  for (int q = 0; q < 9; ++q) A[q] = (double)k[q];
  double c0 = A[4] * A[8] - A[5] * A[7], c1 = A[5] * A[6] - A[3] * A[8], c2 = A[3] * A[7] - A[4] * A[6];
  double det = A[0] * c0 + A[1] * c1 + A[2] * c2;
  double r = 1.0 / det;
  o[0] = (KT)(c0 * r);
  o[1] = (KT)((A[2] * A[7] - A[1] * A[8]) * r);
  o[2] = (KT)((A[1] * A[5] - A[2] * A[4]) * r);
  o[3] = (KT)(c1 * r);
  o[4] = (KT)((A[0] * A[8] - A[2] * A[6]) * r);
  o[5] = (KT)((A[2] * A[3] - A[0] * A[5]) * r);
  o[6] = (KT)(c2 * r);
  o[7] = (KT)((A[1] * A[6] - A[0] * A[7]) * r);
  o[8] = (KT)((A[0] * A[4] - A[1] * A[3]) * r);
}

// Per-element outputs in FP64 (the batch form of internal_forces).
__global__ void k_tet_outputs(TetArgs ta, const double* __restrict__ U,
                              const int32_t* __restrict__ perm, double* __restrict__ f,
                              double* __restrict__ K, double* __restrict__ stress,
                              double* __restrict__ strain) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ta.ne) return;
  int32_t nd[4];
  double n[4][3], u[4][3], V, p1;
  tet_load(ta, e, U, nd, n, V, u, p1);
  TetEval ev;
  tet_eval(ta.mat, ta.p0, p1, n, u, ev);
  const int64_t o = perm[e];
  if (f) {
    for (int a = 0; a < 4; ++a) {
      double fa[3];
      tet_force_row(ev, n, V, a, fa);
      for (int d = 0; d < 3; ++d) f[o * 12 + a * 3 + d] = fa[d];
    }
  }
  if (K) {
    TetBlocks<double> tb;
    tb.load(ev, n, V);
    for (int a = 0; a < 4; ++a)
      for (int b = 0; b < 4; ++b) {
        double blk[9];
        tb.block(a, b, blk);
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j) K[o * 144 + (a * 3 + i) * 12 + b * 3 + j] = blk[i * 3 + j];
      }
  }
  if (ta.mat == ONSAS_MAT_LINEAR) {
    // sigma and eps = sym(grad u)
    if (stress)
      for (int q = 0; q < 9; ++q) stress[o * 9 + q] = ev.S[q];
    if (strain) {
      double H[9];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
          H[i * 3 + j] = u[0][i] * n[0][j] + u[1][i] * n[1][j] + u[2][i] * n[2][j] + u[3][i] * n[3][j];
      for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) strain[o * 9 + i * 3 + j] = 0.5 * (H[i * 3 + j] + H[j * 3 + i]);
    }
  } else {
    // P = F S, C = F'F (Tetrahedrons.jl:217-221)
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) {
        double p = 0, c = 0;
        for (int k = 0; k < 3; ++k) {
          p += ev.F[i * 3 + k] * ev.S[k * 3 + j];
          c += ev.F[k * 3 + i] * ev.F[k * 3 + j];
        }
        if (stress) stress[o * 9 + i * 3 + j] = p;
        if (strain) strain[o * 9 + i * 3 + j] = c;
      }
  }
}

__global__ void k_truss_outputs(TrussArgs ta, const double* __restrict__ U,
                                const int32_t* __restrict__ perm, double* __restrict__ f,
                                double* __restrict__ K, double* __restrict__ stress,
                                double* __restrict__ strain) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ta.ne) return;
  const int64_t ne = ta.ne;
  const int32_t nd0 = ta.conn[e], nd1 = ta.conn[ne + e];
  double d0[3] = {ta.geo[e], ta.geo[ne + e], ta.geo[2 * ne + e]};
  const double l0 = ta.geo[3 * ne + e], A = ta.geo[4 * ne + e];
  double du[3];
  for (int d = 0; d < 3; ++d) du[d] = U[3 * (int64_t)nd1 + d] - U[3 * (int64_t)nd0 + d];
  TrussEval tv;
  truss_eval(ta.strain, ta.E, A, d0, l0, du, tv);
  const int64_t o = perm[e];
  if (f)
    for (int a = 0; a < 2; ++a)
      for (int d = 0; d < 3; ++d) f[o * 6 + a * 3 + d] = (a ? 1.0 : -1.0) * tv.cf * tv.d[d];
  if (K)
    for (int a = 0; a < 2; ++a)
      for (int b = 0; b < 2; ++b)
        for (int i = 0; i < 3; ++i)
          for (int j = 0; j < 3; ++j)
            K[o * 36 + (a * 3 + i) * 6 + b * 3 + j] =
                (a == b ? 1.0 : -1.0) * ((i == j ? tv.cI : 0.0) + tv.cdd * tv.d[i] * tv.d[j]);
  if (stress) {
    for (int q = 0; q < 9; ++q) stress[o * 9 + q] = 0.0;
    stress[o * 9] = tv.sigma;
  }
  if (strain) {
    for (int q = 0; q < 9; ++q) strain[o * 9 + q] = 0.0;
    strain[o * 9] = tv.eps;
  }
}

// f_ext of node i = nodal + sum over its pressure faces (fixed order) of
// p(t) (-n) A / 3 (LocalLoadBoundaryConditions.jl:36-56, TriangularFaces.jl:44-65),
// duplicate dofs summed as in StructuralBoundaryConditions.jl:195-220.
__global__ void k_fext(const double* __restrict__ x, const double* __restrict__ y,
                       const double* __restrict__ z, const int32_t* __restrict__ fconn,
                       const int32_t* __restrict__ fgrp, const int32_t* __restrict__ nf_ptr,
                       const int32_t* __restrict__ nf_idx, const double* __restrict__ pressures,
                       const double* __restrict__ nodal, double* __restrict__ fext, int64_t nn) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= nn) return;
  double acc[3] = {0, 0, 0};
  for (int32_t k = nf_ptr[i]; k < nf_ptr[i + 1]; ++k) {
    const int32_t f = nf_idx[k];
    const int32_t a = fconn[3 * f], b = fconn[3 * f + 1], c = fconn[3 * f + 2];
    const double e1[3] = {x[b] - x[a], y[b] - y[a], z[b] - z[a]};
    const double e2[3] = {x[c] - x[a], y[c] - y[a], z[c] - z[a]};
    const double av[3] = {0.5 * (e1[1] * e2[2] - e1[2] * e2[1]), 0.5 * (e1[2] * e2[0] - e1[0] * e2[2]),
                          0.5 * (e1[0] * e2[1] - e1[1] * e2[0])};
    const double A = sqrt(av[0] * av[0] + av[1] * av[1] + av[2] * av[2]);
    const double p = pressures[fgrp[f]];
    for (int d = 0; d < 3; ++d) acc[d] += p * (-(av[d] / A)) * A / 3;
  }
  for (int d = 0; d < 3; ++d) fext[3 * i + d] = nodal[3 * i + d] + acc[d];
}

// dJ/dG_e = -lambda_e . d f_e / d G_e with
//   NeoHookean: dS/dG = I - C^-1  ->  d f_a / dG = V (F - F^-T) n_a
//   SVK:        dS/dG = 2E        ->  d f_a / dG = 2 V F E n_a
__global__ void k_adjoint_shear(TetArgs ta, const double* __restrict__ U,
                                const double* __restrict__ lam, const int32_t* __restrict__ perm,
                                double* __restrict__ out) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= ta.ne) return;
  int32_t nd[4];
  double n[4][3], u[4][3], V, p1;
  tet_load(ta, e, U, nd, n, V, u, p1);
  TetEval ev;
  tet_eval(ta.mat, ta.p0, p1, n, u, ev);
  if (ta.mat == ONSAS_MAT_NEOHOOKEAN) {
    // F^-T n_a is ev.v[a]
    double s = 0;
    for (int a = 0; a < 4; ++a)
      for (int i = 0; i < 3; ++i) {
        double Fn = ev.F[i * 3] * n[a][0] + ev.F[i * 3 + 1] * n[a][1] + ev.F[i * 3 + 2] * n[a][2];
        s += lam[3 * (int64_t)nd[a] + i] * (Fn - ev.v[a][i]);
      }
    out[perm[e]] = -V * s;
    return;
  }
  // SVK: E = (F'F - I)/2
  double D[9];  // d(F S)/dG = 2 F E
  double E[9];
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j)
      E[i * 3 + j] = 0.5 * (ev.F[i] * ev.F[j] + ev.F[3 + i] * ev.F[3 + j] + ev.F[6 + i] * ev.F[6 + j] -
                            (i == j ? 1.0 : 0.0));
  for (int i = 0; i < 3; ++i)
    for (int j = 0; j < 3; ++j) {
      double s = 0;
      for (int k = 0; k < 3; ++k) s += ev.F[i * 3 + k] * E[k * 3 + j];
      D[i * 3 + j] = 2 * s;
    }
  double s = 0;
  for (int a = 0; a < 4; ++a)
    for (int i = 0; i < 3; ++i) {
      double Dn = D[i * 3] * n[a][0] + D[i * 3 + 1] * n[a][1] + D[i * 3 + 2] * n[a][2];
      s += lam[3 * (int64_t)nd[a] + i] * Dn;
    }
  out[perm[e]] = -V * s;
}

TetArgs tet_args(const ElemGroup& g, int64_t e0, int64_t e1) {
  TetArgs ta;
  ta.conn = g.d_conn;
  ta.geo = g.d_geo;
  ta.slot = g.d_slot;
  ta.G = g.d_G;
  ta.idx = g.d_cidx;
  ta.ne = g.ne;
  ta.e0 = e0;
  ta.e1 = e1;
  ta.mat = g.mat;
  ta.p0 = g.p0;
  ta.p1 = g.p1;
  return ta;
}

TrussArgs truss_args(const ElemGroup& g, int64_t e0, int64_t e1) {
  TrussArgs ta;
  ta.conn = g.d_conn;
  ta.geo = g.d_geo;
  ta.slot = g.d_slot;
  ta.idx = g.d_cidx;
  ta.ne = g.ne;
  ta.e0 = e0;
  ta.e1 = e1;
  ta.strain = g.mat;
  ta.E = g.p0;
  return ta;
}

inline unsigned grid_for(int64_t n, int threads) { return (unsigned)((n + threads - 1) / threads); }

template <typename KT>
int assemble_t(Ctx* c) {
  hipStream_t s = c->stream;
  KT* K = (KT*)c->d_K;
  const int64_t n9 = 9 * c->nnzb;
  hipLaunchKernelGGL(k_fill<KT>, dim3(2048), dim3(256), 0, s, K, n9, KT(0));
  hipLaunchKernelGGL(k_fill<double>, dim3(512), dim3(256), 0, s, c->d_fint, 3 * c->nn, 0.0);
  hipLaunchKernelGGL(k_fixed_identity<KT>, dim3(grid_for(c->nn, 256)), dim3(256), 0, s, K,
                     c->d_diag, c->d_fixmask, c->nn);
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[1], s));
  for (const ElemGroup& g : c->groups) {
    const int ncol = (int)g.color_off.size() - 1;
    for (int k = 0; k < ncol; ++k) {
      const int64_t e0 = g.color_off[k], e1 = g.color_off[k + 1];
      if (e1 <= e0) continue;
      if (g.type == kTet) {
        hipLaunchKernelGGL(k_tet_scatter<KT>, dim3(grid_for(4 * (e1 - e0), kScatterThreads)),
                           dim3(kScatterThreads), 0, s, tet_args(g, e0, e1), c->d_U, c->d_fint, K,
                           c->d_fixmask);
      } else {
        hipLaunchKernelGGL(k_truss_scatter<KT>, dim3(grid_for(2 * (e1 - e0), kScatterThreads)),
                           dim3(kScatterThreads), 0, s, truss_args(g, e0, e1), c->d_U, c->d_fint,
                           K, c->d_fixmask);
      }
    }
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[2], s));
  hipLaunchKernelGGL(k_block_jacobi<KT>, dim3(grid_for(c->nn, 256)), dim3(256), 0, s, K, c->d_diag,
                     (KT*)c->d_Dinv, c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

// LDS of one staged-gather workgroup: the step table, the rows' nodal forces
// and the staged states
size_t stage_lds_bytes(const Ctx* c) {
  const int nst = c->groups[0].hasM ? kTetStateM : kTetStateBase;
  return sizeof(host::StageStep) * (size_t)c->stage_steps_cap +
         3 * sizeof(double) * (size_t)c->stage_pairs_cap +
         (size_t)c->kt * (nst + 1) * (size_t)c->stage_cap;
}

// set_attr: raise the kernel's dynamic LDS limit (setup); else launch it
template <typename KT, int NST, int KS>
int stage_run(Ctx* c, bool set_attr) {
  const size_t lds = stage_lds_bytes(c);
  auto kern = k_gather_lds<KT, NST, kStageThreads, KS>;
  if (set_attr) {
    ONSAS_HIP_TRY(c, hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    return ONSAS_OK;
  }
  const ElemGroup& g = c->groups[0];
  const unsigned gb = (unsigned)((c->nstage_groups + 7) / 8 * 8);  // a multiple of 8 (XCD mapping)
  int rlog2 = 0;
  while ((1 << rlog2) < c->stage_rows) ++rlog2;
  hipLaunchKernelGGL(kern, dim3(gb), dim3(kStageThreads), lds, c->stream, (const KT*)g.d_state,
                     g.d_fe, g.ne, c->d_stage_groups, c->nstage_groups, c->d_stage,
                     c->d_stage_steps, c->d_stage_words, c->d_stage_pairs, c->d_pair_row, rlog2,
                     c->stage_cap, c->stage_pairs_cap, c->stage_steps_cap, c->d_row_ptr,
                     c->d_colmask, c->d_diag, c->d_fixmask, c->d_sell_ptr, c->sell_rstride,
                     (KT*)c->d_Ksell, (KT*)c->d_Dinv, c->d_fint, c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

template <typename KT, int NST>
int stage_dispatch_ks(Ctx* c, bool set_attr) {
  switch (c->stage_ks) {
    case 1: return stage_run<KT, NST, 1>(c, set_attr);
    case 2: return stage_run<KT, NST, 2>(c, set_attr);
    case 3: return stage_run<KT, NST, 3>(c, set_attr);
    case 4: return stage_run<KT, NST, 4>(c, set_attr);
  }
  return set_err(c, ONSAS_ERR_UNSUPPORTED, "staged gather: entries per thread out of range");
}

template <typename KT>
int stage_dispatch(Ctx* c, bool set_attr) {
  return c->groups[0].hasM ? stage_dispatch_ks<KT, kTetStateM>(c, set_attr)
                           : stage_dispatch_ks<KT, kTetStateBase>(c, set_attr);
}

template <typename KT>
int assemble_gather_t(Ctx* c) {
  hipStream_t s = c->stream;
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[1], s));
  GatherArgs ga{};
  ga.ng = (int)c->groups.size();
  for (int gi = 0; gi < ga.ng; ++gi) {
    const ElemGroup& g = c->groups[gi];
    if (g.type == kTet)
      hipLaunchKernelGGL(k_tet_state<KT>, dim3(grid_for(g.ne, 256)), dim3(256), 0, s,
                         tet_args(g, 0, g.ne), c->d_U, c->d_x, c->d_y, c->d_z, (KT*)g.d_state,
                         g.d_fe, (int)g.hasM);
    else
      hipLaunchKernelGGL(k_truss_state<KT>, dim3(grid_for(g.ne, 256)), dim3(256), 0, s,
                         truss_args(g, 0, g.ne), c->d_U, (KT*)g.d_state, g.d_fe);
    ga.g[gi] = GatherGroup{g.d_state, g.d_fe, g.ne, g.goff, g.type, (int)g.hasM};
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[8], s));
  if (c->stage_rows > 0) {
    ONSAS_TRY(stage_dispatch<KT>(c, false));
    ONSAS_HIP_TRY(c, hipEventRecord(c->ev[2], s));
    c->bsr_stale = true;
    return ONSAS_OK;
  }
  const int64_t ngrp = (int64_t)c->sell_groups.size();
  const unsigned gb = (grid_for(ngrp, 4) + 7) / 8 * 8;  // a multiple of 8 (XCD mapping)
  bool hm = false;
  for (const ElemGroup& g : c->groups) hm = hm || g.hasM;
  const ElemGroup& g0 = c->groups[0];
  const bool one_tet = ga.ng == 1 && g0.type == kTet &&
                       (uint64_t)kTetStateM * g0.ne * sizeof(KT) < (uint64_t(1) << 32) &&
                       (uint64_t)12 * g0.ne * sizeof(double) < (uint64_t(1) << 32);
  // two contributions per step (on the 64^3 cube, gather pass in FP32 / FP64:
  // 0.217 / 0.337 ms with 2, 0.227 / 0.378 ms with 4, 0.240 ms with 6 and
  // 0.347 ms with 1; the general kernel 0.275 / 0.388 ms)
  constexpr int kNC = 2;
  if (one_tet && hm)
    hipLaunchKernelGGL((k_gather_tet<KT, true, kNC>), dim3(gb), dim3(256), 0, s,
                       (const KT*)g0.d_state, (const double*)g0.d_fe, (uint32_t)g0.ne,
                       c->d_sell_groups, ngrp, c->d_contrib, c->d_row_ptr, c->d_col, c->d_fixmask,
                       (KT*)c->d_Ksell, (KT*)c->d_Dinv, c->d_fint, c->nn);
  else if (one_tet)
    hipLaunchKernelGGL((k_gather_tet<KT, false, kNC>), dim3(gb), dim3(256), 0, s,
                       (const KT*)g0.d_state, (const double*)g0.d_fe, (uint32_t)g0.ne,
                       c->d_sell_groups, ngrp, c->d_contrib, c->d_row_ptr, c->d_col, c->d_fixmask,
                       (KT*)c->d_Ksell, (KT*)c->d_Dinv, c->d_fint, c->nn);
  else if (hm)
    hipLaunchKernelGGL((k_gather<KT, true>), dim3(gb), dim3(256), 0, s, ga, c->d_sell_groups,
                       ngrp, c->d_contrib, c->d_row_ptr, c->d_col, c->d_fixmask, (KT*)c->d_Ksell,
                       (KT*)c->d_Dinv, c->d_fint, c->nn);
  else
    hipLaunchKernelGGL((k_gather<KT, false>), dim3(gb), dim3(256), 0, s, ga, c->d_sell_groups,
                       ngrp, c->d_contrib, c->d_row_ptr, c->d_col, c->d_fixmask, (KT*)c->d_Ksell,
                       (KT*)c->d_Dinv, c->d_fint, c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  ONSAS_HIP_TRY(c, hipEventRecord(c->ev[2], s));
  c->bsr_stale = true;
  return ONSAS_OK;
}

}  // namespace

int launch_assemble(Ctx* c) {
  if (c->gather)
    return c->precision == ONSAS_PRECISION_FP64 ? assemble_gather_t<double>(c)
                                                : assemble_gather_t<float>(c);
  return c->precision == ONSAS_PRECISION_FP64 ? assemble_t<double>(c) : assemble_t<float>(c);
}

int stage_gather_prepare(Ctx* c) {
  return c->precision == ONSAS_PRECISION_FP64 ? stage_dispatch<double>(c, true)
                                              : stage_dispatch<float>(c, true);
}

size_t stage_gather_lds(const Ctx* c) { return stage_lds_bytes(c); }

int launch_gather_init(Ctx* c) {
  const unsigned grid = grid_for(c->nn, 256);
  if (c->precision == ONSAS_PRECISION_FP64)
    hipLaunchKernelGGL(k_dinv_identity<double>, dim3(grid), dim3(256), 0, c->stream, c->d_diag,
                       (double*)c->d_Dinv, c->nn);
  else
    hipLaunchKernelGGL(k_dinv_identity<float>, dim3(grid), dim3(256), 0, c->stream, c->d_diag,
                       (float*)c->d_Dinv, c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

int launch_tangent_apply_f64(Ctx* c, const double* v, double* y) {
  hipStream_t s = c->stream;
  hipLaunchKernelGGL(k_fill<double>, dim3(512), dim3(256), 0, s, y, 3 * c->nn, 0.0);
  for (const ElemGroup& g : c->groups) {
    const int ncol = (int)g.color_off.size() - 1;
    for (int k = 0; k < ncol; ++k) {
      const int64_t e0 = g.color_off[k], e1 = g.color_off[k + 1];
      if (e1 <= e0) continue;
      if (g.type == kTet)
        hipLaunchKernelGGL(k_tet_apply, dim3(grid_for(4 * (e1 - e0), kScatterThreads)),
                           dim3(kScatterThreads), 0, s, tet_args(g, e0, e1), c->d_U, v, y,
                           c->d_fixmask);
      else
        hipLaunchKernelGGL(k_truss_apply, dim3(grid_for(2 * (e1 - e0), kScatterThreads)),
                           dim3(kScatterThreads), 0, s, truss_args(g, e0, e1), c->d_U, v, y,
                           c->d_fixmask);
    }
  }
  hipLaunchKernelGGL(k_fixed_copy, dim3(grid_for(c->nn, 256)), dim3(256), 0, s, v, y, c->d_fixmask,
                     c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

int launch_locate(Ctx* c, int gi, const double* px, const double* py, const double* pz,
                  int64_t np, double tol, int64_t* elem, double* w) {
  const ElemGroup& g = c->groups[gi];
  const LocateGrid& lg = g.locate;
  hipLaunchKernelGGL(k_locate, dim3(grid_for(np, 256)), dim3(256), 0, c->stream,
                     tet_args(g, 0, g.ne), c->d_x, c->d_y, c->d_z, px, py, pz, np, lg.d_ptr,
                     lg.d_elem, make_double3(lg.lo[0], lg.lo[1], lg.lo[2]),
                     make_double3(lg.inv_h[0], lg.inv_h[1], lg.inv_h[2]),
                     make_int3(lg.nb[0], lg.nb[1], lg.nb[2]), g.d_perm, tol, elem, w);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

int launch_element_outputs(Ctx* c, int gi, const double* U, double* f, double* K, double* stress,
                           double* strain) {
  const ElemGroup& g = c->groups[gi];
  if (g.type == kTet)
    hipLaunchKernelGGL(k_tet_outputs, dim3(grid_for(g.ne, 128)), dim3(128), 0, c->stream,
                       tet_args(g, 0, g.ne), U, g.d_perm, f, K, stress, strain);
  else
    hipLaunchKernelGGL(k_truss_outputs, dim3(grid_for(g.ne, 128)), dim3(128), 0, c->stream,
                       truss_args(g, 0, g.ne), U, g.d_perm, f, K, stress, strain);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

int launch_fext(Ctx* c) {
  hipLaunchKernelGGL(k_fext, dim3(grid_for(c->nn, 256)), dim3(256), 0, c->stream, c->d_x, c->d_y,
                     c->d_z, c->d_face_conn, c->d_face_grp, c->d_nf_ptr, c->d_nf_idx,
                     c->d_pressures, c->d_nodal, c->d_fext, c->nn);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

int launch_adjoint_shear(Ctx* c, int gi, const double* lam, double* out) {
  const ElemGroup& g = c->groups[gi];
  hipLaunchKernelGGL(k_adjoint_shear, dim3(grid_for(g.ne, 128)), dim3(128), 0, c->stream,
                     tet_args(g, 0, g.ne), c->d_U, lam, g.d_perm, out);
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

}  // namespace onsas
