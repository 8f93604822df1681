// amg.hip -- smoothed-aggregation multigrid preconditioner for the PCG (an option
// beside block-Jacobi; the reference's CG is unpreconditioned).  gfx950, wave64.
//
// Hierarchy (symbolic part built once on the host from the block pattern and the
// node coordinates):
//   * greedy aggregation of the node graph: a node whose neighbours are all
//     free becomes a root with its neighbours, left-over nodes join a
//     neighbouring aggregate, isolated ones form their own;
//   * near-nullspace: the 6 rigid-body modes of each node (zero on fixed dofs),
//     orthonormalised per aggregate by modified Gram-Schmidt (twice): the
//     aggregate's rows of Q form the tentative prolongator T (bs x 6 per node),
//     R (6 x 6) is the next level's near-nullspace;
//   * smoothed prolongator P = (I - w D^-1 A) T, w = 1.15 x 4 / (3 lambda) with
//     lambda the power-iteration estimate of lambda_max(D^-1 A)
//     (block-CSR, bs x 6 blocks: row i has a block for every aggregate holding a
//     neighbour of i); next level A_c = P^T (A P);
//   * coarse levels have 6 dofs per node; coarsening goes on until at most 300
//     nodes (1800 dofs) remain, whose matrix is inverted densely (blocked
//     Gauss-Jordan) and applied as one matrix-vector product.
// Numerics (after every assembly, on the device): P, W = A P and P^T W, each
// entry an FP64 sum over a fixed-order list built by the host (deterministic,
// no atomics), 6x6 block-Jacobi inverses, the dense inverse of the coarsest
// matrix, and lambda_max(D^-1 A) per level by power iteration (15 steps the
// first time, then 3 steps warm-started from the last iterate).  A coarse dof
// whose column of T vanished (rank-deficient aggregate) gets a unit diagonal.
// V-cycle: Chebyshev smoothing of degree 2 with the block-Jacobi inverse on
// [1.1 lambda / 30, 1.1 lambda] before and after the coarse correction (the same
// polynomial, so the cycle is symmetric and a valid PCG preconditioner), the
// coarsest level solved exactly.  Each operator product carries the Chebyshev
// update of its rows in its epilogue.  Every cycle kernel returns at once when
// the PCG state says done, so replayed graphs cost nothing after convergence.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

#include <hip/hip_ext.h>

#include "engine.hpp"
#include "host_setup.hpp"

namespace onsas {

namespace {

constexpr int kMaxCoarseNodes = 300;  // coarsest level: <= 1800 dofs, inverted densely
constexpr double kChebRatio = 30.0;  // smoothing interval [hi / 30, hi], hi = 1.1 lambda
constexpr int kPowerFirst = 15;      // power-iteration steps, first estimate
constexpr int kPowerWarm = 3;        // ... re-estimates from the last iterate
// Prolongator smoothing weight w = kProlongWeight * 4 / (3 lambda).  On the
// 64^3 cube (mixed / FP64 PCG iterations of the Newton solve) the factor 0.85,
// 1.0, 1.1, 1.15, 1.2, 1.3, 1.45 gave 417/382, 379/350, 368/321, 366/321,
// 367/321, 373/330, 392/355: past the classical 4 / (3 lambda) the smoothed
// basis functions spread the rigid-body modes further, up to where the
// smoothing starts to amplify the top of the spectrum (w = 2 / lambda at 1.5).
constexpr double kProlongWeight = 1.15;
constexpr int64_t kSell6MinRows = 2048;  // smaller levels: one wave per block row
constexpr int kSell6Window = 256;     // rows sorted by length within windows of this size
constexpr int kRestrictWide = 96;     // P blocks per aggregate above which 2 waves restrict
// LDS of a sliced-ELL coarse product's workgroup for the x values of its column
// window (2,048 nodes in FP32, 1,024 in FP64)
constexpr size_t kX6Lds = 48 * 1024;

inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }
// the same rounded up to a multiple of 8 (grids that xcd_block remaps)
inline unsigned nblocks8(int64_t n, int t) { return (nblocks(n, t) + 7u) / 8u * 8u; }

// Workgroup b of a grid of nb (a multiple of 8) as the logical index that
// gives each of the 8 XCDs (hardware workgroup ids go round-robin over them)
// a contiguous run, so that neighbouring rows' vector gathers share one L2.
__device__ inline int64_t xcd_block(int64_t b, int64_t nb) { return (b & 7) * (nb >> 3) + (b >> 3); }

__device__ inline bool pcg_stopped(const PcgState* st) { return st && (st->done | st->breakdown); }

// The PCG's stop check at the start of the cycle's first kernel (256 threads),
// run by every workgroup alike: ||r||^2 from the update's np partials, summed in
// the order of the 512-wide sum the next k_pcg_A applies to them (virtual
// thread t' adds entries t', t' + 512, ... in order, then the butterfly over
// its wave, then the 8 wave sums in order; here threads t and t + 256 are one
// thread's two accumulators), and the same rule: stop once rr <= tol2 or k
// reached maxiter.  Workgroup 0 records the stop.  True: the PCG has stopped.
__device__ inline bool pcg_stop_check(const double* __restrict__ prr, int np, PcgState* st,
                                      double* sh) {
  if (st->done | st->breakdown) return true;  // the same for every thread
  const int t = threadIdx.x;
  double a = 0, b = 0;
  for (int i = t; i < np; i += 512) a += prr[i];
  for (int i = t + 256; i < np; i += 512) b += prr[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    a += __shfl_xor(a, o, 64);
    b += __shfl_xor(b, o, 64);
  }
  if ((t & 63) == 0) {
    sh[t >> 6] = a;
    sh[4 + (t >> 6)] = b;
  }
  __syncthreads();
  double rr = 0;
#pragma unroll
  for (int w = 0; w < 8; ++w) rr += sh[w];
  const int k = st->k;  // iterations completed
  const bool conv = rr <= st->tol2;
  const bool stop = conv || k >= st->maxiter;
  if (stop && blockIdx.x == 0 && t == 0) {
    st->done = 1;
    st->iters = k;
    st->converged = conv ? 1 : 0;
    st->rr_last = rr;
  }
  return stop;
}

template <typename T>
__device__ inline T pick6(const T (&a)[6], int u) {
  T r = a[0];
  r = u == 1 ? a[1] : r;
  r = u == 2 ? a[2] : r;
  r = u == 3 ? a[3] : r;
  r = u == 4 ? a[4] : r;
  r = u == 5 ? a[5] : r;
  return r;
}

template <typename KT, typename VT>
__device__ inline KT ldv(VT x) {
  return (KT)x;
}

// Value type of the sliced-ELL copy of a large coarse operator, which every
// product of that level's cycle reads: BF16 (stored as uint16, round to nearest
// even; BF16 for the FP32 range, coarse entries can exceed FP16's) in mixed
// mode, FP64 in FP64 mode (which keeps every stored value in FP64).  These
// products load 16 bytes per lane in mixed mode, so half the bytes is half the
// traffic.
template <typename KT>
struct CoarseSmootherType {
  using type = uint16_t;
};
template <>
struct CoarseSmootherType<double> {
  using type = double;
};

template <>
__device__ inline float ldv<float, uint16_t>(uint16_t h) {
  return __uint_as_float((uint32_t)h << 16);
}

template <typename ST, typename KT>
__device__ inline ST to_st(KT x) {
  return (ST)x;
}
template <>
__device__ inline uint16_t to_st<uint16_t, float>(float x) {
  const uint32_t u = __float_as_uint(x);
  if ((u & 0x7f800000u) == 0x7f800000u) return (uint16_t)((u >> 16) | ((u & 0xffffu) ? 0x40u : 0u));
  return (uint16_t)((u + 0x7fffu + ((u >> 16) & 1u)) >> 16);
}

// Value type of the smoothed prolongator P (and its P^T copy): FP16 in mixed
// mode, FP64 in FP64 mode.  P's entries are O(1) (orthonormalised modes); W,
// the coarse operators, restriction and prolongation all use the same rounded
// values, so the coarse operator stays the exact Galerkin product of the P the
// cycle applies.
template <typename KT>
struct ProlongType {
  using type = _Float16;
};
template <>
struct ProlongType<double> {
  using type = double;
};
template <typename KT>
using PT_t = typename ProlongType<KT>::type;

// Sum of a 6-vector over the 64 lanes of a wave (every lane gets the total).
template <typename KT>
__device__ inline void wave_sum6(KT (&v)[6]) {
#pragma unroll
  for (int o = 32; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 6; ++q) v[q] += __shfl_xor(v[q], o, 64);
}

// Chebyshev degree-2 coefficients for lambda_max(D^-1 A) = lam: the smoother is
// d0 = D^-1 b / theta, d1 = c1 d0 + c2 D^-1 (b - A d0), S b = d0 + d1.
struct Cheb {
  double c1, c2, inv_theta;
};
Cheb cheb(double lam) {
  const double hi = 1.1 * lam, lo = hi / kChebRatio;
  const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo);
  const double sigma = theta / delta, rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
  return {rho1 * rho0, 2.0 * rho1 / delta, 1.0 / theta};
}

// ---- numeric setup ---------------------------------------------------------------

// P block p = (i, J), one thread per column c: T_i[:, c] if J is the aggregate of
// i, minus w D_i^-1 times the sum over the operator blocks k of row i whose column
// j lies in J (list pk, row order) of A_k T_j[:, c].  FP64 sums.  Written to P
// (row order) and to PT at rt_pos[p] (the restriction's order).
template <typename KT, int BS>
__global__ __launch_bounds__(256) void k_sa_prolongator(
    const int32_t* __restrict__ p_row, const int32_t* __restrict__ p_col,
    const int32_t* __restrict__ agg, const int64_t* __restrict__ pk_ptr,
    const int32_t* __restrict__ pk, const int32_t* __restrict__ col, const KT* __restrict__ A,
    const KT* __restrict__ Dinv, const KT* __restrict__ T, double w, PT_t<KT>* __restrict__ P,
    const int32_t* __restrict__ rt_pos, PT_t<KT>* __restrict__ PT, int64_t nnz_p) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nnz_p) return;
  const int64_t p = t / 6;
  const int c = (int)(t % 6);
  const int64_t i = p_row[p];
  double s[BS];
#pragma unroll
  for (int r = 0; r < BS; ++r) s[r] = 0;
  for (int64_t e = pk_ptr[p]; e < pk_ptr[p + 1]; ++e) {
    const int64_t k = pk[e];
    if (k < 0) continue;  // own-aggregate block of a row without one in the pattern
    const KT* a = A + (int64_t)BS * BS * k;
    const KT* tj = T + (int64_t)BS * 6 * col[k] + c;
    double tc[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) tc[q] = (double)tj[6 * q];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
      for (int q = 0; q < BS; ++q) s[r] += (double)a[BS * r + q] * tc[q];
  }
  const KT* d = Dinv + (int64_t)BS * BS * i;
  const bool own = p_col[p] == agg[i];
  const KT* ti = T + (int64_t)BS * 6 * i + c;
  PT_t<KT>* o = P + (int64_t)BS * 6 * p + c;
  PT_t<KT>* ot = PT + (int64_t)BS * 6 * rt_pos[p] + c;
#pragma unroll
  for (int r = 0; r < BS; ++r) {
    double ds = 0;
#pragma unroll
    for (int m = 0; m < BS; ++m) ds += (double)d[BS * r + m] * s[m];
    const double pv = (own ? (double)ti[6 * r] : 0.0) - w * ds;
    // FP16 (mixed): rounded through FP32; FP64: stored as computed
    const PT_t<KT> v = std::is_same<PT_t<KT>, double>::value ? (PT_t<KT>)pv : (PT_t<KT>)(float)pv;
    o[6 * r] = v;
    ot[6 * r] = v;
  }
}

// W block (i, J) = (A P)_iJ, one thread per column c: sum over its (k, p) pairs
// (list order) of A_k P_p[:, c].
template <typename KT, int BS>
__global__ __launch_bounds__(256) void k_sa_ap(const int64_t* __restrict__ wl_ptr,
                                               const int32_t* __restrict__ wl_k,
                                               const int32_t* __restrict__ wl_p,
                                               const KT* __restrict__ A,
                                               const PT_t<KT>* __restrict__ P, KT* __restrict__ W,
                                               int64_t nnz_w) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nnz_w) return;
  const int64_t w = t / 6;
  const int c = (int)(t % 6);
  double s[BS];
#pragma unroll
  for (int r = 0; r < BS; ++r) s[r] = 0;
  for (int64_t e = wl_ptr[w]; e < wl_ptr[w + 1]; ++e) {
    const KT* a = A + (int64_t)BS * BS * wl_k[e];
    const PT_t<KT>* pc = P + (int64_t)BS * 6 * wl_p[e] + c;
    double pv[BS];
#pragma unroll
    for (int q = 0; q < BS; ++q) pv[q] = (double)pc[6 * q];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
      for (int q = 0; q < BS; ++q) s[r] += (double)a[BS * r + q] * pv[q];
  }
  KT* o = W + (int64_t)BS * 6 * w + c;
#pragma unroll
  for (int r = 0; r < BS; ++r) o[6 * r] = (KT)s[r];
}

// Next-level block g = (I, J) = (P^T W)_IJ, one thread per column c: sum over its
// (p, w) pairs (list order) of P_p^T W_w[:, c].
template <typename KT, int BS>
__global__ __launch_bounds__(256) void k_sa_ptap(const int64_t* __restrict__ gl_ptr,
                                                 const int32_t* __restrict__ gl_p,
                                                 const int32_t* __restrict__ gl_w,
                                                 const PT_t<KT>* __restrict__ P,
                                                 const KT* __restrict__ W,
                                                 KT* __restrict__ Ac, int64_t nnzb_c) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nnzb_c) return;
  const int64_t g = t / 6;
  const int c = (int)(t % 6);
  double s[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t e = gl_ptr[g]; e < gl_ptr[g + 1]; ++e) {
    if (gl_p[e] < 0) continue;  // diagonal block of a row without a product in it
    const PT_t<KT>* pp = P + (int64_t)BS * 6 * gl_p[e];
    const KT* wc = W + (int64_t)BS * 6 * gl_w[e] + c;
    double wv[BS];
#pragma unroll
    for (int r = 0; r < BS; ++r) wv[r] = (double)wc[6 * r];
#pragma unroll
    for (int r = 0; r < BS; ++r)
#pragma unroll
      for (int u = 0; u < 6; ++u) s[u] += (double)pp[6 * r + u] * wv[r];
  }
  KT* o = Ac + 36 * g + c;
#pragma unroll
  for (int u = 0; u < 6; ++u) o[6 * u] = (KT)s[u];
}

// Diagonal blocks: unit diagonal for unused dofs (written back into A), then the
// FP64 Gauss-Jordan inverse of the (SPD) block.
template <typename KT>
__global__ void k_inv6(KT* __restrict__ A, const int32_t* __restrict__ diag, KT* __restrict__ Dinv,
                       int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  KT* a = A + 36 * (int64_t)diag[i];
  double M[6][12];
  double amax = 0;
#pragma unroll
  for (int q = 0; q < 36; ++q) amax = fmax(amax, fabs((double)a[q]));
#pragma unroll
  for (int r = 0; r < 6; ++r) {
    if (!(fabs((double)a[r * 7]) > 1e-30 * amax) || amax == 0) {
      a[r * 7] = KT(1);  // an unused coarse dof: its row and column are zero
    }
#pragma unroll
    for (int c = 0; c < 6; ++c) {
      M[r][c] = (double)a[r * 6 + c];
      M[r][6 + c] = (r == c) ? 1.0 : 0.0;
    }
  }
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    const double inv = 1.0 / M[k][k];
#pragma unroll
    for (int c = 0; c < 12; ++c) M[k][c] *= inv;
#pragma unroll
    for (int r = 0; r < 6; ++r) {
      if (r == k) continue;
      const double f = M[r][k];
#pragma unroll
      for (int c = 0; c < 12; ++c) M[r][c] -= f * M[k][c];
    }
  }
  KT* o = Dinv + 36 * i;
#pragma unroll
  for (int r = 0; r < 6; ++r)
#pragma unroll
    for (int c = 0; c < 6; ++c) o[r * 6 + c] = (KT)M[r][6 + c];
}

// Coarsest level: block-CSR 6x6 -> dense FP64 (after k_zero_f64; unit diagonal
// for unused dofs), one thread per (node row, component).
__global__ void k_zero_f64(double* __restrict__ D, int64_t n) {
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < n;
       t += (int64_t)gridDim.x * blockDim.x)
    D[t] = 0;
}

template <typename KT>
__global__ void k_dense_scatter(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                const KT* __restrict__ A, double* __restrict__ D, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nc = 6 * n;
  if (t >= nc) return;
  const int64_t I = t / 6;
  const int r = (int)(t % 6);
  double* drow = D + t * nc;
  for (int32_t k = row_ptr[I]; k < row_ptr[I + 1]; ++k) {
    const int64_t J = col[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) drow[6 * J + c] = (double)A[36 * (int64_t)k + r * 6 + c];
  }
  if (drow[t] == 0) drow[t] = 1;
}

// Blocked Gauss-Jordan inverse of the dense SPD coarsest matrix M (n x n, FP64,
// in place; SPD: the pivots are the diagonal).  Per pivot block [k0, k0 + kb):
// A11 <- A11^-1 (k_gj_diag), C <- A21 and A12 <- A11^-1 A12 (k_gj_rows),
// then A22 <- A22 - C A12 and A21 <- -C A11^-1 (k_gj_update).
constexpr int kGJ = 32;  // pivot block

__global__ __launch_bounds__(256) void k_gj_diag(double* __restrict__ M, int64_t n, int64_t k0, int kb) {
  __shared__ double S[kGJ][kGJ + 1];
  __shared__ double colk[kGJ];
  for (int t = threadIdx.x; t < kb * kb; t += blockDim.x) S[t / kb][t % kb] = M[(k0 + t / kb) * n + k0 + t % kb];
  __syncthreads();
  for (int k = 0; k < kb; ++k) {
    if ((int)threadIdx.x < kb) colk[threadIdx.x] = S[threadIdx.x][k];
    __syncthreads();
    const double piv = 1.0 / colk[k];
    if ((int)threadIdx.x < kb) S[k][threadIdx.x] = ((int)threadIdx.x == k) ? piv : S[k][threadIdx.x] * piv;
    __syncthreads();
    for (int t = threadIdx.x; t < kb * kb; t += blockDim.x) {
      const int r = t / kb, c = t % kb;
      if (r == k) continue;
      S[r][c] = (c == k) ? -colk[r] * piv : S[r][c] - colk[r] * S[k][c];
    }
    __syncthreads();
  }
  for (int t = threadIdx.x; t < kb * kb; t += blockDim.x) M[(k0 + t / kb) * n + k0 + t % kb] = S[t / kb][t % kb];
}

// Threads [0, n kGJ): C[i][b] = M[i][k0 + b] for rows i outside the block; then
// one thread per column j outside the block: M[k0 + r][j] = (A11^-1 A12)[r][j].
__global__ __launch_bounds__(256) void k_gj_rows(double* __restrict__ M, double* __restrict__ C, int64_t n,
                                                 int64_t k0, int kb) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t nA = n * kGJ;
  if (t < nA) {
    const int64_t i = t / kGJ;
    const int b = (int)(t % kGJ);
    if (b < kb && (i < k0 || i >= k0 + kb)) C[t] = M[i * n + k0 + b];
    return;
  }
  const int64_t j = t - nA;
  if (j >= n || (j >= k0 && j < k0 + kb)) return;
  // every address below stays inside the block (b, r < kb; clamped otherwise)
  double a[kGJ];
#pragma unroll
  for (int b = 0; b < kGJ; ++b) {
    const int bb = b < kb ? b : 0;
    a[b] = b < kb ? M[(k0 + bb) * n + j] : 0.0;
  }
#pragma unroll
  for (int r = 0; r < kGJ; ++r) {
    if (r >= kb) break;
    const double* ai = M + (k0 + r) * n + k0;  // row r of A11^-1
    double s = 0;
#pragma unroll
    for (int b = 0; b < kGJ; ++b) {
      const int bb = b < kb ? b : 0;
      s += (b < kb ? ai[bb] : 0.0) * a[b];
    }
    M[(k0 + r) * n + j] = s;
  }
}

// Rows i outside the block: M[i][j] = (j in the block ? 0 : M[i][j]) -
// sum_b C[i][b] M[k0 + b][j]; 64 x 64 output tiles, 4 x 4 per thread, C and the
// block rows staged in LDS.
__global__ __launch_bounds__(256) void k_gj_update(double* __restrict__ M, const double* __restrict__ C,
                                                   int64_t n, int64_t k0, int kb) {
  __shared__ double Cs[64][kGJ + 1];
  __shared__ double Ms[kGJ][64 + 1];
  const int64_t i0 = (int64_t)blockIdx.y * 64, j0 = (int64_t)blockIdx.x * 64;
  for (int t = threadIdx.x; t < 64 * kGJ; t += 256) {  // clamped addresses, masked values
    const int r = t / kGJ, b = t % kGJ;
    const int64_t i = i0 + r;
    const bool ok = i < n && b < kb && (i < k0 || i >= k0 + kb);
    Cs[r][b] = ok ? C[(ok ? i : 0) * kGJ + b] : 0.0;
  }
  for (int t = threadIdx.x; t < kGJ * 64; t += 256) {
    const int b = t / 64, cc = t % 64;
    const int64_t j = j0 + cc;
    const bool ok = b < kb && j < n;
    Ms[b][cc] = ok ? M[(k0 + (ok ? b : 0)) * n + (ok ? j : 0)] : 0.0;
  }
  __syncthreads();
  const int ty = threadIdx.x / 16, tx = threadIdx.x % 16;
  double acc[4][4];
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int p = 0; p < 4; ++p) acc[q][p] = 0;
#pragma unroll 8
  for (int b = 0; b < kGJ; ++b) {
    double cv[4], mv[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cv[q] = Cs[4 * ty + q][b];
      mv[q] = Ms[b][4 * tx + q];
    }
#pragma unroll
    for (int q = 0; q < 4; ++q)
#pragma unroll
      for (int p = 0; p < 4; ++p) acc[q][p] += cv[q] * mv[p];
  }
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int64_t i = i0 + 4 * ty + q;
    if (i >= n || (i >= k0 && i < k0 + kb)) continue;
#pragma unroll
    for (int p = 0; p < 4; ++p) {
      const int64_t j = j0 + 4 * tx + p;
      if (j >= n) continue;
      const bool inb = j >= k0 && j < k0 + kb;
      M[i * n + j] = (inb ? 0.0 : M[i * n + j]) - acc[q][p];
    }
  }
}

// d (rows of ldc >= n, zero padded) = s (n x n)
template <typename KT>
__global__ void k_to_kt(const double* __restrict__ s, KT* __restrict__ d, int64_t n, int64_t ldc) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n * ldc) return;
  const int64_t r = t / ldc, c = t % ldc;
  d[t] = c < n ? (KT)s[r * n + c] : KT(0);
}

// ---- V-cycle kernels (all skip once the PCG has stopped) ---------------------------
//
// The operator kernels compute r = b - A v (b null: zero) and then, per row,
//   r_out = r;  d = c1 v + c2 D^-1 r -> d_out;  x_out = x_base + v + d
// (each output skipped when null, x_base null: zero).  b may be r_out and x_base
// may be x_out: every row reads them before it writes them.

// Fine level on the sliced-ELL tangent (one wave per slice); VT = KT for the
// tangent itself, the smoother type for its lower-precision copy.
template <typename KT, typename VT>
__global__ __launch_bounds__(256) void k_sell_cheb(
    const int32_t* __restrict__ sell_ptr, const int32_t* __restrict__ slice_rounds, int64_t rstride,
    const int32_t* __restrict__ scol, const VT* __restrict__ sval, const KT* __restrict__ v,
    const KT* b, KT* r_out, const KT* __restrict__ Dinv, KT c1, KT c2, KT* __restrict__ d_out,
    const KT* x_base, KT* x_out, int64_t nslices, int64_t nn, const PcgState* st,
    double* __restrict__ prz, const double* __restrict__ chk_rr, int chk_np,
    PcgState* __restrict__ chk_st) {
  const int64_t s0 = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (chk_rr) {  // the whole workgroup, before any wave leaves
    __shared__ double csh[8];
    if (pcg_stop_check(chk_rr, chk_np, chk_st, csh)) return;
  }
  // with prz every wave of the workgroup reaches its barrier: a wave past the
  // last slice reads that slice's header, runs no rounds and writes nothing
  if (s0 >= nslices && !prz) return;
  const bool live = s0 < nslices;
  const int64_t s = live ? s0 : nslices - 1;
  // the epilogue's row operands depend on nothing but the row: loaded first,
  // with the stop flag, so their latency hides under the product's
  const int64_t i = kSellRows * s + lane % kSellRows;
  const int64_t ic = i < nn ? i : nn - 1;
  KT bv[3] = {0, 0, 0}, vv3[3] = {0, 0, 0}, xb[3] = {0, 0, 0}, dm[9];
  if (b)
#pragma unroll
    for (int q = 0; q < 3; ++q) bv[q] = b[3 * ic + q];
  if (d_out || x_out) {
#pragma unroll
    for (int q = 0; q < 9; ++q) dm[q] = Dinv[9 * ic + q];
#pragma unroll
    for (int q = 0; q < 3; ++q) vv3[q] = v[3 * ic + q];
    if (x_base)
#pragma unroll
      for (int q = 0; q < 3; ++q) xb[q] = x_base[3 * ic + q];
  }
  const int32_t su = __builtin_amdgcn_readfirstlane((int32_t)s);  // wave-uniform: scalar loads
  const int32_t b0 = sell_ptr[su];
  const int32_t rounds = live ? slice_rounds[su] : 0;
  if (pcg_stopped(st)) return;  // the same for every wave of the launch
  const uint32_t rs = (uint32_t)rstride;
  const int32_t* cp = scol + b0 + lane;
  const VT* vp = sval + 9u * (uint32_t)b0 + lane;
  KT a0 = 0, a1 = 0, a2 = 0;
  for (int m = 0; m < rounds; ++m) {
    const uint32_t j = 3u * (uint32_t)cp[rs * m];
    const VT* vv = vp + 9u * rs * m;
    const KT p0 = v[j], p1 = v[j + 1], p2 = v[j + 2];
    a0 += ldv<KT>(vv[0]) * p0 + ldv<KT>(vv[64]) * p1 + ldv<KT>(vv[128]) * p2;
    a1 += ldv<KT>(vv[192]) * p0 + ldv<KT>(vv[256]) * p1 + ldv<KT>(vv[320]) * p2;
    a2 += ldv<KT>(vv[384]) * p0 + ldv<KT>(vv[448]) * p1 + ldv<KT>(vv[512]) * p2;
  }
#pragma unroll
  for (int o = kSellRows; o < 64; o <<= 1) {
    a0 += __shfl_xor(a0, o, 64);
    a1 += __shfl_xor(a1, o, 64);
    a2 += __shfl_xor(a2, o, 64);
  }
  if (prz) {
    // with x_out: also b.x_out over the workgroup's rows (the PCG's r.z when b
    // is its residual), one partial per workgroup, the 4 wave sums in order
    __shared__ double wz[4];
    double bz = 0;
    if (live && lane < kSellRows && i < nn) {
      KT r[3] = {bv[0] - a0, bv[1] - a1, bv[2] - a2};
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        const KT vq = vv3[q];
        const KT d = c1 * vq + c2 * (dm[3 * q] * r[0] + dm[3 * q + 1] * r[1] + dm[3 * q + 2] * r[2]);
        const KT xo = xb[q] + vq + d;
        x_out[3 * i + q] = xo;
        bz += (double)bv[q] * (double)xo;
      }
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) bz += __shfl_xor(bz, o, 64);
    if (lane == 0) wz[threadIdx.x >> 6] = bz;
    __syncthreads();
    if (threadIdx.x == 0) prz[blockIdx.x] = ((wz[0] + wz[1]) + wz[2]) + wz[3];
    return;
  }
  if (lane >= kSellRows || i >= nn) return;
  KT r[3] = {-a0, -a1, -a2};
  if (b)
#pragma unroll
    for (int q = 0; q < 3; ++q) r[q] += bv[q];
  if (r_out)
#pragma unroll
    for (int q = 0; q < 3; ++q) r_out[3 * i + q] = r[q];
  if (!d_out && !x_out) return;
#pragma unroll
  for (int q = 0; q < 3; ++q) {
    const KT vq = vv3[q];
    const KT d = c1 * vq + c2 * (dm[3 * q] * r[0] + dm[3 * q + 1] * r[1] + dm[3 * q + 2] * r[2]);
    if (d_out) d_out[3 * i + q] = d;
    if (x_out) x_out[3 * i + q] = xb[q] + vq + d;
  }
}

// Small levels (6x6 block-CSR): one wave per block row, lanes over its blocks,
// lanes 0..5 finish one component each.
template <typename KT>
__global__ __launch_bounds__(256) void k_bsr6_cheb(
    const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col, const KT* __restrict__ A,
    const KT* __restrict__ v, const KT* b, KT* r_out, const KT* __restrict__ Dinv, KT c1, KT c2,
    KT* __restrict__ d_out, const KT* x_base, KT* x_out, int64_t n, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t I = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int l = threadIdx.x & 63;
  if (I >= n) return;
  KT acc[6] = {0, 0, 0, 0, 0, 0};
  for (int32_t k = row_ptr[I] + l; k < row_ptr[I + 1]; k += 64) {
    const KT* a = A + 36 * (int64_t)k;
    const KT* xj = v + 6 * (int64_t)col[k];
    KT xv[6];
#pragma unroll
    for (int c = 0; c < 6; ++c) xv[c] = xj[c];
#pragma unroll
    for (int u = 0; u < 6; ++u)
#pragma unroll
      for (int c = 0; c < 6; ++c) acc[u] += a[6 * u + c] * xv[c];
  }
  wave_sum6(acc);
  if (l >= 6) return;
  KT r[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) r[m] = (b ? b[6 * I + m] : KT(0)) - acc[m];
  const KT ru = pick6(r, l);
  if (r_out) r_out[6 * I + l] = ru;
  if (!d_out && !x_out) return;
  const KT* dm = Dinv + 36 * I + 6 * l;
  KT dr = 0;
#pragma unroll
  for (int m = 0; m < 6; ++m) dr += dm[m] * r[m];
  const KT vu = v[6 * I + l];
  const KT d = c1 * vu + c2 * dr;
  if (d_out) d_out[6 * I + l] = d;
  if (x_out) x_out[6 * I + l] = (x_base ? x_base[6 * I + l] : KT(0)) + vu + d;
}

// Value layout of one round of a sliced-ELL level (36 x 64 values: 36 per lane):
// 16-byte chunks of kVPC = 16 / sizeof(VT) values, lane-interleaved (chunk ch
// of lane l at (64 ch + l) kVPC), kFull whole chunks, then the kRem values left
// (BF16: 36 = 4 x 8 + 4) as one narrower chunk at 64 kFull kVPC + kRem l.  Each
// value load of a wave then reads 64 x 16 contiguous bytes (the last one of a
// BF16 round 64 x 8).
template <typename VT>
struct Sell6Layout {
  static constexpr int kVPC = 16 / (int)sizeof(VT);
  static constexpr int kFull = 36 / kVPC;
  static constexpr int kRem = 36 - kFull * kVPC;
  static constexpr int64_t kRound = 36 * 64;
  __host__ __device__ static int64_t offset(int e, int l) {
    return e < kFull * kVPC ? (int64_t)((e / kVPC) * 64 + l) * kVPC + e % kVPC
                            : (int64_t)kFull * kVPC * 64 + (int64_t)l * kRem + (e - kFull * kVPC);
  }
};

// Large levels >= 1 on their sliced-ELL copy: one wave per slice of RPW rows,
// 64 / RPW lanes per row (lane l serves row l % RPW), WV slices per
// workgroup.  The workgroup first copies the x values of its column window
// (host-computed: lo, span <= kX6Lds bytes) into LDS with coalesced 16-byte
// loads, issued with the first pass's column indices, values and row ids, so a
// lane's gathers of x come from LDS (a column outside the window, which only
// an irregular level has, is read from global memory).  A lane issues the
// loads of UR rounds at once, with the round index clamped to the slice's last
// round; rounds past the end multiply by exact zeros, so every sum runs over
// the rounds and the 36 values in order as a one-round loop would.  The row
// operands of the epilogue (b, D^-1, v, x_base) are loaded before the barrier
// and used after the products.  After the reduction over the lanes of a row,
// lanes l / RPW < 6 finish that component of row s6_row[RPW s + l % RPW]
// (rows sorted by length within windows; -1: none).
// WV waves (slices) per workgroup: 4, 8 or 16 with the LDS window (FP32 x), 4
// without it (FP64 x: its 36 values a round need the registers that the
// window's copy would take).
template <typename KT>
constexpr bool sell6_window() { return sizeof(KT) == 4; }
template <typename KT, typename VT, int RPW, int UR>
__device__ inline void sell6_load_pass(const int32_t* __restrict__ s6_col, const VT* __restrict__ A6,
                                       int32_t qb, int32_t qlast, int l, int32_t (&cj)[UR],
                                       VT (&av)[UR][36]) {
  using LY = Sell6Layout<VT>;
#pragma unroll
  for (int k = 0; k < UR; ++k) cj[k] = s6_col[64 * (int64_t)min(qb + k, qlast) + l];
#pragma unroll
  for (int k = 0; k < UR; ++k) {  // unconditional loads (clamped rounds)
    const VT* a = A6 + LY::kRound * (int64_t)min(qb + k, qlast);
#pragma unroll
    for (int ch = 0; ch < LY::kFull; ++ch) {
      const VT* ap = (const VT*)__builtin_assume_aligned(a + LY::offset(ch * LY::kVPC, l), 16);
#pragma unroll
      for (int j = 0; j < LY::kVPC; ++j) av[k][ch * LY::kVPC + j] = ap[j];
    }
    if constexpr (LY::kRem > 0) {
      const VT* ap = (const VT*)__builtin_assume_aligned(a + LY::offset(LY::kFull * LY::kVPC, l),
                                                         LY::kRem * sizeof(VT));
#pragma unroll
      for (int j = 0; j < LY::kRem; ++j) av[k][LY::kFull * LY::kVPC + j] = ap[j];
    }
  }
}

// a clamped round past the slice's end multiplies by exact zeros: the sum is
// unchanged, and no branch lets the loads sink below a test
template <typename KT, typename VT, int UR>
__device__ inline void sell6_accumulate(const VT (&av)[UR][36], const KT (&xv)[UR][6], int32_t qb,
                                        int32_t q1, KT (&acc)[6]) {
#pragma unroll
  for (int k = 0; k < UR; ++k) {
    const bool ok = qb + k < q1;
#pragma unroll
    for (int e = 0; e < 36; ++e) acc[e / 6] += ldv<KT>(av[k][e]) * (ok ? xv[k][e % 6] : KT(0));
  }
}

template <typename KT, typename VT, int RPW, int UR, int WV>
__global__ __launch_bounds__(64 * WV) void k_sell6_cheb(
    const int32_t* __restrict__ s6_ptr, const int32_t* __restrict__ s6_col,
    const int32_t* __restrict__ s6_row, const int32_t* __restrict__ s6_win,
    const VT* __restrict__ A6, const KT* __restrict__ v, const KT* b, KT* r_out,
    const KT* __restrict__ Dinv, KT c1, KT c2, KT* __restrict__ d_out, const KT* x_base,
    KT* x_out, int64_t nslices, int64_t n, const PcgState* st) {
  static_assert(RPW == 2 || RPW == 4 || RPW == 8, "6 components need >= 8 lanes per row");
  static_assert(UR >= 1 && UR <= 4, "1 to 4 rounds in flight (registers: 36 values each)");
  constexpr bool kWin = sell6_window<KT>();
  constexpr int kThreads = 64 * WV;
  constexpr int kWinUnits = kWin ? (int)(kX6Lds / 16) : 1;  // 16-byte units of the window
  constexpr int kPerThread = (kWinUnits + kThreads - 1) / kThreads;
  __shared__ __align__(16) uint4 xs4[kWinUnits];
  const KT* xs = (const KT*)xs4;
  const int64_t s0 = (int64_t)blockIdx.x * WV + (threadIdx.x >> 6);
  const int l = threadIdx.x & 63;
  // every wave of the workgroup reaches the barrier: a wave past the last slice
  // reads that slice's header, runs its rounds as zeros and writes nothing
  const bool live = s0 < nslices;
  const int64_t s = live ? s0 : nslices - 1;
  const int32_t su = __builtin_amdgcn_readfirstlane((int32_t)s);  // wave-uniform: scalar loads
  const int32_t q0 = s6_ptr[su], q1 = live ? s6_ptr[su + 1] : q0;
  const int32_t lo = kWin ? s6_win[2 * blockIdx.x] : 0, span = kWin ? s6_win[2 * blockIdx.x + 1] : 0;
  if (pcg_stopped(st)) return;  // the same for every wave of the launch
  // rounds are clamped to qlast (a slice has at least one round; an empty one
  // reads a neighbouring round and adds exact zeros)
  const int32_t qlast = max(max(q1, q0 + 1) - 1, 0);
  const int u = l / RPW;
  const int uc = u < 6 ? u : 0;
  // optional operands as always-valid pointers (selected on the kernel
  // arguments), so the first pass has no branch that would wait at a join
  const KT* bp = b ? b : v;
  const KT* xbp = x_base ? x_base : v;
  KT acc[6] = {0, 0, 0, 0, 0, 0};
  // first pass: row id, columns and values, then the window (all independent)
  const int64_t I = s6_row[RPW * s + (l & (RPW - 1))];
  int32_t cj[UR];
  VT av[UR][36];
  sell6_load_pass<KT, VT, RPW, UR>(s6_col, A6, q0, qlast, l, cj, av);
  const int nunits = (int)(((int64_t)span * 6 * (int64_t)sizeof(KT)) / 16);
  const uint4* wsrc = (const uint4*)(v + 6 * (int64_t)lo);  // 16-byte aligned: lo even
  uint4 wv[kPerThread];
  // the window's last 8 bytes when its length is not a multiple of 16 bytes
  // (a span that ends at the last node)
  const int tail8 = (int)((((int64_t)span * 6 * (int64_t)sizeof(KT)) % 16) / 8);
  uint2 wt = make_uint2(0, 0);
  if constexpr (kWin) {
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      const int t = (int)threadIdx.x + kThreads * j;
      wv[j] = wsrc[t < nunits ? t : 0];
    }
    if (tail8 && threadIdx.x == 0) wt = ((const uint2*)wsrc)[2 * nunits];
  }
  const int64_t Ic = I >= 0 ? I : 0;
  KT bv[6], dm[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) bv[m] = bp[6 * Ic + m];
#pragma unroll
  for (int m = 0; m < 6; ++m) dm[m] = Dinv[36 * Ic + 6 * uc + m];
  const KT vu = v[6 * Ic + uc];
  const KT xb0 = xbp[6 * Ic + uc];
  if constexpr (kWin) {
#pragma unroll
    for (int j = 0; j < kPerThread; ++j) {
      const int t = (int)threadIdx.x + kThreads * j;
      if (t < nunits) xs4[t] = wv[j];
    }
    if (tail8 && threadIdx.x == 0) ((uint2*)xs4)[2 * nunits] = wt;
    __syncthreads();
  }
  KT xv[UR][6];
#pragma unroll
  for (int k = 0; k < UR; ++k) {
    const uint32_t cw = (uint32_t)(cj[k] - lo);
    const KT* xj = cw < (uint32_t)span ? xs + 6 * cw : v + 6 * (int64_t)cj[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) xv[k][c] = xj[c];
  }
  sell6_accumulate<KT, VT, UR>(av, xv, q0, q1, acc);
  for (int32_t qb = q0 + UR; qb < q1; qb += UR) {
    sell6_load_pass<KT, VT, RPW, UR>(s6_col, A6, qb, qlast, l, cj, av);
#pragma unroll
    for (int k = 0; k < UR; ++k) {
      const uint32_t cw = (uint32_t)(cj[k] - lo);
      const KT* xj = cw < (uint32_t)span ? xs + 6 * cw : v + 6 * (int64_t)cj[k];
#pragma unroll
      for (int c = 0; c < 6; ++c) xv[k][c] = xj[c];
    }
    sell6_accumulate<KT, VT, UR>(av, xv, qb, q1, acc);
  }
  if (!b)
#pragma unroll
    for (int m = 0; m < 6; ++m) bv[m] = KT(0);
  const KT xb = x_base ? xb0 : KT(0);
#pragma unroll
  for (int o = RPW; o < 64; o <<= 1)
#pragma unroll
    for (int w = 0; w < 6; ++w) acc[w] += __shfl_xor(acc[w], o, 64);
  if (!live || u >= 6 || I < 0) return;
  KT r[6];
#pragma unroll
  for (int m = 0; m < 6; ++m) r[m] = bv[m] - acc[m];
  const KT ru = pick6(r, u);
  if (r_out) r_out[6 * I + u] = ru;
  if (!d_out && !x_out) return;
  KT dr = 0;
#pragma unroll
  for (int m = 0; m < 6; ++m) dr += dm[m] * r[m];
  const KT d = c1 * vu + c2 * dr;
  if (d_out) d_out[6 * I + u] = d;
  if (x_out) x_out[6 * I + u] = xb + vu + d;
}

// Block-CSR -> sliced-ELL values (Sell6Layout): one thread per (round slot, value).
template <typename KT>
__global__ void k_to_sell6(const int32_t* __restrict__ s6_src, const KT* __restrict__ A,
                           typename CoarseSmootherType<KT>::type* __restrict__ A6s, int64_t nslots) {
  using ST = typename CoarseSmootherType<KT>::type;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 36 * nslots) return;
  const int64_t slot = t / 36;  // 64 q + l
  const int e = (int)(t % 36);
  const int64_t q = slot >> 6;
  const int l = (int)(slot & 63);
  const int32_t k = s6_src[slot];
  const KT v = k >= 0 ? A[36 * (int64_t)k + e] : KT(0);
  A6s[Sell6Layout<ST>::kRound * q + Sell6Layout<ST>::offset(e, l)] = to_st<ST>(v);
}

// y = s D^-1 v, one thread per (node, component).
template <typename KT, int BS>
__global__ void k_jacobi(const KT* __restrict__ Dinv, const KT* __restrict__ v, KT* __restrict__ y,
                         int64_t n, KT s, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BS * n) return;
  const int64_t i = t / BS;
  const int u = (int)(t % BS);
  const KT* d = Dinv + (int64_t)BS * BS * i + BS * u;
  const KT* vi = v + (int64_t)BS * i;
  KT acc = 0;
#pragma unroll
  for (int c = 0; c < BS; ++c) acc += d[c] * vi[c];
  y[t] = s * acc;
}

// b_c[I] = sum over the P blocks of column I (entries rt_ptr[I].., row order; PT
// holds them contiguously) of P_e^T r_row(e), LPA lanes per aggregate over its
// entries: 32 (two aggregates per wave, half the waves: the fine level, about
// a hundred entries each), 64, 128 (two waves an aggregate, two aggregates a
// workgroup: the levels with a few hundred entries an aggregate; on the 64^3
// cube's level 1 14.3 us against 17.4 us with 256) or 256 (four waves); the
// waves of an aggregate add their sums in wave order through LDS.  With d0
// non-null also d0[I] = s D_c^-1
// b_c[I] (the first Chebyshev step of the coarse level).  An FP32 lane takes
// two of its entries per pass (LPA apart) with clamped, unconditional index
// loads, so both row gathers are in flight together (FP64: one, for the
// registers); a slot past the end adds an exact zero, so the lane's sum keeps
// the order l, l + LPA, ... of a one-entry loop.
template <typename KT, int BS, int LPA>
__global__ __launch_bounds__(256) void k_sa_restrict(
    const int32_t* __restrict__ rt_ptr, const int32_t* __restrict__ rt_row,
    const PT_t<KT>* __restrict__ PT, const KT* __restrict__ rf, KT* __restrict__ bc,
    const KT* __restrict__ Dinv_c, KT s, KT* __restrict__ d0, int64_t nc, const PcgState* st) {
  static_assert(LPA == 32 || LPA == 64 || LPA == 128 || LPA == 256,
                "32, 64, 128 or 256 lanes per aggregate");
  using PV = PT_t<KT>;
  constexpr int kU = sizeof(KT) == 4 ? 2 : 1;
  __shared__ KT sh[4][6];
  // LPA < 256: an XCD-aware run of aggregates (the grid is a multiple of 8)
  const int64_t bl = LPA < 256 ? xcd_block(blockIdx.x, gridDim.x) : (int64_t)blockIdx.x;
  const int64_t Iw = (bl * 256 + threadIdx.x) / LPA;
  const int l = threadIdx.x % LPA;
  if (LPA < 128 && Iw >= nc) return;  // no barrier below for these
  // LPA 128: both aggregates of the workgroup reach the barrier (one past the
  // end reads the last aggregate and writes nothing); LPA 256: Iw < nc
  const bool live = Iw < nc;
  const int64_t I = live ? Iw : nc - 1;
  // the coarse row's D^-1 first (its latency hides under the gathers); an
  // aggregate index uniform over the wave makes its bounds scalar loads
  KT dm[6];
  if (d0)
#pragma unroll
    for (int m = 0; m < 6; ++m) dm[m] = Dinv_c[36 * I + 6 * (l < 6 ? l : 0) + m];
  int32_t e0, e1;
  if (LPA >= 64) {
    const int32_t Iu = __builtin_amdgcn_readfirstlane((int32_t)I);
    e0 = rt_ptr[Iu];
    e1 = rt_ptr[Iu + 1];
  } else {
    e0 = rt_ptr[I];
    e1 = rt_ptr[I + 1];
  }
  if (pcg_stopped(st)) return;
  KT acc[6] = {0, 0, 0, 0, 0, 0};
  for (int32_t e = e0 + l; e < e1; e += kU * LPA) {
    int32_t ek[kU];
    int64_t ik[kU];
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      ek[k] = min(e + k * LPA, e1 - 1);
      ik[k] = rt_row[ek[k]];
    }
    PV pv[kU][BS * 6];
    KT rv[kU][BS];
#pragma unroll
    for (int k = 0; k < kU; ++k) {  // unconditional loads (clamped addresses)
      const PV* pp = (const PV*)__builtin_assume_aligned(PT + (int64_t)BS * 6 * ek[k], 2 * sizeof(PV));
#pragma unroll
      for (int q = 0; q < BS * 6; ++q) pv[k][q] = pp[q];
#pragma unroll
      for (int a = 0; a < BS; ++a) rv[k][a] = rf[BS * ik[k] + a];
    }
#pragma unroll
    for (int k = 0; k < kU; ++k) {
      const bool ok = e + k * LPA < e1;
#pragma unroll
      for (int a = 0; a < BS; ++a) {
        const KT ra = ok ? rv[k][a] : KT(0);
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[u] += (KT)pv[k][6 * a + u] * ra;
      }
    }
  }
  // fixed tree over the aggregate's lanes (within its half of the wave for 32)
#pragma unroll
  for (int o = (LPA < 64 ? LPA : 64) / 2; o > 0; o >>= 1)
#pragma unroll
    for (int q = 0; q < 6; ++q) acc[q] += __shfl_xor(acc[q], o, 64);
  if (LPA >= 128) {  // the aggregate's waves add their sums in wave order
    if ((threadIdx.x & 63) < 6) sh[threadIdx.x >> 6][threadIdx.x & 63] = pick6(acc, threadIdx.x & 63);
    __syncthreads();
    if (l < 6) {
      const int w0 = (int)(threadIdx.x >> 6) & ~(LPA / 64 - 1);  // the aggregate's first wave
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        KT a = sh[w0][u];
#pragma unroll
        for (int w = 1; w < LPA / 64; ++w) a += sh[w0 + w][u];
        acc[u] = a;
      }
    }
  }
  if (l >= 6 || !live) return;
  bc[6 * I + l] = pick6(acc, l);
  if (d0) {
    KT dr = 0;
#pragma unroll
    for (int m = 0; m < 6; ++m) dr += dm[m] * acc[m];
    d0[6 * I + l] = s * dr;
  }
}

// x_f[i] += sum over the P blocks p of row i of P_p x_c[col(p)], one thread per
// (node, dof).  Eight blocks per pass (FP64: four, for the registers) with
// clamped, unconditional index loads: the column indices, then all the pass's
// coarse gathers, are in flight together (a row has about six blocks); a slot
// past the row's end adds an exact zero, so the sum runs in block order as in a
// one-block loop.
template <typename KT, int BS>
__global__ void k_sa_prolong(const int32_t* __restrict__ p_ptr, const int32_t* __restrict__ p_col,
                             const PT_t<KT>* __restrict__ P, const KT* __restrict__ xc,
                             KT* __restrict__ xf, int64_t nf, const PcgState* st) {
  using PV = PT_t<KT>;
  constexpr int kPass = sizeof(KT) == 4 ? 8 : 2;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BS * nf) return;
  const int64_t i = t / BS;
  const int a = (int)(t % BS);
  const KT x0 = xf[t];  // loaded with the row's bounds, ahead of the gathers
  const int32_t p0 = p_ptr[i], p1 = p_ptr[i + 1];
  if (pcg_stopped(st)) return;
  KT acc = 0;
  for (int32_t p = p0; p < p1; p += kPass) {
    int32_t cj[kPass];
#pragma unroll
    for (int k = 0; k < kPass; ++k) cj[k] = p_col[min(p + k, p1 - 1)];
    PV pv[kPass][6];
    KT xv[kPass][6];
#pragma unroll
    for (int k = 0; k < kPass; ++k) {  // unconditional loads (clamped addresses)
      const PV* pp = (const PV*)__builtin_assume_aligned(
          P + (int64_t)BS * 6 * min(p + k, p1 - 1) + 6 * a, 2 * sizeof(PV));
      const KT* xj = (const KT*)__builtin_assume_aligned(xc + 6 * (int64_t)cj[k], 8);
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        pv[k][u] = pp[u];
        xv[k][u] = xj[u];
      }
    }
    __builtin_amdgcn_sched_barrier(0);  // every gather issued before the first product
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const bool ok = p + k < p1;
#pragma unroll
      for (int u = 0; u < 6; ++u) acc += (KT)pv[k][u] * (ok ? xv[k][u] : KT(0));
    }
  }
  xf[t] = x0 + acc;
}

// The same product with two threads per node (all BS dofs): each coarse
// block of x is gathered once for the node's BS rows.  The pair's even thread
// takes the row's blocks 0, 2, 4, ..., the odd one 1, 3, 5, ..., kPass a pass
// each with clamped, unconditional index loads (a slot past the row's end adds
// an exact zero), and the even thread adds the two sums in that order and
// writes the node's values.  One pass covers 2 kPass blocks (a row has about
// seven), so a row's chain of dependent loads is one pass long (on the 64^3
// cube 12.5 us against 16.0 us with one thread and two passes per node).
template <typename KT, int BS, int kPass>
__global__ __launch_bounds__(256) void k_sa_prolong_pair(
    const int32_t* __restrict__ p_ptr, const int32_t* __restrict__ p_col,
    const PT_t<KT>* __restrict__ P, const KT* __restrict__ xc, KT* __restrict__ xf, int64_t nf,
    const PcgState* st) {
  using PV = PT_t<KT>;
  // an XCD-aware run of nodes (the grid is a multiple of 8)
  const int64_t t = xcd_block(blockIdx.x, gridDim.x) * blockDim.x + threadIdx.x;
  const int64_t i = t >> 1;
  const int h = (int)(t & 1);
  // both threads of a pair are in one wave and take the same exits
  if (i >= nf) return;
  KT x0[BS];
#pragma unroll
  for (int a = 0; a < BS; ++a) x0[a] = xf[BS * i + a];
  const int32_t p0 = p_ptr[i], p1 = p_ptr[i + 1];
  if (pcg_stopped(st)) return;
  KT acc[BS];
#pragma unroll
  for (int a = 0; a < BS; ++a) acc[a] = KT(0);
  for (int32_t p = p0 + h; p < p1; p += 2 * kPass) {
    int32_t cj[kPass];
#pragma unroll
    for (int k = 0; k < kPass; ++k) cj[k] = p_col[min(p + 2 * k, p1 - 1)];
    PV pv[kPass][BS * 6];
    KT xv[kPass][6];
#pragma unroll
    for (int k = 0; k < kPass; ++k) {  // unconditional loads (clamped addresses)
      const PV* pp = (const PV*)__builtin_assume_aligned(P + (int64_t)BS * 6 * min(p + 2 * k, p1 - 1),
                                                         2 * sizeof(PV));
      const KT* xj = (const KT*)__builtin_assume_aligned(xc + 6 * (int64_t)cj[k], 8);
#pragma unroll
      for (int q = 0; q < BS * 6; ++q) pv[k][q] = pp[q];
#pragma unroll
      for (int u = 0; u < 6; ++u) xv[k][u] = xj[u];
    }
    __builtin_amdgcn_sched_barrier(0);  // every gather issued before the first product
#pragma unroll
    for (int k = 0; k < kPass; ++k) {
      const bool ok = p + 2 * k < p1;
#pragma unroll
      for (int a = 0; a < BS; ++a)
#pragma unroll
        for (int u = 0; u < 6; ++u) acc[a] += (KT)pv[k][6 * a + u] * (ok ? xv[k][u] : KT(0));
    }
  }
#pragma unroll
  for (int a = 0; a < BS; ++a) {
    const KT other = __shfl_xor(acc[a], 1, 64);
    acc[a] = h == 0 ? acc[a] + other : acc[a];
  }
  if (h == 0)
#pragma unroll
    for (int a = 0; a < BS; ++a) xf[BS * i + a] = x0[a] + acc[a];
}

// x = Ainv b on the coarsest level (dense): one wave per row, lanes over its
// columns, fixed reduction tree.
template <typename KT>
__global__ __launch_bounds__(256) void k_dense_apply(const KT* __restrict__ Ainv,
                                                     const KT* __restrict__ b, KT* __restrict__ x,
                                                     int64_t nc, int64_t ldc, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t r = ((int64_t)blockIdx.x * 256 + threadIdx.x) >> 6;
  const int l = threadIdx.x & 63;
  if (r >= nc) return;
  // rows padded to ldc (a multiple of 4): 4 consecutive values per lane and load
  const KT* a = Ainv + r * ldc;
  KT s = 0;
  for (int64_t c = 4 * l; c < nc; c += 256) {
    const KT* ap = (const KT*)__builtin_assume_aligned(a + c, 4 * sizeof(KT));
#pragma unroll
    for (int j = 0; j < 4; ++j) s += ap[j] * b[c + j < nc ? c + j : nc - 1];  // padding: ap[j] = 0
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (l == 0) x[r] = s;
}

// partial sums of v.v (power iteration), grid-stride, one partial per block
template <typename KT>
__global__ __launch_bounds__(256) void k_norm2(const KT* __restrict__ v, int64_t n,
                                               double* __restrict__ part) {
  __shared__ double sh[4];
  double s = 0;
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    s += (double)v[i] * (double)v[i];
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) part[blockIdx.x] = sh[0] + sh[1] + sh[2] + sh[3];
}

// v = a w
template <typename KT>
__global__ void k_scaled_copy(const KT* __restrict__ w, KT* __restrict__ v, int64_t n, KT a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] = a * w[i];
}

template <typename KT>
__global__ void k_fill_hash(KT* __restrict__ v, int64_t n, const uint8_t* __restrict__ mask, int bs) {
  // deterministic pseudo-random start vector for the power iteration (fixed dofs 0)
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  uint64_t z = (uint64_t)i * 0x9E3779B97F4A7C15ull + 0x632BE59BD9B4E019ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  z ^= z >> 31;
  KT val = (KT)((double)(z >> 11) * (1.0 / 9007199254740992.0) - 0.5);
  if (mask && bs == 3 && ((mask[i / 3] >> (i % 3)) & 1u)) val = KT(0);
  v[i] = val;
}

// ---- host helpers ------------------------------------------------------------------

template <typename T>
int upload(Ctx* c, T** d, const std::vector<T>& h) {
  ONSAS_HIP_TRY(c, hipMalloc((void**)d, std::max<size_t>(h.size(), 1) * sizeof(T)));
  if (!h.empty())
    ONSAS_HIP_TRY(c, hipMemcpy(*d, h.data(), h.size() * sizeof(T), hipMemcpyHostToDevice));
  return ONSAS_OK;
}

int alloc(Ctx* c, void** d, size_t bytes) {
  ONSAS_HIP_TRY(c, hipMalloc(d, std::max<size_t>(bytes, 16)));
  ONSAS_HIP_TRY(c, hipMemset(*d, 0, std::max<size_t>(bytes, 16)));
  return ONSAS_OK;
}

template <typename KT>
int upload_kt(Ctx* c, void** d, const std::vector<double>& h) {
  std::vector<KT> v(h.begin(), h.end());
  return upload(c, (KT**)d, v);
}

// Rows per sliced-ELL slice of a coarse level.  Fewer rows per wave split a
// row over more lanes (fewer rounds per lane) at the price of more padding;
// these launches are bound by latency, not bytes.  Measured on the 64^3 cube's
// level 1 (14,182 rows, 39.7 blocks a row), batch-timed products: 8 rows per
// wave 15.7 us (BF16, LDS window) and 33.1 us (FP64), 4 rows 21.3 / 34.3 us,
// 2 rows 25.3 / 37.3 us.  ONSAS_SELL6_RPW (8, 4 or 2) overrides.
int sell6_rows_per_wave() {
  if (const char* e = std::getenv("ONSAS_SELL6_RPW")) {
    const int r = std::atoi(e);
    if (r == 2 || r == 4 || r == 8) return r;
  }
  return 8;
}

// Rounds of a sliced-ELL coarse level whose loads a lane issues together
// (k_sell6_cheb): 2 with BF16 values, 1 with FP64 values (36 doubles a round);
// measured best at 8 rows per wave (BF16, 8 waves: 13.6 us at 2 rounds, 15.9 us
// at 1, 13.9 us at 4).  ONSAS_SELL6_ROUNDS (1, 2 or 4) overrides.
int sell6_rounds_in_flight(bool f64) {
  if (const char* e = std::getenv("ONSAS_SELL6_ROUNDS")) {
    const int r = std::atoi(e);
    if (r == 1 || r == 2 || r == 4) return r;
  }
  return f64 ? 1 : 2;
}

// Waves (slices) per workgroup of a sliced-ELL coarse level's product: each
// FP32 workgroup copies its window of x into LDS, so larger groups share one
// copy; smaller groups spread a small level over more CUs.  Level 1 of the
// 64^3 cube (1,773 slices of 8 rows), batch-timed, 8 rows per wave, 2 rounds:
// 16 waves 15.8 us (111 workgroups: fewer than the 256 CUs), 8 waves 13.6 us,
// 4 waves 14.1 us (a smaller window, more gathers from global memory).  FP64:
// 4 (no window).  ONSAS_SELL6_WAVES (4, 8 or 16) overrides in FP32.
int sell6_waves_per_group(bool f64) {
  if (f64) return 4;
  if (const char* e = std::getenv("ONSAS_SELL6_WAVES")) {
    const int w = std::atoi(e);
    if (w == 4 || w == 8 || w == 16) return w;
  }
  return 8;
}

}  // namespace

// Build the hierarchy's symbolic part (host) and allocate its device arrays.
int amg_build(Ctx* c) {
  Amg& A = c->amg;
  const int64_t gen = A.gen;  // monotonic across rebuilds (graph cache key)
  A = Amg{};
  A.gen = gen + 1;
  {
    // degree 1 on the finest level: a damped-Jacobi step (w = 1/theta) before
    // and after the coarse correction, two tangent products per cycle instead
    // of four, for ~4 % more PCG iterations at 64^3 (12 % less solve time);
    // ONSAS_AMG_FINE_DEGREE=2 selects the degree-2 Chebyshev smoother there
    const char* e = std::getenv("ONSAS_AMG_FINE_DEGREE");
    A.fine_degree = (e && std::atoi(e) == 2) ? 2 : 1;
    // the coarser levels' smoothing: degree 2 (four products of a level per
    // cycle) or, with ONSAS_AMG_COARSE_DEGREE=1, the damped-Jacobi step (two)
    const char* ec = std::getenv("ONSAS_AMG_COARSE_DEGREE");
    A.coarse_degree = (ec && std::atoi(ec) == 1) ? 1 : 2;
  }
  const int64_t nn = c->nn;
  const bool f64 = c->precision == ONSAS_PRECISION_FP64;
  AmgLevel L0;
  L0.bs = 3;
  L0.n = nn;
  L0.nnzb = (int64_t)c->col.size();
  A.lev.push_back(L0);
  // level-0 near-nullspace: translations and rotations
  std::vector<double> Bcur(nn * 3 * 6, 0.0);
  for (int64_t i = 0; i < nn; ++i)
    for (int d = 0; d < 3; ++d) Bcur[i * 18 + d * 6 + d] = 1.0;
  // rigid-body rotations about ONE point (the node centroid) for the whole
  // mesh: each aggregate's R then holds its coordinates of the same global
  // modes, which the next level's tentative prolongator needs (rotations
  // about a per-aggregate point span the same local space, but their R
  // would describe a different, non-rigid field in every aggregate)
  double cx = 0, cy = 0, cz = 0;
  for (int64_t i = 0; i < nn; ++i) {
    cx += c->X[i];
    cy += c->Y[i];
    cz += c->Z[i];
  }
  cx /= (double)nn;
  cy /= (double)nn;
  cz /= (double)nn;
  for (int64_t i = 0; i < nn; ++i) {
    const double x = c->X[i] - cx, y = c->Y[i] - cy, z = c->Z[i] - cz;
    double* b = &Bcur[i * 18];
    // rotations about x, y, z: columns 3, 4, 5 (rows: ux, uy, uz)
    b[0 * 6 + 3] = 0;
    b[1 * 6 + 3] = -z;
    b[2 * 6 + 3] = y;
    b[0 * 6 + 4] = z;
    b[1 * 6 + 4] = 0;
    b[2 * 6 + 4] = -x;
    b[0 * 6 + 5] = -y;
    b[1 * 6 + 5] = x;
    b[2 * 6 + 5] = 0;
    for (int d = 0; d < 3; ++d)
      if ((c->fixmask[i] >> d) & 1u)
        for (int q = 0; q < 6; ++q) b[d * 6 + q] = 0;
  }
  std::vector<int32_t> rp(c->row_ptr), cl(c->col);
  int bs = 3;
  int64_t n = nn;
  for (;;) {
    host::SymbolicLevel S;
    std::string err;
    if (!host::symbolic_level(rp, cl, n, bs, Bcur, kMaxCoarseNodes, kSell6MinRows, kSell6Window,
                              sell6_rows_per_wave(), S, &err))
      return set_err(c, ONSAS_ERR_UNSUPPORTED, err);
    const int64_t na = S.na;
    AmgLevel Lc;
    Lc.bs = 6;
    Lc.n = na;
    Lc.row_ptr = std::move(S.row_ptr);
    Lc.col = std::move(S.col);
    Lc.nnzb = (int64_t)Lc.col.size();
    if (S.rpw > 0) {
      Lc.rpw6 = S.rpw;
      Lc.ur6 = sell6_rounds_in_flight(f64);
      Lc.waves6 = sell6_waves_per_group(f64);
      Lc.nslices6 = (int64_t)S.s6_ptr.size() - 1;
      Lc.nrounds6 = S.s6_ptr.back();
    }

    // device arrays of the transfer (finer level) and of the coarse level
    AmgLevel& F = A.lev.back();
    F.nnz_p = S.nnz_p;
    F.nnz_w = S.nnz_w;
    F.restrict_wide = S.nnz_p > (int64_t)kRestrictWide * na;
    ONSAS_TRY(upload(c, &F.d_agg, S.agg));
    ONSAS_TRY(f64 ? upload_kt<double>(c, &F.d_T, S.T) : upload_kt<float>(c, &F.d_T, S.T));
    ONSAS_TRY(upload(c, &F.d_p_ptr, S.p_ptr));
    ONSAS_TRY(upload(c, &F.d_p_col, S.p_col));
    ONSAS_TRY(upload(c, &F.d_p_row, S.p_row));
    ONSAS_TRY(upload(c, &F.d_pk_ptr, S.pk_ptr));
    ONSAS_TRY(upload(c, &F.d_pk, S.pk));
    ONSAS_TRY(upload(c, &F.d_rt_ptr, S.rt_ptr));
    ONSAS_TRY(upload(c, &F.d_rt_row, S.rt_row));
    ONSAS_TRY(upload(c, &F.d_rt_pos, S.rt_pos));
    ONSAS_TRY(upload(c, &F.d_wl_ptr, S.wl_ptr));
    ONSAS_TRY(upload(c, &F.d_wl_k, S.wl_k));
    ONSAS_TRY(upload(c, &F.d_wl_p, S.wl_p));
    ONSAS_TRY(upload(c, &F.d_gl_ptr, S.gl_ptr));
    ONSAS_TRY(upload(c, &F.d_gl_p, S.gl_p));
    ONSAS_TRY(upload(c, &F.d_gl_w, S.gl_w));
    const size_t pbytes = f64 ? sizeof(double) : sizeof(_Float16);  // P's value type
    ONSAS_TRY(alloc(c, &F.d_P, pbytes * bs * 6 * (size_t)S.nnz_p));
    ONSAS_TRY(alloc(c, &F.d_PT, pbytes * bs * 6 * (size_t)S.nnz_p));
    ONSAS_TRY(alloc(c, &F.d_W, c->kt * bs * 6 * (size_t)S.nnz_w));
    ONSAS_TRY(upload(c, &Lc.d_row_ptr, Lc.row_ptr));
    ONSAS_TRY(upload(c, &Lc.d_col, Lc.col));
    ONSAS_TRY(upload(c, &Lc.d_diag, S.diag));
    ONSAS_TRY(alloc(c, &Lc.d_A, c->kt * 36 * (size_t)Lc.nnzb));
    if (Lc.nslices6 > 0) {
      ONSAS_TRY(upload(c, &Lc.d_s6_ptr, S.s6_ptr));
      ONSAS_TRY(upload(c, &Lc.d_s6_col, S.s6_col));
      ONSAS_TRY(upload(c, &Lc.d_s6_src, S.s6_src));
      ONSAS_TRY(upload(c, &Lc.d_s6_row, S.s6_row));
      if (!f64)  // the FP32 product reads x through an LDS window per workgroup
        ONSAS_TRY(upload(c, &Lc.d_s6_win,
                         host::sell6_windows(S.s6_ptr, S.s6_col, S.s6_src, na, Lc.waves6,
                                             (int32_t)(kX6Lds / (6 * c->kt)))));
      ONSAS_TRY(alloc(c, &Lc.d_A6s, (f64 ? sizeof(double) : sizeof(uint16_t)) * 36 * 64 *
                                        (size_t)Lc.nrounds6));
    }
    ONSAS_TRY(alloc(c, &Lc.d_Dinv, c->kt * 36 * (size_t)na));
    for (void** v : {&Lc.d_b, &Lc.d_x, &Lc.d_r, &Lc.d_d0, &Lc.d_d1, &Lc.d_ev})
      ONSAS_TRY(alloc(c, v, c->kt * 6 * (size_t)na));
    const bool last = na <= kMaxCoarseNodes;
    if (!last && (double)na > (double)n / 1.05)
      return set_err(c, ONSAS_ERR_UNSUPPORTED,
                     "multigrid coarsening stalled at " + std::to_string(na) + " nodes");
    A.lev.push_back(std::move(Lc));
    if (last) break;
    rp = A.lev.back().row_ptr;
    cl = A.lev.back().col;
    n = na;
    bs = 6;
    Bcur.swap(S.Bc);
  }
  for (void** v : {&A.lev[0].d_x, &A.lev[0].d_r, &A.lev[0].d_d0, &A.lev[0].d_d1, &A.lev[0].d_ev})
    ONSAS_TRY(alloc(c, v, c->kt * 3 * (size_t)nn));
  A.nc = 6 * (int)A.lev.back().n;
  ONSAS_TRY(alloc(c, (void**)&A.d_dense, sizeof(double) * (size_t)A.nc * A.nc));
  ONSAS_TRY(alloc(c, (void**)&A.d_gjc, sizeof(double) * (size_t)A.nc * kGJ));
  A.ldc = (A.nc + 3) & ~3;
  ONSAS_TRY(alloc(c, &A.d_Ainv, c->kt * (size_t)A.nc * A.ldc));
  ONSAS_TRY(alloc(c, (void**)&A.d_part, sizeof(double) * 1024));
  ONSAS_TRY(alloc(c, (void**)&A.d_rzs, sizeof(double) * 2 * (size_t)c->nslices));
  A.built = true;
  A.stale = true;
  return ONSAS_OK;
}

namespace {

template <typename KT, typename VT, int R, int UR, int WV>
void launch_sell6_ruw(const AmgLevel& L, const void* A6, const KT* v, const KT* b, KT* r_out,
                      double c1, double c2, KT* d_out, const KT* x_base, KT* x_out,
                      const PcgState* st, hipStream_t s) {
  hipLaunchKernelGGL((k_sell6_cheb<KT, VT, R, UR, WV>), dim3(nblocks(L.nslices6, WV)),
                     dim3(64 * WV), 0, s, L.d_s6_ptr, L.d_s6_col, L.d_s6_row, L.d_s6_win,
                     (const VT*)A6, v, b, r_out, (const KT*)L.d_Dinv, (KT)c1, (KT)c2, d_out,
                     x_base, x_out, L.nslices6, L.n, st);
}

template <typename KT, typename VT, int R, int UR>
void launch_sell6_ru(const AmgLevel& L, const void* A6, const KT* v, const KT* b, KT* r_out,
                     double c1, double c2, KT* d_out, const KT* x_base, KT* x_out,
                     const PcgState* st, hipStream_t s) {
  if constexpr (sell6_window<KT>()) {
    if (L.waves6 == 16)
      return launch_sell6_ruw<KT, VT, R, UR, 16>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out,
                                                 st, s);
    if (L.waves6 == 8)
      return launch_sell6_ruw<KT, VT, R, UR, 8>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out,
                                                st, s);
  }
  launch_sell6_ruw<KT, VT, R, UR, 4>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
}

template <typename KT, typename VT, int R>
void launch_sell6_r(const AmgLevel& L, const void* A6, const KT* v, const KT* b, KT* r_out,
                    double c1, double c2, KT* d_out, const KT* x_base, KT* x_out,
                    const PcgState* st, hipStream_t s) {
  if (L.ur6 == 1)
    launch_sell6_ru<KT, VT, R, 1>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
  else if (L.ur6 == 2)
    launch_sell6_ru<KT, VT, R, 2>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
  else
    launch_sell6_ru<KT, VT, R, 4>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
}

template <typename KT, typename VT>
void launch_sell6(const AmgLevel& L, const void* A6, const KT* v, const KT* b, KT* r_out,
                  double c1, double c2, KT* d_out, const KT* x_base, KT* x_out,
                  const PcgState* st, hipStream_t s) {
  if (L.rpw6 == 2)
    launch_sell6_r<KT, VT, 2>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
  else if (L.rpw6 == 4)
    launch_sell6_r<KT, VT, 4>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
  else
    launch_sell6_r<KT, VT, 8>(L, A6, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
}

// r = b - A_l v and the per-row Chebyshev epilogue (see the operator kernels);
// prz (fine level, with b and x_out and without r_out / d_out): b.x_out per slice.
// The fine level reads the tangent itself; coarse levels with a sliced-ELL copy
// use it for every product.
template <typename KT>
void op_cheb(Ctx* c, int l, const KT* v, const KT* b, KT* r_out, double c1, double c2, KT* d_out,
             const KT* x_base, KT* x_out, const PcgState* st, hipEvent_t ea = nullptr,
             hipEvent_t eb = nullptr, double* prz = nullptr, const double* chk_rr = nullptr) {
  hipStream_t s = c->stream;
  if (l == 0) {
    // chk_rr: the PCG's stop check runs in this kernel's prologue
    PcgState* wst = chk_rr ? const_cast<PcgState*>(st) : nullptr;
    const int64_t rstride = c->ell_rounds ? 64 * c->nslices : 64;
    const dim3 grid(nblocks(c->nslices, 4));
    if (ea)  // sampled launch: the dispatch itself stamps the start and stop events
      hipExtLaunchKernelGGL((k_sell_cheb<KT, KT>), grid, dim3(256), 0, s, ea, eb, 0, c->d_sell_ptr,
                            c->d_slice_rounds, rstride, c->d_sell_col, (const KT*)c->d_Ksell, v, b,
                            r_out, (const KT*)c->d_Dinv, (KT)c1, (KT)c2, d_out, x_base, x_out,
                            c->nslices, c->nn, st, prz, chk_rr, c->np, wst);
    else
      hipLaunchKernelGGL((k_sell_cheb<KT, KT>), grid, dim3(256), 0, s, c->d_sell_ptr,
                         c->d_slice_rounds, rstride, c->d_sell_col, (const KT*)c->d_Ksell, v, b,
                         r_out, (const KT*)c->d_Dinv, (KT)c1, (KT)c2, d_out, x_base, x_out,
                         c->nslices, c->nn, st, prz, chk_rr, c->np, wst);
  } else {
    const AmgLevel& L = c->amg.lev[l];
    using CT = typename CoarseSmootherType<KT>::type;
    // a level's products all use one operator (its sliced-ELL copy, BF16 in
    // mixed mode), so its smoothing and residuals stay consistent and the
    // cycle symmetric positive definite
    if (L.nslices6 > 0)
      launch_sell6<KT, CT>(L, L.d_A6s, v, b, r_out, c1, c2, d_out, x_base, x_out, st, s);
    else
      hipLaunchKernelGGL(k_bsr6_cheb<KT>, dim3(nblocks(L.n, 4)), dim3(256), 0, s, L.d_row_ptr,
                         L.d_col, (const KT*)L.d_A, v, b, r_out, (const KT*)L.d_Dinv, (KT)c1,
                         (KT)c2, d_out, x_base, x_out, L.n, st);
  }
}

// lambda_max(D^-1 A) of level l by power iteration (host-driven), from a hashed
// start vector or from the last iterate
template <typename KT>
int estimate_lambda(Ctx* c, int l, int iters, bool fresh) {
  Amg& A = c->amg;
  AmgLevel& L = A.lev[l];
  hipStream_t s = c->stream;
  const int64_t nd = (int64_t)L.bs * L.n;
  KT* v = (KT*)L.d_ev;
  KT* w = (KT*)L.d_d1;
  if (fresh)
    hipLaunchKernelGGL(k_fill_hash<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, v, nd,
                       l == 0 ? c->d_fixmask : nullptr, L.bs);
  std::vector<double> h(512);
  double lam = L.lam > 0 ? L.lam : 1.0;
  for (int it = 0; it < iters; ++it) {
    op_cheb<KT>(c, l, v, nullptr, nullptr, 0.0, 1.0, w, nullptr, nullptr, nullptr);  // w = -D^-1 A v
    hipLaunchKernelGGL(k_norm2<KT>, dim3(256), dim3(256), 0, s, w, nd, A.d_part);
    hipLaunchKernelGGL(k_norm2<KT>, dim3(256), dim3(256), 0, s, v, nd, A.d_part + 256);
    ONSAS_HIP_TRY(c, hipMemcpyAsync(h.data(), A.d_part, 512 * sizeof(double), hipMemcpyDeviceToHost, s));
    ONSAS_HIP_TRY(c, hipStreamSynchronize(s));
    double nw = 0, nv = 0;
    for (int k = 0; k < 256; ++k) {
      nw += h[k];
      nv += h[256 + k];
    }
    if (!(nw > 0) || !(nv > 0)) break;
    lam = std::sqrt(nw / nv);
    hipLaunchKernelGGL(k_scaled_copy<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, w, v, nd,
                       (KT)(1.0 / std::sqrt(nw)));
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  L.lam = lam;
  return ONSAS_OK;
}

template <typename KT, int BS>
void launch_products(Ctx* c, AmgLevel& F, AmgLevel& C, const KT* Af, const KT* Df,
                     const int32_t* colf) {
  hipStream_t s = c->stream;
  const double w = kProlongWeight * 4.0 / (3.0 * F.lam);
  hipLaunchKernelGGL((k_sa_prolongator<KT, BS>), dim3(nblocks(6 * F.nnz_p, 256)), dim3(256), 0, s,
                     F.d_p_row, F.d_p_col, F.d_agg, F.d_pk_ptr, F.d_pk, colf, Af, Df,
                     (const KT*)F.d_T, w, (PT_t<KT>*)F.d_P, F.d_rt_pos, (PT_t<KT>*)F.d_PT,
                     F.nnz_p);
  hipLaunchKernelGGL((k_sa_ap<KT, BS>), dim3(nblocks(6 * F.nnz_w, 256)), dim3(256), 0, s,
                     F.d_wl_ptr, F.d_wl_k, F.d_wl_p, Af, (const PT_t<KT>*)F.d_P, (KT*)F.d_W,
                     F.nnz_w);
  hipLaunchKernelGGL((k_sa_ptap<KT, BS>), dim3(nblocks(6 * C.nnzb, 256)), dim3(256), 0, s,
                     F.d_gl_ptr, F.d_gl_p, F.d_gl_w, (const PT_t<KT>*)F.d_P, (const KT*)F.d_W,
                     (KT*)C.d_A, C.nnzb);
}

template <typename KT>
int amg_numeric_t(Ctx* c) {
  Amg& A = c->amg;
  hipStream_t s = c->stream;
  const int nl = (int)A.lev.size();
  ONSAS_TRY(ensure_bsr(c));  // level-0 values in block-CSR order
  for (int l = 0; l + 1 < nl; ++l) {
    AmgLevel& F = A.lev[l];
    AmgLevel& C = A.lev[l + 1];
    ONSAS_TRY(estimate_lambda<KT>(c, l, A.lam_set ? kPowerWarm : kPowerFirst, !A.lam_set));
    if (l == 0)
      launch_products<KT, 3>(c, F, C, (const KT*)c->d_K, (const KT*)c->d_Dinv, c->d_col);
    else
      launch_products<KT, 6>(c, F, C, (const KT*)F.d_A, (const KT*)F.d_Dinv, F.d_col);
    if (l + 2 < nl) {
      hipLaunchKernelGGL(k_inv6<KT>, dim3(nblocks(C.n, 128)), dim3(128), 0, s, (KT*)C.d_A, C.d_diag,
                         (KT*)C.d_Dinv, C.n);
      if (C.nslices6 > 0)
        hipLaunchKernelGGL(k_to_sell6<KT>, dim3(nblocks(36 * 64 * C.nrounds6, 256)), dim3(256), 0,
                           s, C.d_s6_src, (const KT*)C.d_A,
                           (typename CoarseSmootherType<KT>::type*)C.d_A6s, 64 * C.nrounds6);
    }
    ONSAS_HIP_TRY(c, hipGetLastError());
  }
  const AmgLevel& Lc = A.lev[nl - 1];
  const int64_t nc = A.nc;
  hipLaunchKernelGGL(k_zero_f64, dim3(256), dim3(256), 0, s, A.d_dense, nc * nc);
  hipLaunchKernelGGL(k_dense_scatter<KT>, dim3(nblocks(nc, 128)), dim3(128), 0, s, Lc.d_row_ptr,
                     Lc.d_col, (const KT*)Lc.d_A, A.d_dense, Lc.n);
  for (int64_t k0 = 0; k0 < nc; k0 += kGJ) {
    const int kb = (int)std::min<int64_t>(kGJ, nc - k0);
    hipLaunchKernelGGL(k_gj_diag, dim3(1), dim3(256), 0, s, A.d_dense, nc, k0, kb);
    hipLaunchKernelGGL(k_gj_rows, dim3(nblocks(nc * kGJ + nc, 256)), dim3(256), 0, s, A.d_dense,
                       A.d_gjc, nc, k0, kb);
    hipLaunchKernelGGL(k_gj_update, dim3(nblocks(nc, 64), nblocks(nc, 64)), dim3(256), 0, s,
                       A.d_dense, (const double*)A.d_gjc, nc, k0, kb);
  }
  hipLaunchKernelGGL(k_to_kt<KT>, dim3(nblocks((int64_t)A.nc * A.ldc, 256)), dim3(256), 0, s,
                     A.d_dense, (KT*)A.d_Ainv, (int64_t)A.nc, (int64_t)A.ldc);
  ONSAS_HIP_TRY(c, hipGetLastError());
  A.lam_set = true;
  A.stale = false;
  ++A.gen;  // new smoothing coefficients: graphs captured before are stale
  return ONSAS_OK;
}

template <typename KT, int BS>
void launch_restrict(Ctx* c, const AmgLevel& L, const AmgLevel& C, const KT* r, KT sc, KT* d0c,
                     const PcgState* st) {
  if (L.restrict_wide)
    hipLaunchKernelGGL((k_sa_restrict<KT, BS, 128>), dim3(nblocks8(C.n, 2)), dim3(256), 0, c->stream,
                       L.d_rt_ptr, L.d_rt_row, (const PT_t<KT>*)L.d_PT, r, (KT*)C.d_b,
                       (const KT*)C.d_Dinv, sc, d0c, C.n, st);
  else
    hipLaunchKernelGGL((k_sa_restrict<KT, BS, 32>), dim3(nblocks8(C.n, 8)), dim3(256), 0, c->stream,
                       L.d_rt_ptr, L.d_rt_row, (const PT_t<KT>*)L.d_PT, r, (KT*)C.d_b,
                       (const KT*)C.d_Dinv, sc, d0c, C.n, st);
}

template <typename KT>
int vcycle_t(Ctx* c, const KT* b0, KT* z0, const PcgState* st, bool d0_ready, hipEvent_t* ev,
             double* prz, const double* chk_rr) {
  auto evs = [&](int l, int j) { return l == 0 && ev ? ev[j] : nullptr; };
  Amg& A = c->amg;
  hipStream_t s = c->stream;
  const int nl = (int)A.lev.size();
  if (!d0_ready) {  // first Chebyshev step of the fine level; coarser ones come with
                    // the restriction
    const AmgLevel& L = A.lev[0];
    hipLaunchKernelGGL((k_jacobi<KT, 3>), dim3(nblocks(3 * L.n, 256)), dim3(256), 0, s,
                       (const KT*)c->d_Dinv, b0, (KT*)L.d_d0, L.n, (KT)cheb(L.lam).inv_theta, st);
  }
  auto degree1 = [&](int l) { return (l == 0 ? A.fine_degree : A.coarse_degree) == 1; };
  // down: pre-smoothing, residual, restriction
  for (int l = 0; l + 1 < nl; ++l) {
    AmgLevel& L = A.lev[l];
    const Cheb ch = cheb(L.lam);
    const KT* b = l == 0 ? b0 : (const KT*)L.d_b;
    KT* r = (KT*)L.d_r;
    if (degree1(l)) {
      // degree 1: x = d0 itself (prolonged into in place), the exact r = b - A d0
      op_cheb<KT>(c, l, (const KT*)L.d_d0, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, st,
                  evs(l, 0), evs(l, 1), nullptr, l == 0 ? chk_rr : nullptr);
    } else {
      KT* x = (KT*)L.d_x;
      // x = d0 + d1, then r = b - A x
      op_cheb<KT>(c, l, (const KT*)L.d_d0, b, nullptr, ch.c1, ch.c2, (KT*)L.d_d1, nullptr, x, st,
                  nullptr, nullptr, nullptr, l == 0 ? chk_rr : nullptr);
      op_cheb<KT>(c, l, x, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, st, evs(l, 0), evs(l, 1));
    }
    AmgLevel& C = A.lev[l + 1];
    const bool coarsest = l + 2 == nl;
    const KT sc = coarsest ? KT(0) : (KT)cheb(C.lam).inv_theta;
    KT* d0c = coarsest ? nullptr : (KT*)C.d_d0;
    if (l == 0)
      launch_restrict<KT, 3>(c, L, C, r, sc, d0c, st);
    else
      launch_restrict<KT, 6>(c, L, C, r, sc, d0c, st);
  }
  // coarsest: exact
  const AmgLevel& Lc = A.lev[nl - 1];
  hipLaunchKernelGGL(k_dense_apply<KT>, dim3(nblocks(A.nc, 4)), dim3(256), 0, s,
                     (const KT*)A.d_Ainv, (const KT*)Lc.d_b, (KT*)Lc.d_x, (int64_t)A.nc,
                     (int64_t)A.ldc, st);
  // up: prolongation, post-smoothing of the residual
  for (int l = nl - 2; l >= 0; --l) {
    AmgLevel& L = A.lev[l];
    const AmgLevel& C = A.lev[l + 1];
    const Cheb ch = cheb(L.lam);
    const KT* b = l == 0 ? b0 : (const KT*)L.d_b;
    const bool deg1 = degree1(l);
    KT* d0 = (KT*)L.d_d0;
    KT* x = deg1 ? d0 : (KT*)L.d_x;
    KT* r = (KT*)L.d_r;
    // the level's result goes to x (coarser levels) for the next prolongation;
    // the degree-1 step reads its input at neighbouring rows, so it writes
    // there out of place
    KT* z = l == 0 ? z0 : (KT*)L.d_x;
    if (l == 0)
      hipLaunchKernelGGL((k_sa_prolong_pair<KT, 3, (sizeof(KT) == 4 ? 4 : 2)>),
                         dim3(nblocks8(2 * L.n, 256)), dim3(256), 0, s, L.d_p_ptr, L.d_p_col,
                         (const PT_t<KT>*)L.d_P, (const KT*)C.d_x, x, L.n, st);
    else
      hipLaunchKernelGGL((k_sa_prolong<KT, 6>), dim3(nblocks(6 * L.n, 256)), dim3(256), 0, s,
                         L.d_p_ptr, L.d_p_col, (const PT_t<KT>*)L.d_P, (const KT*)C.d_x, x, L.n, st);
    if (deg1) {
      // z = x + D^-1 (b - A x) / theta: the degree-1 step in the product's epilogue
      op_cheb<KT>(c, l, x, b, nullptr, 0.0, ch.inv_theta, nullptr, nullptr, z, st, evs(l, 2),
                  evs(l, 3), l == 0 ? prz : nullptr);
    } else {
      // r = b - A x and d0, then z = x + d0 + d1
      op_cheb<KT>(c, l, x, b, r, 0.0, ch.inv_theta, d0, nullptr, nullptr, st, evs(l, 2), evs(l, 3));
      op_cheb<KT>(c, l, d0, r, nullptr, ch.c1, ch.c2, nullptr, x, z, st);
    }
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

}  // namespace

int amg_numeric(Ctx* c) {
  return c->precision == ONSAS_PRECISION_FP64 ? amg_numeric_t<double>(c) : amg_numeric_t<float>(c);
}

int launch_vcycle(Ctx* c, const void* b, void* x, const PcgState* st, bool d0_ready,
                  hipEvent_t* ev, double* prz, const double* chk_rr) {
  if (prz && c->amg.fine_degree != 1)
    return set_err(c, ONSAS_ERR_STATE, "r.z partials come from the degree-1 fine smoothing only");
  if (chk_rr && !d0_ready)
    return set_err(c, ONSAS_ERR_STATE, "the stop check rides on the cycle's first tangent product");
  return c->precision == ONSAS_PRECISION_FP64
             ? vcycle_t<double>(c, (const double*)b, (double*)x, st, d0_ready, ev, prz, chk_rr)
             : vcycle_t<float>(c, (const float*)b, (float*)x, st, d0_ready, ev, prz, chk_rr);
}

double amg_fine_smoother(Ctx* c, void** d0) {
  *d0 = c->amg.lev[0].d_d0;
  return cheb(c->amg.lev[0].lam).inv_theta;
}

// ---- hierarchy inspection (tests, diagnostics) -------------------------------------

// the hierarchy has values once a solve has run its numeric setup (lam_set)
namespace {
// The cycle's products of level l, in cycle order, on the last solve's vectors
// (no stop flag), between two events.  Level 0 alternates its two products
// (degree 1: r = b - A d0 and z = d0 + D^-1 (b - A d0) / theta); a sliced-ELL
// level >= 1 the same two with the degree-1 smoother, else its four (x = d0 +
// d1, r = b - A x, r and d0 = D^-1 r / theta, x += d0 + d1).
template <typename KT>
int amg_time_level_products_t(Ctx* c, int l, int launches, double* ms) {
  Amg& A = c->amg;
  AmgLevel& L = A.lev[l];
  const Cheb ch = cheb(L.lam);
  hipStream_t s = c->stream;
  hipEvent_t e0 = c->ev_spmv[2], e1 = c->ev_spmv[3];
  KT *x = (KT*)L.d_x, *r = (KT*)L.d_r, *d0 = (KT*)L.d_d0, *d1 = (KT*)L.d_d1;
  ONSAS_HIP_TRY(c, hipEventRecord(e0, s));
  for (int i = 0; i < launches; ++i) {
    if (l == 0) {
      const KT* b = (const KT*)c->d_pr;
      if (A.fine_degree == 1) {
        if (i & 1)
          op_cheb<KT>(c, 0, d0, b, nullptr, 0.0, ch.inv_theta, nullptr, nullptr, x, nullptr);
        else
          op_cheb<KT>(c, 0, d0, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, nullptr);
      } else {  // r = b - A x;  r and d = D^-1 r / theta
        if (i & 1)
          op_cheb<KT>(c, 0, x, b, r, 0.0, ch.inv_theta, d1, nullptr, nullptr, nullptr);
        else
          op_cheb<KT>(c, 0, x, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, nullptr);
      }
    } else if (A.coarse_degree == 1) {
      const KT* b = (const KT*)L.d_b;
      if (i & 1)
        op_cheb<KT>(c, l, d0, b, nullptr, 0.0, ch.inv_theta, nullptr, nullptr, x, nullptr);
      else
        op_cheb<KT>(c, l, d0, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, nullptr);
    } else {
      const KT* b = (const KT*)L.d_b;
      switch (i & 3) {
        case 0: op_cheb<KT>(c, l, d0, b, nullptr, ch.c1, ch.c2, d1, nullptr, x, nullptr); break;
        case 1: op_cheb<KT>(c, l, x, b, r, 0.0, 0.0, nullptr, nullptr, nullptr, nullptr); break;
        case 2: op_cheb<KT>(c, l, x, b, r, 0.0, ch.inv_theta, d0, nullptr, nullptr, nullptr); break;
        default: op_cheb<KT>(c, l, d0, r, nullptr, ch.c1, ch.c2, nullptr, x, x, nullptr); break;
      }
    }
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  ONSAS_HIP_TRY(c, hipEventRecord(e1, s));
  ONSAS_HIP_TRY(c, hipEventSynchronize(e1));
  float t = 0;
  ONSAS_HIP_TRY(c, hipEventElapsedTime(&t, e0, e1));
  *ms = (double)t / launches;
  return ONSAS_OK;
}
}  // namespace

int amg_time_level_products(Ctx* c, int level, int launches, double* ms) {
  if (!c->amg.built || !c->amg.lam_set)
    return set_err(c, ONSAS_ERR_STATE, "no multigrid hierarchy: run a PCG solve with ONSAS_PRECOND_AMG first");
  if (launches < 2 || launches > 10000) return set_err(c, ONSAS_ERR_ARG, "launches must be in [2, 10000]");
  const int nl = (int)c->amg.lev.size();
  if (level < 0 || level + 1 >= nl)
    return set_err(c, ONSAS_ERR_ARG, "level must be a smoothed level (0 .. levels - 2)");
  if (level > 0 && c->amg.lev[level].nslices6 == 0)
    return set_err(c, ONSAS_ERR_ARG, "this level has no sliced-ELL operator");
  return c->precision == ONSAS_PRECISION_FP64
             ? amg_time_level_products_t<double>(c, level, launches, ms)
             : amg_time_level_products_t<float>(c, level, launches, ms);
}

int amg_nlevels(Ctx* c) { return c->amg.built && c->amg.lam_set ? (int)c->amg.lev.size() : 0; }

int amg_export_info(Ctx* c, int l, int64_t* n, int32_t* bs, int64_t* nnzb, int64_t* nnz_p,
                    double* lam) {
  const Amg& A = c->amg;
  if (!A.built || !A.lam_set || l < 0 || l >= (int)A.lev.size())
    return set_err(c, ONSAS_ERR_ARG, "no such multigrid level");
  const AmgLevel& L = A.lev[l];
  if (n) *n = L.n;
  if (bs) *bs = L.bs;
  if (nnzb) *nnzb = l == 0 ? c->nnzb : L.nnzb;
  if (nnz_p) *nnz_p = l + 1 < (int)A.lev.size() ? L.nnz_p : 0;
  if (lam) *lam = L.lam;
  return ONSAS_OK;
}

namespace {
template <typename T>
int copy_back(Ctx* c, std::vector<T>& h, const void* d, size_t count) {
  h.resize(count);
  if (count)
    ONSAS_HIP_TRY(c, hipMemcpyAsync(h.data(), d, count * sizeof(T), hipMemcpyDeviceToHost, c->stream));
  ONSAS_HIP_TRY(c, hipStreamSynchronize(c->stream));
  return ONSAS_OK;
}
}  // namespace

int amg_export_operator(Ctx* c, int l, int32_t* row_ptr, int32_t* col, double* vals) {
  const Amg& A = c->amg;
  if (!A.built || !A.lam_set || l < 1 || l >= (int)A.lev.size())
    return set_err(c, ONSAS_ERR_ARG, "multigrid operators are exported for levels >= 1");
  const AmgLevel& L = A.lev[l];
  if (row_ptr) std::copy(L.row_ptr.begin(), L.row_ptr.end(), row_ptr);
  if (col) std::copy(L.col.begin(), L.col.end(), col);
  if (vals) {
    const size_t m = 36 * (size_t)L.nnzb;
    if (c->precision == ONSAS_PRECISION_FP64) {
      std::vector<double> h;
      ONSAS_TRY(copy_back(c, h, L.d_A, m));
      std::copy(h.begin(), h.end(), vals);
    } else {
      std::vector<float> h;
      ONSAS_TRY(copy_back(c, h, L.d_A, m));
      std::copy(h.begin(), h.end(), vals);
    }
  }
  return ONSAS_OK;
}

int amg_export_prolongator(Ctx* c, int l, int32_t* p_ptr, int32_t* p_col, double* vals) {
  const Amg& A = c->amg;
  if (!A.built || !A.lam_set || l < 0 || l + 1 >= (int)A.lev.size())
    return set_err(c, ONSAS_ERR_ARG, "no prolongator below this multigrid level");
  const AmgLevel& L = A.lev[l];
  if (p_ptr) {
    std::vector<int32_t> h;
    ONSAS_TRY(copy_back(c, h, L.d_p_ptr, L.n + 1));
    std::copy(h.begin(), h.end(), p_ptr);
  }
  if (p_col) {
    std::vector<int32_t> h;
    ONSAS_TRY(copy_back(c, h, L.d_p_col, L.nnz_p));
    std::copy(h.begin(), h.end(), p_col);
  }
  if (vals) {
    const size_t m = (size_t)L.bs * 6 * L.nnz_p;
    if (c->precision == ONSAS_PRECISION_FP64) {
      std::vector<double> h;
      ONSAS_TRY(copy_back(c, h, L.d_P, m));
      std::copy(h.begin(), h.end(), vals);
    } else {
      std::vector<_Float16> h;
      ONSAS_TRY(copy_back(c, h, L.d_P, m));
      for (size_t k = 0; k < m; ++k) vals[k] = (double)(float)h[k];
    }
  }
  return ONSAS_OK;
}

void amg_free(Ctx* c) {
  Amg& A = c->amg;
  for (AmgLevel& L : A.lev)
    for (void* p : {(void*)L.d_row_ptr, (void*)L.d_col, (void*)L.d_diag, L.d_A, L.d_Dinv, L.d_b,
                    L.d_x, L.d_r, L.d_d0, L.d_d1, L.d_ev, (void*)L.d_agg, L.d_T, (void*)L.d_p_ptr,
                    (void*)L.d_p_col, (void*)L.d_p_row, L.d_P, (void*)L.d_pk_ptr, (void*)L.d_pk,
                    (void*)L.d_rt_ptr, (void*)L.d_rt_row, (void*)L.d_rt_pos, L.d_PT, L.d_W, (void*)L.d_wl_ptr, (void*)L.d_wl_k,
                    (void*)L.d_wl_p, (void*)L.d_gl_ptr, (void*)L.d_gl_p, (void*)L.d_gl_w,
                    (void*)L.d_s6_ptr, (void*)L.d_s6_col, (void*)L.d_s6_src,
                    (void*)L.d_s6_row, (void*)L.d_s6_win, L.d_A6s})
      if (p) (void)hipFree(p);
  for (void* p : {(void*)A.d_dense, (void*)A.d_gjc, A.d_Ainv, (void*)A.d_part,
                  (void*)A.d_rzs})
    if (p) (void)hipFree(p);
  A = Amg{};
}

}  // namespace onsas
