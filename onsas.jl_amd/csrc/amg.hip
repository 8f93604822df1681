// amg.hip -- aggregation multigrid preconditioner for the PCG (an option beside
// block-Jacobi; the reference's CG is unpreconditioned).  gfx950, wave64.
//
// Hierarchy (built once on the host, from the block pattern and the node
// coordinates):
//   * greedy aggregation of the node graph: a node whose neighbours are all
//     free becomes a root with its neighbours, left-over nodes join a
//     neighbouring aggregate, isolated ones form their own;
//   * near-nullspace: the 6 rigid-body modes of each node (zero on fixed dofs),
//     orthonormalised per aggregate by modified Gram-Schmidt (twice): the
//     aggregate's rows of Q form the tentative prolongator P (bs x 6 per node),
//     R (6 x 6) is the next level's near-nullspace;
//   * coarse levels have 6 dofs per node; coarsening goes on until at most 12
//     nodes (72 dofs) remain, whose matrix is inverted densely in LDS.
// Numerics (every assembly, on the device): Galerkin A_c = P^T A P summed per
// coarse block over fixed-order lists of the finer blocks that land in it
// (deterministic, no atomics), 6x6 block-Jacobi inverses, and the dense
// inverse of the coarsest matrix.  A coarse dof whose column of Q vanished
// (rank-deficient aggregate) gets a unit diagonal.
// V-cycle: one damped block-Jacobi sweep (omega = 1 / (1.1 lambda_max(D^-1 A)),
// lambda by power iteration) before and after the coarse correction, the
// coarsest level solved exactly; the cycle is symmetric, so it is a valid PCG
// preconditioner.  Every cycle kernel returns at once when the PCG state says
// done, so replayed graphs cost nothing after convergence.
#include <algorithm>
#include <cmath>
#include <numeric>
#include <vector>

#include "engine.hpp"

namespace onsas {

namespace {

constexpr int kMaxCoarseNodes = 12;  // coarsest level: <= 72 dofs, inverted in LDS

inline unsigned nblocks(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

__device__ inline bool pcg_stopped(const PcgState* st) { return st && (st->done | st->breakdown); }

// ---- numeric setup ---------------------------------------------------------------

// One thread per coarse block: sum over its finer blocks (i, j) of P_i^T A_ij P_j,
// in FP64.  Finer values at off + stride * (row-major entry).
template <typename KT, int BSF>
__global__ void k_galerkin(const int64_t* __restrict__ gptr, const int64_t* __restrict__ goff,
                           const int32_t* __restrict__ gi, const int32_t* __restrict__ gj,
                           const KT* __restrict__ Af, int64_t fstride, const KT* __restrict__ Pf,
                           KT* __restrict__ Ac, int64_t ncb) {
  const int64_t cb = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (cb >= ncb) return;
  double acc[36];
#pragma unroll
  for (int q = 0; q < 36; ++q) acc[q] = 0;
  for (int64_t t = gptr[cb]; t < gptr[cb + 1]; ++t) {
    const KT* a = Af + goff[t];
    const KT* pi = Pf + (int64_t)gi[t] * (BSF * 6);
    const KT* pj = Pf + (int64_t)gj[t] * (BSF * 6);
    double T[BSF][6];  // A_ij P_j
#pragma unroll
    for (int r = 0; r < BSF; ++r)
#pragma unroll
      for (int c = 0; c < 6; ++c) T[r][c] = 0;
#pragma unroll
    for (int r = 0; r < BSF; ++r)
#pragma unroll
      for (int s = 0; s < BSF; ++s) {
        const double v = (double)a[fstride * (r * BSF + s)];
#pragma unroll
        for (int c = 0; c < 6; ++c) T[r][c] += v * (double)pj[s * 6 + c];
      }
#pragma unroll
    for (int r = 0; r < BSF; ++r)
#pragma unroll
      for (int u = 0; u < 6; ++u) {
        const double pv = (double)pi[r * 6 + u];
#pragma unroll
        for (int c = 0; c < 6; ++c) acc[u * 6 + c] += pv * T[r][c];
      }
  }
  KT* o = Ac + 36 * cb;
#pragma unroll
  for (int q = 0; q < 36; ++q) o[q] = (KT)acc[q];
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

// Coarsest level: block-CSR 6x6 -> dense FP64 (unit diagonal for unused dofs).
template <typename KT>
__global__ void k_dense_scatter(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                                const KT* __restrict__ A, double* __restrict__ D, int64_t n) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;  // (node row, comp)
  const int64_t nc = 6 * n;
  if (t >= nc) return;
  const int64_t I = t / 6;
  const int r = (int)(t % 6);
  double* drow = D + t * nc;
  for (int64_t c = 0; c < nc; ++c) drow[c] = 0;
  for (int32_t k = row_ptr[I]; k < row_ptr[I + 1]; ++k) {
    const int64_t J = col[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) drow[6 * J + c] = (double)A[36 * (int64_t)k + r * 6 + c];
  }
  if (drow[t] == 0) drow[t] = 1;
}

// In-place Gauss-Jordan inverse of an SPD dense matrix (n <= 72) in LDS, one
// workgroup; the pivots are the diagonal (SPD: no pivoting needed).
__global__ __launch_bounds__(256) void k_dense_inverse(double* __restrict__ D, int64_t n) {
  __shared__ double M[72 * 72];
  __shared__ double colk[72];
  const int64_t nn = n * n;
  for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) M[t] = D[t];
  __syncthreads();
  for (int64_t k = 0; k < n; ++k) {
    for (int64_t r = threadIdx.x; r < n; r += blockDim.x) colk[r] = M[r * n + k];
    __syncthreads();
    const double piv = 1.0 / colk[k];
    for (int64_t c = threadIdx.x; c < n; c += blockDim.x)
      M[k * n + c] = (c == k) ? piv : M[k * n + c] * piv;
    __syncthreads();
    for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) {
      const int64_t r = t / n, c = t - r * n;
      if (r == k) continue;
      M[t] = (c == k) ? -colk[r] * piv : M[t] - colk[r] * M[k * n + c];
    }
    __syncthreads();
  }
  for (int64_t t = threadIdx.x; t < nn; t += blockDim.x) D[t] = M[t];
}

template <typename KT>
__global__ void k_to_kt(const double* __restrict__ s, KT* __restrict__ d, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) d[i] = (KT)s[i];
}

// ---- V-cycle kernels (all skip once the PCG has stopped) ---------------------------

// r = b - K x on the fine sliced-ELL tangent (one wave per slice).
template <typename KT>
__global__ __launch_bounds__(256) void k_sell_resid(const int32_t* __restrict__ sell_ptr,
                                                    const int32_t* __restrict__ slice_rounds,
                                                    int64_t rstride, const int32_t* __restrict__ scol,
                                                    const KT* __restrict__ sval,
                                                    const KT* __restrict__ x, const KT* __restrict__ b,
                                                    KT* __restrict__ r, int64_t nslices, int64_t nn,
                                                    const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t s = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (s >= nslices) return;
  const int32_t b0 = sell_ptr[s];
  const int32_t rounds = slice_rounds[s];
  const uint32_t rs = (uint32_t)rstride;
  const int32_t* cp = scol + b0 + lane;
  const KT* vp = sval + 9u * (uint32_t)b0 + lane;
  KT a0 = 0, a1 = 0, a2 = 0;
  for (int m = 0; m < rounds; ++m) {
    const uint32_t j = 3u * (uint32_t)cp[rs * m];
    const KT* v = vp + 9u * rs * m;
    const KT p0 = x[j], p1 = x[j + 1], p2 = x[j + 2];
    a0 += v[0] * p0 + v[64] * p1 + v[128] * p2;
    a1 += v[192] * p0 + v[256] * p1 + v[320] * p2;
    a2 += v[384] * p0 + v[448] * p1 + v[512] * p2;
  }
#pragma unroll
  for (int o = kSellRows; o < 64; o <<= 1) {
    a0 += __shfl_xor(a0, o, 64);
    a1 += __shfl_xor(a1, o, 64);
    a2 += __shfl_xor(a2, o, 64);
  }
  const int64_t i = kSellRows * s + lane % kSellRows;
  if (lane < kSellRows && i < nn) {
    r[3 * i] = b[3 * i] - a0;
    r[3 * i + 1] = b[3 * i + 1] - a1;
    r[3 * i + 2] = b[3 * i + 2] - a2;
  }
}

// r = b - A x on a 6x6 block-CSR level: one thread per (row, component).
template <typename KT>
__global__ void k_bsr6_resid(const int32_t* __restrict__ row_ptr, const int32_t* __restrict__ col,
                             const KT* __restrict__ A, const KT* __restrict__ x,
                             const KT* __restrict__ b, KT* __restrict__ r, int64_t n,
                             const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * n) return;
  const int64_t I = t / 6;
  const int u = (int)(t % 6);
  KT acc = 0;
  for (int32_t k = row_ptr[I]; k < row_ptr[I + 1]; ++k) {
    const KT* a = A + 36 * (int64_t)k + 6 * u;
    const KT* xj = x + 6 * (int64_t)col[k];
#pragma unroll
    for (int c = 0; c < 6; ++c) acc += a[c] * xj[c];
  }
  r[t] = b[t] - acc;
}

// x = omega D^-1 b (first) or x += omega D^-1 r, one thread per (node, component).
template <typename KT, int BS>
__global__ void k_smooth(const KT* __restrict__ Dinv, const KT* __restrict__ v, KT* __restrict__ x,
                         int64_t n, KT omega, int accumulate, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BS * n) return;
  const int64_t i = t / BS;
  const int u = (int)(t % BS);
  const KT* d = Dinv + (int64_t)BS * BS * i + BS * u;
  const KT* vi = v + (int64_t)BS * i;
  KT s = 0;
#pragma unroll
  for (int c = 0; c < BS; ++c) s += d[c] * vi[c];
  x[t] = accumulate ? x[t] + omega * s : omega * s;
}

// b_c[I] = sum over the member nodes i of aggregate I of P_i^T r_i.
template <typename KT, int BSF>
__global__ void k_restrict(const int64_t* __restrict__ aptr, const int32_t* __restrict__ anodes,
                           const KT* __restrict__ P, const KT* __restrict__ rf, KT* __restrict__ bc,
                           int64_t nc, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= 6 * nc) return;
  const int64_t I = t / 6;
  const int u = (int)(t % 6);
  KT acc = 0;
  for (int64_t k = aptr[I]; k < aptr[I + 1]; ++k) {
    const int64_t i = anodes[k];
    const KT* p = P + (int64_t)BSF * 6 * i;
#pragma unroll
    for (int a = 0; a < BSF; ++a) acc += p[a * 6 + u] * rf[BSF * i + a];
  }
  bc[t] = acc;
}

// x_f[i] += P_i x_c[agg(i)].
template <typename KT, int BSF>
__global__ void k_prolong(const int32_t* __restrict__ agg, const KT* __restrict__ P,
                          const KT* __restrict__ xc, KT* __restrict__ xf, int64_t nf,
                          const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= BSF * nf) return;
  const int64_t i = t / BSF;
  const int a = (int)(t % BSF);
  const KT* p = P + (int64_t)BSF * 6 * i + 6 * a;
  const KT* xi = xc + 6 * (int64_t)agg[i];
  KT s = 0;
#pragma unroll
  for (int c = 0; c < 6; ++c) s += p[c] * xi[c];
  xf[t] += s;
}

// x = Ainv b on the coarsest level (dense), one thread per row.
template <typename KT>
__global__ void k_dense_apply(const KT* __restrict__ Ainv, const KT* __restrict__ b,
                              KT* __restrict__ x, int64_t nc, const PcgState* st) {
  if (pcg_stopped(st)) return;
  const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (r >= nc) return;
  const KT* a = Ainv + r * nc;
  KT s = 0;
  for (int64_t c = 0; c < nc; ++c) s += a[c] * b[c];
  x[r] = s;
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

template <typename KT>
__global__ void k_scale(KT* __restrict__ v, int64_t n, KT a) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= a;
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

// greedy aggregation of a node graph (CSR, self included)
int64_t aggregate(const std::vector<int32_t>& rp, const std::vector<int32_t>& cl, int64_t n,
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

}  // namespace

// Build the hierarchy (host) and allocate its device arrays.
int amg_build(Ctx* c) {
  Amg& A = c->amg;
  A = Amg{};
  const int64_t nn = c->nn;
  // level 0: the fine tangent (3 dofs per node); its near-nullspace
  AmgLevel L0;
  L0.bs = 3;
  L0.n = nn;
  A.lev.push_back(L0);
  std::vector<double> B(nn * 3 * 6, 0.0);  // node-major, bs x 6
  for (int64_t i = 0; i < nn; ++i) {
    double* b = &B[i * 18];
    for (int d = 0; d < 3; ++d) b[d * 6 + d] = 1.0;
  }
  // graph of the current level
  std::vector<int32_t> rp(c->row_ptr), cl(c->col);
  int bs = 3;
  int64_t n = nn;
  std::vector<double> Bcur = B;  // rotations filled per aggregate (relative to its centroid)
  for (int lev = 0;; ++lev) {
    std::vector<int32_t> agg;
    const int64_t na = aggregate(rp, cl, n, agg);
    // members per aggregate
    std::vector<int64_t> aptr(na + 1, 0);
    for (int64_t i = 0; i < n; ++i) aptr[agg[i] + 1]++;
    for (int64_t a = 0; a < na; ++a) aptr[a + 1] += aptr[a];
    std::vector<int32_t> anodes(n);
    {
      std::vector<int64_t> pos(aptr.begin(), aptr.end() - 1);
      for (int64_t i = 0; i < n; ++i) anodes[pos[agg[i]]++] = (int32_t)i;
    }
    if (lev == 0) {  // rigid-body rotations about each aggregate's centroid
      for (int64_t a = 0; a < na; ++a) {
        double cx = 0, cy = 0, cz = 0;
        const int64_t m = aptr[a + 1] - aptr[a];
        for (int64_t k = aptr[a]; k < aptr[a + 1]; ++k) {
          cx += c->X[anodes[k]];
          cy += c->Y[anodes[k]];
          cz += c->Z[anodes[k]];
        }
        cx /= m;
        cy /= m;
        cz /= m;
        for (int64_t k = aptr[a]; k < aptr[a + 1]; ++k) {
          const int64_t i = anodes[k];
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
      }
    }
    // per aggregate: Q (rows of the member nodes) and R (the coarse near-nullspace)
    std::vector<double> P(n * bs * 6, 0.0), Bc(na * 36, 0.0);
    std::vector<double> col;
    for (int64_t a = 0; a < na; ++a) {
      const int64_t m = aptr[a + 1] - aptr[a];
      const int64_t rows = m * bs;
      std::vector<double> V(rows * 6), Q(rows * 6, 0.0);
      for (int64_t k = 0; k < m; ++k)
        for (int r = 0; r < bs; ++r)
          for (int q = 0; q < 6; ++q) V[(k * bs + r) * 6 + q] = Bcur[anodes[aptr[a] + k] * bs * 6 + r * 6 + q];
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
          for (int q = 0; q < 6; ++q) P[anodes[aptr[a] + k] * bs * 6 + r * 6 + q] = Q[(k * bs + r) * 6 + q];
      for (int q = 0; q < 36; ++q) Bc[a * 36 + q] = R[q];
    }
    // coarse pattern and Galerkin lists (finer blocks in their storage order)
    const int64_t nb_f = rp[n];
    std::vector<std::vector<int32_t>> crow(na);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t k = rp[i]; k < rp[i + 1]; ++k) crow[agg[i]].push_back(agg[cl[k]]);
    AmgLevel Lc;
    Lc.bs = 6;
    Lc.n = na;
    Lc.row_ptr.assign(na + 1, 0);
    for (int64_t a = 0; a < na; ++a) {
      std::sort(crow[a].begin(), crow[a].end());
      crow[a].erase(std::unique(crow[a].begin(), crow[a].end()), crow[a].end());
      Lc.row_ptr[a + 1] = Lc.row_ptr[a] + (int32_t)crow[a].size();
    }
    Lc.nnzb = Lc.row_ptr[na];
    Lc.col.resize(Lc.nnzb);
    Lc.diag.assign(na, -1);
    for (int64_t a = 0; a < na; ++a) {
      std::copy(crow[a].begin(), crow[a].end(), Lc.col.begin() + Lc.row_ptr[a]);
      auto it = std::lower_bound(crow[a].begin(), crow[a].end(), (int32_t)a);
      Lc.diag[a] = Lc.row_ptr[a] + (int32_t)(it - crow[a].begin());
    }
    std::vector<int64_t> gptr(Lc.nnzb + 1, 0), goff(nb_f);
    std::vector<int32_t> gi(nb_f), gj(nb_f);
    std::vector<int32_t> cblk(nb_f);
    for (int64_t i = 0; i < n; ++i)
      for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
        const int32_t I = agg[i], J = agg[cl[k]];
        auto b0 = Lc.col.begin() + Lc.row_ptr[I], b1 = Lc.col.begin() + Lc.row_ptr[I + 1];
        cblk[k] = (int32_t)(std::lower_bound(b0, b1, J) - Lc.col.begin());
        gptr[cblk[k] + 1]++;
      }
    for (int64_t b = 0; b < Lc.nnzb; ++b) gptr[b + 1] += gptr[b];
    {
      std::vector<int64_t> pos(gptr.begin(), gptr.end() - 1);
      const int64_t rstride = c->ell_rounds ? 64 * c->nslices : 64;
      for (int64_t i = 0; i < n; ++i)
        for (int32_t k = rp[i]; k < rp[i + 1]; ++k) {
          const int64_t t = pos[cblk[k]]++;
          if (lev == 0) {  // the finer block lives in the sliced-ELL copy
            const int64_t s = i / kSellRows, kk = k - rp[i];
            const int64_t lane = kSellRows * (kk % kSellLanes) + i % kSellRows;
            goff[t] = 9 * (c->sell_ptr[s] + rstride * (kk / kSellLanes)) + lane;
          } else {
            goff[t] = 36 * (int64_t)k;
          }
          gi[t] = (int32_t)i;
          gj[t] = cl[k];
        }
    }
    // device arrays: transfer of the finer level, Galerkin lists of the coarse one
    AmgLevel& F = A.lev.back();
    std::vector<int32_t> agg32(agg.begin(), agg.end());
    ONSAS_TRY(upload(c, &F.d_agg, agg32));
    ONSAS_TRY(upload(c, &F.d_aptr, aptr));
    ONSAS_TRY(upload(c, &F.d_anodes, anodes));
    {
      std::vector<float> Pf(P.begin(), P.end());
      if (c->precision == ONSAS_PRECISION_FP64) ONSAS_TRY(upload(c, (double**)&F.d_P, P));
      else ONSAS_TRY(upload(c, (float**)&F.d_P, Pf));
    }
    ONSAS_TRY(upload(c, &Lc.d_gptr, gptr));
    ONSAS_TRY(upload(c, &Lc.d_goff, goff));
    ONSAS_TRY(upload(c, &Lc.d_gi, gi));
    ONSAS_TRY(upload(c, &Lc.d_gj, gj));
    ONSAS_TRY(upload(c, &Lc.d_row_ptr, Lc.row_ptr));
    ONSAS_TRY(upload(c, &Lc.d_col, Lc.col));
    ONSAS_TRY(upload(c, &Lc.d_diag, Lc.diag));
    ONSAS_TRY(alloc(c, &Lc.d_A, c->kt * 36 * (size_t)Lc.nnzb));
    ONSAS_TRY(alloc(c, &Lc.d_Dinv, c->kt * 36 * (size_t)na));
    for (void** v : {&Lc.d_b, &Lc.d_x, &Lc.d_r}) ONSAS_TRY(alloc(c, v, c->kt * 6 * (size_t)na));
    const bool last = na <= kMaxCoarseNodes;
    if (!last && (double)na > (double)n / 1.05)
      return set_err(c, ONSAS_ERR_UNSUPPORTED,
                     "multigrid coarsening stalled at " + std::to_string(na) + " nodes");
    A.lev.push_back(std::move(Lc));
    if (last) break;
    // the coarse level becomes the finer one
    rp = A.lev.back().row_ptr;
    cl = A.lev.back().col;
    n = na;
    bs = 6;
    Bcur.swap(Bc);
  }
  ONSAS_TRY(alloc(c, &A.lev[0].d_r, c->kt * 3 * (size_t)nn));
  A.nc = 6 * A.lev.back().n;
  ONSAS_TRY(alloc(c, (void**)&A.d_dense, sizeof(double) * (size_t)A.nc * A.nc));
  ONSAS_TRY(alloc(c, &A.d_Ainv, c->kt * (size_t)A.nc * A.nc));
  ONSAS_TRY(alloc(c, (void**)&A.d_part, sizeof(double) * 1024));
  A.built = true;
  A.stale = true;
  return ONSAS_OK;
}

namespace {

template <typename KT>
int apply_level_op(Ctx* c, int l, const KT* x, const KT* b, KT* r, const PcgState* st) {
  Amg& A = c->amg;
  hipStream_t s = c->stream;
  if (l == 0) {
    const int64_t rstride = c->ell_rounds ? 64 * c->nslices : 64;
    hipLaunchKernelGGL(k_sell_resid<KT>, dim3(nblocks(c->nslices, 4)), dim3(256), 0, s,
                       c->d_sell_ptr, c->d_slice_rounds, rstride, c->d_sell_col,
                       (const KT*)c->d_Ksell, x, b, r, c->nslices, c->nn, st);
  } else {
    const AmgLevel& L = A.lev[l];
    hipLaunchKernelGGL(k_bsr6_resid<KT>, dim3(nblocks(6 * L.n, 256)), dim3(256), 0, s, L.d_row_ptr,
                       L.d_col, (const KT*)L.d_A, x, b, r, L.n, st);
  }
  return ONSAS_OK;
}

template <typename KT>
void smooth(Ctx* c, int l, const KT* v, KT* x, int accumulate, const PcgState* st) {
  Amg& A = c->amg;
  const AmgLevel& L = A.lev[l];
  const KT om = (KT)L.omega;
  if (l == 0)
    hipLaunchKernelGGL((k_smooth<KT, 3>), dim3(nblocks(3 * L.n, 256)), dim3(256), 0, c->stream,
                       (const KT*)c->d_Dinv, v, x, L.n, om, accumulate, st);
  else
    hipLaunchKernelGGL((k_smooth<KT, 6>), dim3(nblocks(6 * L.n, 256)), dim3(256), 0, c->stream,
                       (const KT*)L.d_Dinv, v, x, L.n, om, accumulate, st);
}

// largest eigenvalue of D^-1 A on level l by power iteration (host-driven)
template <typename KT>
int estimate_omega(Ctx* c, int l) {
  Amg& A = c->amg;
  AmgLevel& L = A.lev[l];
  hipStream_t s = c->stream;
  const int64_t nd = (int64_t)L.bs * L.n;
  KT* v = (KT*)(l == 0 ? c->d_pq : L.d_b);  // scratch: no PCG is running
  KT* w = (KT*)(l == 0 ? c->d_pp1 : L.d_x);
  KT* r = (KT*)L.d_r;
  KT* zero = (KT*)(l == 0 ? c->d_pp0 : L.d_r);
  hipLaunchKernelGGL(k_fill_hash<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, v, nd,
                     l == 0 ? c->d_fixmask : nullptr, L.bs);
  std::vector<double> h(1024);
  double lam = 1.0;
  for (int it = 0; it < 12; ++it) {
    // w = D^-1 A v  (A v = 0 - (0 - A v): the residual kernel with b = 0 and x = -v)
    hipLaunchKernelGGL(k_scale<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, v, nd, KT(-1));
    if (l == 0) hipMemsetAsync(zero, 0, nd * sizeof(KT), s);
    else hipMemsetAsync(zero, 0, nd * sizeof(KT), s);
    ONSAS_TRY(apply_level_op<KT>(c, l, v, zero, r, nullptr));  // r = 0 - A(-v) = A v
    hipLaunchKernelGGL(k_scale<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, v, nd, KT(-1));
    const double om_save = L.omega;
    L.omega = 1.0;
    smooth<KT>(c, l, r, w, 0, nullptr);  // w = D^-1 A v
    L.omega = om_save;
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
    // v = w / |w|
    ONSAS_HIP_TRY(c, hipMemcpyAsync(v, w, nd * sizeof(KT), hipMemcpyDeviceToDevice, s));
    hipLaunchKernelGGL(k_scale<KT>, dim3(nblocks(nd, 256)), dim3(256), 0, s, v, nd,
                       (KT)(1.0 / std::sqrt(nw)));
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  L.omega = 1.0 / (1.1 * lam);
  return ONSAS_OK;
}

template <typename KT>
int amg_numeric_t(Ctx* c) {
  Amg& A = c->amg;
  hipStream_t s = c->stream;
  const int nl = (int)A.lev.size();
  for (int l = 1; l < nl; ++l) {
    AmgLevel& L = A.lev[l];
    const AmgLevel& F = A.lev[l - 1];
    if (l == 1)
      hipLaunchKernelGGL((k_galerkin<KT, 3>), dim3(nblocks(L.nnzb, 128)), dim3(128), 0, s, L.d_gptr,
                         L.d_goff, L.d_gi, L.d_gj, (const KT*)c->d_Ksell, (int64_t)64,
                         (const KT*)F.d_P, (KT*)L.d_A, L.nnzb);
    else
      hipLaunchKernelGGL((k_galerkin<KT, 6>), dim3(nblocks(L.nnzb, 128)), dim3(128), 0, s, L.d_gptr,
                         L.d_goff, L.d_gi, L.d_gj, (const KT*)F.d_A, (int64_t)1, (const KT*)F.d_P,
                         (KT*)L.d_A, L.nnzb);
    if (l < nl - 1)
      hipLaunchKernelGGL(k_inv6<KT>, dim3(nblocks(L.n, 128)), dim3(128), 0, s, (KT*)L.d_A, L.d_diag,
                         (KT*)L.d_Dinv, L.n);
  }
  const AmgLevel& Lc = A.lev[nl - 1];
  hipLaunchKernelGGL(k_dense_scatter<KT>, dim3(nblocks(A.nc, 128)), dim3(128), 0, s, Lc.d_row_ptr,
                     Lc.d_col, (const KT*)Lc.d_A, A.d_dense, Lc.n);
  hipLaunchKernelGGL(k_dense_inverse, dim3(1), dim3(256), 0, s, A.d_dense, (int64_t)A.nc);
  hipLaunchKernelGGL(k_to_kt<KT>, dim3(nblocks((int64_t)A.nc * A.nc, 256)), dim3(256), 0, s,
                     A.d_dense, (KT*)A.d_Ainv, (int64_t)A.nc * A.nc);
  ONSAS_HIP_TRY(c, hipGetLastError());
  if (!A.omega_set) {  // lambda_max(D^-1 A) varies little with the state
    for (int l = 0; l < nl - 1; ++l) ONSAS_TRY(estimate_omega<KT>(c, l));
    A.omega_set = true;
  }
  A.stale = false;
  return ONSAS_OK;
}

template <typename KT>
int vcycle_t(Ctx* c, const KT* b0, KT* x0, const PcgState* st) {
  Amg& A = c->amg;
  hipStream_t s = c->stream;
  const int nl = (int)A.lev.size();
  // down: pre-smooth, residual, restrict
  for (int l = 0; l < nl - 1; ++l) {
    AmgLevel& L = A.lev[l];
    const KT* b = l == 0 ? b0 : (const KT*)L.d_b;
    KT* x = l == 0 ? x0 : (KT*)L.d_x;
    smooth<KT>(c, l, b, x, 0, st);
    ONSAS_TRY(apply_level_op<KT>(c, l, x, b, (KT*)L.d_r, st));
    const AmgLevel& C = A.lev[l + 1];
    if (l == 0)
      hipLaunchKernelGGL((k_restrict<KT, 3>), dim3(nblocks(6 * C.n, 256)), dim3(256), 0, s, L.d_aptr,
                         L.d_anodes, (const KT*)L.d_P, (const KT*)L.d_r, (KT*)C.d_b, C.n, st);
    else
      hipLaunchKernelGGL((k_restrict<KT, 6>), dim3(nblocks(6 * C.n, 256)), dim3(256), 0, s, L.d_aptr,
                         L.d_anodes, (const KT*)L.d_P, (const KT*)L.d_r, (KT*)C.d_b, C.n, st);
  }
  // coarsest: exact
  const AmgLevel& Lc = A.lev[nl - 1];
  hipLaunchKernelGGL(k_dense_apply<KT>, dim3(nblocks(A.nc, 64)), dim3(64), 0, s,
                     (const KT*)A.d_Ainv, (const KT*)Lc.d_b, (KT*)Lc.d_x, (int64_t)A.nc, st);
  // up: prolong, post-smooth
  for (int l = nl - 2; l >= 0; --l) {
    AmgLevel& L = A.lev[l];
    const KT* b = l == 0 ? b0 : (const KT*)L.d_b;
    KT* x = l == 0 ? x0 : (KT*)L.d_x;
    const AmgLevel& C = A.lev[l + 1];
    if (l == 0)
      hipLaunchKernelGGL((k_prolong<KT, 3>), dim3(nblocks(3 * L.n, 256)), dim3(256), 0, s, L.d_agg,
                         (const KT*)L.d_P, (const KT*)C.d_x, x, L.n, st);
    else
      hipLaunchKernelGGL((k_prolong<KT, 6>), dim3(nblocks(6 * L.n, 256)), dim3(256), 0, s, L.d_agg,
                         (const KT*)L.d_P, (const KT*)C.d_x, x, L.n, st);
    ONSAS_TRY(apply_level_op<KT>(c, l, x, b, (KT*)L.d_r, st));
    smooth<KT>(c, l, (const KT*)L.d_r, x, 1, st);
  }
  ONSAS_HIP_TRY(c, hipGetLastError());
  return ONSAS_OK;
}

}  // namespace

int amg_numeric(Ctx* c) {
  return c->precision == ONSAS_PRECISION_FP64 ? amg_numeric_t<double>(c) : amg_numeric_t<float>(c);
}

int launch_vcycle(Ctx* c, const void* b, void* x, const PcgState* st) {
  return c->precision == ONSAS_PRECISION_FP64
             ? vcycle_t<double>(c, (const double*)b, (double*)x, st)
             : vcycle_t<float>(c, (const float*)b, (float*)x, st);
}

void amg_free(Ctx* c) {
  Amg& A = c->amg;
  for (AmgLevel& L : A.lev)
    for (void* p : {(void*)L.d_row_ptr, (void*)L.d_col, (void*)L.d_diag, L.d_A, L.d_Dinv, L.d_b, L.d_x,
                    L.d_r, (void*)L.d_agg, L.d_P, (void*)L.d_aptr, (void*)L.d_anodes,
                    (void*)L.d_gptr, (void*)L.d_goff, (void*)L.d_gi, (void*)L.d_gj})
      if (p) (void)hipFree(p);
  for (void* p : {(void*)A.d_dense, A.d_Ainv, (void*)A.d_part})
    if (p) (void)hipFree(p);
  A = Amg{};
}

}  // namespace onsas
