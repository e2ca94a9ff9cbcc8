// Batched LOBPCG for K x = lambda M x on node-block matrices (FP64), and a batched Jacobi symmetric
// eigensolver for the small Rayleigh-Ritz problems.
//
// Variants share one sparsity pattern (bptr/bcol); each has its own K values (and optionally its own
// M values and block-Jacobi preconditioner).  Per variant and iteration (Knyazev's LOBPCG):
//   K1 k_lob_resid   R = K X - M X diag(lambda), W = T R, column norms of R and M X (partials);
//                    with the two-level preconditioner k_lob_resid_agg, which also forms P^T R
//   K2 k_lob_spmm    K W and M W in one pass over the shared pattern
//   K3 k_lob_gram    partial Gram matrices S^T K S, S^T M S of the basis S = [X W P]
//   K4 k_lob_gsum    fixed-order sum of the partials
//   K5 k_lob_rr      one workgroup: convergence test, then Rayleigh-Ritz with a rank-revealing
//                    M-orthonormalisation of S (eigen-decomposition of the scaled S^T M S, directions
//                    below a relative threshold dropped), both small eigenproblems by parallel cyclic
//                    Jacobi in LDS
//   K6 k_lob_update_rows X' = S C, P' = [W P] C_WP, with the K- and M-images recombined the same way
//   K7 k_lob_flip    swap the double-buffered blocks, advance the iteration
// Block vectors are stored row-major [n][m] (the m columns of one DOF row contiguous), so every
// kernel reads whole rows with coalesced loads.  The host polls a status word one chunk behind.
#include <algorithm>

#include "fe_internal.h"

namespace {

constexpr int kMaxBlock = 32;
constexpr int kMaxBasis = 3 * kMaxBlock;
constexpr int kThreads = 256;
constexpr int kRRThreads = 1024;
constexpr int kRefresh = 32;  // iterations between explicit K X, M X, K P, M P
constexpr int kNumBlocks = 15;  // X0 X1 KX0 KX1 MX0 MX1 P0 P1 KP0 KP1 MP0 MP1 W KW MW
enum { B_X = 0, B_KX = 2, B_MX = 4, B_P = 6, B_KP = 8, B_MP = 10, B_W = 12, B_KW = 13, B_MW = 14 };

struct LobState {
  int32_t it, done, cur, init;
  int32_t ns, kept, nconv, restart;  // restart: drop P from the next iteration's basis
  double lambda[kMaxBlock];
  double resid[kMaxBlock];
};

struct LobStatus {
  int32_t active, max_it, failed, unconverged;
};

struct LobParams {
  int64_t nb, n, nnzb, nm;  // node rows, DOF rows, blocks, n*m
  int32_t m, n_wanted, largest, max_iter;
  double tol, drop;
  int32_t nchunk_g, nchunk_r, nvar, pad;
};

__device__ inline double* lob_block(double* base, const LobParams& p, int v, int idx) {
  return base + ((int64_t)v * kNumBlocks + idx) * p.nm;
}

// --------------------------------------------------------------------------------------------------
// K1: residual block, preconditioned residual, column norms
// --------------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kThreads) void k_lob_resid(LobParams p, const double* __restrict__ T,
                                                        int64_t t_stride,
                                                        const uint8_t* __restrict__ fixed,
                                                        const LobState* st, double* blk,
                                                        double* __restrict__ rpart) {
  const int v = blockIdx.y;
  __shared__ int s_skip, s_cur;
  __shared__ double s_lam[kMaxBlock];
  __shared__ int s_lock[kMaxBlock];
  __shared__ double s_rr[kThreads], s_bb[kThreads];
  if (threadIdx.x == 0) {
    s_skip = st[v].done;
    s_cur = st[v].cur;
  }
  if (threadIdx.x < p.m) {
    s_lam[threadIdx.x] = st[v].lambda[threadIdx.x];
    // soft locking: a wanted pair already converged ten times below the tolerance (at the last
    // check) contributes no search direction; it stays in X and in the Rayleigh-Ritz problem
    s_lock[threadIdx.x] = !st[v].init && threadIdx.x < p.n_wanted &&
                          st[v].resid[threadIdx.x] <= 0.1 * p.tol;
  }
  __syncthreads();
  if (s_skip) return;
  const int m = p.m;
  const int npg = kThreads / m;
  const int g = threadIdx.x / m, j = threadIdx.x % m;
  const double* KX = lob_block(blk, p, v, B_KX + s_cur);
  const double* MX = lob_block(blk, p, v, B_MX + s_cur);
  double* W = lob_block(blk, p, v, B_W);
  const double* Tv = T ? T + (int64_t)v * t_stride : nullptr;
  const int64_t per = ceil_div(p.nb, (int64_t)gridDim.x);
  const int64_t lo = min(p.nb, (int64_t)blockIdx.x * per), hi = min(p.nb, lo + per);
  double rr = 0.0, bb = 0.0;
  if (g < npg) {
    const double lam = s_lam[j];
    for (int64_t i = lo + g; i < hi; i += npg) {
      double R[D];
#pragma unroll
      for (int e = 0; e < D; ++e) {
        const int64_t at = (i * D + e) * m + j;
        const double mx = MX[at];
        R[e] = KX[at] - lam * mx;
        rr += R[e] * R[e];
        bb += mx * mx;
      }
#pragma unroll
      for (int e = 0; e < D; ++e) {
        double w = R[e];
        if (Tv) {
          w = 0.0;
#pragma unroll
          for (int f = 0; f < D; ++f) w += Tv[(i * D + e) * D + f] * R[f];
        }
        if ((fixed && fixed[i * D + e]) || s_lock[j]) w = 0.0;
        W[(i * D + e) * m + j] = w;
      }
    }
  }
  s_rr[threadIdx.x] = rr;
  s_bb[threadIdx.x] = bb;
  __syncthreads();
  if (threadIdx.x < m) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < npg; ++q) {
      a += s_rr[q * m + threadIdx.x];
      b += s_bb[q * m + threadIdx.x];
    }
    double* out = rpart + ((int64_t)v * gridDim.x + blockIdx.x) * 2 * m;
    out[threadIdx.x] = a;
    out[m + threadIdx.x] = b;
  }
}

// --------------------------------------------------------------------------------------------------
// K1b (two-level preconditioner): W += P A_c^-1 P^T R, for every column of the block.  P rows of
// constrained DOFs are zero, so W stays zero there.
// --------------------------------------------------------------------------------------------------
// K1 and K1b's restriction in one pass (two-level preconditioner): chunk b of the residual partials
// takes aggregates [b n_agg / chunks, (b+1) n_agg / chunks) and walks their nodes (every node lies in
// one aggregate), forming R, W = T R and the column norms as k_lob_resid does and r_c = P^T R (thread
// (group g, column j) sums an aggregate's nodes g, g + npg, ..., then the groups are added in a
// fixed order), so K X and M X are read once per iteration.
template <int D, int NC>
__global__ __launch_bounds__(kThreads) void k_lob_resid_agg(LobParams p, fe_coarse C, int64_t ld,
                                                            const double* __restrict__ T, int64_t t_stride,
                                                            const uint8_t* __restrict__ fixed,
                                                            const LobState* st, double* blk,
                                                            double* __restrict__ rpart,
                                                            double* __restrict__ rc) {
  const int v = blockIdx.y;
  __shared__ int s_skip, s_cur;
  __shared__ double s_lam[kMaxBlock];
  __shared__ int s_lock[kMaxBlock];
  __shared__ double s_rr[kThreads], s_bb[kThreads];
  __shared__ double s_acc[kThreads][NC];
  if (threadIdx.x == 0) {
    s_skip = st[v].done;
    s_cur = st[v].cur;
  }
  if (threadIdx.x < p.m) {
    s_lam[threadIdx.x] = st[v].lambda[threadIdx.x];
    s_lock[threadIdx.x] = !st[v].init && threadIdx.x < p.n_wanted &&
                          st[v].resid[threadIdx.x] <= 0.1 * p.tol;
  }
  __syncthreads();
  if (s_skip) return;
  const int m = p.m;
  const int npg = kThreads / m;
  const int g = threadIdx.x / m, j = threadIdx.x % m;
  const double* KX = lob_block(blk, p, v, B_KX + s_cur);
  const double* MX = lob_block(blk, p, v, B_MX + s_cur);
  double* W = lob_block(blk, p, v, B_W);
  const double* Tv = T ? T + (int64_t)v * t_stride : nullptr;
  const int64_t alo = C.n_agg * blockIdx.x / gridDim.x, ahi = C.n_agg * (blockIdx.x + 1) / gridDim.x;
  const double lam = s_lam[j];
  double rr = 0.0, bb = 0.0;
  for (int64_t agg = alo; agg < ahi; ++agg) {
    double acc[NC];
#pragma unroll
    for (int c = 0; c < NC; ++c) acc[c] = 0.0;
    if (g < npg) {
      const int32_t a0 = C.agg_ptr[agg], a1 = C.agg_ptr[agg + 1];
      for (int32_t k = a0 + g; k < a1; k += npg) {
        const int64_t i = C.agg_nodes[k];
        double R[D];
#pragma unroll
        for (int e = 0; e < D; ++e) {
          const int64_t at = (i * D + e) * m + j;
          const double mx = MX[at];
          R[e] = KX[at] - lam * mx;
          rr += R[e] * R[e];
          bb += mx * mx;
          const double* Pr = C.P + (i * D + e) * NC;
#pragma unroll
          for (int c = 0; c < NC; ++c) acc[c] += Pr[c] * R[e];
        }
#pragma unroll
        for (int e = 0; e < D; ++e) {
          double w = R[e];
          if (Tv) {
            w = 0.0;
#pragma unroll
            for (int f = 0; f < D; ++f) w += Tv[(i * D + e) * D + f] * R[f];
          }
          if ((fixed && fixed[i * D + e]) || s_lock[j]) w = 0.0;
          W[(i * D + e) * m + j] = w;
        }
      }
    }
#pragma unroll
    for (int c = 0; c < NC; ++c) s_acc[threadIdx.x][c] = acc[c];
    __syncthreads();
    if (threadIdx.x < m * NC) {
      const int jj = threadIdx.x % m, c = threadIdx.x / m;
      double sum = 0.0;
      for (int q = 0; q < npg; ++q) sum += s_acc[q * m + jj][c];
      rc[((int64_t)v * ld + agg * NC + c) * m + jj] = sum;
    }
    __syncthreads();
  }
  s_rr[threadIdx.x] = rr;
  s_bb[threadIdx.x] = bb;
  __syncthreads();
  if (threadIdx.x < m) {
    double a = 0.0, b = 0.0;
    for (int q = 0; q < npg; ++q) {
      a += s_rr[q * m + threadIdx.x];
      b += s_bb[q * m + threadIdx.x];
    }
    double* out = rpart + ((int64_t)v * gridDim.x + blockIdx.x) * 2 * m;
    out[threadIdx.x] = a;
    out[m + threadIdx.x] = b;
  }
}

// z_c = A_c^-1 r_c per variant: a workgroup computes 64 rows x m columns, staging 32-wide slabs of
// A_c^-1 and r_c in LDS.
constexpr int kCgRows = 64, kCgK = 32;
__global__ __launch_bounds__(kThreads) void k_lob_cgemm(LobParams p, int64_t ld, const LobState* st,
                                                        const double* __restrict__ Ainv,
                                                        int64_t ainv_stride,
                                                        const double* __restrict__ rc,
                                                        double* __restrict__ zc) {
  const int v = blockIdx.y;
  if (st[v].done) return;
  __shared__ double s_a[kCgRows][kCgK + 1];
  __shared__ double s_r[kCgK][kMaxBlock];
  const int m = p.m;
  const int64_t r0 = (int64_t)blockIdx.x * kCgRows;
  const double* A = Ainv + (int64_t)v * ainv_stride;
  const double* R = rc + (int64_t)v * ld * m;
  constexpr int kPer = (kCgRows * kMaxBlock + kThreads - 1) / kThreads;  // outputs per thread (max)
  double acc[kPer];
#pragma unroll
  for (int q = 0; q < kPer; ++q) acc[q] = 0.0;
  for (int64_t k0 = 0; k0 < ld; k0 += kCgK) {
    for (int t = threadIdx.x; t < kCgRows * kCgK; t += kThreads) {
      const int rr = t / kCgK, kk = t % kCgK;
      s_a[rr][kk] = A[(r0 + rr) * ld + k0 + kk];
    }
    for (int t = threadIdx.x; t < kCgK * m; t += kThreads) s_r[t / m][t % m] = R[(k0 + t / m) * m + t % m];
    __syncthreads();
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int o = threadIdx.x + q * kThreads;
      if (o < kCgRows * m) {
        const int rr = o / m, jj = o % m;
        double s = acc[q];
#pragma unroll 8
        for (int kk = 0; kk < kCgK; ++kk) s += s_a[rr][kk] * s_r[kk][jj];
        acc[q] = s;
      }
    }
    __syncthreads();
  }
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int o = threadIdx.x + q * kThreads;
    if (o < kCgRows * m) zc[((int64_t)v * ld + r0 + o / m) * m + o % m] = acc[q];
  }
}

// W += P z_c: one thread per (DOF row, column).
template <int D, int NC>
__global__ __launch_bounds__(kThreads) void k_lob_cprolong(LobParams p, fe_coarse C, int64_t ld,
                                                           const LobState* st,
                                                           const double* __restrict__ zc, double* blk) {
  const int v = blockIdx.y;
  if (st[v].done) return;
  // a thread owns 4 consecutive columns of one DOF row (m is a multiple of 4): P's row is read once
  // per 4 outputs, z and W as 16-byte pairs
  const int m = p.m, tpr = m / 4, npg = kThreads / tpr;
  const int g = threadIdx.x / tpr, j0 = (threadIdx.x - g * tpr) * 4;
  const int64_t row = (int64_t)blockIdx.x * npg + g;
  if (g >= npg || row >= p.n) return;
  const int64_t agg = C.node_agg[row / D];
  const double* Pr = C.P + row * NC;
  const double* z = zc + ((int64_t)v * ld + agg * NC) * m + j0;
  double s[4] = {0.0, 0.0, 0.0, 0.0};
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    const double pc = Pr[c];
    const double2 a = reinterpret_cast<const double2*>(z + c * m)[0];
    const double2 b = reinterpret_cast<const double2*>(z + c * m)[1];
    s[0] += pc * a.x;
    s[1] += pc * a.y;
    s[2] += pc * b.x;
    s[3] += pc * b.y;
  }
  double2* Wr = reinterpret_cast<double2*>(lob_block(blk, p, v, B_W) + row * m + j0);
  const double2 w0 = Wr[0], w1 = Wr[1];
  Wr[0] = make_double2(w0.x + s[0], w0.y + s[1]);
  Wr[1] = make_double2(w1.x + s[2], w1.y + s[3]);
}

// --------------------------------------------------------------------------------------------------
// K2: K Y and M Y for the block Y = X (start) or W, one thread per (node row, column)
// --------------------------------------------------------------------------------------------------
// LAYOUT: FE_LAYOUT_BSR keeps a block's d*d values together, so the lanes of a row walk their blocks
// through contiguous memory (the planes layout spreads each step over d*d lines of the L1).
template <int D, int LAYOUT>
__global__ __launch_bounds__(kThreads) void k_lob_spmm(LobParams p, const int32_t* __restrict__ bptr,
                                                       const int32_t* __restrict__ bcol,
                                                       const double* __restrict__ Kv, int64_t k_stride,
                                                       const double* __restrict__ Mv, int64_t m_stride,
                                                       const LobState* st, double* blk, int src) {
  const int v = blockIdx.y;
  __shared__ int s_skip, s_cur;
  if (threadIdx.x == 0) {
    s_skip = st[v].done;
    s_cur = st[v].cur;
  }
  __syncthreads();
  if (s_skip) return;
  // a thread owns 4 consecutive columns of one node row (m is a multiple of 4)
  const int m = p.m, tpr = m / 4;
  const int npg = kThreads / tpr;
  const int g = threadIdx.x / tpr, j0 = (threadIdx.x % tpr) * 4;
  const int64_t i = (int64_t)blockIdx.x * npg + g;
  if (g >= npg || i >= p.nb) return;
  // src 0: W -> KW, MW; 1: the current X -> KX, MX; 2: the current P -> KP, MP
  const int yb = src == 0 ? B_W : (src == 1 ? B_X : B_P) + s_cur;
  const int kb = src == 0 ? B_KW : (src == 1 ? B_KX : B_KP) + s_cur;
  const int mb = src == 0 ? B_MW : (src == 1 ? B_MX : B_MP) + s_cur;
  const double* Y = lob_block(blk, p, v, yb);
  double* KY = lob_block(blk, p, v, kb);
  double* MY = lob_block(blk, p, v, mb);
  const double* K = Kv + (int64_t)v * k_stride;
  const double* M = Mv + (int64_t)v * m_stride;
  double ak[D][4], am[D][4];
#pragma unroll
  for (int e = 0; e < D; ++e)
#pragma unroll
    for (int u = 0; u < 4; ++u) ak[e][u] = am[e][u] = 0.0;
  const int32_t q0 = bptr[i], q1 = bptr[i + 1];
  for (int32_t q = q0; q < q1; ++q) {
    const int64_t c = bcol[q];
    double y[D][4];
#pragma unroll
    for (int f = 0; f < D; ++f) {
      const double2* src = reinterpret_cast<const double2*>(Y + (c * D + f) * m + j0);
      const double2 a = src[0], b = src[1];
      y[f][0] = a.x;
      y[f][1] = a.y;
      y[f][2] = b.x;
      y[f][3] = b.y;
    }
#pragma unroll
    for (int e = 0; e < D; ++e)
#pragma unroll
      for (int f = 0; f < D; ++f) {
        const int64_t at = LAYOUT == FE_LAYOUT_PLANES ? (int64_t)(e * D + f) * p.nnzb + q
                                                      : (int64_t)q * D * D + e * D + f;
        const double kv = K[at], mv = M[at];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ak[e][u] += kv * y[f][u];
          am[e][u] += mv * y[f][u];
        }
      }
  }
#pragma unroll
  for (int e = 0; e < D; ++e) {
    double2* dk = reinterpret_cast<double2*>(KY + (i * D + e) * m + j0);
    double2* dm = reinterpret_cast<double2*>(MY + (i * D + e) * m + j0);
    dk[0] = make_double2(ak[e][0], ak[e][1]);
    dk[1] = make_double2(ak[e][2], ak[e][3]);
    dm[0] = make_double2(am[e][0], am[e][1]);
    dm[1] = make_double2(am[e][2], am[e][3]);
  }
}

// The same product for the block-contiguous layout with the workgroup's K blocks staged in LDS.
// Its rows' blocks are one contiguous range of the value array, so a 128-thread workgroup loads them
// with coalesced 8-byte loads (each block once) instead of every lane of a row loading each value of
// every block (d^2 scattered lines per wavefront load).  M, one matrix shared by every variant, is
// read directly (it stays in the caches); staging it too doubled the LDS per workgroup and cost more
// in occupancy than it saved.  Blocks beyond the LDS capacity (rows of unusually high degree) are
// read from HBM; the per-row sum order is the same, so the result is bitwise that of k_lob_spmm.
constexpr int kSpmmThreads = 128;
constexpr int kSpmmLdsVals = 2448;  // 272 blocks of 3 x 3 (19.1 KB)

template <int D>
__device__ __forceinline__ void spmm_block(const double* __restrict__ kb, const double* __restrict__ mb,
                                           const double (&y)[D][4], double (&ak)[D][4],
                                           double (&am)[D][4]) {
#pragma unroll
  for (int e = 0; e < D; ++e)
#pragma unroll
    for (int f = 0; f < D; ++f) {
      const double kv = kb[e * D + f], mv = mb[e * D + f];
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        ak[e][u] += kv * y[f][u];
        am[e][u] += mv * y[f][u];
      }
    }
}

template <int D>
__device__ __forceinline__ void spmm_load_y(const double* __restrict__ Y, int64_t c, int m, int j0,
                                            double (&y)[D][4]) {
#pragma unroll
  for (int f = 0; f < D; ++f) {
    const double2* src = reinterpret_cast<const double2*>(Y + (c * D + f) * m + j0);
    const double2 a = src[0], b = src[1];
    y[f][0] = a.x;
    y[f][1] = a.y;
    y[f][2] = b.x;
    y[f][3] = b.y;
  }
}

template <int D>
__global__ __launch_bounds__(kSpmmThreads) void k_lob_spmm_lds(LobParams p, const int32_t* __restrict__ bptr,
                                                               const int32_t* __restrict__ bcol,
                                                               const double* __restrict__ Kv,
                                                               int64_t k_stride,
                                                               const double* __restrict__ Mv,
                                                               int64_t m_stride, const LobState* st,
                                                               double* blk, int src) {
  constexpr int DD = D * D;
  constexpr int kCap = kSpmmLdsVals / DD;
  __shared__ double sK[kCap * DD];
  __shared__ int32_t sCol[kCap];
  __shared__ int s_skip, s_cur;
  const int v = blockIdx.y;
  if (threadIdx.x == 0) {
    s_skip = st[v].done;
    s_cur = st[v].cur;
  }
  __syncthreads();
  if (s_skip) return;
  const int m = p.m, tpr = m / 4;
  const int npg = kSpmmThreads / tpr;
  const int64_t r0 = (int64_t)blockIdx.x * npg;
  const int64_t r1 = r0 + npg < p.nb ? r0 + npg : p.nb;
  const int32_t qa = bptr[r0], qb = bptr[r1];
  const int nst = qb - qa < kCap ? qb - qa : kCap;
  const double* K = Kv + (int64_t)v * k_stride;
  const double* M = Mv + (int64_t)v * m_stride;
  for (int t = threadIdx.x; t < nst * DD; t += kSpmmThreads) {
    sK[t] = K[(int64_t)qa * DD + t];
  }
  for (int t = threadIdx.x; t < nst; t += kSpmmThreads) sCol[t] = bcol[qa + t];
  __syncthreads();
  const int g = threadIdx.x / tpr, j0 = (threadIdx.x % tpr) * 4;
  const int64_t i = r0 + g;
  if (g >= npg || i >= p.nb) return;
  const int yb = src == 0 ? B_W : (src == 1 ? B_X : B_P) + s_cur;
  const int kb = src == 0 ? B_KW : (src == 1 ? B_KX : B_KP) + s_cur;
  const int mb = src == 0 ? B_MW : (src == 1 ? B_MX : B_MP) + s_cur;
  const double* Y = lob_block(blk, p, v, yb);
  double* KY = lob_block(blk, p, v, kb);
  double* MY = lob_block(blk, p, v, mb);
  double ak[D][4], am[D][4];
#pragma unroll
  for (int e = 0; e < D; ++e)
#pragma unroll
    for (int u = 0; u < 4; ++u) ak[e][u] = am[e][u] = 0.0;
  const int32_t q0 = bptr[i], q1 = bptr[i + 1];
  const int32_t qs = q1 < qa + nst ? q1 : qa + nst;  // end of the row's staged blocks
  for (int32_t q = q0; q < qs; ++q) {
    const int l = q - qa;
    double y[D][4];
    spmm_load_y<D>(Y, sCol[l], m, j0, y);
    spmm_block<D>(sK + l * DD, M + (int64_t)q * DD, y, ak, am);
  }
  for (int32_t q = q0 > qs ? q0 : qs; q < q1; ++q) {
    double y[D][4];
    spmm_load_y<D>(Y, bcol[q], m, j0, y);
    spmm_block<D>(K + (int64_t)q * DD, M + (int64_t)q * DD, y, ak, am);
  }
#pragma unroll
  for (int e = 0; e < D; ++e) {
    double2* dk = reinterpret_cast<double2*>(KY + (i * D + e) * m + j0);
    double2* dm = reinterpret_cast<double2*>(MY + (i * D + e) * m + j0);
    dk[0] = make_double2(ak[e][0], ak[e][1]);
    dk[1] = make_double2(ak[e][2], ak[e][3]);
    dm[0] = make_double2(am[e][0], am[e][1]);
    dm[1] = make_double2(am[e][2], am[e][3]);
  }
}

// --------------------------------------------------------------------------------------------------
// K3: partial Gram matrices over a range of DOF rows on the FP64 matrix cores.  The basis S = [X W P]
// (ns <= 96 columns) is cut into 16-column tiles and every upper tile pair (I <= J) of S^T (K S) and
// S^T (M S) gets one 16x16x4 MFMA accumulator.  A workgroup of four wavefronts takes one row range:
// wavefronts 0, 1 form the K Gram, 2, 3 the M Gram; the first of each pair takes the tile pairs with
// J <= 3 (10), the second those with J >= 4 (11).  For v_mfma_f64_16x16x4f64, lane l holds
// A[l % 16][l / 16] and B[l / 16][l % 16], and component c of the result is D[4c + l / 16][l % 16];
// with A = S^T and B = (K S) both operands are the element (row l / 16, column 16 T + l % 16) of a
// 4-row step, so each lane loads its operands straight from HBM (16 consecutive doubles of a row per
// 16 lanes) and nothing is staged in LDS.  Diagonal tiles are stored whole, the others above the
// diagonal only.
// --------------------------------------------------------------------------------------------------
__device__ inline const double* basis_col_src(const double* X, const double* W, const double* P,
                                              int c, int m, int& cc) {
  if (c < m) {
    cc = c;
    return X;
  }
  if (c < 2 * m) {
    cc = c - m;
    return W;
  }
  cc = c - 2 * m;
  return P;
}

constexpr int kGramTile = 16;
constexpr int kGramTiles = kMaxBasis / kGramTile;  // 6
constexpr int kGramThreads = 256;
constexpr int kGramSteps = 3;
constexpr int kGramStepsW = 2;  // the W-column Gram: 2 steps in flight keep it at 3 waves per SIMD

typedef double GramAcc __attribute__((ext_vector_type(4)));

// Tiles of one Gram wavefront: S tiles 0 .. NI-1 and image tiles J0 .. J0+NJ-1.
template <int H>
struct GramHalf {
  static constexpr int NI = H ? 6 : 4, J0 = H ? 4 : 0, NJ = H ? 2 : 4;
};

template <int H>
__device__ __forceinline__ void gram_rows(const double* __restrict__ base, int ns, int nT, int m,
                                          const int (&bS)[3], const int (&bY)[3], int64_t nm,
                                          int64_t lo, int64_t hi, double* __restrict__ out, int NS) {
  constexpr int NI = GramHalf<H>::NI, J0 = GramHalf<H>::J0, NJ = GramHalf<H>::NJ;
  const int lane = threadIdx.x & 63, rr = lane >> 4, cl = lane & 15;
  // offsets inside the variant's blocks of this lane's column in each tile; a column past the basis
  // is masked (never read)
  int64_t offS[NI], offY[NJ];
  bool okS[NI], okY[NJ];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int c = i * kGramTile + cl;
    okS[i] = c < ns;
    offS[i] = okS[i] ? bS[c / m] * nm + c % m : 0;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int c = (J0 + j) * kGramTile + cl;
    okY[j] = c < ns;
    offY[j] = okY[j] ? bY[c / m] * nm + c % m : 0;
  }
  GramAcc acc[NI][NJ];
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = GramAcc{0.0, 0.0, 0.0, 0.0};
  // kGramSteps 4-row MFMA steps per iteration; the next iteration's operands are loaded before this
  // one's products, so their loads are in flight while the matrix cores work
  double sa[kGramSteps][NI], yb[kGramSteps][NJ], sn[kGramSteps][NI], yn[kGramSteps][NJ];
  auto load = [&](int64_t r0, double(&s)[kGramSteps][NI], double(&y)[kGramSteps][NJ])
                  __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < kGramSteps; ++k) {
      const int64_t row = r0 + 4 * k + rr;
      const bool rok = row < hi;
      const int64_t ro = row * m;
#pragma unroll
      for (int i = 0; i < NI; ++i) s[k][i] = (rok && okS[i]) ? base[offS[i] + ro] : 0.0;
#pragma unroll
      for (int j = 0; j < NJ; ++j) y[k][j] = (rok && okY[j]) ? base[offY[j] + ro] : 0.0;
    }
  };
  load(lo, sa, yb);
  for (int64_t r0 = lo; r0 < hi; r0 += 4 * kGramSteps) {
    load(r0 + 4 * kGramSteps, sn, yn);  // the next iteration's operands (zeros past hi)
#pragma unroll
    for (int k = 0; k < kGramSteps; ++k)
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          if (i <= J0 + j && J0 + j < nT)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[k][i], yb[k][j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int k = 0; k < kGramSteps; ++k) {
#pragma unroll
      for (int i = 0; i < NI; ++i) sa[k][i] = sn[k][i];
#pragma unroll
      for (int j = 0; j < NJ; ++j) yb[k][j] = yn[k][j];
    }
  }
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      if (i > J0 + j || J0 + j >= nT) continue;
      const int col = (J0 + j) * kGramTile + cl;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int row = i * kGramTile + 4 * c + rr;
        if (row < NS && col < NS) out[row * NS + col] = acc[i][j][c];
      }
    }
}

__global__ __launch_bounds__(kGramThreads, 2) void k_lob_gram(LobParams p, const LobState* st,
                                                           const double* blk,
                                                           double* __restrict__ gpart) {
  const int v = blockIdx.y, chunk = blockIdx.x;
  if (st[v].done) return;
  const int m = p.m, ns = st[v].ns, cur = st[v].cur;
  const int nT = (ns + kGramTile - 1) / kGramTile;
  const int wave = threadIdx.x >> 6, img = wave >> 1, half = wave & 1;  // img 0: K, 1: M
  if (half && nT <= GramHalf<1>::J0) return;  // no tile pair with J >= 4
  const double* base = blk + (int64_t)v * kNumBlocks * p.nm;
  const int bS[3] = {B_X + cur, B_W, B_P + cur};
  const int bY[3] = {(img ? B_MX : B_KX) + cur, img ? B_MW : B_KW, (img ? B_MP : B_KP) + cur};
  const int64_t per = ceil_div(ceil_div(p.n, (int64_t)p.nchunk_g), 4 * kGramSteps) * 4 * kGramSteps;
  const int64_t lo = min(p.n, (int64_t)chunk * per), hi = min(p.n, lo + per);
  const int NS = 3 * m;
  double* out = gpart + (((int64_t)v * p.nchunk_g + chunk) * 2 + img) * NS * NS;
  if (half)
    gram_rows<1>(base, ns, nT, m, bS, bY, p.nm, lo, hi, out, NS);
  else
    gram_rows<0>(base, ns, nT, m, bS, bY, p.nm, lo, hi, out, NS);
}

// K4: fixed-order sum of the Gram partials into G[v][2][NS][NS] (both triangles).
__global__ __launch_bounds__(kThreads) void k_lob_gsum(LobParams p, const LobState* st,
                                                       const double* __restrict__ gpart,
                                                       double* __restrict__ G) {
  const int v = blockIdx.y;
  if (st[v].done) return;
  const int NS = 3 * p.m, ns = st[v].ns;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= 2 * NS * NS) return;
  const int mat = e / (NS * NS), a = (e % (NS * NS)) / NS, b = e % NS;
  if (a >= ns || b >= ns) return;
  // below the diagonal only the diagonal tiles hold entries; elsewhere read the upper entry
  const int at = a / kGramTile, bt = b / kGramTile;
  const int ra = (at > bt) ? b : a, rb = (at > bt) ? a : b;
  double s = 0.0;
  const double* src = gpart + (int64_t)v * p.nchunk_g * 2 * NS * NS + mat * NS * NS + ra * NS + rb;
  for (int c = 0; c < p.nchunk_g; ++c) s += src[(int64_t)c * 2 * NS * NS];
  G[((int64_t)v * 2 + mat) * NS * NS + a * NS + b] = s;
}

// --------------------------------------------------------------------------------------------------
// K3' / K4': the Gram matrices of S = [X W P] between refreshes.  X = S_old C and P = S_old C0 (C0:
// C with its X rows zeroed) are combinations of the previous basis, so their blocks X^T K X,
// X^T K P, P^T K P (and the M ones) are C^T G_old C, C^T G_old C0, C0^T G_old C0, which k_lob_rr
// forms from the previous Gram (Gp: [2][2m][2m], rows and columns X then P).  Only the columns of
// the new W are summed over the rows here: S^T (K W) and S^T (M W), reading S, KW and MW -- five
// blocks instead of the nine (S and both images) of the full Gram.  Same MFMA layout as k_lob_gram;
// wavefronts 0, 1 form K W, 2, 3 M W, each taking three 16-column tiles of S against the two tiles
// of W's image.  Every kRefresh iterations (when the images of X and P are recomputed) the full
// Gram runs instead, so the recurrence restarts from explicit products.
// --------------------------------------------------------------------------------------------------
__device__ __forceinline__ void gram_rows_w(const double* __restrict__ base, int ns, int m, int half,
                                            const int (&bS)[3], int bY, int64_t nm, int64_t lo,
                                            int64_t hi, double* __restrict__ out) {
  constexpr int NI = 3, NJ = 2;
  const int lane = threadIdx.x & 63, rr = lane >> 4, cl = lane & 15;
  int64_t offS[NI], offY[NJ];
  bool okS[NI], okY[NJ];
#pragma unroll
  for (int i = 0; i < NI; ++i) {
    const int c = (half * NI + i) * kGramTile + cl;
    okS[i] = c < ns;
    offS[i] = okS[i] ? bS[c / m] * nm + c % m : 0;
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int c = j * kGramTile + cl;
    okY[j] = c < m;
    offY[j] = okY[j] ? (int64_t)bY * nm + c : 0;
  }
  GramAcc acc[NI][NJ];
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) acc[i][j] = GramAcc{0.0, 0.0, 0.0, 0.0};
  double sa[kGramStepsW][NI], yb[kGramStepsW][NJ], sn[kGramStepsW][NI], yn[kGramStepsW][NJ];
  auto load = [&](int64_t r0, double(&sv)[kGramStepsW][NI], double(&yv)[kGramStepsW][NJ])
                  __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < kGramStepsW; ++k) {
      const int64_t row = r0 + 4 * k + rr;
      const bool rok = row < hi;
      const int64_t ro = row * m;
#pragma unroll
      for (int i = 0; i < NI; ++i) sv[k][i] = (rok && okS[i]) ? base[offS[i] + ro] : 0.0;
#pragma unroll
      for (int j = 0; j < NJ; ++j) yv[k][j] = (rok && okY[j]) ? base[offY[j] + ro] : 0.0;
    }
  };
  load(lo, sa, yb);
  for (int64_t r0 = lo; r0 < hi; r0 += 4 * kGramStepsW) {
    load(r0 + 4 * kGramStepsW, sn, yn);
#pragma unroll
    for (int k = 0; k < kGramStepsW; ++k)
#pragma unroll
      for (int i = 0; i < NI; ++i)
#pragma unroll
        for (int j = 0; j < NJ; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(sa[k][i], yb[k][j], acc[i][j], 0, 0, 0);
#pragma unroll
    for (int k = 0; k < kGramStepsW; ++k) {
#pragma unroll
      for (int i = 0; i < NI; ++i) sa[k][i] = sn[k][i];
#pragma unroll
      for (int j = 0; j < NJ; ++j) yb[k][j] = yn[k][j];
    }
  }
#pragma unroll
  for (int i = 0; i < NI; ++i)
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      const int col = j * kGramTile + cl;
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int row = (half * NI + i) * kGramTile + 4 * c + rr;
        if (row < ns && col < m) out[row * m + col] = acc[i][j][c];
      }
    }
}

__global__ __launch_bounds__(kGramThreads, 3) void k_lob_gram_w(LobParams p, const LobState* st,
                                                             const double* blk,
                                                             double* __restrict__ gpart) {
  const int v = blockIdx.y, chunk = blockIdx.x;
  if (st[v].done) return;
  const int m = p.m, ns = st[v].ns, cur = st[v].cur;
  const int wave = threadIdx.x >> 6, img = wave >> 1, half = wave & 1;  // img 0: K, 1: M
  if (half * 3 * kGramTile >= ns) return;
  const double* base = blk + (int64_t)v * kNumBlocks * p.nm;
  const int bS[3] = {B_X + cur, B_W, B_P + cur};
  const int64_t per = ceil_div(ceil_div(p.n, (int64_t)p.nchunk_g), 4 * kGramSteps) * 4 * kGramSteps;
  const int64_t lo = min(p.n, (int64_t)chunk * per), hi = min(p.n, lo + per);
  const int NS = 3 * m;
  double* out = gpart + (((int64_t)v * p.nchunk_g + chunk) * 2 + img) * NS * m;
  gram_rows_w(base, ns, m, half, bS, img ? B_MW : B_KW, p.nm, lo, hi, out);
}

// G[v][2][NS][NS] from the W-column partials (fixed chunk order) and the recurrence blocks Gp.
__global__ __launch_bounds__(kThreads) void k_lob_gsum_w(LobParams p, const LobState* st,
                                                         const double* __restrict__ gpart,
                                                         const double* __restrict__ Gp,
                                                         double* __restrict__ G) {
  const int v = blockIdx.y;
  if (st[v].done) return;
  const int m = p.m, NS = 3 * m, ns = st[v].ns;
  const int e = blockIdx.x * kThreads + threadIdx.x;
  if (e >= 2 * NS * NS) return;
  const int mat = e / (NS * NS), a = (e % (NS * NS)) / NS, b = e % NS;
  if (a >= ns || b >= ns) return;
  const bool aw = a >= m && a < 2 * m, bw = b >= m && b < 2 * m;
  double g;
  if (aw || bw) {
    const int r = bw ? a : b, c = bw ? b - m : a - m;  // (row of S, column of W)
    const double* src = gpart + ((int64_t)v * p.nchunk_g * 2 + mat) * NS * m + r * m + c;
    g = 0.0;
    for (int k = 0; k < p.nchunk_g; ++k) g += src[(int64_t)k * 2 * NS * m];
  } else {
    const int ia = a < m ? a : a - m, ib = b < m ? b : b - m;  // X: 0..m-1, P: m..2m-1
    const int lo_ = ia < ib ? ia : ib, hi_ = ia < ib ? ib : ia;
    g = Gp[(((int64_t)v * 2 + mat) * 2 * m + lo_) * 2 * m + hi_];
  }
  G[((int64_t)v * 2 + mat) * NS * NS + a * NS + b] = g;
}

// --------------------------------------------------------------------------------------------------
// Parallel cyclic Jacobi on an r x r symmetric matrix in LDS (row-major, leading dimension r).
// On return the diagonal of A holds the eigenvalues and the columns of V the eigenvectors.
// Round-robin pairing: every step rotates r/2 disjoint planes at once (rows, then columns).
// --------------------------------------------------------------------------------------------------
// Stops when the off-diagonal squared norm is at most off2 times the diagonal's.
__device__ void jacobi_eig(double* A, double* V, int r, int max_sweeps, double off2 = 1e-30) {
  __shared__ double s_c[kMaxBasis / 2 + 1], s_s[kMaxBasis / 2 + 1];
  __shared__ int s_p[kMaxBasis / 2 + 1], s_q[kMaxBasis / 2 + 1];
  __shared__ double s_off[kRRThreads / 64], s_dia[kRRThreads / 64];
  __shared__ int s_stop;
  const int tid = threadIdx.x, nth = blockDim.x;
  for (int idx = tid; idx < r * r; idx += nth) V[idx] = (idx / r == idx % r) ? 1.0 : 0.0;
  const int re = r + (r & 1), npair = re / 2;
  __syncthreads();
  for (int sweep = 0; sweep < max_sweeps; ++sweep) {
    double off = 0.0, dia = 0.0;
    for (int idx = tid; idx < r * r; idx += nth) {
      const double a = A[idx];
      if (idx / r == idx % r)
        dia += a * a;
      else
        off += a * a;
    }
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      off += __shfl_xor(off, o, 64);
      dia += __shfl_xor(dia, o, 64);
    }
    if ((tid & 63) == 0) {
      s_off[tid >> 6] = off;
      s_dia[tid >> 6] = dia;
    }
    __syncthreads();
    if (tid == 0) {
      double o2 = 0.0, d2 = 0.0;
      for (int w = 0; w < (nth + 63) / 64; ++w) {
        o2 += s_off[w];
        d2 += s_dia[w];
      }
      s_stop = !(o2 > off2 * d2);
    }
    __syncthreads();
    if (s_stop) break;
    for (int step = 0; step < re - 1; ++step) {
      if (tid < npair) {
        auto pos = [&](int i) { return i == 0 ? 0 : ((i - 1 + step) % (re - 1)) + 1; };
        int pp = pos(tid), qq = pos(re - 1 - tid);
        if (pp > qq) {
          const int t = pp;
          pp = qq;
          qq = t;
        }
        double c = 1.0, s = 0.0;
        if (qq < r) {
          const double apq = A[pp * r + qq];
          if (apq != 0.0) {
            const double app = A[pp * r + pp], aqq = A[qq * r + qq];
            const double tau = (aqq - app) / (2.0 * apq);
            double t;
            if (fabs(tau) > 1e150)
              t = 0.5 / tau;
            else
              t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            c = 1.0 / sqrt(1.0 + t * t);
            s = t * c;
          }
        }
        s_p[tid] = pp;
        s_q[tid] = qq;
        s_c[tid] = c;
        s_s[tid] = s;
      }
      __syncthreads();
      // A <- J^T A J in one pass: each 2 x 2 block (rows {p_k, q_k}, columns {p_l, q_l}) gets the row
      // rotation of pair k, then the column rotation of pair l (the same operations, in the same
      // order, as rotating all rows and then all columns); V <- V J alongside
      for (int idx = tid; idx < npair * npair; idx += nth) {
        const int k = idx / npair, l = idx - k * npair;
        const int p1 = s_p[k], q1 = s_q[k], p2 = s_p[l], q2 = s_q[l];
        const bool hq1 = q1 < r, hq2 = q2 < r;
        const double c1 = s_c[k], s1 = s_s[k], c2 = s_c[l], s2 = s_s[l];
        const double a_pp = A[p1 * r + p2], a_pq = hq2 ? A[p1 * r + q2] : 0.0;
        const double a_qp = hq1 ? A[q1 * r + p2] : 0.0;
        const double a_qq = (hq1 && hq2) ? A[q1 * r + q2] : 0.0;
        // rows
        const double r_pp = c1 * a_pp - s1 * a_qp, r_pq = c1 * a_pq - s1 * a_qq;
        const double r_qp = s1 * a_pp + c1 * a_qp, r_qq = s1 * a_pq + c1 * a_qq;
        // columns
        A[p1 * r + p2] = c2 * r_pp - s2 * r_pq;
        if (hq2) A[p1 * r + q2] = s2 * r_pp + c2 * r_pq;
        if (hq1) {
          A[q1 * r + p2] = c2 * r_qp - s2 * r_qq;
          if (hq2) A[q1 * r + q2] = s2 * r_qp + c2 * r_qq;
        }
      }
      for (int idx = tid; idx < r * npair; idx += nth) {
        const int k = idx / r, i = idx - k * r;
        const int qq = s_q[k];
        const double s = s_s[k];
        if (qq >= r || s == 0.0) continue;
        const int pp = s_p[k];
        const double c = s_c[k];
        const double vp = V[i * r + pp], vq = V[i * r + qq];
        V[i * r + pp] = c * vp - s * vq;
        V[i * r + qq] = s * vp + c * vq;
      }
      __syncthreads();
    }
  }
}

constexpr int kJacobiSweeps = 30;
// The Rayleigh-Ritz problems stop at an off-diagonal norm of 1e-12 of the diagonal's: the Ritz values
// then carry errors of order 1e-24, far below what the residual tolerance resolves.
constexpr double kRROff2 = 1e-24;

// Batched symmetric eigen-decomposition: one workgroup per matrix.
__global__ __launch_bounds__(kRRThreads) void k_sym_eig(int r, const double* __restrict__ A,
                                                        double* __restrict__ w,
                                                        double* __restrict__ V) {
  extern __shared__ double lds[];
  double* sA = lds;
  double* sV = lds + r * r;
  const int64_t mat = blockIdx.x;
  for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) {
    const int i = idx / r, j = idx % r;
    sA[idx] = 0.5 * (A[mat * r * r + i * r + j] + A[mat * r * r + j * r + i]);
  }
  __syncthreads();
  jacobi_eig(sA, sV, r, kJacobiSweeps);
  for (int idx = threadIdx.x; idx < r * r; idx += blockDim.x) V[mat * r * r + idx] = sV[idx];
  for (int i = threadIdx.x; i < r; i += blockDim.x) w[mat * r + i] = sA[i * r + i];
}

// --------------------------------------------------------------------------------------------------
// K5: convergence test and Rayleigh-Ritz, one workgroup per variant.
// --------------------------------------------------------------------------------------------------
__global__ __launch_bounds__(kRRThreads) void k_lob_rr(LobParams p, LobState* st,
                                                       const double* __restrict__ rpart,
                                                       const double* __restrict__ G,
                                                       double* __restrict__ Zg,
                                                       double* __restrict__ Tg,
                                                       double* __restrict__ Cg,
                                                       double* __restrict__ Gp) {
  extern __shared__ double lds[];
  const int v = blockIdx.x;
  LobState* s = st + v;
  __shared__ int s_skip, s_ns, s_init, s_keep;
  __shared__ int s_kept[kMaxBasis];
  __shared__ int s_sel[kMaxBlock];
  __shared__ double s_scale[kMaxBasis];
  __shared__ double s_mu[kMaxBasis];
  const int m = p.m, NS = 3 * m, tid = threadIdx.x;
  if (tid == 0) {
    s_skip = s->done;
    s_ns = s->ns;
    s_init = s->init;
  }
  __syncthreads();
  if (s_skip) return;
  const int ns = s_ns;
  // ---- convergence of the current block (its residuals were formed by k_lob_resid) ----
  if (!s_init) {
    if (tid < m) {
      double a = 0.0, b = 0.0;
      const double* src = rpart + (int64_t)v * p.nchunk_r * 2 * m;
      for (int c = 0; c < p.nchunk_r; ++c) {
        a += src[(int64_t)c * 2 * m + tid];
        b += src[(int64_t)c * 2 * m + m + tid];
      }
      const double lam = fabs(s->lambda[tid]);
      const double den = (lam > 0.0 ? lam : 1.0) * sqrt(b);
      s->resid[tid] = den > 0.0 ? sqrt(a) / den : sqrt(a);
    }
    __syncthreads();
    if (tid == 0) {
      int nc = 0;
      for (int j = 0; j < p.n_wanted; ++j) nc += s->resid[j] <= p.tol;
      s->nconv = nc;
      if (nc == p.n_wanted)
        s->done = 1;
      else if (s->it >= p.max_iter)
        s->done = 2;
      s_skip = s->done;
    }
    __syncthreads();
    if (s_skip) return;
  }
  double* sA = lds;
  double* sV = lds + NS * NS;  // both used with leading dimension = the current size
  const double* GK = G + (int64_t)v * 2 * NS * NS;
  const double* GM = GK + NS * NS;
  double* Z = Zg + (int64_t)v * NS * NS;
  double* T1 = Tg + (int64_t)v * NS * NS;
  // ---- rank-revealing M-orthonormalisation of the basis: eig of D GM D, D = diag(GM)^-1/2 ----
  if (tid < ns) {
    const double d = GM[tid * NS + tid];
    s_scale[tid] = d > 0.0 ? 1.0 / sqrt(d) : 0.0;
  }
  __syncthreads();
  for (int idx = tid; idx < ns * ns; idx += blockDim.x) {
    const int a = idx / ns, b = idx % ns;
    sA[idx] = 0.5 * (GM[a * NS + b] + GM[b * NS + a]) * s_scale[a] * s_scale[b];
  }
  __syncthreads();
  jacobi_eig(sA, sV, ns, kJacobiSweeps, kRROff2);
  if (tid == 0) {
    double mx = 0.0;
    for (int k = 0; k < ns; ++k) mx = fmax(mx, sA[k * ns + k]);
    int kept = 0;
    for (int k = 0; k < ns; ++k) {
      const double mu = sA[k * ns + k];
      if (mu > p.drop * mx) {
        s_kept[kept] = k;
        s_mu[kept] = mu;
        ++kept;
      }
    }
    s_keep = kept;
    s->kept = kept;
    // directions dropped for near-dependence (soft-locked W columns aside): the next iteration
    // restarts the search from [X W] without P
    int zero = 0;
    for (int k = 0; k < ns; ++k) zero += !(s_scale[k] > 0.0);
    s->restart = (kept + zero < ns);
  }
  __syncthreads();
  const int rk = s_keep;
  if (rk < m) {
    if (tid == 0) s->done = 3;
    return;
  }
  // Z = D Q_kept diag(mu)^-1/2 (ns x rk), leading dimension NS
  for (int idx = tid; idx < ns * rk; idx += blockDim.x) {
    const int a = idx / rk, k = idx % rk;
    Z[a * NS + k] = s_scale[a] * sV[a * ns + s_kept[k]] / sqrt(s_mu[k]);
  }
  __syncthreads();
  // T1 = GK Z (ns x rk), then H = Z^T T1 (rk x rk) into sA
  for (int idx = tid; idx < ns * rk; idx += blockDim.x) {
    const int a = idx / rk, k = idx % rk;
    double t = 0.0;
    for (int b = 0; b < ns; ++b) t += 0.5 * (GK[a * NS + b] + GK[b * NS + a]) * Z[b * NS + k];
    T1[a * NS + k] = t;
  }
  __syncthreads();
  for (int idx = tid; idx < rk * rk; idx += blockDim.x) {
    const int k = idx / rk, l = idx % rk;
    double h = 0.0;
    for (int a = 0; a < ns; ++a) h += Z[a * NS + k] * T1[a * NS + l];
    sA[idx] = h;
  }
  __syncthreads();
  for (int idx = tid; idx < rk * rk; idx += blockDim.x) {
    const int k = idx / rk, l = idx % rk;
    if (k < l) {
      const double h = 0.5 * (sA[k * rk + l] + sA[l * rk + k]);
      sA[k * rk + l] = h;
      sA[l * rk + k] = h;
    }
  }
  __syncthreads();
  jacobi_eig(sA, sV, rk, kJacobiSweeps, kRROff2);
  // select the m smallest (largest) Ritz values in order: thread k ranks value k (ties by index)
  if (tid < rk) {
    const double t = sA[tid * rk + tid];
    int rank = 0;
    for (int l = 0; l < rk; ++l) {
      const double u = sA[l * rk + l];
      const bool before = p.largest ? (u > t) : (u < t);
      rank += before || (u == t && l < tid);
    }
    if (rank < m) {
      s_sel[rank] = tid;
      s->lambda[rank] = t;
    }
  }
  __syncthreads();
  // C = Z Y_sel (ns x m)
  double* C = Cg + (int64_t)v * NS * m;
  for (int idx = tid; idx < ns * m; idx += blockDim.x) {
    const int a = idx / m, j = idx % m;
    double c = 0.0;
    for (int k = 0; k < rk; ++k) c += Z[a * NS + k] * sV[k * rk + s_sel[j]];
    C[idx] = c;
  }
  __syncthreads();
  // Gram blocks of the next basis' X' = S C and P' = S C0 (C0: the rows of X zeroed), for the
  // iterations that sum only the new W's columns: T = G C, T0 = G C0 (into T1), then
  // Gp = [C C0]^T [T T0] (upper triangle used), for G_K and G_M
  const int m2 = 2 * m;
  for (int mat = 0; mat < 2; ++mat) {
    const double* Gm = mat ? GM : GK;
    for (int idx = tid; idx < ns * m; idx += blockDim.x) {
      const int a = idx / m, j = idx % m;
      double t = 0.0, t0 = 0.0;
      for (int b = 0; b < ns; ++b) {
        const double gab = 0.5 * (Gm[a * NS + b] + Gm[b * NS + a]) * C[b * m + j];
        t += gab;
        if (b >= m) t0 += gab;
      }
      T1[idx] = t;
      T1[ns * m + idx] = t0;
    }
    __syncthreads();
    double* out = Gp + ((int64_t)v * 2 + mat) * m2 * m2;
    for (int idx = tid; idx < m2 * m2; idx += blockDim.x) {
      const int i = idx / m2, j = idx % m2;
      if (j < i) continue;
      const double* Tj = T1 + (j < m ? 0 : ns * m);
      const int jc = j < m ? j : j - m, ic = i < m ? i : i - m;
      double g = 0.0;
      for (int a = (i < m ? 0 : m); a < ns; ++a) g += C[a * m + ic] * Tj[a * m + jc];
      out[idx] = g;
    }
    __syncthreads();
  }
}

// --------------------------------------------------------------------------------------------------
// K6, one thread per DOF row holding all MB columns: the coefficients, uniform over the workgroup,
// are read as scalars (the constant address space), so no coefficient goes through LDS.  Per image
// (basis, K-image, M-image) the steps are W, P (their sum is P' = W C_W + P C_P, stored before X is
// added) and X (X' = P' + X C_X); W and P are skipped while they are not in the basis.
using ConstF64 = const __attribute__((address_space(4))) double*;

template <int MB>
__device__ __forceinline__ void upd_load(const double* __restrict__ Y, double (&y)[MB]) {
#pragma unroll
  for (int a = 0; a < MB; a += 2) {
    const double2 t = *reinterpret_cast<const double2*>(Y + a);
    y[a] = t.x;
    y[a + 1] = t.y;
  }
}

template <int MB>
__device__ __forceinline__ void upd_store(double* __restrict__ Y, const double (&acc)[MB]) {
#pragma unroll
  for (int j = 0; j < MB; j += 2) reinterpret_cast<double2*>(Y + j)[0] = make_double2(acc[j], acc[j + 1]);
}

template <int MB>
__global__ __launch_bounds__(kThreads) void k_lob_update_rows(LobParams p, const LobState* st,
                                                              double* blk,
                                                              const double* __restrict__ Cg) {
  const int v = blockIdx.y;
  if (st[v].done) return;
  const int cur = st[v].cur, ns = st[v].ns;
  const int64_t row = (int64_t)blockIdx.x * kThreads + threadIdx.x;
  if (row >= p.n) return;
  const ConstF64 C = (ConstF64)(Cg + (int64_t)v * 3 * MB * MB);
  double* base = blk + (int64_t)v * kNumBlocks * p.nm + row * MB;
  const int nk = ns > 2 * MB ? 3 : ns > MB ? 2 : 1;  // steps per image: [W [P]] X
  const int nsteps = 3 * nk;
  // step s: image s / nk, kind 0 W, 1 P, 2 X (the last step of an image is always X)
  auto kind_of = [nk](int s) { const int k = s % nk; return k == nk - 1 ? 2 : k; };
  auto src_of = [&](int s) -> const double* {
    const int which = s / nk, k = kind_of(s);
    const int b = k == 0 ? B_W + which : k == 1 ? B_P + 2 * which + cur : B_X + 2 * which + cur;
    return base + (int64_t)b * p.nm;
  };
  // no register prefetch of the next step's row: 125 instead of 172 VGPRs (block 28) doubles the
  // waves per SIMD, which hides the loads better than the prefetch did
  double acc[MB], y[MB];
#pragma unroll
  for (int j = 0; j < MB; ++j) acc[j] = 0.0;
#pragma unroll 1
  for (int s = 0; s < nsteps; ++s) {
    upd_load<MB>(src_of(s), y);
    const int which = s / nk, k = kind_of(s);
    if (k == 2) upd_store<MB>(base + (int64_t)(B_P + 2 * which + 1 - cur) * p.nm, acc);
    const ConstF64 Ck = C + (k == 2 ? 0 : k == 0 ? MB * MB : 2 * MB * MB);
#pragma unroll
    for (int a = 0; a < MB; ++a)
#pragma unroll
      for (int j = 0; j < MB; ++j) acc[j] = fma(y[a], Ck[a * MB + j], acc[j]);
    if (k == 2) {
      upd_store<MB>(base + (int64_t)(B_X + 2 * which + 1 - cur) * p.nm, acc);
#pragma unroll
      for (int j = 0; j < MB; ++j) acc[j] = 0.0;
    }
  }
}

void launch_update(hipStream_t st, const LobParams& p, const LobState* s, double* blk, const double* C) {
  const dim3 grid((unsigned)ceil_div(p.n, kThreads), p.nvar);
  switch (p.m) {
#define FE_LOB_UPD(MB) \
  case MB:             \
    hipLaunchKernelGGL(k_lob_update_rows<MB>, grid, dim3(kThreads), 0, st, p, s, blk, C); \
    break;
    FE_LOB_UPD(4) FE_LOB_UPD(8) FE_LOB_UPD(12) FE_LOB_UPD(16)
    FE_LOB_UPD(20) FE_LOB_UPD(24) FE_LOB_UPD(28) FE_LOB_UPD(32)
#undef FE_LOB_UPD
  }
}

// K7: the new blocks become current; the basis gains W (and P after the first iteration).
__global__ void k_lob_flip(LobParams p, LobState* st) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= p.nvar || st[v].done) return;
  LobState* s = st + v;
  s->cur ^= 1;
  if (s->init) {
    s->init = 0;
    s->ns = 2 * p.m;
  } else {
    s->it += 1;
    s->ns = s->restart ? 2 * p.m : 3 * p.m;
  }
}

__global__ void k_lob_status(LobParams p, const LobState* st, LobStatus* out) {
  __shared__ int s_act[kThreads], s_max[kThreads], s_fail[kThreads], s_unc[kThreads];
  int act = 0, mx = 0, fail = 0, unc = 0;
  for (int v = threadIdx.x; v < p.nvar; v += blockDim.x) {
    act += st[v].done == 0;
    mx = max(mx, st[v].it);
    fail += st[v].done == 3;
    unc += st[v].done == 2;
  }
  s_act[threadIdx.x] = act;
  s_max[threadIdx.x] = mx;
  s_fail[threadIdx.x] = fail;
  s_unc[threadIdx.x] = unc;
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int t = 1; t < (int)blockDim.x; ++t) {
      act += s_act[t];
      mx = max(mx, s_max[t]);
      fail += s_fail[t];
      unc += s_unc[t];
    }
    out->active = act;
    out->max_it = mx;
    out->failed = fail;
    out->unconverged = unc;
  }
}

__global__ void k_lob_init(LobParams p, LobState* st) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v >= p.nvar) return;
  LobState* s = st + v;
  s->it = 0;
  s->done = 0;
  s->cur = 0;
  s->init = 1;
  s->ns = p.m;
  s->kept = 0;
  s->nconv = 0;
  s->restart = 0;
  for (int j = 0; j < kMaxBlock; ++j) {
    s->lambda[j] = 0.0;
    s->resid[j] = HUGE_VAL;  // not yet measured: no pair is locked
  }
}

// Copy the current X of every variant to the caller's block and export lambda / residuals.
__global__ void k_lob_export(LobParams p, const LobState* st, double* blk, double* X, double* lam,
                             double* res, int32_t* iters, int32_t* status) {
  const int v = blockIdx.y;
  const int cur = st[v].cur;
  const double* src = lob_block(blk, p, v, B_X + cur);
  for (int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; t < p.nm;
       t += (int64_t)gridDim.x * blockDim.x)
    X[(int64_t)v * p.nm + t] = src[t];
  if (blockIdx.x == 0 && threadIdx.x < p.m) {
    lam[v * p.m + threadIdx.x] = st[v].lambda[threadIdx.x];
    res[v * p.m + threadIdx.x] = st[v].resid[threadIdx.x];
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    iters[v] = st[v].it;
    status[v] = st[v].done;
  }
}

// --------------------------------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------------------------------
constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct LobWork {
  LobState* st;
  LobStatus* status;
  double* blk;
  double* gpart;
  double* rpart;
  double* G;
  double* C;
  double* Z;
  double* T1;
  double* Gp;  // [nvar][2][2m][2m] Gram blocks of the next X and P (upper triangles)
  double* rc;  // two-level: [nvar][ld][m] restricted residuals and coarse solutions
  double* zc;
  int64_t bytes;
};

LobParams make_params(const fe_bsr* K, const fe_lobpcg_opts* o) {
  LobParams p{};
  p.nb = K->n_brows;
  p.n = K->n_brows * K->dof;
  p.nnzb = K->nnzb;
  p.m = o->block;
  p.nm = p.n * p.m;
  p.n_wanted = o->n_wanted;
  p.largest = o->largest;
  p.max_iter = o->max_iter;
  p.tol = o->tol;
  p.drop = 1e-11;
  p.nvar = o->n_variants;
  // Gram chunks: about 2,048 four-wavefront workgroups over all variants, at least 8 and at most
  // 512 per variant (the partials of one variant, 2 (3m)^2 doubles per chunk, are summed each
  // iteration)
  int64_t chunks = std::min<int64_t>(512, 2048 / std::max(1, o->n_variants));
  chunks = std::max<int64_t>(8, std::min<int64_t>(chunks, ceil_div(p.n, 64)));
  p.nchunk_g = (int)std::max<int64_t>(1, chunks);
  p.nchunk_r = (int)std::max<int64_t>(1, std::min<int64_t>(256, ceil_div(p.nb, 64)));
  return p;
}

LobWork layout_lob(char* base, const LobParams& p, int64_t ld = 0) {
  LobWork w{};
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* q = base ? base + off : nullptr;
    off += align_up(bytes);
    return q;
  };
  const int64_t NS = 3 * p.m, nv = p.nvar;
  w.st = (LobState*)take(nv * (int64_t)sizeof(LobState));
  w.status = (LobStatus*)take(sizeof(LobStatus));
  w.blk = (double*)take(nv * kNumBlocks * p.nm * (int64_t)sizeof(double));
  w.gpart = (double*)take(nv * p.nchunk_g * 2 * NS * NS * (int64_t)sizeof(double));
  w.rpart = (double*)take(nv * p.nchunk_r * 2 * p.m * (int64_t)sizeof(double));
  w.G = (double*)take(nv * 2 * NS * NS * (int64_t)sizeof(double));
  w.C = (double*)take(nv * NS * p.m * (int64_t)sizeof(double));
  w.Z = (double*)take(nv * NS * NS * (int64_t)sizeof(double));
  w.T1 = (double*)take(nv * NS * NS * (int64_t)sizeof(double));
  w.Gp = (double*)take(nv * 2 * 4 * (int64_t)p.m * p.m * (int64_t)sizeof(double));
  if (ld > 0) {
    w.rc = (double*)take(nv * ld * p.m * (int64_t)sizeof(double));
    w.zc = (double*)take(nv * ld * p.m * (int64_t)sizeof(double));
  }
  w.bytes = off;
  return w;
}

int check_opts(const fe_bsr* K, const fe_lobpcg_opts* o) {
  FE_REQUIRE(K && o, "fe_lobpcg: null argument");
  FE_REQUIRE(K->dof >= 1 && K->dof <= 3, "fe_lobpcg: dof must be 1..3");
  FE_REQUIRE(K->row_align == 0, "fe_lobpcg: the matrix's own pattern expected (row_align 0)");
  FE_REQUIRE(K->precision == 0 && (K->layout == FE_LAYOUT_PLANES || K->layout == FE_LAYOUT_BSR),
             "fe_lobpcg: matrices must be FP64 in FE_LAYOUT_PLANES or FE_LAYOUT_BSR");
  FE_REQUIRE(o->block >= 4 && o->block <= kMaxBlock && o->block % 4 == 0,
             "fe_lobpcg: block must be a multiple of 4 in [4, %d]", kMaxBlock);
  FE_REQUIRE(o->n_wanted >= 1 && o->n_wanted <= o->block, "fe_lobpcg: need 1 <= n_wanted <= block");
  FE_REQUIRE(o->n_variants >= 1 && o->n_variants <= 65535, "fe_lobpcg: n_variants out of range");
  FE_REQUIRE(o->max_iter >= 1 && o->tol > 0, "fe_lobpcg: max_iter and tol must be positive");
  FE_REQUIRE(K->n_brows * K->dof >= 3 * (int64_t)o->block,
             "fe_lobpcg: the system must have at least 3 * block DOFs");
  return FE_OK;
}

// Two-level preconditioner of the iteration (NULL C: block Jacobi T only).
struct LobCoarse {
  const fe_coarse* C;
  const double* Ainv;
  int64_t ainv_stride;
  int64_t ld;
};

// The two-level preconditioner's part of an iteration after the residual pass (k_lob_resid_agg,
// which also formed r_c = P^T R): z_c = A_c^-1 r_c, then W += P z_c.
template <int D, int NC>
void launch_lob_coarse(hipStream_t st, const LobParams& p, const LobCoarse& cc, const LobWork& w) {
  const int nv = p.nvar;
  hipLaunchKernelGGL(k_lob_cgemm, dim3((unsigned)(cc.ld / kCgRows), nv), dim3(kThreads), 0, st, p, cc.ld,
                     w.st, cc.Ainv, cc.ainv_stride, w.rc, w.zc);
  hipLaunchKernelGGL((k_lob_cprolong<D, NC>), dim3((unsigned)ceil_div(p.n, kThreads / (p.m / 4)), nv), dim3(kThreads),
                     0, st, p, *cc.C, cc.ld, w.st, w.zc, w.blk);
}

template <int D, int LAYOUT>
int run_lobpcg(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T,
               const uint8_t* fixed, double* X, double* lam, double* res, int32_t* iters,
               int32_t* vstatus, const LobWork& w, const LobParams& p, const fe_lobpcg_opts* o,
               fe_lobpcg_stats* stats, const LobCoarse& cc) {
  hipStream_t st = h->stream;
  const int nv = p.nvar;
  if (cc.C) FE_CHECK_HIP(hipMemsetAsync(w.rc, 0, (size_t)nv * cc.ld * p.m * sizeof(double), st));
  const dim3 gvar((unsigned)ceil_div(nv, 64));
  hipLaunchKernelGGL(k_lob_init, gvar, dim3(64), 0, st, p, w.st);
  FE_CHECK_LAUNCH();
  // caller's initial block into X0 of every variant
  for (int v = 0; v < nv; ++v)
    FE_CHECK_HIP(hipMemcpyAsync(w.blk + (int64_t)v * kNumBlocks * p.nm, X + (int64_t)v * p.nm,
                                p.nm * sizeof(double), hipMemcpyDeviceToDevice, st));
  // P blocks start at zero so that unused basis columns read finite values
  for (int v = 0; v < nv; ++v)
    for (int b = B_P; b < B_W; ++b)
      FE_CHECK_HIP(hipMemsetAsync(w.blk + ((int64_t)v * kNumBlocks + b) * p.nm, 0,
                                  p.nm * sizeof(double), st));
  const double* Kv = (const double*)K->vals;
  const double* Mv = (const double*)M->vals;
  // block SpMM: K staged in LDS for the block-contiguous layout
  const int npg = (LAYOUT == FE_LAYOUT_BSR ? kSpmmThreads : kThreads) / (p.m / 4);
  const dim3 gspmm((unsigned)ceil_div(p.nb, npg), nv);
  auto spmm = [&](int src) {
    if (LAYOUT == FE_LAYOUT_BSR)
      hipLaunchKernelGGL((k_lob_spmm_lds<D>), gspmm, dim3(kSpmmThreads), 0, st, p, K->bptr, K->bcol, Kv,
                         o->k_stride, Mv, o->m_stride, w.st, w.blk, src);
    else
      hipLaunchKernelGGL((k_lob_spmm<D, LAYOUT>), gspmm, dim3(kThreads), 0, st, p, K->bptr, K->bcol, Kv,
                         o->k_stride, Mv, o->m_stride, w.st, w.blk, src);
  };
  const dim3 gres((unsigned)p.nchunk_r, nv);
  const dim3 ggram((unsigned)p.nchunk_g, nv);
  const int NS = 3 * p.m;
  const dim3 ggsum((unsigned)ceil_div(2 * NS * NS, kThreads), nv);
  const size_t rr_lds = 2 * (size_t)NS * NS * sizeof(double);
  FE_CHECK_HIP(hipFuncSetAttribute((const void*)k_lob_rr, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)rr_lds));
  // full: the Gram of the whole basis from its images; otherwise only the new W's columns, the
  // blocks of X and P coming from the previous Rayleigh-Ritz step
  auto rr_update = [&](bool full) -> int {
    if (full) {
      hipLaunchKernelGGL(k_lob_gram, ggram, dim3(kGramThreads), 0, st, p, w.st, w.blk, w.gpart);
      hipLaunchKernelGGL(k_lob_gsum, ggsum, dim3(kThreads), 0, st, p, w.st, w.gpart, w.G);
    } else {
      hipLaunchKernelGGL(k_lob_gram_w, ggram, dim3(kGramThreads), 0, st, p, w.st, w.blk, w.gpart);
      hipLaunchKernelGGL(k_lob_gsum_w, ggsum, dim3(kThreads), 0, st, p, w.st, w.gpart, w.Gp, w.G);
    }
    hipLaunchKernelGGL(k_lob_rr, dim3(nv), dim3(kRRThreads), rr_lds, st, p, w.st, w.rpart, w.G, w.Z,
                       w.T1, w.C, w.Gp);
    launch_update(st, p, w.st, w.blk, w.C);
    hipLaunchKernelGGL(k_lob_flip, gvar, dim3(64), 0, st, p, w.st);
    FE_CHECK_LAUNCH();
    return FE_OK;
  };
  // start: K X, M X, Rayleigh-Ritz on X alone
  spmm(1);
  FE_CHECK_LAUNCH();
  int rc = rr_update(true);
  if (rc) return rc;
  const int chunk = o->check_every > 0 ? o->check_every : 8;
  LobStatus* pinned = (LobStatus*)((char*)h->pinned + 1024);
  hipEvent_t ev[2];
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  // The K- and M-images of X and P are carried by recombination; every kRefresh iterations they are
  // recomputed from X and P so that rounding cannot accumulate into a residual floor.
  int64_t host_it = 0;
  auto enqueue = [&](int slot) -> int {
    for (int it = 0; it < chunk; ++it, ++host_it) {
      if (host_it % kRefresh == kRefresh - 1) {
        spmm(1);
        spmm(2);
      }
      constexpr int NC = D == 3 ? 6 : D == 2 ? 3 : 1;
      if (cc.C) {
        hipLaunchKernelGGL((k_lob_resid_agg<D, NC>), gres, dim3(kThreads), 0, st, p, *cc.C, cc.ld, T,
                           o->t_stride, fixed, w.st, w.blk, w.rpart, w.rc);
        launch_lob_coarse<D, NC>(st, p, cc, w);
      } else {
        hipLaunchKernelGGL((k_lob_resid<D>), gres, dim3(kThreads), 0, st, p, T, o->t_stride, fixed,
                           w.st, w.blk, w.rpart);
      }
      spmm(0);
      int rc2 = rr_update(host_it % kRefresh == kRefresh - 1);
      if (rc2) return rc2;
    }
    hipLaunchKernelGGL(k_lob_status, dim3(1), dim3(kThreads), 0, st, p, w.st, w.status);
    FE_CHECK_HIP(hipMemcpyAsync(&pinned[slot], w.status, sizeof(LobStatus), hipMemcpyDeviceToHost, st));
    FE_CHECK_HIP(hipEventRecord(ev[slot], st));
    return FE_OK;
  };
  // one chunk always queued ahead of the polled one; the last residual check needs max_iter + 1
  // residual passes, so the bound below covers every variant reaching max_iter
  const int64_t max_chunks = (int64_t)o->max_iter / chunk + 4;
  LobStatus last{};
  rc = enqueue(0);
  if (rc == FE_OK) rc = enqueue(1);
  for (int64_t c = 0; rc == FE_OK && c < max_chunks; ++c) {
    const int slot = (int)(c & 1);
    FE_CHECK_HIP(hipEventSynchronize(ev[slot]));
    last = pinned[slot];
    if (last.active == 0) break;
    rc = enqueue(slot);
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  hipLaunchKernelGGL(k_lob_export, dim3((unsigned)std::min<int64_t>(1024, ceil_div(p.nm, kThreads)), nv),
                     dim3(kThreads), 0, st, p, w.st, w.blk, X, lam, res, iters, vstatus);
  hipLaunchKernelGGL(k_lob_status, dim3(1), dim3(kThreads), 0, st, p, w.st, w.status);
  FE_CHECK_HIP(hipMemcpyAsync(&pinned[0], w.status, sizeof(LobStatus), hipMemcpyDeviceToHost, st));
  FE_CHECK_HIP(hipStreamSynchronize(st));
  last = pinned[0];
  stats->max_iterations = last.max_it;
  stats->unconverged = last.unconverged + last.active;
  stats->failed = last.failed;
  stats->status = last.failed                       ? FE_ERR_BREAKDOWN
                  : (last.unconverged + last.active) ? FE_ERR_NOT_CONVERGED
                                                     : FE_OK;
  return FE_OK;
}


int check_lob_coarse(const fe_bsr* K, const fe_coarse* C, const char* who) {
  FE_REQUIRE(C && C->P && C->node_agg && C->agg_ptr && C->agg_nodes, "%s: null coarse space", who);
  FE_REQUIRE(C->dof == K->dof && C->n_nodes == K->n_brows, "%s: coarse space and K disagree", who);
  FE_REQUIRE(C->nc == (K->dof == 3 ? 6 : K->dof == 2 ? 3 : 1), "%s: nc must be 6, 3 or 1 for dof 3, 2, 1",
             who);
  FE_REQUIRE(C->n_agg >= 1 && C->n_agg < (int64_t)1 << 31, "%s: bad aggregate count", who);
  return FE_OK;
}

int lobpcg_common(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T,
                  const uint8_t* fixed, double* X, double* lambda, double* resid, int32_t* iters,
                  int32_t* vstatus, void* work, int64_t work_bytes, const fe_lobpcg_opts* o,
                  fe_lobpcg_stats* stats, const LobCoarse& cc, const char* who) {
  int rc = check_opts(K, o);
  if (rc) return rc;
  FE_REQUIRE(M && X && lambda && resid && iters && vstatus && work && stats, "%s: null argument", who);
  FE_REQUIRE(M->precision == 0 && M->layout == K->layout && M->dof == K->dof &&
                 M->n_brows == K->n_brows && M->nnzb == K->nnzb && M->row_align == 0,
             "%s: M must be FP64 in the layout of K, on the pattern of K", who);
  FE_REQUIRE(o->k_stride >= 0 && o->m_stride >= 0 && o->t_stride >= 0, "%s: strides must be non-negative",
             who);
  *stats = fe_lobpcg_stats{};
  const LobParams p = make_params(K, o);
  const LobWork w = layout_lob((char*)work, p, cc.ld);
  FE_REQUIRE(work_bytes >= w.bytes, "%s: workspace too small (%lld < %lld)", who, (long long)work_bytes,
             (long long)w.bytes);
#define FE_RUN(DD)                                                                                   \
  return K->layout == FE_LAYOUT_BSR                                                                  \
             ? run_lobpcg<DD, FE_LAYOUT_BSR>(h, K, M, T, fixed, X, lambda, resid, iters, vstatus, w, p, \
                                             o, stats, cc)                                           \
             : run_lobpcg<DD, FE_LAYOUT_PLANES>(h, K, M, T, fixed, X, lambda, resid, iters, vstatus, w, \
                                                p, o, stats, cc)
  switch (K->dof) {
    case 1: FE_RUN(1);
    case 2: FE_RUN(2);
    default: FE_RUN(3);
  }
#undef FE_RUN
}

}  // namespace

extern "C" int fe_lobpcg_workspace(fe_handle* h, const fe_bsr* K, const fe_lobpcg_opts* o,
                                   int64_t* bytes) {
  FE_ENTER(h);
  int rc = check_opts(K, o);
  if (rc) return rc;
  FE_REQUIRE(bytes, "fe_lobpcg_workspace: null argument");
  *bytes = layout_lob(nullptr, make_params(K, o)).bytes;
  return FE_OK;
}

extern "C" int fe_lobpcg(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T,
                         const uint8_t* fixed, double* X, double* lambda, double* resid,
                         int32_t* iters, int32_t* vstatus, void* work, int64_t work_bytes,
                         const fe_lobpcg_opts* o, fe_lobpcg_stats* stats) {
  FE_ENTER(h);
  return lobpcg_common(h, K, M, T, fixed, X, lambda, resid, iters, vstatus, work, work_bytes, o, stats,
                       LobCoarse{nullptr, nullptr, 0, 0}, "fe_lobpcg");
}

extern "C" int fe_lobpcg_coarse_workspace(fe_handle* h, const fe_bsr* K, const fe_coarse* C,
                                          const fe_lobpcg_opts* o, int64_t* bytes) {
  FE_ENTER(h);
  int rc = check_opts(K, o);
  if (rc) return rc;
  rc = check_lob_coarse(K, C, "fe_lobpcg_coarse_workspace");
  if (rc) return rc;
  FE_REQUIRE(bytes, "fe_lobpcg_coarse_workspace: null argument");
  *bytes = layout_lob(nullptr, make_params(K, o), fe_coarse_ld(C->n_agg * C->nc)).bytes;
  return FE_OK;
}

extern "C" int fe_lobpcg_coarse(fe_handle* h, const fe_bsr* K, const fe_bsr* M, const double* T,
                                const uint8_t* fixed, const fe_coarse* C, const double* Ainv,
                                int64_t ainv_stride, double* X, double* lambda, double* resid,
                                int32_t* iters, int32_t* vstatus, void* work, int64_t work_bytes,
                                const fe_lobpcg_opts* o, fe_lobpcg_stats* stats) {
  FE_ENTER(h);
  int rc = check_opts(K, o);
  if (rc) return rc;
  rc = check_lob_coarse(K, C, "fe_lobpcg_coarse");
  if (rc) return rc;
  const int64_t ld = fe_coarse_ld(C->n_agg * C->nc);
  FE_REQUIRE(Ainv, "fe_lobpcg_coarse: null coarse inverse");
  FE_REQUIRE(o->n_variants == 1 || ainv_stride >= ld * ld,
             "fe_lobpcg_coarse: ainv_stride must be >= ld * ld for several variants");
  return lobpcg_common(h, K, M, T, fixed, X, lambda, resid, iters, vstatus, work, work_bytes, o, stats,
                       LobCoarse{C, Ainv, ainv_stride, ld}, "fe_lobpcg_coarse");
}

extern "C" int fe_sym_eig(fe_handle* h, int n_mat, int r, const double* A, double* w, double* V) {
  FE_ENTER(h);
  FE_REQUIRE(A && w && V, "fe_sym_eig: null argument");
  FE_REQUIRE(n_mat >= 0 && r >= 1 && r <= kMaxBasis, "fe_sym_eig: need 1 <= r <= %d", kMaxBasis);
  if (n_mat == 0) return FE_OK;
  const size_t lds = 2 * (size_t)r * r * sizeof(double);
  FE_CHECK_HIP(hipFuncSetAttribute((const void*)k_sym_eig, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds));
  hipLaunchKernelGGL(k_sym_eig, dim3((unsigned)n_mat), dim3(kRRThreads), lds, h->stream, r, A, w, V);
  FE_CHECK_LAUNCH();
  return FE_OK;
}
