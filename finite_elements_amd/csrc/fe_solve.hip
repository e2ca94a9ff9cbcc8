// Sparse matrix-vector products, preconditioner setup and the device-driven PCG.
//
// PCG layout: three kernels per iteration (Chronopoulos-Gear recurrences, see below): a streaming
// vector update, the SpMV with the iteration's three dot products as per-workgroup partials, and a
// one-workgroup kernel that sums the partials in a fixed order (run-to-run bitwise reproducible).
// Every chunk of iterations ends with a conditional reliable update (true FP64 residual r = b - A x):
// it runs when the recursive residual reached the tolerance (so convergence is always judged on the
// true FP64 residual), when the iteration limit is hit, or (mixed precision) when the residual fell
// by ru_delta since the last update.  Its reductions use the last-workgroup-sums pattern (fixed
// order as well).  The host only polls a status word between chunks.
#include "fe_internal.h"

namespace {

constexpr int kBlock = 256;

// Value index of entry k (= di*D + dj) of block q.
template <int D, int LAYOUT>
__device__ __host__ inline int64_t vix(int64_t q, int k, int64_t nnzb) {
  return LAYOUT == FE_LAYOUT_PLANES ? (int64_t)k * nnzb + q : q * D * D + k;
}

// --------------------------------------------------------------------------------------------------
// reductions
// --------------------------------------------------------------------------------------------------
__device__ inline double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sum NV values over the workgroup; result valid in thread 0.
template <int NV>
__device__ inline void block_sum(double (&v)[NV]) {
  __shared__ double red[NV][kBlock / 64];
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = wave_sum(v[k]);
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) red[k][wid] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      double s = 0.0;
      for (int w = 0; w < kBlock / 64; ++w) s += red[k][w];
      v[k] = s;
    }
  __syncthreads();
}

// Write this workgroup's partials; return true in every thread of the last workgroup to finish,
// which then holds the grid totals in v (thread 0).  The counter is reset for the next use.
template <int NV>
__device__ inline bool grid_sum(double (&v)[NV], double* partials, unsigned* counter) {
  __shared__ bool am_last;
  block_sum<NV>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) partials[k * gridDim.x + blockIdx.x] = v[k];
    __threadfence();
    const unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return false;
  __threadfence();
  double t[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    double s = 0.0;
    for (int b = threadIdx.x; b < (int)gridDim.x; b += kBlock)
      s += __hip_atomic_load(&partials[k * gridDim.x + b], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    t[k] = s;
  }
  block_sum<NV>(t);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < NV; ++k) v[k] = t[k];
    *counter = 0u;
  }
  return true;
}

// --------------------------------------------------------------------------------------------------
// SpMV on node-block rows.  G lanes per block row; each lane walks the row's blocks with stride G.
// --------------------------------------------------------------------------------------------------
template <typename T, int D, int G, int LAYOUT>
__device__ inline void bsr_row(const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                               const T* __restrict__ vals, int64_t nnzb, const T* __restrict__ x,
                               int64_t i, int lane, T (&acc)[D]) {
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = T(0);
  const int32_t s = bptr[i], e = bptr[i + 1];
  for (int32_t q = s + lane; q < e; q += G) {
    const int64_t j = bcol[q];
    T xj[D];
#pragma unroll
    for (int d = 0; d < D; ++d) xj[d] = x[j * D + d];
#pragma unroll
    for (int di = 0; di < D; ++di)
#pragma unroll
      for (int dj = 0; dj < D; ++dj) acc[di] += vals[vix<D, LAYOUT>(q, di * D + dj, nnzb)] * xj[dj];
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], off, G);
}

template <typename T, int D>
__device__ inline T pick(const T (&acc)[D], int lane) {
  T out = acc[0];
#pragma unroll
  for (int d = 1; d < D; ++d)
    if (lane == d) out = acc[d];
  return out;
}

template <typename T, int D, int G, int LAYOUT>
__global__ __launch_bounds__(kBlock) void k_spmv_bsr(int64_t nb, int64_t nnzb,
                                                     const int32_t* __restrict__ bptr,
                                                     const int32_t* __restrict__ bcol,
                                                     const T* __restrict__ vals,
                                                     const T* __restrict__ x, T* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t i = gid / G;
  const int lane = (int)(gid % G);
  if (i >= nb) return;
  T acc[D];
  bsr_row<T, D, G, LAYOUT>(bptr, bcol, vals, nnzb, x, i, lane, acc);
  if (lane < D) y[i * D + lane] = pick<T, D>(acc, lane);
}

template <int G>
__global__ __launch_bounds__(kBlock) void k_spmv_csr(int64_t n, const int32_t* __restrict__ rowptr,
                                                     const int32_t* __restrict__ colind,
                                                     const double* __restrict__ vals,
                                                     const double* __restrict__ x,
                                                     double* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t i = gid / G;
  const int lane = (int)(gid % G);
  if (i >= n) return;
  double acc = 0.0;
  for (int32_t k = rowptr[i] + lane; k < rowptr[i + 1]; k += G) acc += vals[k] * x[colind[k]];
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1) acc += __shfl_xor(acc, off, G);
  if (lane == 0) y[i] = acc;
}

template <int D>
constexpr int lanes_for() {
  return D == 3 ? 16 : 8;
}

inline int lanes_for_dof(int d) { return d == 3 ? 16 : 8; }

// --------------------------------------------------------------------------------------------------
// value conversion between precisions and layouts
// --------------------------------------------------------------------------------------------------
template <typename S, typename T, int D, int LS, int LT>
__global__ void k_bsr_convert(int64_t nnzb, const S* __restrict__ src, T* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nnzb * D * D) return;
  // enumerate in destination order so that the stores are coalesced
  int64_t q;
  int k;
  if (LT == FE_LAYOUT_PLANES) {
    k = (int)(t / nnzb);
    q = t - (int64_t)k * nnzb;
  } else {
    q = t / (D * D);
    k = (int)(t - q * D * D);
  }
  dst[vix<D, LT>(q, k, nnzb)] = (T)src[vix<D, LS>(q, k, nnzb)];
}

// --------------------------------------------------------------------------------------------------
// preconditioner setup: inverse diagonal / inverse diagonal blocks (FP64)
// --------------------------------------------------------------------------------------------------
template <int D, int LAYOUT>
__global__ void k_precond_setup(int64_t nb, int64_t nnzb, const int32_t* __restrict__ bptr,
                                const int32_t* __restrict__ bcol, const double* __restrict__ vals,
                                int kind, double* __restrict__ out) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int64_t qd = -1;
  for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q)
    if (bcol[q] == i) qd = q;
  double blk[D * D];
#pragma unroll
  for (int k = 0; k < D * D; ++k)
    blk[k] = qd >= 0 ? vals[vix<D, LAYOUT>(qd, k, nnzb)] : ((k % (D + 1)) == 0 ? 1.0 : 0.0);
  if (kind == FE_PC_JACOBI) {
#pragma unroll
    for (int d = 0; d < D; ++d) {
      const double v = blk[d * D + d];
      out[i * D + d] = (v != 0.0) ? 1.0 / v : 1.0;
    }
    return;
  }
  double inv[D * D];
  if constexpr (D == 1) {
    inv[0] = blk[0] != 0.0 ? 1.0 / blk[0] : 1.0;
  } else if constexpr (D == 2) {
    const double det = blk[0] * blk[3] - blk[1] * blk[2];
    const double r = 1.0 / det;
    inv[0] = blk[3] * r;
    inv[1] = -blk[1] * r;
    inv[2] = -blk[2] * r;
    inv[3] = blk[0] * r;
  } else {
    const double* m = blk;
    const double c00 = m[4] * m[8] - m[5] * m[7];
    const double c01 = m[5] * m[6] - m[3] * m[8];
    const double c02 = m[3] * m[7] - m[4] * m[6];
    const double det = m[0] * c00 + m[1] * c01 + m[2] * c02;
    const double r = 1.0 / det;
    inv[0] = c00 * r;
    inv[1] = (m[2] * m[7] - m[1] * m[8]) * r;
    inv[2] = (m[1] * m[5] - m[2] * m[4]) * r;
    inv[3] = c01 * r;
    inv[4] = (m[0] * m[8] - m[2] * m[6]) * r;
    inv[5] = (m[2] * m[3] - m[0] * m[5]) * r;
    inv[6] = c02 * r;
    inv[7] = (m[1] * m[6] - m[0] * m[7]) * r;
    inv[8] = (m[0] * m[4] - m[1] * m[3]) * r;
  }
#pragma unroll
  for (int k = 0; k < D * D; ++k) out[i * D * D + k] = inv[k];
}

// --------------------------------------------------------------------------------------------------
// PCG (Chronopoulos-Gear form): one global reduction point per iteration.
//   K_a  k_cg_vec  : p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M r
//                    (streaming, one thread per DOF, coalesced; the node's r staged in LDS for M)
//   K_b  k_cg_spmv : w = A u, gamma = r.u, delta = w.u, rr = r.r  -> alpha, beta for the next K_a
// s = A p is carried by recurrence, so each iteration reads the matrix once and gathers one vector.
// Reliable updates (k_ru_residual, k_ru_prep, k_ru_spmv): r = b - A64 x in FP64, u = M r, then one
// matrix pass computes w = A u and s = A p together and the exact alpha of the continued direction
// from (u,w), (u,s), (p,s) -- the Krylov direction is kept, not restarted.
// --------------------------------------------------------------------------------------------------
struct PcgState {
  double alpha, beta, gamma, gamma_prev;
  double rr, rr_ref, rr_true, bb;
  double tol2bb, ru_delta2, rr_best, pad0;
  int32_t iter, max_iter;
  int32_t done;     // 0 running, 1 converged, 2 iteration limit / stagnation, 3 breakdown
  int32_t need_ru;  // 1: reliable update / true-residual check requested
  int32_t n_ru;
  int32_t first;
  int32_t mixed;
  int32_t stall;    // reliable updates since the true residual last improved
  int32_t restarts;
  int32_t pad1;
  unsigned counter[4];
};

struct PcgStatus {
  int32_t done, iter, n_ru, need_ru;
  double rr, rr_true, bb;
  int32_t restarts, pad;
};

constexpr int kStallLimit = 10;

// Contiguous node range of logical workgroup v (XCD-aware ids).
__device__ inline void block_range(int64_t nb, int64_t& lo, int64_t& hi) {
  const int64_t v = xcd_remap(blockIdx.x, gridDim.x);
  const int64_t per = ceil_div(nb, (int64_t)gridDim.x);
  lo = min(nb, v * per);
  hi = min(nb, lo + per);
}

// u_k = (M r)_k for DOF k = node*D + d, with the node's D residuals in LDS (s_r) for block Jacobi.
template <typename T, int D, int PC>
__device__ inline T precond_dof(const T* __restrict__ M, int64_t node, int d, int64_t k,
                                const T* s_rnode, T rk) {
  if (PC == FE_PC_JACOBI) return M[k] * rk;
  const T* m = M + node * D * D + d * D;
  T s = T(0);
#pragma unroll
  for (int j = 0; j < D; ++j) s += m[j] * s_rnode[j];
  return s;
}

constexpr int kVecNodes = 64;  // nodes per workgroup of the streaming kernels (64*D threads)

template <typename T, int D, int PC>
__global__ __launch_bounds__(kVecNodes * D) void k_cg_vec(int64_t nb, double* __restrict__ x,
                                                          T* __restrict__ p, T* __restrict__ s,
                                                          T* __restrict__ r, T* __restrict__ u,
                                                          const T* __restrict__ w,
                                                          const T* __restrict__ M,
                                                          const PcgState* st) {
  __shared__ T s_r[kVecNodes * D];
  __shared__ int s_skip;
  __shared__ double s_ab[2];
  if (threadIdx.x == 0) {
    s_skip = st->done || st->need_ru;
    s_ab[0] = st->alpha;
    s_ab[1] = st->beta;
  }
  __syncthreads();
  if (s_skip) return;
  const double alpha = s_ab[0];
  const T alpha_t = (T)alpha, beta = (T)s_ab[1];
  const int64_t n = nb * D;
  const int64_t k = (int64_t)blockIdx.x * kVecNodes * D + threadIdx.x;
  T rk = T(0);
  if (k < n) {
    const T pk = u[k] + beta * p[k];
    const T sk = w[k] + beta * s[k];
    p[k] = pk;
    s[k] = sk;
    x[k] += alpha * (double)pk;
    rk = r[k] - alpha_t * sk;
    r[k] = rk;
  }
  s_r[threadIdx.x] = rk;
  __syncthreads();
  if (k < n) {
    const int d = threadIdx.x % D;
    u[k] = precond_dof<T, D, PC>(M, k / D, d, k, &s_r[threadIdx.x - d], rk);
  }
}

// Each workgroup owns kSpmvPasses * (kBlock / G) consecutive node rows (one row per group of G lanes
// and pass) and writes its partial dots (no fences, no atomics); k_cg_reduce sums them in a fixed
// order and computes alpha, beta.  The row count per workgroup keeps the partials few (nb / 128 for
// D = 3) so the single-workgroup reduction stays short.
constexpr int kSpmvPasses = 8;

template <int G>
constexpr int64_t cg_rows_per_block() { return (int64_t)kSpmvPasses * (kBlock / G); }

template <typename T, int D, int G>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(int64_t nb, int64_t nnzb,
                                                    const int32_t* __restrict__ bptr,
                                                    const int32_t* __restrict__ bcol,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ u, const T* __restrict__ r,
                                                    T* __restrict__ w, const PcgState* st,
                                                    double* __restrict__ partials) {
  __shared__ int s_skip;
  if (threadIdx.x == 0) s_skip = st->done || st->need_ru;
  __syncthreads();
  if (s_skip) return;
  const int lane = threadIdx.x % G;
  const int64_t row0 = (int64_t)blockIdx.x * cg_rows_per_block<G>() + threadIdx.x / G;
  double v[3] = {0.0, 0.0, 0.0};
  for (int pass = 0; pass < kSpmvPasses; ++pass) {
    const int64_t i = row0 + (int64_t)pass * (kBlock / G);
    if (i >= nb) break;
    T acc[D];
    bsr_row<T, D, G, FE_LAYOUT_PLANES>(bptr, bcol, vals, nnzb, u, i, lane, acc);
    if (lane < D) {
      const int64_t k = i * D + lane;
      const T wk = pick<T, D>(acc, lane);
      w[k] = wk;
      const double uk = u[k], rk = r[k];
      v[0] += rk * uk;
      v[1] += (double)wk * uk;
      v[2] += rk * rk;
    }
  }
  block_sum<3>(v);
  if (threadIdx.x == 0) {
#pragma unroll
    for (int k = 0; k < 3; ++k) partials[(int64_t)k * gridDim.x + blockIdx.x] = v[k];
  }
}

constexpr int kRedThreads = 1024;

// Fixed-order sum of the SpMV partials, then the Chronopoulos-Gear scalars for the next iteration.
// The partial loads are issued before the state word is examined (they are plain reads), the three
// sums share each loop trip so their loads are in flight together, and wave 0 finishes the sum.
__global__ __launch_bounds__(kRedThreads) void k_cg_reduce(int nparts,
                                                           const double* __restrict__ partials,
                                                           PcgState* st) {
  __shared__ double red[3][kRedThreads / 64];
  double a0 = 0.0, a1 = 0.0, a2 = 0.0;
  for (int b = threadIdx.x; b < nparts; b += kRedThreads) {
    a0 += partials[b];
    a1 += partials[nparts + b];
    a2 += partials[2 * nparts + b];
  }
  const int skip = st->done || st->need_ru;
  if (skip) return;
  double v[3] = {wave_sum(a0), wave_sum(a1), wave_sum(a2)};
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) red[k][wid] = v[k];
  __syncthreads();
  if (wid != 0) return;
#pragma unroll
  for (int k = 0; k < 3; ++k) v[k] = wave_sum(lane < kRedThreads / 64 ? red[k][lane] : 0.0);
  if (lane != 0) return;
  const double g = v[0], dl = v[1];
  st->iter += 1;
  st->rr = v[2];
  if (!isfinite(g) || !isfinite(dl) || !isfinite(v[2]) || g < 0.0) {
    st->done = 3;
    return;
  }
  st->gamma_prev = st->gamma;
  st->gamma = g;
  // convergence claims, the iteration limit and (mixed precision) the reliable-update threshold
  // hand over to the true-residual check at the end of the chunk
  if (v[2] <= st->tol2bb || st->iter >= st->max_iter ||
      (st->mixed && v[2] < st->ru_delta2 * st->rr_ref)) {
    st->need_ru = 1;
    return;
  }
  const double beta = g / st->gamma_prev;
  const double den = dl - beta * g / st->alpha;
  if (!(den > 0.0)) {
    // the recurrences lost positivity (rounding, most often in FP32): restart from the true
    // residual; a genuine breakdown shows up there as p.Ap <= 0
    st->need_ru = 1;
    st->first = 1;
    st->restarts += 1;
    return;
  }
  st->beta = beta;
  st->alpha = g / den;
}

// Reliable update, part 1: r64 = b - A64 x, partial ||r64||^2.
template <int D, int G>
__global__ __launch_bounds__(kBlock) void k_ru_residual(int64_t nb, int64_t nnzb,
                                                        const int32_t* __restrict__ bptr,
                                                        const int32_t* __restrict__ bcol,
                                                        const double* __restrict__ vals,
                                                        const double* __restrict__ b,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ r64, PcgState* st,
                                                        double* partials) {
  __shared__ int s_skip;
  if (threadIdx.x == 0) s_skip = st->done || !st->need_ru;
  __syncthreads();
  if (s_skip) return;
  int64_t lo, hi;
  block_range(nb, lo, hi);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  double rr = 0.0;
  for (int64_t i = lo + grp; i < hi; i += kBlock / G) {
    double acc[D];
    bsr_row<double, D, G, FE_LAYOUT_PLANES>(bptr, bcol, vals, nnzb, x, i, lane, acc);
    if (lane < D) {
      const int64_t k = i * D + lane;
      const double ri = b[k] - pick<double, D>(acc, lane);
      r64[k] = ri;
      rr += ri * ri;
    }
  }
  double v[1] = {rr};
  if (grid_sum<1>(v, partials, &st->counter[1]) && threadIdx.x == 0) {
    st->rr_true = v[0];
    // decide here: converged, out of iterations, or continue with the replaced residual
    if (v[0] <= st->tol2bb) {
      st->done = 1;
    } else if (st->iter >= st->max_iter) {
      st->done = 2;
    } else {
      if (v[0] < 0.81 * st->rr_best) {
        st->rr_best = v[0];
        st->stall = 0;
      } else if (++st->stall >= kStallLimit) {
        st->done = 2;
      }
      // the recursive residual drifted far below the true one: the direction is stale, restart
      if (!st->first && v[0] > 4.0 * st->rr) {
        st->first = 1;
        st->restarts += 1;
      }
    }
    if (st->done) {
      st->rr = v[0];
      st->n_ru += 1;
      st->need_ru = 0;
    }
  }
}

// Reliable update, part 2: r = r64, u = M r (and p, s cleared on a (re)start).
template <typename T, int D, int PC>
__global__ __launch_bounds__(kVecNodes * D) void k_ru_prep(int64_t nb, const double* __restrict__ r64,
                                                           T* __restrict__ r, T* __restrict__ u,
                                                           T* __restrict__ p, T* __restrict__ s,
                                                           const T* __restrict__ M,
                                                           const PcgState* st) {
  __shared__ T s_r[kVecNodes * D];
  __shared__ int s_skip, s_first;
  if (threadIdx.x == 0) {
    s_skip = st->done || !st->need_ru;
    s_first = st->first;
  }
  __syncthreads();
  if (s_skip) return;
  const int64_t n = nb * D;
  const int64_t k = (int64_t)blockIdx.x * kVecNodes * D + threadIdx.x;
  T rk = T(0);
  if (k < n) {
    rk = (T)r64[k];
    r[k] = rk;
    if (s_first) {
      p[k] = T(0);
      s[k] = T(0);
    }
  }
  s_r[threadIdx.x] = rk;
  __syncthreads();
  if (k < n) {
    const int d = threadIdx.x % D;
    u[k] = precond_dof<T, D, PC>(M, k / D, d, k, &s_r[threadIdx.x - d], rk);
  }
}

// Reliable update, part 3: one matrix pass for w = A u and s = A p; dots (r,u), (u,w), (u,s),
// (p,s), (r,r); beta = gamma/gamma_prev (0 on a (re)start) and the exact alpha of the direction
// p' = u + beta p: p'.Ap' = (u,w) + 2 beta (u,s) + beta^2 (p,s).
template <typename T, int D, int G>
__global__ __launch_bounds__(kBlock) void k_ru_spmv(int64_t nb, int64_t nnzb,
                                                    const int32_t* __restrict__ bptr,
                                                    const int32_t* __restrict__ bcol,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ u, const T* __restrict__ p,
                                                    const T* __restrict__ r, T* __restrict__ w,
                                                    T* __restrict__ s, PcgState* st,
                                                    double* partials) {
  __shared__ int s_skip;
  if (threadIdx.x == 0) s_skip = st->done || !st->need_ru;
  __syncthreads();
  if (s_skip) return;
  int64_t lo, hi;
  block_range(nb, lo, hi);
  const int grp = threadIdx.x / G, lane = threadIdx.x % G;
  double acc5[5] = {0, 0, 0, 0, 0};
  for (int64_t i = lo + grp; i < hi; i += kBlock / G) {
    T au[D], ap[D];
#pragma unroll
    for (int d = 0; d < D; ++d) au[d] = ap[d] = T(0);
    const int32_t b0 = bptr[i], b1 = bptr[i + 1];
    for (int32_t q = b0 + lane; q < b1; q += G) {
      const int64_t j = bcol[q];
      T uj[D], pj[D];
#pragma unroll
      for (int d = 0; d < D; ++d) {
        uj[d] = u[j * D + d];
        pj[d] = p[j * D + d];
      }
#pragma unroll
      for (int di = 0; di < D; ++di)
#pragma unroll
        for (int dj = 0; dj < D; ++dj) {
          const T a = vals[vix<D, FE_LAYOUT_PLANES>(q, di * D + dj, nnzb)];
          au[di] += a * uj[dj];
          ap[di] += a * pj[dj];
        }
    }
#pragma unroll
    for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
      for (int d = 0; d < D; ++d) {
        au[d] += __shfl_xor(au[d], off, G);
        ap[d] += __shfl_xor(ap[d], off, G);
      }
    if (lane < D) {
      const int64_t k = i * D + lane;
      const T wk = pick<T, D>(au, lane), sk = pick<T, D>(ap, lane);
      w[k] = wk;
      s[k] = sk;
      const double uk = u[k], pk = p[k], rk = r[k];
      acc5[0] += rk * uk;
      acc5[1] += uk * (double)wk;
      acc5[2] += uk * (double)sk;
      acc5[3] += pk * (double)sk;
      acc5[4] += rk * rk;
    }
  }
  if (grid_sum<5>(acc5, partials, &st->counter[2]) && threadIdx.x == 0) {
    const double g = acc5[0];
    if (!isfinite(g) || g < 0.0 || !isfinite(acc5[1])) {
      st->done = 3;
      return;
    }
    // gamma_prev holds r.u of the iteration before the one whose residual was replaced
    const double beta = st->first ? 0.0 : g / st->gamma_prev;
    const double pap = acc5[1] + 2.0 * beta * acc5[2] + beta * beta * acc5[3];
    if (!(pap > 0.0)) {
      st->done = 3;
      return;
    }
    st->beta = beta;
    st->alpha = g / pap;
    st->gamma = g;
    st->first = 0;
    st->rr = acc5[4];
    st->rr_ref = st->rr_true;
    st->n_ru += 1;
    st->need_ru = 0;
  }
}

__global__ void k_pcg_init(int64_t n, const double* __restrict__ b, PcgState* st, double* partials,
                           double tol, int max_iter, int mixed, double ru_delta) {
  double s = 0.0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n;
       i += (int64_t)gridDim.x * kBlock)
    s += b[i] * b[i];
  double v[1] = {s};
  if (grid_sum<1>(v, partials, &st->counter[3]) && threadIdx.x == 0) {
    st->bb = v[0];
    st->tol2bb = tol * tol * v[0];
    st->ru_delta2 = ru_delta * ru_delta;
    st->max_iter = max_iter;
    st->mixed = mixed;
    st->need_ru = 1;
    st->first = 1;
    st->iter = 0;
    st->done = 0;
    st->n_ru = 0;
    st->alpha = 1.0;
    st->beta = 0.0;
    st->gamma = 1.0;
    st->gamma_prev = 1.0;
    st->rr = 0.0;
    st->rr_best = 1.0e308;
    st->stall = 0;
    st->restarts = 0;
  }
}

__global__ void k_pcg_status(const PcgState* st, PcgStatus* out) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out->done = st->done;
    out->iter = st->iter;
    out->n_ru = st->n_ru;
    out->need_ru = st->need_ru;
    out->rr = st->rr;
    out->rr_true = st->rr_true;
    out->bb = st->bb;
    out->restarts = st->restarts;
  }
}

// --------------------------------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------------------------------
struct Work {
  PcgState* st;
  PcgStatus* status;
  double* partials;
  double* spmv_partials;
  void* r;
  void* u;
  void* w;
  void* p;
  void* s;
  double* r64;
  int64_t bytes;
};

constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

// Fixed grid of the reduction kernels: 8 workgroups per CU (a multiple of the 8 XCDs), fewer for
// small systems.
int pcg_grid(const fe_handle* h, int64_t nb) {
  int64_t g = (int64_t)h->num_cus * 8;
  if (g < 8) g = 8;
  g = (g + 7) / 8 * 8;
  const int64_t need = ceil_div(nb, 16);
  while (g > 8 && g / 2 >= need) g /= 2;
  return (int)g;
}

Work layout_work(char* base, int64_t nb, int dof, int elem, int grid) {
  const int64_t n = nb * dof;
  const int64_t spmv_blocks = ceil_div(nb, (int64_t)kSpmvPasses * (kBlock / lanes_for_dof(dof)));
  Work w{};
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  w.st = (PcgState*)take(sizeof(PcgState));
  w.status = (PcgStatus*)take(sizeof(PcgStatus));
  w.partials = (double*)take((int64_t)5 * grid * sizeof(double));
  w.spmv_partials = (double*)take(3 * spmv_blocks * (int64_t)sizeof(double));
  w.r = take(n * elem);
  w.u = take(n * elem);
  w.w = take(n * elem);
  w.p = take(n * elem);
  w.s = take(n * elem);
  w.r64 = (double*)take(n * (int64_t)sizeof(double));
  w.bytes = off;
  return w;
}

template <typename T, int D, int PC>
int run_pcg(fe_handle* h, const fe_bsr* A64, const fe_bsr* Ait, const T* M, const double* b,
            double* x, const Work& w, int grid, const fe_pcg_opts* o, fe_pcg_stats* s) {
  constexpr int G = lanes_for<D>();
  const int64_t nb = A64->n_brows, n = nb * D, nnzb = A64->nnzb;
  hipStream_t st = h->stream;
  FE_CHECK_HIP(hipMemsetAsync(w.st, 0, sizeof(PcgState), st));
  FE_CHECK_HIP(hipMemsetAsync(w.p, 0, n * sizeof(T), st));
  FE_CHECK_HIP(hipMemsetAsync(w.s, 0, n * sizeof(T), st));
  hipLaunchKernelGGL(k_pcg_init, dim3(grid), dim3(kBlock), 0, st, n, b, w.st, w.partials, o->tol,
                     o->max_iter, o->mixed, o->ru_delta);
  FE_CHECK_LAUNCH();
  const int32_t* bptr = A64->bptr;
  const int32_t* bcol = A64->bcol;
  const double* v64 = (const double*)A64->vals;
  const T* vit = (const T*)Ait->vals;
  T* r = (T*)w.r;
  T* u = (T*)w.u;
  T* wv = (T*)w.w;
  T* p = (T*)w.p;
  T* sv = (T*)w.s;
  const dim3 gvec((unsigned)ceil_div(nb, kVecNodes)), bvec(kVecNodes * D);
  const dim3 gspmv((unsigned)ceil_div(nb, cg_rows_per_block<G>()));
  auto launch_ru = [&]() -> int {
    hipLaunchKernelGGL((k_ru_residual<D, G>), dim3(grid), dim3(kBlock), 0, st, nb, nnzb, bptr, bcol,
                       v64, b, x, w.r64, w.st, w.partials);
    hipLaunchKernelGGL((k_ru_prep<T, D, PC>), gvec, bvec, 0, st, nb, w.r64, r, u, p, sv, M, w.st);
    hipLaunchKernelGGL((k_ru_spmv<T, D, G>), dim3(grid), dim3(kBlock), 0, st, nb, nnzb, bptr, bcol, vit,
                       u, p, r, wv, sv, w.st, w.partials);
    FE_CHECK_LAUNCH();
    return FE_OK;
  };
  int rc = launch_ru();
  if (rc) return rc;
  const int chunk = o->check_every > 0 ? o->check_every : 16;
  PcgStatus* pinned = (PcgStatus*)h->pinned;
  hipEvent_t ev[2];
  hipEvent_t pev[2][2];  // profile: [slot][start/stop] around the chunk's first SpMV kernel
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  const bool prof = o->profile != 0;
  if (prof)
    for (int a = 0; a < 2; ++a)
      for (int b2 = 0; b2 < 2; ++b2) FE_CHECK_HIP(hipEventCreate(&pev[a][b2]));
  auto enqueue_chunk = [&](int slot) -> int {
    for (int it = 0; it < chunk; ++it) {
      hipLaunchKernelGGL((k_cg_vec<T, D, PC>), gvec, bvec, 0, st, nb, x, p, sv, r, u, wv, M, w.st);
      if (prof && it == 0) FE_CHECK_HIP(hipEventRecord(pev[slot][0], st));
      hipLaunchKernelGGL((k_cg_spmv<T, D, G>), gspmv, dim3(kBlock), 0, st, nb, nnzb, bptr, bcol, vit, u,
                         r, wv, w.st, w.spmv_partials);
      if (prof && it == 0) FE_CHECK_HIP(hipEventRecord(pev[slot][1], st));
      hipLaunchKernelGGL(k_cg_reduce, dim3(1), dim3(kRedThreads), 0, st, (int)gspmv.x, w.spmv_partials,
                         w.st);
    }
    int rc2 = launch_ru();
    if (rc2) return rc2;
    hipLaunchKernelGGL(k_pcg_status, dim3(1), dim3(64), 0, st, w.st, w.status);
    FE_CHECK_HIP(hipMemcpyAsync(&pinned[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, st));
    FE_CHECK_HIP(hipEventRecord(ev[slot], st));
    return FE_OK;
  };
  // One chunk is always queued ahead of the one being polled, so the device never waits for the
  // host; chunks queued after convergence find the done flag set and return immediately.
  const int64_t max_chunks = 4 * ((int64_t)o->max_iter / chunk + 2) + 8;
  PcgStatus last{};
  int finished = 0;
  double prof_ms = 0.0;
  int prof_n = 0;
  int prev_running = 1;  // the chunk before the polled one left the solver running (no pause)
  rc = enqueue_chunk(0);
  if (rc == FE_OK) rc = enqueue_chunk(1);
  for (int64_t c = 0; rc == FE_OK && c < max_chunks; ++c) {
    const int slot = (int)(c & 1);
    FE_CHECK_HIP(hipEventSynchronize(ev[slot]));
    last = pinned[slot];
    if (prof && prev_running) {
      // the sampled SpMV ran in full only if the solver was running when this chunk started
      float ms = 0.f;
      if (hipEventElapsedTime(&ms, pev[slot][0], pev[slot][1]) == hipSuccess) {
        prof_ms += ms;
        prof_n += 1;
      }
    }
    prev_running = !last.done && !last.need_ru;
    if (last.done) {
      finished = 1;
      break;
    }
    rc = enqueue_chunk(slot);
  }
  if (prof)
    for (int a = 0; a < 2; ++a)
      for (int b2 = 0; b2 < 2; ++b2) (void)hipEventDestroy(pev[a][b2]);
  s->spmv_ms = prof_n ? prof_ms / prof_n : 0.0;
  s->spmv_samples = prof_n;
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  if (rc) return rc;
  FE_CHECK_HIP(hipStreamSynchronize(st));
  if (!finished) {
    hipLaunchKernelGGL(k_pcg_status, dim3(1), dim3(64), 0, st, w.st, w.status);
    FE_CHECK_HIP(hipMemcpyAsync(&pinned[0], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, st));
    FE_CHECK_HIP(hipStreamSynchronize(st));
    last = pinned[0];
  }
  s->iterations = last.iter;
  s->reliable_updates = last.n_ru;
  s->restarts = last.restarts;
  s->bnorm = sqrt(last.bb);
  s->relres = last.bb > 0 ? sqrt(last.rr_true / last.bb) : 0.0;
  s->relres_rec = last.bb > 0 ? sqrt(last.rr / last.bb) : 0.0;
  if (last.done == 1)
    s->status = FE_OK;
  else if (last.done == 3)
    s->status = FE_ERR_BREAKDOWN;
  else
    s->status = FE_ERR_NOT_CONVERGED;
  return FE_OK;
}

template <typename T, int D>
int dispatch_pc(fe_handle* h, const fe_bsr* A64, const fe_bsr* Ait, const void* M, const double* b,
                double* x, const Work& w, int grid, const fe_pcg_opts* o, fe_pcg_stats* s) {
  if (o->precond == FE_PC_JACOBI)
    return run_pcg<T, D, FE_PC_JACOBI>(h, A64, Ait, (const T*)M, b, x, w, grid, o, s);
  return run_pcg<T, D, FE_PC_BLOCK_JACOBI>(h, A64, Ait, (const T*)M, b, x, w, grid, o, s);
}

template <typename T>
int dispatch_dof(fe_handle* h, const fe_bsr* A64, const fe_bsr* Ait, const void* M, const double* b,
                 double* x, const Work& w, int grid, const fe_pcg_opts* o, fe_pcg_stats* s) {
  switch (A64->dof) {
    case 1: return dispatch_pc<T, 1>(h, A64, Ait, M, b, x, w, grid, o, s);
    case 2: return dispatch_pc<T, 2>(h, A64, Ait, M, b, x, w, grid, o, s);
    case 3: return dispatch_pc<T, 3>(h, A64, Ait, M, b, x, w, grid, o, s);
    default: return fe_set_error(FE_ERR_ARG, "fe_pcg: dof must be 1..3");
  }
}

template <typename T, int D, int LAYOUT>
void launch_spmv(fe_handle* h, const fe_bsr* A, const void* x, void* y) {
  constexpr int G = lanes_for<D>();
  const int64_t threads = A->n_brows * G;
  hipLaunchKernelGGL((k_spmv_bsr<T, D, G, LAYOUT>), dim3((unsigned)ceil_div(threads, kBlock)),
                     dim3(kBlock), 0, h->stream, A->n_brows, A->nnzb, A->bptr, A->bcol,
                     (const T*)A->vals, (const T*)x, (T*)y);
}

template <typename T, int D>
void launch_spmv_l(fe_handle* h, const fe_bsr* A, const void* x, void* y) {
  if (A->layout == FE_LAYOUT_PLANES)
    launch_spmv<T, D, FE_LAYOUT_PLANES>(h, A, x, y);
  else
    launch_spmv<T, D, FE_LAYOUT_BSR>(h, A, x, y);
}

int check_bsr(const fe_bsr* A, const char* who) {
  if (!A) return fe_set_error(FE_ERR_ARG, "%s: null matrix", who);
  if (A->dof < 1 || A->dof > 3) return fe_set_error(FE_ERR_ARG, "%s: dof must be 1..3", who);
  if (A->precision != 0 && A->precision != 1)
    return fe_set_error(FE_ERR_ARG, "%s: precision must be 0 (FP64) or 1 (FP32)", who);
  if (A->layout != FE_LAYOUT_BSR && A->layout != FE_LAYOUT_PLANES)
    return fe_set_error(FE_ERR_ARG, "%s: layout must be FE_LAYOUT_BSR or FE_LAYOUT_PLANES", who);
  if (A->n_brows < 0 || A->nnzb < 0) return fe_set_error(FE_ERR_ARG, "%s: negative sizes", who);
  if (A->n_brows > 0 && (!A->bptr || !A->bcol || !A->vals))
    return fe_set_error(FE_ERR_ARG, "%s: null matrix array", who);
  return FE_OK;
}

template <typename S, typename T, int D>
void launch_convert(fe_handle* h, const fe_bsr* A, const fe_bsr* B) {
  const int64_t n = A->nnzb * D * D;
  dim3 grid((unsigned)ceil_div(n, 256)), block(256);
  const S* src = (const S*)A->vals;
  T* dst = (T*)B->vals;
  if (A->layout == FE_LAYOUT_PLANES) {
    if (B->layout == FE_LAYOUT_PLANES)
      hipLaunchKernelGGL((k_bsr_convert<S, T, D, FE_LAYOUT_PLANES, FE_LAYOUT_PLANES>), grid, block, 0,
                         h->stream, A->nnzb, src, dst);
    else
      hipLaunchKernelGGL((k_bsr_convert<S, T, D, FE_LAYOUT_PLANES, FE_LAYOUT_BSR>), grid, block, 0,
                         h->stream, A->nnzb, src, dst);
  } else {
    if (B->layout == FE_LAYOUT_PLANES)
      hipLaunchKernelGGL((k_bsr_convert<S, T, D, FE_LAYOUT_BSR, FE_LAYOUT_PLANES>), grid, block, 0,
                         h->stream, A->nnzb, src, dst);
    else
      hipLaunchKernelGGL((k_bsr_convert<S, T, D, FE_LAYOUT_BSR, FE_LAYOUT_BSR>), grid, block, 0,
                         h->stream, A->nnzb, src, dst);
  }
}

template <typename S, typename T>
void launch_convert_d(fe_handle* h, const fe_bsr* A, const fe_bsr* B) {
  if (A->dof == 1) launch_convert<S, T, 1>(h, A, B);
  else if (A->dof == 2) launch_convert<S, T, 2>(h, A, B);
  else launch_convert<S, T, 3>(h, A, B);
}

}  // namespace

extern "C" int fe_spmv(fe_handle* h, const fe_bsr* A, const void* x, void* y) {
  FE_ENTER(h);
  int rc = check_bsr(A, "fe_spmv");
  if (rc) return rc;
  if (A->n_brows == 0) return FE_OK;
  FE_REQUIRE(x && y && x != y, "fe_spmv: x and y must be distinct non-null vectors");
  if (A->precision == 0) {
    if (A->dof == 1) launch_spmv_l<double, 1>(h, A, x, y);
    else if (A->dof == 2) launch_spmv_l<double, 2>(h, A, x, y);
    else launch_spmv_l<double, 3>(h, A, x, y);
  } else {
    if (A->dof == 1) launch_spmv_l<float, 1>(h, A, x, y);
    else if (A->dof == 2) launch_spmv_l<float, 2>(h, A, x, y);
    else launch_spmv_l<float, 3>(h, A, x, y);
  }
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_spmv_csr(fe_handle* h, int64_t n_rows, const int32_t* rowptr,
                           const int32_t* colind, const double* vals, const double* x, double* y) {
  FE_ENTER(h);
  FE_REQUIRE(n_rows >= 0, "fe_spmv_csr: negative n_rows");
  if (n_rows == 0) return FE_OK;
  FE_REQUIRE(rowptr && colind && vals && x && y && x != y, "fe_spmv_csr: bad pointers");
  constexpr int G = 8;
  hipLaunchKernelGGL(k_spmv_csr<G>, dim3((unsigned)ceil_div(n_rows * G, kBlock)), dim3(kBlock), 0,
                     h->stream, n_rows, rowptr, colind, vals, x, y);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_bsr_convert(fe_handle* h, const fe_bsr* src, const fe_bsr* dst) {
  FE_ENTER(h);
  int rc = check_bsr(src, "fe_bsr_convert(src)");
  if (rc) return rc;
  rc = check_bsr(dst, "fe_bsr_convert(dst)");
  if (rc) return rc;
  FE_REQUIRE(src->dof == dst->dof && src->nnzb == dst->nnzb && src->n_brows == dst->n_brows,
             "fe_bsr_convert: src and dst must share the pattern");
  FE_REQUIRE(src->vals != dst->vals, "fe_bsr_convert: must be out-of-place");
  if (src->nnzb == 0) return FE_OK;
  if (src->precision == 0 && dst->precision == 0) launch_convert_d<double, double>(h, src, dst);
  else if (src->precision == 0) launch_convert_d<double, float>(h, src, dst);
  else if (dst->precision == 0) launch_convert_d<float, double>(h, src, dst);
  else launch_convert_d<float, float>(h, src, dst);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_precond_setup(fe_handle* h, const fe_bsr* A, int kind, double* out) {
  FE_ENTER(h);
  int rc = check_bsr(A, "fe_precond_setup");
  if (rc) return rc;
  FE_REQUIRE(out, "fe_precond_setup: null output");
  FE_REQUIRE(A->precision == 0, "fe_precond_setup: FP64 matrix expected");
  FE_REQUIRE(kind == FE_PC_JACOBI || kind == FE_PC_BLOCK_JACOBI, "fe_precond_setup: bad kind");
  if (A->n_brows == 0) return FE_OK;
  dim3 grid((unsigned)ceil_div(A->n_brows, 256)), block(256);
  const double* v = (const double*)A->vals;
#define FE_PC_LAUNCH(D)                                                                           \
  if (A->layout == FE_LAYOUT_PLANES)                                                              \
    hipLaunchKernelGGL((k_precond_setup<D, FE_LAYOUT_PLANES>), grid, block, 0, h->stream,         \
                       A->n_brows, A->nnzb, A->bptr, A->bcol, v, kind, out);                      \
  else                                                                                            \
    hipLaunchKernelGGL((k_precond_setup<D, FE_LAYOUT_BSR>), grid, block, 0, h->stream, A->n_brows, \
                       A->nnzb, A->bptr, A->bcol, v, kind, out);
  if (A->dof == 1) {
    FE_PC_LAUNCH(1)
  } else if (A->dof == 2) {
    FE_PC_LAUNCH(2)
  } else {
    FE_PC_LAUNCH(3)
  }
#undef FE_PC_LAUNCH
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_pcg_workspace(fe_handle* h, const fe_bsr* A, const fe_pcg_opts* opts,
                                int64_t* bytes) {
  FE_ENTER(h);
  FE_REQUIRE(A && opts && bytes, "fe_pcg_workspace: null argument");
  const int grid = pcg_grid(h, A->n_brows);
  const int elem = opts->mixed ? (int)sizeof(float) : (int)sizeof(double);
  *bytes = layout_work(nullptr, A->n_brows, A->dof, elem, grid).bytes;
  return FE_OK;
}

extern "C" int fe_pcg(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                      const void* M32, const double* b, double* x, void* work, int64_t work_bytes,
                      const fe_pcg_opts* o, fe_pcg_stats* s) {
  FE_ENTER(h);
  FE_REQUIRE(b && x && work && o && s, "fe_pcg: null argument");
  int rc = check_bsr(A64, "fe_pcg(A64)");
  if (rc) return rc;
  FE_REQUIRE(A64->precision == 0 && A64->layout == FE_LAYOUT_PLANES,
             "fe_pcg: A64 must be FP64 in FE_LAYOUT_PLANES");
  FE_REQUIRE(o->precond == FE_PC_JACOBI || o->precond == FE_PC_BLOCK_JACOBI, "fe_pcg: bad precond");
  FE_REQUIRE(o->tol > 0 && o->max_iter > 0, "fe_pcg: tol and max_iter must be positive");
  if (o->mixed) {
    rc = check_bsr(A32, "fe_pcg(A32)");
    if (rc) return rc;
    FE_REQUIRE(M32, "fe_pcg: mixed precision needs the FP32 preconditioner");
    FE_REQUIRE(A32->precision == 1 && A32->layout == FE_LAYOUT_PLANES && A32->dof == A64->dof &&
                   A32->n_brows == A64->n_brows && A32->nnzb == A64->nnzb,
               "fe_pcg: A32 must be the FP32 FE_LAYOUT_PLANES copy of A64");
  } else {
    FE_REQUIRE(M64, "fe_pcg: missing preconditioner");
  }
  *s = fe_pcg_stats{};
  if (A64->n_brows == 0) return FE_OK;
  const int grid = pcg_grid(h, A64->n_brows);
  const int elem = o->mixed ? (int)sizeof(float) : (int)sizeof(double);
  const int64_t n = A64->n_brows * A64->dof;
  Work w = layout_work((char*)work, A64->n_brows, A64->dof, elem, grid);
  FE_REQUIRE(work_bytes >= w.bytes, "fe_pcg: workspace too small (%lld < %lld)",
             (long long)work_bytes, (long long)w.bytes);
  if (o->mixed) return dispatch_dof<float>(h, A64, A32, M32, b, x, w, grid, o, s);
  return dispatch_dof<double>(h, A64, A64, M64, b, x, w, grid, o, s);
}
