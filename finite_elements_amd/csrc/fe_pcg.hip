// Device-driven PCG on independent segments of one block-diagonal system (one segment = the whole
// system for fe_pcg; one segment per plate / design variant for fe_pcg_batched).
//
// Chronopoulos-Gear recurrences, one global reduction point per iteration, three kernels:
//   k_cg_vec    p = u + beta p, s = w + beta s, x += alpha p, r -= alpha s, u = M r
//               (streaming, one thread per DOF; the node's r staged in LDS for block Jacobi)
//   k_cg_spmv   w = A u and the partial dots r.u, w.u, r.r of one tile
//   k_cg_reduce per segment: fixed-order sum of its tiles' partials, then alpha, beta
// s = A p is carried by recurrence, so an iteration reads the matrix once and gathers one vector.
// Work is cut into tiles: a tile is a run of at most kSpmvPasses * (kBlock / G) consecutive node rows
// of ONE segment, so every partial sum belongs to one segment and segments never mix.  Each
// segment carries its own scalars and stops on its own; kernels skip the tiles of finished segments.
//
// Reliable updates (FP64 true residual) end every chunk of iterations for the segments that asked for
// one: when the recursive residual reached the tolerance (convergence is always judged on the true
// FP64 residual), at the iteration limit, or (mixed precision) when the residual fell by ru_delta
// since the last update.  The first two stop the segment's iterations until the update; the third
// only marks it pending and the segment iterates on to the end of the chunk, so no launched
// iteration idles.  r = b - A64 x, u = M r, then one matrix pass computes w = A u and s = A p
// together and the exact alpha of the continued direction from (u,w), (u,s), (p,s): the Krylov
// direction is kept, not restarted.  When the true residual disagrees with the recursive one the
// segment restarts from the true residual; ten updates without progress stop it as not converged.
// Every reduction is a fixed-order sum (run-to-run bitwise reproducible).  The host polls a status
// word one chunk behind; a chunk can be replayed from a captured HIP graph.
#include <algorithm>
#include <vector>

#include "fe_sparse.h"

namespace {

using namespace fe_sparse;

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
  int32_t ru_pending;  // mixed: the residual fell by ru_delta; update at the end of the chunk
};

struct PcgStatus {
  int32_t active, max_iter, n_ru, pad;
  double rr_ratio;  // largest rr / (tol^2 b.b) over the running segments (the distance to the target)
};

struct Tile {
  int32_t seg, row0, row1, pad;
};

// int32 words between consecutive segments' states (the coarse correction's per-segment checks)
constexpr int kStateStride = (int)(sizeof(PcgState) / sizeof(int32_t));
static_assert(sizeof(PcgState) % sizeof(int32_t) == 0, "PcgState is whole int32 words");

constexpr int kStallLimit = 10;
constexpr int kMaxRestarts32 = 3;   // FP32: more restarts than this and the segment goes to FP64
constexpr int kMaxRestarts64 = 50;
constexpr int kVecNodes = 64;  // nodes per pass of the streaming kernels (64 * D threads)
constexpr int kSpmvPasses = 8;
constexpr int kRed = 256;      // threads of the per-segment reduction kernels

template <int G>
constexpr int tile_rows() { return kSpmvPasses * (kBlock / G); }
inline int tile_rows_dof(int d) { return kSpmvPasses * (kBlock / lanes_for_dof(d)); }

// Logical tile of this workgroup: consecutive workgroups are dispatched round-robin over the 8 XCDs,
// so the remap gives every XCD (and its L2) one contiguous run of tiles.  The tile count is padded
// to a multiple of 8 with empty tiles.
__device__ inline int tile_id() { return (int)xcd_remap(blockIdx.x, gridDim.x); }

__device__ inline bool seg_running(const PcgState* s) { return !s->done && !s->need_ru; }

// This workgroup's tile: with one segment (n_rows >= 0, its node rows) the tiles are regular -- tile t
// holds rows [t R, min((t + 1) R, n_rows)), the padding tiles none -- and are formed without a load.
template <int D>
__device__ inline Tile tile_of(const Tile* __restrict__ tiles, int t, int32_t n_rows) {
  constexpr int R = tile_rows<lanes_for<D>()>();
  if (n_rows < 0) return tiles[t];
  const int64_t r0 = min((int64_t)t * R, (int64_t)n_rows);
  return Tile{0, (int32_t)r0, (int32_t)min(r0 + R, (int64_t)n_rows), 0};
}

// The running test with both words loaded before either is tested (short-circuit evaluation would
// make them two dependent round trips; the empty asm keeps the loads together).
__device__ inline bool seg_running_now(const PcgState* s) {
  const int32_t dn = s->done, nr = s->need_ru;
  asm volatile("" ::"s"(dn), "s"(nr));
  return (dn | nr) == 0;
}
__device__ inline bool seg_updating(const PcgState* s) { return !s->done && s->need_ru; }

// Sum NV partial arrays (each ntiles long) over the tiles [t0, t1) of one segment; result in thread 0.
// A thread adds its tiles t0 + tid + j kRed in order of j; the loads of kSumBatch consecutive j are
// issued together (a 64^3 cube has 2,152 tiles: one round of loads instead of nine).
constexpr int kSumBatch = 8;
template <int NV>
__device__ inline void seg_sum(const double* __restrict__ partials, int ntiles, int t0, int t1,
                               double (&v)[NV]) {
  __shared__ double red[NV][kRed / 64];
#pragma unroll
  for (int k = 0; k < NV; ++k) v[k] = 0.0;
  for (int t = t0 + (int)threadIdx.x; t < t1; t += kSumBatch * kRed) {
    double a[kSumBatch][NV];
#pragma unroll
    for (int j = 0; j < kSumBatch; ++j)
#pragma unroll
      for (int k = 0; k < NV; ++k)
        a[j][k] = t + j * kRed < t1 ? partials[(int64_t)k * ntiles + t + j * kRed] : 0.0;
#pragma unroll
    for (int j = 0; j < kSumBatch; ++j)
      if (t + j * kRed < t1)
#pragma unroll
        for (int k = 0; k < NV; ++k) v[k] += a[j][k];
  }
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
      for (int w = 0; w < kRed / 64; ++w) s += red[k][w];
      v[k] = s;
    }
}

// u_k = (M r)_k for DOF k = node*D + d, with the node's D residuals in LDS (s_rnode) for block Jacobi.
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

// --------------------------------------------------------------------------------------------------
// iteration kernels
// --------------------------------------------------------------------------------------------------
// SPEC (one segment, whose tiles all run until the solve ends): the tile's vectors and preconditioner
// rows are loaded before the state test, so that its latency overlaps theirs; a launch that finds the
// segment finished then drops them.  A batch (SPEC false) loads only for running segments.
template <typename T, int D, int PC, bool SPEC>
__global__ __launch_bounds__(kVecNodes * D) void k_cg_vec(const Tile* __restrict__ tiles,
                                                          double* __restrict__ x, T* __restrict__ p,
                                                          T* __restrict__ s, T* __restrict__ r,
                                                          T* __restrict__ u, const T* __restrict__ w,
                                                          const T* __restrict__ M,
                                                          const PcgState* st, int32_t n_rows) {
  constexpr int kPasses = tile_rows<lanes_for<D>()>() / kVecNodes;
  static_assert(kPasses * kVecNodes == tile_rows<lanes_for<D>()>(), "tile = whole passes");
  constexpr int NM = PC == FE_PC_JACOBI ? 1 : D;  // preconditioner entries per DOF
  __shared__ T s_r[kPasses][kVecNodes * D];
  __shared__ int s_skip;
  __shared__ double s_ab[2];
  asm volatile("" ::"s"(n_rows), "s"(tiles), "s"(st));  // the arguments in one round of loads
  const int tid_ = tile_id();
  const Tile tl = tile_of<D>(tiles, tid_, n_rows);
  const int64_t kend = (int64_t)tl.row1 * D;
  const int d = threadIdx.x % D;
  // the tile's passes unrolled: every pass's loads are issued before the one barrier that publishes
  // the new residuals for the block-Jacobi product
  T uk[kPasses], pk[kPasses], wk[kPasses], sk[kPasses], rk[kPasses], mk[kPasses][NM];
  double xk[kPasses];
  auto load = [&]() {
#pragma unroll
    for (int ps = 0; ps < kPasses; ++ps) {
      const int64_t k = ((int64_t)tl.row0 + ps * kVecNodes) * D + threadIdx.x;
      uk[ps] = pk[ps] = wk[ps] = sk[ps] = rk[ps] = T(0);
      xk[ps] = 0.0;
#pragma unroll
      for (int j = 0; j < NM; ++j) mk[ps][j] = T(0);
      if (k < kend) {
        uk[ps] = u[k];
        pk[ps] = p[k];
        wk[ps] = w[k];
        sk[ps] = s[k];
        rk[ps] = r[k];
        xk[ps] = x[k];
        const T* m = PC == FE_PC_JACOBI ? M + k : M + (k / D) * D * D + d * D;
#pragma unroll
        for (int j = 0; j < NM; ++j) mk[ps][j] = m[j];
      }
    }
  };
  if constexpr (SPEC) load();
  if (threadIdx.x == 0) {
    const PcgState* ss = st + tl.seg;
    const double al = ss->alpha, be = ss->beta;
    s_skip = !seg_running_now(ss);
    s_ab[0] = al;
    s_ab[1] = be;
  }
  __syncthreads();
  if (s_skip) return;
  if constexpr (!SPEC) load();
  const double alpha = s_ab[0];
  const T alpha_t = (T)alpha, beta = (T)s_ab[1];
#pragma unroll
  for (int ps = 0; ps < kPasses; ++ps) {
    const int64_t k = ((int64_t)tl.row0 + ps * kVecNodes) * D + threadIdx.x;
    T rn = T(0);
    if (k < kend) {
      const T pn = uk[ps] + beta * pk[ps];
      const T sn = wk[ps] + beta * sk[ps];
      p[k] = pn;
      s[k] = sn;
      x[k] = xk[ps] + alpha * (double)pn;
      rn = rk[ps] - alpha_t * sn;
      r[k] = rn;
    }
    rk[ps] = rn;
    s_r[ps][threadIdx.x] = rn;
  }
  __syncthreads();
#pragma unroll
  for (int ps = 0; ps < kPasses; ++ps) {
    const int64_t k = ((int64_t)tl.row0 + ps * kVecNodes) * D + threadIdx.x;
    if (k >= kend) continue;
    T un;
    if (PC == FE_PC_JACOBI) {
      un = mk[ps][0] * rk[ps];
    } else {
      const T* rn = &s_r[ps][threadIdx.x - d];
      un = T(0);
#pragma unroll
      for (int j = 0; j < NM; ++j) un += mk[ps][j] * rn[j];
    }
    u[k] = un;
  }
}

// VEC: the matrix is on a padded pattern (row_align 4) and G lanes walk a row with 16-byte loads.
template <typename T, int D, int G, bool VEC>
__global__ __launch_bounds__(kBlock) void k_cg_spmv(const Tile* __restrict__ tiles, int ntiles,
                                                    int64_t nnzb, const int32_t* __restrict__ bptr,
                                                    const int32_t* __restrict__ bcol,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ u, const T* __restrict__ r,
                                                    T* __restrict__ w, const PcgState* st,
                                                    double* __restrict__ partials, int force,
                                                    int32_t n_rows) {
  __shared__ int s_skip;
  asm volatile("" ::"s"(n_rows), "s"(tiles), "s"(st), "s"(force));  // the arguments in one round
  const int tid_ = tile_id();
  const Tile tl = tile_of<D>(tiles, tid_, n_rows);
  constexpr int kRowsPerPass = kBlock / G;
  constexpr int kPasses = tile_rows<lanes_for<D>()>() / kRowsPerPass;
  // VEC: every pass's row bounds are loaded before the state test (their latency overlaps it)
  int32_t rs[VEC ? kPasses : 1], re[VEC ? kPasses : 1];
  if constexpr (VEC) {
#pragma unroll
    for (int pass = 0; pass < kPasses; ++pass) {
      const int64_t i = tl.row0 + pass * kRowsPerPass + threadIdx.x / G;
      const bool ok = i < tl.row1;
      rs[pass] = ok ? bptr[i] : 0;
      re[pass] = ok ? bptr[i + 1] : 0;
    }
  }
  if (threadIdx.x == 0) s_skip = !seg_running_now(st + tl.seg) && !force;
  __syncthreads();
  if (s_skip) return;
  const int lane = threadIdx.x % G;
  double v[3] = {0.0, 0.0, 0.0};
  auto row_out = [&](int64_t i, const T (&acc)[D]) {
    const int64_t k = i * D + lane;
    const T wk = pick<T, D>(acc, lane);
    w[k] = wk;
    const double uk = u[k], rk = r[k];
    v[0] += rk * uk;
    v[1] += (double)wk * uk;
    v[2] += rk * rk;
  };
  if constexpr (VEC) {
    // a tile is a fixed number of passes: unrolled, so that the loads of every pass are in flight
    // together
    static_assert(kPasses * kRowsPerPass == tile_rows<lanes_for<D>()>(), "tile = whole passes");
#pragma unroll
    for (int pass = 0; pass < kPasses; ++pass) {
      const int64_t i = tl.row0 + pass * kRowsPerPass + threadIdx.x / G;
      const bool ok = i < tl.row1;
      T acc[D];
      bsr_row_vec_range<T, D, G>(bcol, vals, nnzb, u, rs[pass], re[pass], lane, acc);
      if (ok && lane < D) row_out(i, acc);
    }
  } else {
    for (int64_t i = tl.row0 + threadIdx.x / G; i < tl.row1; i += kBlock / G) {
      T acc[D];
      bsr_row<T, D, G, FE_LAYOUT_PLANES>(bptr, bcol, vals, nnzb, u, i, lane, acc);
      if (lane < D) row_out(i, acc);
    }
  }
  block_sum<3>(v);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 3; ++k) partials[(int64_t)k * ntiles + tid_] = v[k];
}

// One workgroup per segment: the Chronopoulos-Gear scalars for the next iteration.
// last: this is the chunk's last iteration (a pending reliable update is requested now).
__global__ __launch_bounds__(kRed) void k_cg_reduce(const int32_t* __restrict__ seg_tiles, int ntiles,
                                                    const double* __restrict__ partials,
                                                    PcgState* st, int last) {
  PcgState* s = st + blockIdx.x;
  // the tile range, the running test and the fields the update reads are loaded together (only
  // thread 0 writes the state, after the sums); the test is uniform
  const int t0 = seg_tiles[blockIdx.x], t1 = seg_tiles[blockIdx.x + 1];
  const int32_t done = s->done, need_ru = s->need_ru;
  const double gamma0 = s->gamma, alpha0 = s->alpha, tol2bb = s->tol2bb;
  const double ru_delta2 = s->ru_delta2, rr_ref = s->rr_ref;
  const int32_t iter = s->iter + 1, max_iter = s->max_iter, mixed = s->mixed;
  int32_t ru_pending = s->ru_pending;
  // an empty asm that takes the values keeps their loads ahead of the exit (the compiler would
  // otherwise sink each one to its first use, a memory round trip apiece)
  asm volatile("" ::"s"(t0), "s"(t1), "s"(gamma0), "s"(alpha0), "s"(tol2bb), "s"(ru_delta2), "s"(rr_ref),
               "s"(iter), "s"(max_iter), "s"(mixed), "s"(ru_pending), "s"(ntiles), "s"(partials));
  if (done | need_ru) return;
  double v[3];
  seg_sum<3>(partials, ntiles, t0, t1, v);
  if (threadIdx.x != 0) return;
  const double g = v[0], dl = v[1];
  s->iter = iter;
  s->rr = v[2];
  if (!isfinite(g) || !isfinite(dl) || !isfinite(v[2]) || g < 0.0) {
    s->done = 3;
    return;
  }
  s->gamma_prev = gamma0;
  s->gamma = g;
  // convergence claims and the iteration limit hand over to the true-residual check at the end of
  // the chunk; (mixed precision) the reliable-update threshold marks the update pending
  if (v[2] <= tol2bb || iter >= max_iter) {
    s->need_ru = 1;
    return;
  }
  if (mixed && v[2] < ru_delta2 * rr_ref) s->ru_pending = ru_pending = 1;
  const double beta = g / gamma0;
  const double den = dl - beta * g / alpha0;
  if (!(den > 0.0)) {
    // the recurrences lost positivity (rounding, most often in FP32): restart from the true
    // residual; a genuine breakdown shows up there as p.Ap <= 0
    s->need_ru = 1;
    s->first = 1;
    s->restarts += 1;
    return;
  }
  s->beta = beta;
  s->alpha = g / den;
  // the update replaces r, u, w, s and recomputes beta and alpha from gamma_prev, as at a hand-over
  if (last && ru_pending) s->need_ru = 1;
}

// --------------------------------------------------------------------------------------------------
// reliable update
// --------------------------------------------------------------------------------------------------
// Part 1 (tiles): r64 = b - A64 x, partial ||r64||^2.
template <int D, int G>
__global__ __launch_bounds__(kBlock) void k_ru_residual(const Tile* __restrict__ tiles, int ntiles,
                                                        int64_t nnzb, const int32_t* __restrict__ bptr,
                                                        const int32_t* __restrict__ bcol,
                                                        const double* __restrict__ vals,
                                                        const double* __restrict__ b,
                                                        const double* __restrict__ x,
                                                        double* __restrict__ r64, const PcgState* st,
                                                        double* __restrict__ partials) {
  __shared__ int s_skip;
  const int tid_ = tile_id();
  const Tile tl = tiles[tid_];
  if (threadIdx.x == 0) s_skip = !seg_updating(st + tl.seg);
  __syncthreads();
  if (s_skip) return;
  const int lane = threadIdx.x % G;
  double v[1] = {0.0};
  for (int64_t i = tl.row0 + threadIdx.x / G; i < tl.row1; i += kBlock / G) {
    double acc[D];
    bsr_row<double, D, G, FE_LAYOUT_PLANES>(bptr, bcol, vals, nnzb, x, i, lane, acc);
    if (lane < D) {
      const int64_t k = i * D + lane;
      const double ri = b[k] - pick<double, D>(acc, lane);
      r64[k] = ri;
      v[0] += ri * ri;
    }
  }
  block_sum<1>(v);
  if (threadIdx.x == 0) partials[tid_] = v[0];
}

// Part 2 (per segment): judge the true residual -- converged, out of iterations, stagnating, or
// continue with the replaced residual (restarting the direction when it drifted).
__global__ __launch_bounds__(kRed) void k_ru_decide(const int32_t* __restrict__ seg_tiles, int ntiles,
                                                    const double* __restrict__ partials,
                                                    PcgState* st) {
  PcgState* s = st + blockIdx.x;
  const int t0 = seg_tiles[blockIdx.x], t1 = seg_tiles[blockIdx.x + 1];
  if (!seg_updating(s)) return;
  double v[1];
  seg_sum<1>(partials, ntiles, t0, t1, v);
  if (threadIdx.x != 0) return;
  s->rr_true = v[0];
  if (v[0] <= s->tol2bb) {
    s->done = 1;
  } else if (s->iter >= s->max_iter) {
    s->done = 2;
  } else {
    if (v[0] < 0.81 * s->rr_best) {
      s->rr_best = v[0];
      s->stall = 0;
    } else if (++s->stall >= kStallLimit) {
      s->done = 2;
    }
    // the recursive residual drifted far below the true one: the direction is stale, restart
    if (!s->first && v[0] > 4.0 * s->rr) {
      s->first = 1;
      s->restarts += 1;
    }
    if (s->restarts > (s->mixed ? kMaxRestarts32 : kMaxRestarts64)) s->done = 2;
  }
  if (s->done) {
    s->rr = v[0];
    s->n_ru += 1;
    s->need_ru = 0;
  }
}

// Part 3 (tiles): r = r64, u = M r (and p, s cleared on a (re)start).
template <typename T, int D, int PC>
__global__ __launch_bounds__(kVecNodes * D) void k_ru_prep(const Tile* __restrict__ tiles,
                                                           const double* __restrict__ r64,
                                                           T* __restrict__ r, T* __restrict__ u,
                                                           T* __restrict__ p, T* __restrict__ s,
                                                           const T* __restrict__ M,
                                                           const PcgState* st) {
  __shared__ T s_r[kVecNodes * D];
  __shared__ int s_skip, s_first;
  const int tid_ = tile_id();
  const Tile tl = tiles[tid_];
  if (threadIdx.x == 0) {
    s_skip = !seg_updating(st + tl.seg);
    s_first = st[tl.seg].first;
  }
  __syncthreads();
  if (s_skip) return;
  const int64_t kend = (int64_t)tl.row1 * D;
  for (int64_t base = (int64_t)tl.row0 * D; base < kend; base += kVecNodes * D) {
    const int64_t k = base + threadIdx.x;
    T rk = T(0);
    if (k < kend) {
      rk = (T)r64[k];
      r[k] = rk;
      if (s_first) {
        p[k] = T(0);
        s[k] = T(0);
      }
    }
    s_r[threadIdx.x] = rk;
    __syncthreads();
    if (k < kend) {
      const int d = threadIdx.x % D;
      u[k] = precond_dof<T, D, PC>(M, k / D, d, k, &s_r[threadIdx.x - d], rk);
    }
    __syncthreads();
  }
}

// Part 4 (tiles): one matrix pass for w = A u and s = A p; partial (r,u), (u,w), (u,s), (p,s), (r,r).
template <typename T, int D, int G>
__global__ __launch_bounds__(kBlock) void k_ru_spmv(const Tile* __restrict__ tiles, int ntiles,
                                                    int64_t nnzb, const int32_t* __restrict__ bptr,
                                                    const int32_t* __restrict__ bcol,
                                                    const T* __restrict__ vals,
                                                    const T* __restrict__ u, const T* __restrict__ p,
                                                    const T* __restrict__ r, T* __restrict__ w,
                                                    T* __restrict__ s, const PcgState* st,
                                                    double* __restrict__ partials) {
  __shared__ int s_skip;
  const int tid_ = tile_id();
  const Tile tl = tiles[tid_];
  if (threadIdx.x == 0) s_skip = !seg_updating(st + tl.seg);
  __syncthreads();
  if (s_skip) return;
  const int lane = threadIdx.x % G;
  double acc5[5] = {0, 0, 0, 0, 0};
  for (int64_t i = tl.row0 + threadIdx.x / G; i < tl.row1; i += kBlock / G) {
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
  block_sum<5>(acc5);
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < 5; ++k) partials[(int64_t)k * ntiles + tid_] = acc5[k];
}

// Part 5 (per segment): beta = gamma/gamma_prev (0 on a (re)start) and the exact alpha of the
// direction p' = u + beta p: p'.Ap' = (u,w) + 2 beta (u,s) + beta^2 (p,s).
__global__ __launch_bounds__(kRed) void k_ru_finish(const int32_t* __restrict__ seg_tiles, int ntiles,
                                                    const double* __restrict__ partials,
                                                    PcgState* st) {
  PcgState* s = st + blockIdx.x;
  const int t0 = seg_tiles[blockIdx.x], t1 = seg_tiles[blockIdx.x + 1];
  if (!seg_updating(s)) return;
  double v[5];
  seg_sum<5>(partials, ntiles, t0, t1, v);
  if (threadIdx.x != 0) return;
  const double g = v[0];
  if (!isfinite(g) || g < 0.0 || !isfinite(v[1])) {
    s->done = 3;
    return;
  }
  // gamma_prev holds r.u of the iteration before the one whose residual was replaced
  const double beta = s->first ? 0.0 : g / s->gamma_prev;
  const double pap = v[1] + 2.0 * beta * v[2] + beta * beta * v[3];
  if (!(pap > 0.0)) {
    s->done = 3;
    return;
  }
  s->beta = beta;
  s->alpha = g / pap;
  s->gamma = g;
  s->first = 0;
  s->rr = v[4];
  s->rr_ref = s->rr_true;
  s->n_ru += 1;
  s->need_ru = 0;
  s->ru_pending = 0;
}

// --------------------------------------------------------------------------------------------------
// set-up, status, export
// --------------------------------------------------------------------------------------------------
template <int D>
__global__ __launch_bounds__(kBlock) void k_tile_norm(const Tile* __restrict__ tiles,
                                                      const double* __restrict__ b,
                                                      double* __restrict__ partials) {
  const int tid_ = tile_id();
  const Tile tl = tiles[tid_];
  double v[1] = {0.0};
  for (int64_t k = (int64_t)tl.row0 * D + threadIdx.x; k < (int64_t)tl.row1 * D; k += kBlock)
    v[0] += b[k] * b[k];
  block_sum<1>(v);
  if (threadIdx.x == 0) partials[tid_] = v[0];
}

__global__ __launch_bounds__(kRed) void k_pcg_init(const int32_t* __restrict__ seg_tiles, int ntiles,
                                                   const double* __restrict__ partials, PcgState* st,
                                                   double tol, int max_iter, int mixed,
                                                   double ru_delta) {
  double v[1];
  seg_sum<1>(partials, ntiles, seg_tiles[blockIdx.x], seg_tiles[blockIdx.x + 1], v);
  if (threadIdx.x != 0) return;
  PcgState* s = st + blockIdx.x;
  *s = PcgState{};
  s->bb = v[0];
  s->tol2bb = tol * tol * v[0];
  s->ru_delta2 = ru_delta * ru_delta;
  s->max_iter = max_iter;
  s->mixed = mixed;
  s->need_ru = 1;
  s->first = 1;
  s->alpha = 1.0;
  s->gamma = 1.0;
  s->gamma_prev = 1.0;
  s->rr_best = 1.0e308;
}

__global__ __launch_bounds__(1024) void k_pcg_status(int nseg, const PcgState* st, PcgStatus* out) {
  __shared__ int s_act[1024 / 64], s_max[1024 / 64], s_ru[1024 / 64];
  __shared__ double s_ratio[1024 / 64];
  int act = 0, mx = 0, ru = 0;
  double ratio = 0.0;
  for (int g = threadIdx.x; g < nseg; g += blockDim.x) {
    act += st[g].done == 0;
    mx = max(mx, st[g].iter);
    ru += st[g].n_ru;
    if (st[g].done == 0 && st[g].tol2bb > 0.0) ratio = fmax(ratio, st[g].rr / st[g].tol2bb);
  }
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) {
    act += __shfl_xor(act, off, 64);
    mx = max(mx, __shfl_xor(mx, off, 64));
    ru += __shfl_xor(ru, off, 64);
    ratio = fmax(ratio, __shfl_xor(ratio, off, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    s_act[threadIdx.x >> 6] = act;
    s_max[threadIdx.x >> 6] = mx;
    s_ru[threadIdx.x >> 6] = ru;
    s_ratio[threadIdx.x >> 6] = ratio;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int w = 1; w < (int)blockDim.x / 64; ++w) {
      act += s_act[w];
      mx = max(mx, s_max[w]);
      ru += s_ru[w];
      ratio = fmax(ratio, s_ratio[w]);
    }
    out->active = act;
    out->max_iter = mx;
    out->n_ru = ru;
    out->rr_ratio = ratio;
  }
}

__global__ void k_pcg_export(int nseg, const PcgState* st, fe_pcg_seg_stats* out) {
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= nseg) return;
  const PcgState& s = st[g];
  fe_pcg_seg_stats o{};
  o.iterations = s.iter;
  o.reliable_updates = s.n_ru;
  o.restarts = s.restarts;
  o.fallback = 0;
  o.status = s.done == 1 ? FE_OK : s.done == 3 ? FE_ERR_BREAKDOWN : FE_ERR_NOT_CONVERGED;
  o.bnorm = sqrt(s.bb);
  o.relres = s.bb > 0 ? sqrt(s.rr_true / s.bb) : 0.0;
  o.relres_rec = s.bb > 0 ? sqrt(s.rr / s.bb) : 0.0;
  out[g] = o;
}

// --------------------------------------------------------------------------------------------------
// host side
// --------------------------------------------------------------------------------------------------
constexpr int64_t kAlign = 256;
inline int64_t align_up(int64_t v) { return (v + kAlign - 1) / kAlign * kAlign; }

struct Work {
  PcgState* st;
  PcgStatus* status;
  int32_t* seg_tiles;
  Tile* tiles;
  double* partials;
  fe_pcg_seg_stats* seg_stats;
  void* r;
  void* u;
  void* w;
  void* p;
  void* s;
  double* r64;
  void* rc;  // coarse restriction / solution (two-level only)
  void* zc;
  int64_t bytes;
};

// Tiles of every segment (host): segment g owns node rows [seg_nodes[g], seg_nodes[g+1]).
int64_t count_tiles(int nseg, const int32_t* seg_nodes, int rows) {
  int64_t t = 0;
  for (int g = 0; g < nseg; ++g) t += ceil_div((int64_t)seg_nodes[g + 1] - seg_nodes[g], rows);
  return t == 0 ? 0 : (t + 7) / 8 * 8;  // padded with empty tiles to a multiple of the 8 XCDs
}

Work layout_work(char* base, int64_t nb, int dof, int elem, int nseg, int64_t ntiles,
                 int64_t n_coarse = 0) {
  const int64_t n = nb * dof;
  Work w{};
  int64_t off = 0;
  auto take = [&](int64_t bytes) {
    char* p = base ? base + off : nullptr;
    off += align_up(bytes);
    return p;
  };
  w.st = (PcgState*)take(nseg * (int64_t)sizeof(PcgState));
  w.status = (PcgStatus*)take(sizeof(PcgStatus));
  w.seg_tiles = (int32_t*)take((nseg + 1) * (int64_t)sizeof(int32_t));
  w.tiles = (Tile*)take(std::max<int64_t>(1, ntiles) * (int64_t)sizeof(Tile));
  w.partials = (double*)take(5 * std::max<int64_t>(1, ntiles) * (int64_t)sizeof(double));
  w.seg_stats = (fe_pcg_seg_stats*)take(nseg * (int64_t)sizeof(fe_pcg_seg_stats));
  w.r = take(n * elem);
  w.u = take(n * elem);
  w.w = take(n * elem);
  w.p = take(n * elem);
  w.s = take(n * elem);
  w.r64 = (double*)take(n * (int64_t)sizeof(double));
  w.rc = n_coarse ? take(n_coarse * (int64_t)sizeof(double)) : nullptr;
  w.zc = n_coarse ? take(n_coarse * (int64_t)sizeof(double)) : nullptr;
  w.bytes = off;
  return w;
}

struct Problem {
  const fe_bsr* A64;
  const fe_bsr* Ait;
  const void* M;
  const double* b;
  double* x;
  int nseg;
  int64_t ntiles;
  const fe_coarse* C;  // two-level: coarse space and the inverse coarse operator (iteration precision)
  const void* Ainv;
};

template <typename T, int D, int PC>
int run_pcg(fe_handle* h, const Problem& pb, const Work& w, const fe_pcg_opts* o, fe_pcg_stats* s) {
  constexpr int G = lanes_for<D>();
  const fe_bsr* A64 = pb.A64;
  const int64_t nnzb = A64->nnzb, n = A64->n_brows * D;
  // the iteration matrix may be on a padded pattern of its own
  const fe_bsr* Ait = pb.Ait;
  const int64_t nnzb_it = Ait->nnzb;
  const int32_t* bptr_it = Ait->bptr;
  const int32_t* bcol_it = Ait->bcol;
  const bool vec = row_aligned(Ait);
  constexpr int GV = G / Vec16<T>::W;
  const int nt = (int)pb.ntiles, nseg = pb.nseg;
  hipStream_t st = h->stream;
  const T* M = (const T*)pb.M;
  FE_CHECK_HIP(hipMemsetAsync(w.p, 0, n * sizeof(T), st));
  FE_CHECK_HIP(hipMemsetAsync(w.s, 0, n * sizeof(T), st));
  const dim3 gt((unsigned)nt), gseg((unsigned)nseg);
  hipLaunchKernelGGL((k_tile_norm<D>), gt, dim3(kBlock), 0, st, w.tiles, pb.b, w.partials);
  hipLaunchKernelGGL(k_pcg_init, gseg, dim3(kRed), 0, st, w.seg_tiles, nt, w.partials, w.st, o->tol,
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
  // one segment: the kernels form their regular tiles from the row count instead of loading them
  const int32_t one_rows = nseg == 1 ? (int32_t)A64->n_brows : -1;
  auto launch_spmv = [&](hipStream_t q, int force) {
    if (vec)
      hipLaunchKernelGGL((k_cg_spmv<T, D, GV, true>), gt, dim3(kBlock), 0, q, w.tiles, nt, nnzb_it,
                         bptr_it, bcol_it, vit, u, r, wv, w.st, w.partials, force, one_rows);
    else
      hipLaunchKernelGGL((k_cg_spmv<T, D, G, false>), gt, dim3(kBlock), 0, q, w.tiles, nt, nnzb_it,
                         bptr_it, bcol_it, vit, u, r, wv, w.st, w.partials, force, one_rows);
  };
  const dim3 bvec(kVecNodes * D);
  auto launch_ru = [&](hipStream_t q) -> int {
    hipLaunchKernelGGL((k_ru_residual<D, G>), gt, dim3(kBlock), 0, q, w.tiles, nt, nnzb, bptr, bcol,
                       v64, pb.b, pb.x, w.r64, w.st, w.partials);
    hipLaunchKernelGGL(k_ru_decide, gseg, dim3(kRed), 0, q, w.seg_tiles, nt, w.partials, w.st);
    hipLaunchKernelGGL((k_ru_prep<T, D, PC>), gt, bvec, 0, q, w.tiles, w.r64, r, u, p, sv, M, w.st);
    if (pb.C) {
      int rc3 = fe_coarse_correction(q, pb.C, (int)(sizeof(T) == 4), pb.Ainv, r, u, w.rc, w.zc,
                                     &w.st[0].done, kStateStride, 1);
      if (rc3) return rc3;
    }
    hipLaunchKernelGGL((k_ru_spmv<T, D, G>), gt, dim3(kBlock), 0, q, w.tiles, nt, nnzb_it, bptr_it,
                       bcol_it, vit, u, p, r, wv, sv, w.st, w.partials);
    hipLaunchKernelGGL(k_ru_finish, gseg, dim3(kRed), 0, q, w.seg_tiles, nt, w.partials, w.st);
    return FE_OK;
  };
  {
    int rc0 = launch_ru(st);
    if (rc0) return rc0;
  }
  FE_CHECK_LAUNCH();
  const int chunk = o->check_every > 0 ? o->check_every : 16;
  PcgStatus* pinned = (PcgStatus*)h->pinned;
  hipEvent_t ev[2];
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[0], hipEventDisableTiming));
  FE_CHECK_HIP(hipEventCreateWithFlags(&ev[1], hipEventDisableTiming));
  hipGraphExec_t gexec[2] = {nullptr, nullptr};
  int rc = FE_OK;
  // graph mode: chunks run on a stream of their own (capture needs a non-default stream), ordered
  // after everything queued so far on the handle's stream
  hipStream_t cs = st;
  hipEvent_t e_start = nullptr;
  if (o->graph) {
    FE_CHECK_HIP(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking));
    FE_CHECK_HIP(hipEventCreateWithFlags(&e_start, hipEventDisableTiming));
    FE_CHECK_HIP(hipEventRecord(e_start, st));
    FE_CHECK_HIP(hipStreamWaitEvent(cs, e_start, 0));
  }
  // one chunk of iterations + the reliable update + the status copy into pinned slot `slot`
  auto chunk_body = [&](int slot) -> int {
    for (int it = 0; it < chunk; ++it) {
      if (nseg == 1)
        hipLaunchKernelGGL((k_cg_vec<T, D, PC, true>), gt, bvec, 0, cs, w.tiles, pb.x, p, sv, r, u, wv, M,
                           w.st, one_rows);
      else
        hipLaunchKernelGGL((k_cg_vec<T, D, PC, false>), gt, bvec, 0, cs, w.tiles, pb.x, p, sv, r, u, wv,
                           M, w.st, one_rows);
      if (pb.C) {
        int rc3 = fe_coarse_correction(cs, pb.C, (int)(sizeof(T) == 4), pb.Ainv, r, u, w.rc, w.zc,
                                       &w.st[0].done, kStateStride, 0);
        if (rc3) return rc3;
      }
      launch_spmv(cs, 0);
      hipLaunchKernelGGL(k_cg_reduce, gseg, dim3(kRed), 0, cs, w.seg_tiles, nt, w.partials, w.st,
                         (int)(it == chunk - 1));
    }
    {
      int rc4 = launch_ru(cs);
      if (rc4) return rc4;
    }
    hipLaunchKernelGGL(k_pcg_status, dim3(1), dim3(1024), 0, cs, nseg, w.st, w.status);
    FE_CHECK_HIP(hipMemcpyAsync(&pinned[slot], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, cs));
    return FE_OK;
  };
  if (o->graph) {
    for (int slot = 0; slot < 2 && rc == FE_OK; ++slot) {
      hipGraph_t graph = nullptr;
      FE_CHECK_HIP(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal));
      rc = chunk_body(slot);
      hipError_t e = hipStreamEndCapture(cs, &graph);
      if (rc == FE_OK && e != hipSuccess)
        rc = fe_set_error((int)e, "hipStreamEndCapture: %s", hipGetErrorString(e));
      if (rc == FE_OK) {
        e = hipGraphInstantiate(&gexec[slot], graph, nullptr, nullptr, 0);
        if (e != hipSuccess) rc = fe_set_error((int)e, "hipGraphInstantiate: %s", hipGetErrorString(e));
      }
      if (graph) (void)hipGraphDestroy(graph);
    }
  }
  auto enqueue_chunk = [&](int slot) -> int {
    if (o->graph) {
      FE_CHECK_HIP(hipGraphLaunch(gexec[slot], cs));
    } else {
      int rc2 = chunk_body(slot);
      if (rc2) return rc2;
      FE_CHECK_LAUNCH();
    }
    FE_CHECK_HIP(hipEventRecord(ev[slot], cs));
    return FE_OK;
  };
  // Two chunks are kept in flight, so the device never waits for the host, except when the residual
  // trend of the last two polls predicts that the chunk already queued ends the solve: then the host
  // waits for it before queueing another (a chunk queued after convergence finds every segment done
  // and returns at once, but its launches still cost device time).
  const int64_t max_chunks = 4 * ((int64_t)o->max_iter / chunk + 2) + 8;
  PcgStatus last{};
  int finished = 0;
  int64_t queued = 0, polled = 0;
  double prev_ratio = -1.0;
  if (rc == FE_OK) rc = enqueue_chunk((int)(queued++ & 1));
  if (rc == FE_OK) rc = enqueue_chunk((int)(queued++ & 1));
  while (rc == FE_OK && polled < queued && polled < max_chunks) {
    const int slot = (int)(polled & 1);
    hipError_t e = hipEventSynchronize(ev[slot]);
    if (e != hipSuccess) {
      rc = fe_set_error((int)e, "hipEventSynchronize: %s", hipGetErrorString(e));
      break;
    }
    last = pinned[slot];
    ++polled;
    if (last.active == 0) {
      finished = 1;
      break;
    }
    const double ratio = last.rr_ratio;
    const bool ends_next = prev_ratio > 0.0 && ratio > 0.0 && ratio < prev_ratio &&
                           ratio * (ratio / prev_ratio) <= 1.0;
    prev_ratio = ratio;
    const int64_t ahead = ends_next ? 1 : 2;
    while (rc == FE_OK && queued - polled < ahead && queued < max_chunks)
      rc = enqueue_chunk((int)(queued++ & 1));
  }
  (void)hipEventDestroy(ev[0]);
  (void)hipEventDestroy(ev[1]);
  // drain the queued chunk before the graphs go away
  hipError_t se = hipStreamSynchronize(cs);
  for (int a = 0; a < 2; ++a)
    if (gexec[a]) (void)hipGraphExecDestroy(gexec[a]);
  if (o->graph) {
    (void)hipEventDestroy(e_start);
    (void)hipStreamDestroy(cs);
  }
  if (rc) return rc;
  if (se != hipSuccess) return fe_set_error((int)se, "hipStreamSynchronize: %s", hipGetErrorString(se));
  (void)finished;
  if (o->profile) {
    // the iteration's SpMV kernel timed back to back on the final state (same matrix, vectors and
    // grid; the forced launches only rewrite w and the partials after the solve has ended)
    constexpr int kProfLaunches = 20;
    hipEvent_t p0, p1;
    FE_CHECK_HIP(hipEventCreate(&p0));
    FE_CHECK_HIP(hipEventCreate(&p1));
    FE_CHECK_HIP(hipEventRecord(p0, st));
    for (int k = 0; k < kProfLaunches; ++k)
      launch_spmv(st, 1);
    FE_CHECK_HIP(hipEventRecord(p1, st));
    FE_CHECK_HIP(hipEventSynchronize(p1));
    float ms = 0.f;
    FE_CHECK_HIP(hipEventElapsedTime(&ms, p0, p1));
    s->spmv_ms = ms / kProfLaunches;
    s->spmv_samples = kProfLaunches;
    (void)hipEventDestroy(p0);
    (void)hipEventDestroy(p1);
  }
  hipLaunchKernelGGL(k_pcg_export, dim3((unsigned)ceil_div(nseg, 256)), dim3(256), 0, st, nseg, w.st,
                     w.seg_stats);
  hipLaunchKernelGGL(k_pcg_status, dim3(1), dim3(1024), 0, st, nseg, w.st, w.status);
  FE_CHECK_LAUNCH();
  FE_CHECK_HIP(hipMemcpyAsync(&pinned[0], w.status, sizeof(PcgStatus), hipMemcpyDeviceToHost, st));
  FE_CHECK_HIP(hipStreamSynchronize(st));
  last = pinned[0];
  s->iterations = last.max_iter;
  s->reliable_updates = last.n_ru;
  return FE_OK;
}

template <typename T, int D>
int dispatch_pc(fe_handle* h, const Problem& pb, const Work& w, const fe_pcg_opts* o, fe_pcg_stats* s) {
  if (o->precond == FE_PC_JACOBI) return run_pcg<T, D, FE_PC_JACOBI>(h, pb, w, o, s);
  return run_pcg<T, D, FE_PC_BLOCK_JACOBI>(h, pb, w, o, s);
}

template <typename T>
int dispatch_dof(fe_handle* h, const Problem& pb, const Work& w, const fe_pcg_opts* o, fe_pcg_stats* s) {
  switch (pb.A64->dof) {
    case 1: return dispatch_pc<T, 1>(h, pb, w, o, s);
    case 2: return dispatch_pc<T, 2>(h, pb, w, o, s);
    default: return dispatch_pc<T, 3>(h, pb, w, o, s);
  }
}

int check_pcg_args(const fe_bsr* A64, const fe_bsr* A32, const void* M64, const void* M32,
                   const fe_pcg_opts* o, const char* who) {
  int rc = check_bsr(A64, who);
  if (rc) return rc;
  FE_REQUIRE(A64->precision == 0 && A64->layout == FE_LAYOUT_PLANES,
             "%s: A64 must be FP64 in FE_LAYOUT_PLANES", who);
  FE_REQUIRE(o->precond == FE_PC_JACOBI || o->precond == FE_PC_BLOCK_JACOBI, "%s: bad precond", who);
  FE_REQUIRE(o->tol > 0 && o->max_iter > 0, "%s: tol and max_iter must be positive", who);
  if (o->mixed) {
    rc = check_bsr(A32, who);
    if (rc) return rc;
    FE_REQUIRE(M32, "%s: mixed precision needs the FP32 preconditioner", who);
    FE_REQUIRE(A32->precision == 1 && A32->layout == FE_LAYOUT_PLANES && A32->dof == A64->dof &&
                   A32->n_brows == A64->n_brows && (A32->nnzb == A64->nnzb || A32->row_align == 4),
               "%s: A32 must be the FP32 FE_LAYOUT_PLANES copy of A64", who);
  } else {
    FE_REQUIRE(M64, "%s: missing preconditioner", who);
    if (A32) {
      rc = check_bsr(A32, who);
      if (rc) return rc;
      FE_REQUIRE(A32->precision == 0 && A32->layout == FE_LAYOUT_PLANES && A32->dof == A64->dof &&
                     A32->n_brows == A64->n_brows && (A32->nnzb == A64->nnzb || A32->row_align == 4),
                 "%s: without mixed precision A32 must be NULL or an FP64 FE_LAYOUT_PLANES copy of A64",
                 who);
    }
  }
  return FE_OK;
}

int check_segments(const fe_bsr* A, int nseg, const int32_t* seg_nodes, const char* who) {
  FE_REQUIRE(nseg >= 1 && seg_nodes, "%s: need at least one segment", who);
  FE_REQUIRE(seg_nodes[0] == 0 && seg_nodes[nseg] == A->n_brows,
             "%s: segments must cover node rows 0..n_brows", who);
  for (int g = 0; g < nseg; ++g)
    FE_REQUIRE(seg_nodes[g + 1] >= seg_nodes[g], "%s: segment offsets must not decrease", who);
  return FE_OK;
}

int pcg_common(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64, const void* M32,
               int nseg, const int32_t* seg_nodes, const double* b, double* x, void* work,
               int64_t work_bytes, const fe_pcg_opts* o, fe_pcg_stats* s, fe_pcg_seg_stats* seg_out,
               const char* who, const fe_coarse* C = nullptr, const void* Ainv64 = nullptr,
               const void* Ainv32 = nullptr) {
  const int elem = (int)sizeof(double);  // FP64-sized vectors: the FP64 fallback reuses them
  const int rows = tile_rows_dof(A64->dof);
  const int64_t ntiles = count_tiles(nseg, seg_nodes, rows);
  FE_REQUIRE(ntiles < (int64_t)1 << 31, "%s: too many tiles", who);
  Work w = layout_work((char*)work, A64->n_brows, A64->dof, elem, nseg, ntiles,
                       C ? C->n_agg * C->nc : 0);
  FE_REQUIRE(work_bytes >= w.bytes, "%s: workspace too small (%lld < %lld)", who,
             (long long)work_bytes, (long long)w.bytes);
  // tile list and per-segment tile offsets
  std::vector<Tile> tiles((size_t)std::max<int64_t>(1, ntiles));
  std::vector<int32_t> seg_tiles((size_t)nseg + 1);
  int64_t t = 0;
  for (int g = 0; g < nseg; ++g) {
    seg_tiles[g] = (int32_t)t;
    for (int64_t r0 = seg_nodes[g]; r0 < seg_nodes[g + 1]; r0 += rows)
      tiles[t++] = Tile{g, (int32_t)r0, (int32_t)std::min<int64_t>(r0 + rows, seg_nodes[g + 1]), 0};
  }
  // padding tiles: empty row ranges of the last segment (they contribute zero partials)
  for (; t < ntiles; ++t) tiles[t] = Tile{nseg - 1, seg_nodes[nseg], seg_nodes[nseg], 0};
  seg_tiles[nseg] = (int32_t)t;
  FE_CHECK_HIP(hipMemcpyAsync(w.tiles, tiles.data(), tiles.size() * sizeof(Tile),
                              hipMemcpyHostToDevice, h->stream));
  FE_CHECK_HIP(hipMemcpyAsync(w.seg_tiles, seg_tiles.data(), seg_tiles.size() * sizeof(int32_t),
                              hipMemcpyHostToDevice, h->stream));
  // the block-diagonal coarse space applies its FP64 blocks at every precision
  Problem pb{A64, A32 ? A32 : A64, o->mixed ? M32 : M64, b, x, nseg, ntiles, C,
             (o->mixed && !(C && C->n_seg)) ? Ainv32 : Ainv64};
  *s = fe_pcg_stats{};
  int rc = ntiles == 0 ? FE_OK
           : o->mixed  ? dispatch_dof<float>(h, pb, w, o, s)
                       : dispatch_dof<double>(h, pb, w, o, s);
  if (rc) {
    (void)hipStreamSynchronize(h->stream);  // the host tile arrays must outlive their copies
    return rc;
  }
  std::vector<fe_pcg_seg_stats> seg((size_t)nseg);
  if (ntiles > 0) {
    // on the solver's stream: a plain hipMemcpy would also wait for every other blocking stream
    FE_CHECK_HIP(hipMemcpyAsync(seg.data(), w.seg_stats, nseg * sizeof(fe_pcg_seg_stats),
                                hipMemcpyDeviceToHost, h->stream));
    FE_CHECK_HIP(hipStreamSynchronize(h->stream));
  } else {
    for (auto& q : seg) q = fe_pcg_seg_stats{};
  }
  // plain FP64 PCG finishes every segment whose FP32 iteration failed, from the x it reached;
  // segments that already converged stop at the first true-residual check
  int failed = 0;
  for (const auto& q : seg) failed += q.status != FE_OK;
  if (o->mixed && M64 && failed && ntiles > 0 && (!C || Ainv64)) {
    // a segment that ended worse than the zero vector restarts from zero
    const int D = A64->dof;
    for (int g = 0; g < nseg; ++g)
      if (seg[g].status != FE_OK && !(seg[g].relres < 1.0) && seg_nodes[g + 1] > seg_nodes[g])
        FE_CHECK_HIP(hipMemsetAsync(x + (int64_t)D * seg_nodes[g], 0,
                                    (size_t)D * (seg_nodes[g + 1] - seg_nodes[g]) * sizeof(double),
                                    h->stream));
    fe_pcg_opts o64 = *o;
    o64.mixed = 0;
    o64.profile = 0;
    Problem pb64{A64, A64, M64, b, x, nseg, ntiles, C, Ainv64};
    fe_pcg_stats s64{};
    rc = dispatch_dof<double>(h, pb64, w, &o64, &s64);
    if (rc) return rc;
    std::vector<fe_pcg_seg_stats> seg64((size_t)nseg);
    FE_CHECK_HIP(hipMemcpyAsync(seg64.data(), w.seg_stats, nseg * sizeof(fe_pcg_seg_stats),
                                hipMemcpyDeviceToHost, h->stream));
    FE_CHECK_HIP(hipStreamSynchronize(h->stream));
    for (int g = 0; g < nseg; ++g) {
      if (seg[g].status == FE_OK) continue;
      fe_pcg_seg_stats q = seg64[g];
      q.iterations += seg[g].iterations;
      q.reliable_updates += seg[g].reliable_updates;
      q.restarts += seg[g].restarts;
      q.fallback = 1;
      seg[g] = q;
    }
    s->fallbacks = failed;
  }
  // aggregate: worst status, largest residuals, most iterations
  s->status = FE_OK;
  s->relres = 0.0;
  s->relres_rec = 0.0;
  s->restarts = 0;
  double bb = 0.0;
  for (int g = 0; g < nseg; ++g) {
    const fe_pcg_seg_stats& q = seg[g];
    if (q.status == FE_ERR_BREAKDOWN || (q.status != FE_OK && s->status == FE_OK)) s->status = q.status;
    s->relres = std::max(s->relres, q.relres);
    s->relres_rec = std::max(s->relres_rec, q.relres_rec);
    s->restarts += q.restarts;
    s->iterations = std::max(s->iterations, q.iterations);
    bb += q.bnorm * q.bnorm;
  }
  s->bnorm = sqrt(bb);
  if (seg_out) std::copy(seg.begin(), seg.end(), seg_out);
  return FE_OK;
}

}  // namespace

extern "C" int fe_pcg_workspace(fe_handle* h, const fe_bsr* A, const fe_pcg_opts* opts,
                                int64_t* bytes) {
  FE_ENTER(h);
  FE_REQUIRE(A && opts && bytes, "fe_pcg_workspace: null argument");
  const int32_t seg[2] = {0, (int32_t)A->n_brows};
  return fe_pcg_batched_workspace(h, A, 1, seg, opts, bytes);
}

extern "C" int fe_pcg_batched_workspace(fe_handle* h, const fe_bsr* A, int32_t n_seg,
                                        const int32_t* seg_nodes, const fe_pcg_opts* opts,
                                        int64_t* bytes) {
  FE_ENTER(h);
  FE_REQUIRE(A && opts && bytes, "fe_pcg_batched_workspace: null argument");
  FE_REQUIRE(A->dof >= 1 && A->dof <= 3, "fe_pcg_batched_workspace: dof must be 1..3");
  int rc = check_segments(A, n_seg, seg_nodes, "fe_pcg_batched_workspace");
  if (rc) return rc;
  const int elem = (int)sizeof(double);
  const int64_t ntiles = count_tiles(n_seg, seg_nodes, tile_rows_dof(A->dof));
  *bytes = layout_work(nullptr, A->n_brows, A->dof, elem, n_seg, ntiles).bytes;
  return FE_OK;
}

extern "C" int fe_pcg(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                      const void* M32, const double* b, double* x, void* work, int64_t work_bytes,
                      const fe_pcg_opts* o, fe_pcg_stats* s) {
  FE_ENTER(h);
  FE_REQUIRE(b && x && work && o && s, "fe_pcg: null argument");
  int rc = check_pcg_args(A64, A32, M64, M32, o, "fe_pcg");
  if (rc) return rc;
  FE_REQUIRE(A64->n_brows < ((int64_t)1 << 31), "fe_pcg: too many node rows");
  *s = fe_pcg_stats{};
  if (A64->n_brows == 0) return FE_OK;
  const int32_t seg[2] = {0, (int32_t)A64->n_brows};
  return pcg_common(h, A64, A32, M64, M32, 1, seg, b, x, work, work_bytes, o, s, nullptr, "fe_pcg");
}

extern "C" int fe_pcg_batched(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                              const void* M32, int32_t n_seg, const int32_t* seg_nodes,
                              const double* b, double* x, void* work, int64_t work_bytes,
                              const fe_pcg_opts* o, fe_pcg_stats* s, fe_pcg_seg_stats* seg_stats) {
  FE_ENTER(h);
  FE_REQUIRE(b && x && work && o && s, "fe_pcg_batched: null argument");
  int rc = check_pcg_args(A64, A32, M64, M32, o, "fe_pcg_batched");
  if (rc) return rc;
  rc = check_segments(A64, n_seg, seg_nodes, "fe_pcg_batched");
  if (rc) return rc;
  return pcg_common(h, A64, A32, M64, M32, n_seg, seg_nodes, b, x, work, work_bytes, o, s, seg_stats,
                    "fe_pcg_batched");
}

extern "C" int fe_pcg_coarse_workspace(fe_handle* h, const fe_bsr* A, const fe_coarse* C,
                                       const fe_pcg_opts* opts, int64_t* bytes) {
  FE_ENTER(h);
  FE_REQUIRE(A && C && opts && bytes, "fe_pcg_coarse_workspace: null argument");
  FE_REQUIRE(A->dof >= 1 && A->dof <= 3, "fe_pcg_coarse_workspace: dof must be 1..3");
  const int32_t seg[2] = {0, (int32_t)A->n_brows};
  const int64_t ntiles = count_tiles(1, seg, tile_rows_dof(A->dof));
  *bytes = layout_work(nullptr, A->n_brows, A->dof, (int)sizeof(double), 1, ntiles, C->n_agg * C->nc)
               .bytes;
  return FE_OK;
}

extern "C" int fe_pcg_coarse(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                             const void* M32, const fe_coarse* C, const void* Ainv64,
                             const void* Ainv32, const double* b, double* x, void* work,
                             int64_t work_bytes, const fe_pcg_opts* o, fe_pcg_stats* s) {
  FE_ENTER(h);
  FE_REQUIRE(b && x && work && o && s && C, "fe_pcg_coarse: null argument");
  int rc = check_pcg_args(A64, A32, M64, M32, o, "fe_pcg_coarse");
  if (rc) return rc;
  FE_REQUIRE(C->dof == A64->dof && C->n_nodes == A64->n_brows, "fe_pcg_coarse: coarse space and A disagree");
  FE_REQUIRE(C->n_seg == 0, "fe_pcg_coarse: a block-diagonal coarse space goes with fe_pcg_batched_coarse");
  FE_REQUIRE(o->mixed ? (Ainv32 && C->P32) : (Ainv64 != nullptr),
             "fe_pcg_coarse: missing coarse inverse (FP32 inverse and P32 for mixed precision)");
  FE_REQUIRE(A64->n_brows < ((int64_t)1 << 31), "fe_pcg_coarse: too many node rows");
  *s = fe_pcg_stats{};
  if (A64->n_brows == 0) return FE_OK;
  const int32_t seg[2] = {0, (int32_t)A64->n_brows};
  return pcg_common(h, A64, A32, M64, M32, 1, seg, b, x, work, work_bytes, o, s, nullptr,
                    "fe_pcg_coarse", C, Ainv64, Ainv32);
}

extern "C" int fe_pcg_batched_coarse_workspace(fe_handle* h, const fe_bsr* A, int32_t n_seg,
                                               const int32_t* seg_nodes, const fe_coarse* C,
                                               const fe_pcg_opts* opts, int64_t* bytes) {
  FE_ENTER(h);
  FE_REQUIRE(A && C && opts && bytes, "fe_pcg_batched_coarse_workspace: null argument");
  FE_REQUIRE(A->dof >= 1 && A->dof <= 3, "fe_pcg_batched_coarse_workspace: dof must be 1..3");
  int rc = check_segments(A, n_seg, seg_nodes, "fe_pcg_batched_coarse_workspace");
  if (rc) return rc;
  const int64_t ntiles = count_tiles(n_seg, seg_nodes, tile_rows_dof(A->dof));
  *bytes = layout_work(nullptr, A->n_brows, A->dof, (int)sizeof(double), n_seg, ntiles, C->n_agg * C->nc)
               .bytes;
  return FE_OK;
}

extern "C" int fe_pcg_batched_coarse(fe_handle* h, const fe_bsr* A64, const fe_bsr* A32, const void* M64,
                                     const void* M32, int32_t n_seg, const int32_t* seg_nodes,
                                     const fe_coarse* C, const void* Ainv64, const void* Ainv32,
                                     const double* b, double* x, void* work, int64_t work_bytes,
                                     const fe_pcg_opts* o, fe_pcg_stats* s, fe_pcg_seg_stats* seg_stats) {
  FE_ENTER(h);
  FE_REQUIRE(b && x && work && o && s && C, "fe_pcg_batched_coarse: null argument");
  int rc = check_pcg_args(A64, A32, M64, M32, o, "fe_pcg_batched_coarse");
  if (rc) return rc;
  rc = check_segments(A64, n_seg, seg_nodes, "fe_pcg_batched_coarse");
  if (rc) return rc;
  FE_REQUIRE(C->dof == A64->dof && C->n_nodes == A64->n_brows,
             "fe_pcg_batched_coarse: coarse space and A disagree");
  FE_REQUIRE(C->n_seg == n_seg && C->agg_seg && C->seg_of_agg && C->blk_ofs && C->xhat && C->T &&
                 C->fixed_bits && (!o->mixed || (C->xhat32 && C->T32)),
             "fe_pcg_batched_coarse: a block-diagonal coarse space with one block per segment expected");
  FE_REQUIRE(Ainv64 != nullptr, "fe_pcg_batched_coarse: missing coarse inverse (FP64 blocks)");
  return pcg_common(h, A64, A32, M64, M32, n_seg, seg_nodes, b, x, work, work_bytes, o, s, seg_stats,
                    "fe_pcg_batched_coarse", C, Ainv64, Ainv32);
}
