// Two-level preconditioner, coarse part: the coarse operator A_c = P^T A P of aggregated rigid-body
// modes, its inverse, and the coarse correction u += P A_c^-1 P^T r.
//
//   k_coarse_ac       one workgroup per coarse block pair (I, J): the sum over the fine blocks (i, j)
//                     with agg(i) = I, agg(j) = J of P_i^T A_ij P_j, in a fixed order (no atomics);
//                     k_coarse_ac_rm forms the same from x^ and T (T_I^T (sum Z_i^T A_ij Z_j) T_J)
//   k_gj_*            in-place blocked Gauss-Jordan inversion of the dense SPD A_c (FP64, 64 x 64
//                     blocks, no pivoting: every pivot block is a Schur complement of an SPD matrix)
//   k_coarse_restrict one workgroup per aggregate: r_c = P^T r
//   k_coarse_gemv     one wavefront per coarse row: z = A_c^-1 r_c
//   k_coarse_prolong  one thread per DOF: u += P z
// With P = Z(x^) T given by the scaled node coordinates x^ and the aggregates' transforms T, the
// restriction sums the rigid-body moments of r (force and moment about the aggregate's centre) and
// applies T^T, and the prolongation applies T to z and evaluates the rigid-body motion at the node
// (k_coarse_restrict_rm, k_coarse_prolong_rm): 15 bytes per node are read instead of P's d*nc values.
// The dense inverse has a leading dimension padded to a multiple of 64 (identity in the padding).
#include <algorithm>
#include <mutex>

#include "fe_sparse.h"

namespace {

using namespace fe_sparse;

constexpr int kNB = 64;  // Gauss-Jordan block size

// ---------------------------------------------------------------------------------------------------
// coarse operator
// ---------------------------------------------------------------------------------------------------
constexpr int kAcThreads = 256;

// One workgroup per coarse pair (I, J); its threads stride over the pair's fine blocks (ascending
// order), each accumulating Z_i^T A_ij Z_j, then a fixed-order workgroup reduction.
template <int D, int NC>
__global__ __launch_bounds__(kAcThreads) void k_coarse_ac(fe_coarse C, int64_t nnzb,
                                                          const int32_t* __restrict__ bcol,
                                                          const double* __restrict__ vals, int64_t ld,
                                                          double* __restrict__ Ac) {
  __shared__ double red[NC * NC][kAcThreads / 64];
  const int64_t pr = blockIdx.x;
  const int32_t I = C.pair_ij[2 * pr], J = C.pair_ij[2 * pr + 1];
  double acc[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) acc[k] = 0.0;
  for (int32_t t = C.pair_ptr[pr] + threadIdx.x; t < C.pair_ptr[pr + 1]; t += kAcThreads) {
    const int64_t q = C.pair_q[t];
    const int64_t i = C.pair_row[t], j = bcol[q];
    double a[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) a[k] = vals[(int64_t)k * nnzb + q];
    // T = A_ij Z_j (D x NC)
    double tz[D][NC];
#pragma unroll
    for (int e = 0; e < D; ++e)
#pragma unroll
      for (int c = 0; c < NC; ++c) tz[e][c] = 0.0;
#pragma unroll
    for (int f = 0; f < D; ++f) {
      double zj[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) zj[c] = C.P[(j * D + f) * NC + c];
#pragma unroll
      for (int e = 0; e < D; ++e)
#pragma unroll
        for (int c = 0; c < NC; ++c) tz[e][c] += a[e * D + f] * zj[c];
    }
    // acc += Z_i^T T
#pragma unroll
    for (int e = 0; e < D; ++e) {
      double zi[NC];
#pragma unroll
      for (int c = 0; c < NC; ++c) zi[c] = C.P[(i * D + e) * NC + c];
#pragma unroll
      for (int x = 0; x < NC; ++x)
#pragma unroll
        for (int y = 0; y < NC; ++y) acc[x * NC + y] += zi[x] * tz[e][y];
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NC * NC) {
    const int k = threadIdx.x, x = k / NC, y = k % NC;
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kAcThreads / 64; ++w) v += red[k][w];
    // a coarse column that is zero (rank-deficient aggregate) gets a decoupled unit diagonal
    if (I == J && x == y && v == 0.0) v = 1.0;
    Ac[((int64_t)I * NC + x) * ld + (int64_t)J * NC + y] = v;
  }
}

// Rigid-body modes at a node's scaled coordinates x: column c < D is the translation e_c, the
// others the rotations omega x x (3D: about x, y, z; 2D: about z).
template <int D, int NC>
__device__ inline void rigid_modes_at(const double x[3], double (&Z)[D][NC]) {
#pragma unroll
  for (int e = 0; e < D; ++e)
#pragma unroll
    for (int c = 0; c < D; ++c) Z[e][c] = (e == c) ? 1.0 : 0.0;
  if constexpr (D == 3) {
    Z[0][3] = 0.0, Z[1][3] = -x[2], Z[2][3] = x[1];
    Z[0][4] = x[2], Z[1][4] = 0.0, Z[2][4] = -x[0];
    Z[0][5] = -x[1], Z[1][5] = x[0], Z[2][5] = 0.0;
  } else if constexpr (D == 2) {
    Z[0][2] = -x[1], Z[1][2] = x[0];
  }
}

// A_c = P^T A P with P = Z(x^) T: one workgroup per coarse pair (I, J) sums W = Z_i^T A_ij Z_j over the
// pair's fine blocks (ascending order; constrained rows and columns of A_ij read as zero), then
// forms T_I^T W T_J.  Reads the block values and the nodes' x^ and constraint bits, not P.
template <int D, int NC>
__global__ __launch_bounds__(kAcThreads) void k_coarse_ac_rm(fe_coarse C, int64_t nnzb,
                                                             const int32_t* __restrict__ bcol,
                                                             const double* __restrict__ vals, int64_t ld,
                                                             double* __restrict__ Ac) {
  __shared__ double red[NC * NC][kAcThreads / 64];
  __shared__ double s_W[NC * NC];
  const int64_t pr = blockIdx.x;
  const int32_t I = C.pair_ij[2 * pr], J = C.pair_ij[2 * pr + 1];
  double acc[NC * NC];
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) acc[k] = 0.0;
  // A thread adds its blocks t, t + kAcThreads, ... in that order, two per pass with every load of
  // the pair in flight together: ids, then the columns, values, row bits and row coordinates, then
  // the column bits and coordinates.  Past the pair's end the second block is the first again with
  // every entry read as constrained: its product is exact zeros, which leave the sums unchanged.
  const int32_t t_end = C.pair_ptr[pr + 1];
  for (int32_t t = C.pair_ptr[pr] + threadIdx.x; t < t_end; t += 2 * kAcThreads) {
    const bool two = t + kAcThreads < t_end;
    const int32_t tt[2] = {t, two ? t + kAcThreads : t};
    int32_t q0 = C.pair_q[tt[0]], q1 = C.pair_q[tt[1]];
    int32_t r0 = C.pair_row[tt[0]], r1 = C.pair_row[tt[1]];
    asm volatile("" : "+v"(q0), "+v"(q1), "+v"(r0), "+v"(r1));  // first use: all four in flight
    const int64_t q[2] = {q0, q1}, i[2] = {r0, r1};
    int32_t c0 = bcol[q[0]], c1 = bcol[q[1]];
    double av[2][D * D], xi[2][3];
    int bi[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
#pragma unroll
      for (int k = 0; k < D * D; ++k) av[h][k] = vals[(int64_t)k * nnzb + q[h]];
      bi[h] = C.fixed_bits[i[h]];
#pragma unroll
      for (int c = 0; c < 3; ++c) xi[h][c] = C.xhat[i[h] * 3 + c];
    }
    asm volatile("" : "+v"(c0), "+v"(c1));  // the columns first; the values stay in flight
    const int64_t j[2] = {c0, c1};
    double xj[2][3];
    int bj[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      bj[h] = C.fixed_bits[j[h]];
#pragma unroll
      for (int c = 0; c < 3; ++c) xj[h][c] = C.xhat[j[h] * 3 + c];
    }
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int brow = (h && !two) ? 7 : bi[h];
      double a[D][D];
#pragma unroll
      for (int e = 0; e < D; ++e)
#pragma unroll
        for (int f = 0; f < D; ++f) a[e][f] = (((brow >> e) | (bj[h] >> f)) & 1) ? 0.0 : av[h][e * D + f];
      double Zi[D][NC], Zj[D][NC];
      rigid_modes_at<D, NC>(xi[h], Zi);
      rigid_modes_at<D, NC>(xj[h], Zj);
      // tz = A_ij Z_j, then acc += Z_i^T tz
      double tz[D][NC];
#pragma unroll
      for (int e = 0; e < D; ++e)
#pragma unroll
        for (int c = 0; c < NC; ++c) {
          double v = 0.0;
#pragma unroll
          for (int f = 0; f < D; ++f) v += a[e][f] * Zj[f][c];
          tz[e][c] = v;
        }
#pragma unroll
      for (int x = 0; x < NC; ++x)
#pragma unroll
        for (int y = 0; y < NC; ++y) {
          double v = 0.0;
#pragma unroll
          for (int e = 0; e < D; ++e) v += Zi[e][x] * tz[e][y];
          acc[x * NC + y] += v;
        }
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NC * NC; ++k) {
    const double v = wave_sum(acc[k]);
    if (lane == 0) red[k][wid] = v;
  }
  __syncthreads();
  if (threadIdx.x < NC * NC) {
    double v = 0.0;
#pragma unroll
    for (int w = 0; w < kAcThreads / 64; ++w) v += red[threadIdx.x][w];
    s_W[threadIdx.x] = v;
  }
  __syncthreads();
  if (threadIdx.x < NC * NC) {
    const int x = threadIdx.x / NC, y = threadIdx.x % NC;
    const double* TI = C.T + (int64_t)I * NC * NC;
    const double* TJ = C.T + (int64_t)J * NC * NC;
    double v = 0.0;
    for (int k = 0; k < NC; ++k) {
      double wk = 0.0;
      for (int l = 0; l < NC; ++l) wk += s_W[k * NC + l] * TJ[l * NC + y];
      v += TI[k * NC + x] * wk;
    }
    // a coarse column that is zero (rank-deficient aggregate) gets a decoupled unit diagonal
    if (I == J && x == y && v == 0.0) v = 1.0;
    if (C.n_seg) {
      // block-diagonal space: the pair's segment block, row-major n_s x n_s (a pair across two
      // segments would be a coupling the space does not represent: it is dropped)
      const int sg = C.seg_of_agg[I], a0 = C.agg_seg[sg];
      const int64_t ns = (int64_t)NC * (C.agg_seg[sg + 1] - a0);
      if (C.seg_of_agg[J] == sg)
        Ac[C.blk_ofs[sg] + ((int64_t)(I - a0) * NC + x) * ns + (int64_t)(J - a0) * NC + y] = v;
    } else {
      Ac[((int64_t)I * NC + x) * ld + (int64_t)J * NC + y] = v;
    }
  }
}

__global__ void k_identity_pad(int64_t n, int64_t ld, double* __restrict__ A) {
  const int64_t k = n + blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k < ld) A[k * ld + k] = 1.0;
}

// ---------------------------------------------------------------------------------------------------
// blocked Gauss-Jordan inversion (in place, FP64)
// ---------------------------------------------------------------------------------------------------
// C (64 x 64, global) = alpha * X * Y (+ C if accumulate), X and Y 64 x 64 tiles in LDS.
__device__ void tile_gemm_store(double (*sX)[kNB + 1], double (*sY)[kNB + 1], double* __restrict__ Cg,
                                int64_t ld, double alpha, bool accumulate) {
  // 256 threads, 4 x 4 outputs each
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
  double acc[4][4] = {};
  for (int kk = 0; kk < kNB; ++kk) {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      x[a] = sX[tr + a][kk];
      y[a] = sY[kk][tc + a];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      double* c = Cg + (int64_t)(tr + a) * ld + tc + b;
      *c = (accumulate ? *c : 0.0) + alpha * acc[a][b];
    }
}

// acc = X * Y for the 4 x 4 outputs (rows tr.., columns tc..) of this thread, 256 threads.
__device__ inline void tile_gemm_acc(double (*sX)[kNB + 1], double (*sY)[kNB + 1], double (&acc)[4][4]) {
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) acc[a][b] = 0.0;
  for (int kk = 0; kk < kNB; ++kk) {
    double x[4], y[4];
#pragma unroll
    for (int a = 0; a < 4; ++a) {
      x[a] = sX[tr + a][kk];
      y[a] = sY[kk][tc + a];
    }
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int b = 0; b < 4; ++b) acc[a][b] += x[a] * y[b];
  }
}

constexpr int kGjThreads = 256;

// A 64 x 64 tile into LDS by the 256 threads: each thread's eight 16-byte loads are issued together
// (the rows are 16-byte aligned: ld and the tile offsets are multiples of 64).
__device__ void load_tile(const double* __restrict__ A, int64_t ld, int bi, int bj,
                          double (*s)[kNB + 1]) {
  constexpr int kPer = kNB * kNB / 2 / kGjThreads;
  double2 v[kPer];
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int idx = 2 * ((int)threadIdx.x + it * kGjThreads);
    const int r = idx / kNB, c = idx % kNB;
    v[it] = *reinterpret_cast<const double2*>(A + (int64_t)(bi * kNB + r) * ld + bj * kNB + c);
  }
#pragma unroll
  for (int it = 0; it < kPer; ++it) {
    const int idx = 2 * ((int)threadIdx.x + it * kGjThreads);
    const int r = idx / kNB, c = idx % kNB;
    s[r][c] = v[it].x;
    s[r][c + 1] = v[it].y;
  }
}

// In-place Gauss-Jordan inversion of a 64 x 64 tile by 256 threads: thread (g, c) keeps rows
// 16g..16g+15 of column c in a[].  Each elimination step broadcasts the pivot row, the pivot column and
// the pivot's reciprocal (formed once, by the pivot's owner) through LDS buffers that alternate
// between steps, so one barrier per step separates their writes from the reads.  No pivoting: the
// diagonal blocks met by the blocked elimination of an SPD matrix are SPD.
struct GjLds {
  double row[2][kNB], col[2][kNB], piv[2];
};

__device__ void gj_invert_tile(double (&a)[16], GjLds& L) {
  const int c = threadIdx.x % kNB, g = threadIdx.x / kNB;
  for (int pb = 0; pb < kNB / 16; ++pb) {
#pragma unroll
    for (int q = 0; q < 16; ++q) {
      const int p = pb * 16 + q, buf = q & 1;
      if (g == pb) {
        L.row[buf][c] = a[q];
        if (c == p) L.piv[buf] = 1.0 / a[q];
      }
      if (c == p)
#pragma unroll
        for (int i = 0; i < 16; ++i) L.col[buf][g * 16 + i] = a[i];
      __syncthreads();
      const double prow = (c == p ? 1.0 : L.row[buf][c]) * L.piv[buf];
#pragma unroll
      for (int i = 0; i < 16; ++i) {
        const int r = g * 16 + i;
        a[i] = (r == p) ? prow : (c == p ? 0.0 : a[i]) - L.col[buf][r] * prow;
      }
    }
  }
}

// S = inv(A_00) (the first pivot block; later ones are inverted inside k_gj_update).
__global__ __launch_bounds__(kGjThreads) void k_gj_inv(int64_t ld, const double* __restrict__ A,
                                                       double* __restrict__ S) {
  __shared__ GjLds L;
  const int c = threadIdx.x % kNB, g = threadIdx.x / kNB;
  double a[16];
#pragma unroll
  for (int i = 0; i < 16; ++i) a[i] = A[(int64_t)(g * 16 + i) * ld + c];
  gj_invert_tile(a, L);
#pragma unroll
  for (int i = 0; i < 16; ++i) S[(g * 16 + i) * kNB + c] = a[i];
}

// Step k of the blocked elimination reads A (the matrix after step k-1) and writes B; S = inv(A_kk).
// Part 1 (workgroup per block column j): B_kj = S A_kj, B_kk = S.
__global__ __launch_bounds__(kGjThreads) void k_gj_row(int64_t ld, int k, const double* __restrict__ A,
                                                       double* __restrict__ B,
                                                       const double* __restrict__ S) {
  __shared__ double sP[kNB][kNB + 1], sT[kNB][kNB + 1];
  const int j = blockIdx.x;
  if (j == k) {
    for (int idx = threadIdx.x; idx < kNB * kNB; idx += blockDim.x)
      B[(int64_t)(k * kNB + idx / kNB) * ld + k * kNB + idx % kNB] = S[idx];
    return;
  }
  for (int idx = threadIdx.x; idx < kNB * kNB; idx += blockDim.x) sP[idx / kNB][idx % kNB] = S[idx];
  load_tile(A, ld, k, j, sT);
  __syncthreads();
  tile_gemm_store(sP, sT, B + (int64_t)(k * kNB) * ld + j * kNB, ld, 1.0, false);
}

// Part 2 (workgroup per block (i, j), i != k): B_ij = A_ij - A_ik B_kj (j != k), B_ik = -A_ik S.
// The workgroup of block (k+1, k+1) then inverts its new block into Snext, the next step's S.
__global__ __launch_bounds__(kGjThreads) void k_gj_update(int64_t ld, int k, const double* __restrict__ A,
                                                          double* __restrict__ B,
                                                          const double* __restrict__ S,
                                                          double* __restrict__ Snext) {
  __shared__ double sX[kNB][kNB + 1], sY[kNB][kNB + 1];
  __shared__ GjLds L;
  const int nb = (int)(ld / kNB);
  const int i = blockIdx.x / nb, j = blockIdx.x % nb;
  if (i == k) return;
  load_tile(A, ld, i, k, sX);
  if (j == k) {
    for (int idx = threadIdx.x; idx < kNB * kNB; idx += blockDim.x) sY[idx / kNB][idx % kNB] = S[idx];
    __syncthreads();
    tile_gemm_store(sX, sY, B + (int64_t)(i * kNB) * ld + k * kNB, ld, -1.0, false);
    return;
  }
  load_tile(B, ld, k, j, sY);
  // this thread's 4 x 4 outputs of A_ij, loaded before the product so that their latency overlaps it
  const int tr = (threadIdx.x / 16) * 4, tc = (threadIdx.x % 16) * 4;
  double aij[4][4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const double* row = A + (int64_t)(i * kNB + tr + a) * ld + j * kNB + tc;
    const double2 lo = *reinterpret_cast<const double2*>(row);
    const double2 hi = *reinterpret_cast<const double2*>(row + 2);
    aij[a][0] = lo.x, aij[a][1] = lo.y, aij[a][2] = hi.x, aij[a][3] = hi.y;
  }
  __syncthreads();
  double acc[4][4];
  tile_gemm_acc(sX, sY, acc);
  const bool pivot_next = (i == k + 1 && j == k + 1);
  if (pivot_next) __syncthreads();  // sX is reused for the new block below
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    double v[4];
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      v[b] = aij[a][b] - acc[a][b];
      if (pivot_next) sX[tr + a][tc + b] = v[b];
    }
    double* row = B + (int64_t)(i * kNB + tr + a) * ld + j * kNB + tc;
    *reinterpret_cast<double2*>(row) = double2{v[0], v[1]};
    *reinterpret_cast<double2*>(row + 2) = double2{v[2], v[3]};
  }
  if (!pivot_next) return;
  // the next pivot block (this workgroup's own output): invert it into Snext
  __syncthreads();
  const int c = threadIdx.x % kNB, g = threadIdx.x / kNB;
  double t[16];
#pragma unroll
  for (int q = 0; q < 16; ++q) t[q] = sX[g * 16 + q][c];
  gj_invert_tile(t, L);
#pragma unroll
  for (int q = 0; q < 16; ++q) Snext[(g * 16 + q) * kNB + c] = t[q];
}

// Block-diagonal space: one workgroup per segment inverts its block (n_s <= 128, SPD, no pivoting) by
// in-place Gauss-Jordan in LDS, then writes the symmetrised inverse and its FP32 copy.
constexpr int kSegMax = 128;
constexpr int kSegThreads = 1024;

__global__ __launch_bounds__(kSegThreads) void k_coarse_inv_seg(fe_coarse C, double* __restrict__ A,
                                                               float* __restrict__ A32) {
  extern __shared__ double sA[];  // [n][n]
  __shared__ double s_row[kSegMax], s_col[kSegMax];
  const int sg = blockIdx.x;
  const int n = C.nc * (C.agg_seg[sg + 1] - C.agg_seg[sg]);
  if (n == 0) return;
  double* Ag = A + C.blk_ofs[sg];
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) sA[idx] = Ag[idx];
  __syncthreads();
  for (int p = 0; p < n; ++p) {
    const double piv = 1.0 / sA[p * n + p];
    for (int j = threadIdx.x; j < n; j += blockDim.x) {
      s_row[j] = (j == p ? 1.0 : sA[p * n + j]) * piv;
      s_col[j] = sA[j * n + p];
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
      const int i = idx / n, j = idx - i * n;
      sA[idx] = (i == p) ? s_row[j] : (j == p ? 0.0 : sA[idx]) - s_col[i] * s_row[j];
    }
    __syncthreads();
  }
  for (int idx = threadIdx.x; idx < n * n; idx += blockDim.x) {
    const int i = idx / n, j = idx - i * n;
    const double v = 0.5 * (sA[idx] + sA[j * n + i]);
    Ag[idx] = v;
    if (A32) A32[C.blk_ofs[sg] + idx] = (float)v;
  }
}

// Symmetrise the inverse (average with its transpose) and make the FP32 copy.
__global__ void k_sym_cast(int64_t n, int64_t ld, double* __restrict__ A, float* __restrict__ A32) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= n * n) return;
  const int64_t r = t / n, c = t % n;
  if (r > c) return;
  const double v = 0.5 * (A[r * ld + c] + A[c * ld + r]);
  A[r * ld + c] = v;
  A[c * ld + r] = v;
  if (A32) {
    A32[r * ld + c] = (float)v;
    A32[c * ld + r] = (float)v;
  }
}

// ---------------------------------------------------------------------------------------------------
// coarse correction
// ---------------------------------------------------------------------------------------------------
// mode 0: run while the segment iterates; mode 1: run during its reliable update; mode 2: always.
__device__ inline bool coarse_active(const int32_t* state, int mode) {
  if (mode == 2) return true;
  const int done = state[0], need_ru = state[1];
  return mode == 0 ? (!done && !need_ru) : (!done && need_ru);
}

constexpr int kRestrictThreads = 1024;

template <typename T, int D, int NC>
__global__ __launch_bounds__(kRestrictThreads) void k_coarse_restrict(fe_coarse C, const T* __restrict__ P,
                                                                      const T* __restrict__ r,
                                                                      T* __restrict__ rc,
                                                                      const int32_t* state, int mode) {
  __shared__ double red[NC][kRestrictThreads / 64];
  __shared__ int s_skip;
  if (threadIdx.x == 0) s_skip = state && !coarse_active(state, mode);
  __syncthreads();
  if (s_skip) return;
  const int g = blockIdx.x;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int32_t t = C.agg_ptr[g] + threadIdx.x; t < C.agg_ptr[g + 1]; t += blockDim.x) {
    const int64_t i = C.agg_nodes[t];
#pragma unroll
    for (int e = 0; e < D; ++e) {
      const double re = (double)r[i * D + e];
#pragma unroll
      for (int c = 0; c < NC; ++c) acc[c] += (double)P[(i * D + e) * NC + c] * re;
    }
  }
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    acc[c] = wave_sum(acc[c]);
    if (lane == 0) red[c][wid] = acc[c];
  }
  __syncthreads();
  if (threadIdx.x < NC) {
    const int c = threadIdx.x;
    double v = 0.0;
    for (int q = 0; q < kRestrictThreads / 64; ++q) v += red[c][q];
    rc[(int64_t)g * NC + c] = (T)v;
  }
}

template <typename T>
struct Vec4;
template <>
struct Vec4<float> {
  using type = float4;
};
template <>
struct Vec4<double> {
  using type = double4;
};

// One wavefront per row; 4-wide vector loads of the row (rows start 64-aligned: ld % 64 == 0).
// A lane adds its 4-wide pieces c = 4 lane + 256 j in order of j; the loads of kGemvBatch of them are
// issued together, the first batch before the state test (whose latency it overlaps).  Pieces past
// the row's end load the row's last piece again (a cache hit) and are not added: unconditional loads
// keep the batch in flight together.
constexpr int kGemvBatch = 8;
template <typename T>
__global__ __launch_bounds__(256) void k_coarse_gemv(int64_t n, int64_t ld, const T* __restrict__ Ainv,
                                                     const T* __restrict__ x, T* __restrict__ y,
                                                     const int32_t* state, int mode) {
  using V = typename Vec4<T>::type;
  __shared__ int s_skip;
  const int64_t row = (int64_t)blockIdx.x * 4 + threadIdx.x / 64;
  const int lane = threadIdx.x & 63;
  const bool live = row < n;
  const T* a = Ainv + (live ? row : 0) * ld;
  const int64_t n4 = live ? n & ~(int64_t)3 : 0;
  V av[kGemvBatch], xv[kGemvBatch];
  auto load = [&](int64_t c0) {
    if (n4 == 0) return;  // uniform over the wavefront
#pragma unroll
    for (int j = 0; j < kGemvBatch; ++j) {
      const int64_t c = min(c0 + 256 * j, n4 - 4);
      av[j] = *reinterpret_cast<const V*>(a + c);
      xv[j] = V{x[c], x[c + 1], x[c + 2], x[c + 3]};
    }
  };
  load(4 * lane);
  if (threadIdx.x == 0) s_skip = state && !coarse_active(state, mode);
  __syncthreads();
  if (s_skip || !live) return;
  double acc = 0.0;
  for (int64_t c0 = 4 * lane; c0 < n4; c0 += 256 * kGemvBatch) {
    if (c0 != 4 * lane) load(c0);
#pragma unroll
    for (int j = 0; j < kGemvBatch; ++j) {
      if (c0 + 256 * j < n4)
        acc += (double)av[j].x * (double)xv[j].x + (double)av[j].y * (double)xv[j].y +
               (double)av[j].z * (double)xv[j].z + (double)av[j].w * (double)xv[j].w;
    }
  }
  for (int64_t c = n4 + lane; c < n; c += 64) acc += (double)a[c] * (double)x[c];
  acc = wave_sum(acc);
  if (lane == 0) y[row] = (T)acc;
}

// Block-diagonal space: z_s = A_s^-1 r_s, one workgroup per segment; thread i forms row i from column i
// of the symmetric block (coalesced over i).  The blocks are read in FP64 at every precision: a thin
// mesh's coarse block is too ill-conditioned for an FP32 inverse.
template <typename T>
__global__ __launch_bounds__(kSegMax) void k_coarse_gemv_seg(fe_coarse C, const double* __restrict__ Ainv,
                                                            const T* __restrict__ x, T* __restrict__ y,
                                                            const int32_t* state, int state_stride,
                                                            int mode) {
  __shared__ T sx[kSegMax];
  const int sg = blockIdx.x;
  if (state && !coarse_active(state + (int64_t)sg * state_stride, mode)) return;  // uniform
  const int a0 = C.agg_seg[sg];
  const int n = C.nc * (C.agg_seg[sg + 1] - a0);
  const int64_t c0 = (int64_t)a0 * C.nc;
  if (threadIdx.x < n) sx[threadIdx.x] = x[c0 + threadIdx.x];
  __syncthreads();
  const int i = threadIdx.x;
  if (i >= n) return;
  const double* Ag = Ainv + C.blk_ofs[sg];
  double acc = 0.0;
  for (int j = 0; j < n; ++j) acc += Ag[(int64_t)j * n + i] * (double)sx[j];
  y[c0 + i] = (T)acc;
}

// Rigid-body moments of a nodal vector r (constrained DOFs read as zero) at scaled coordinates x:
// the columns of Z are the translations and the rotations omega x x about the axes.
template <int D>
__device__ inline void rigid_moments(const double x[3], const double (&rv)[D], double* w) {
  if constexpr (D == 3) {
    w[0] += rv[0];
    w[1] += rv[1];
    w[2] += rv[2];
    w[3] += x[1] * rv[2] - x[2] * rv[1];
    w[4] += x[2] * rv[0] - x[0] * rv[2];
    w[5] += x[0] * rv[1] - x[1] * rv[0];
  } else if constexpr (D == 2) {
    w[0] += rv[0];
    w[1] += rv[1];
    w[2] += x[0] * rv[1] - x[1] * rv[0];
  } else {
    w[0] += rv[0];
  }
}

// r_c = P^T r with P = Z(x^) T: one workgroup per aggregate sums the moments w, then r_c = T^T w.
// NT threads per workgroup, sized to the aggregates (1,000-node aggregates of one large system,
// ~100-node aggregates of a batch of small meshes).  A thread adds its nodes t, t + NT, ... in that
// order; two of them are loaded at a time.  The aggregate's bounds and T columns are loaded before the
// state test, whose latency they overlap.
template <typename T, int D, int NC, int NT>
__global__ __launch_bounds__(NT) void k_coarse_restrict_rm(fe_coarse C, const T* __restrict__ xh,
                                                                         const T* __restrict__ Tm,
                                                                         const T* __restrict__ r,
                                                                         T* __restrict__ rc,
                                                                         const int32_t* state,
                                                                         int state_stride, int mode) {
  __shared__ double red[NC][NT / 64];
  __shared__ int s_skip;
  const int g = blockIdx.x;
  const int tid = threadIdx.x;
  const int32_t a0 = C.agg_ptr[g], a1 = C.agg_ptr[g + 1];
  // T's column tid for the threads tid < NC of the first wavefront (the lanes above load column 0:
  // unconditional loads stay in flight together)
  T tcol[NC];
  if (tid < 64) {
    const int c = tid < NC ? tid : 0;
#pragma unroll
    for (int k = 0; k < NC; ++k) tcol[k] = Tm[((int64_t)g * NC + k) * NC + c];
  }
  if (tid == 0) {
    const int64_t sg = C.n_seg ? C.seg_of_agg[g] : 0;  // the aggregate's segment
    s_skip = state && !coarse_active(state + sg * state_stride, mode);
  }
  __syncthreads();
  if (s_skip) return;
  double acc[NC];
#pragma unroll
  for (int c = 0; c < NC; ++c) acc[c] = 0.0;
  for (int32_t t = a0 + tid; t < a1; t += 2 * NT) {
    // the second node of the pair is the first again, with every DOF read as fixed, when t + NT is
    // past the aggregate: its moments are then exact zeros, which leave the sums' bits unchanged, and
    // the pair's loads are unconditional, in one block, in flight together
    const bool two = t + NT < a1;
    int32_t n0 = C.agg_nodes[t];
    int32_t n1 = C.agg_nodes[two ? t + NT : t];
    // both node ids in flight before either is used: the empty asm is their first use (the compiler
    // would otherwise widen the first id, and wait for it, before issuing the second load)
    asm volatile("" : "+v"(n0), "+v"(n1));
    const int64_t i0 = n0, i1 = n1;
    int braw[2];
    T xr[2][3], rr[2][D];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int64_t i = h ? i1 : i0;
      braw[h] = C.fixed_bits[i];
#pragma unroll
      for (int c = 0; c < 3; ++c) xr[h][c] = xh[i * 3 + c];
#pragma unroll
      for (int e = 0; e < D; ++e) rr[h][e] = r[i * D + e];
    }
    __builtin_amdgcn_sched_barrier(0);  // both nodes' data in flight before any is used
    double x[2][3], rv[2][D];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int bits = (h && !two) ? 7 : braw[h];
#pragma unroll
      for (int c = 0; c < 3; ++c) x[h][c] = (double)xr[h][c];
#pragma unroll
      for (int e = 0; e < D; ++e) rv[h][e] = ((bits >> e) & 1) ? 0.0 : (double)rr[h][e];
    }
    rigid_moments<D>(x[0], rv[0], acc);
    rigid_moments<D>(x[1], rv[1], acc);
  }
  const int lane = tid & 63, wid = tid >> 6;
#pragma unroll
  for (int c = 0; c < NC; ++c) {
    acc[c] = wave_sum(acc[c]);
    if (lane == 0) red[c][wid] = acc[c];
  }
  __syncthreads();
  if (tid < NC) {
    // every one of the NC threads sums all NC moments (same order), then applies its column of T
    double wsum[NC];
#pragma unroll
    for (int k = 0; k < NC; ++k) {
      double v = 0.0;
      for (int q = 0; q < NT / 64; ++q) v += red[k][q];
      wsum[k] = v;
    }
    double v = 0.0;
#pragma unroll
    for (int k = 0; k < NC; ++k) v += (double)tcol[k] * wsum[k];
    rc[(int64_t)g * NC + tid] = (T)v;
  }
}

// u += P z with P = Z(x^) T: one thread per node applies T to its aggregate's z and adds the rigid
// motion at x^ to the free DOFs.
template <typename T, int D, int NC>
__global__ __launch_bounds__(256) void k_coarse_prolong_rm(fe_coarse C, const T* __restrict__ xh,
                                                           const T* __restrict__ Tm,
                                                           const T* __restrict__ z, T* __restrict__ u,
                                                           const int32_t* state, int state_stride,
                                                           int mode) {
  __shared__ int s_skip;
  // the node's aggregate, coordinates and fixed flags are loaded before the state test (whose
  // latency they overlap); one system: one state for the workgroup; a batch: the node's segment
  // (ranges of a workgroup may belong to several)
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  const bool live = i < C.n_nodes;
  const int64_t il = live ? i : 0;
  const int64_t g = C.node_agg[il];
  const int bits = C.fixed_bits[il];
  const T x = xh[il * 3], yy = xh[il * 3 + 1], zz = xh[il * 3 + 2];
  T uo[D];
#pragma unroll
  for (int e = 0; e < D; ++e) uo[e] = u[il * D + e];
  if (threadIdx.x == 0) s_skip = !C.n_seg && state && !coarse_active(state, mode);
  __syncthreads();
  if (s_skip || !live) return;
  if (C.n_seg && state && !coarse_active(state + (int64_t)C.seg_of_agg[g] * state_stride, mode)) return;
  const T* Tg = Tm + g * NC * NC;
  const T* zg = z + g * NC;
  T y[NC];
#pragma unroll
  for (int k = 0; k < NC; ++k) {
    T v = T(0);
#pragma unroll
    for (int c = 0; c < NC; ++c) v += Tg[k * NC + c] * zg[c];
    y[k] = v;
  }
  T du[D];
  if constexpr (D == 3) {
    du[0] = y[0] + zz * y[4] - yy * y[5];
    du[1] = y[1] - zz * y[3] + x * y[5];
    du[2] = y[2] + yy * y[3] - x * y[4];
  } else if constexpr (D == 2) {
    du[0] = y[0] - yy * y[2];
    du[1] = y[1] + x * y[2];
  } else {
    du[0] = y[0];
  }
#pragma unroll
  for (int e = 0; e < D; ++e)
    if (!((bits >> e) & 1)) u[i * D + e] = uo[e] + du[e];
}

template <typename T, int D, int NC>
__global__ __launch_bounds__(256) void k_coarse_prolong(fe_coarse C, const T* __restrict__ P,
                                                        const T* __restrict__ z, T* __restrict__ u,
                                                        const int32_t* state, int mode) {
  __shared__ int s_skip;
  if (threadIdx.x == 0) s_skip = state && !coarse_active(state, mode);
  __syncthreads();
  if (s_skip) return;
  const int64_t k = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (k >= C.n_nodes * D) return;
  const int64_t i = k / D;
  const T* zg = z + (int64_t)C.node_agg[i] * NC;
  const T* pk = P + k * NC;
  T s = T(0);
#pragma unroll
  for (int c = 0; c < NC; ++c) s += pk[c] * zg[c];
  u[k] += s;
}

int check_coarse(const fe_coarse* C, const char* who) {
  FE_REQUIRE(C, "%s: null coarse space", who);
  FE_REQUIRE(C->dof >= 1 && C->dof <= 3, "%s: dof must be 1..3", who);
  FE_REQUIRE(C->nc == (C->dof == 1 ? 1 : C->dof == 2 ? 3 : 6),
             "%s: nc must be 1, 3 or 6 rigid-body modes for dof 1, 2 or 3", who);
  FE_REQUIRE(C->n_agg >= 1 && C->n_pairs >= 1 && C->n_nodes >= 1, "%s: empty coarse space", who);
  FE_REQUIRE(C->node_agg && C->agg_ptr && C->agg_nodes && C->pair_ptr && C->pair_ij && C->pair_q &&
                 C->pair_row && C->P,
             "%s: null coarse array", who);
  if (C->n_seg) {
    FE_REQUIRE(C->n_seg > 0 && C->agg_seg && C->seg_of_agg && C->blk_ofs && C->xhat && C->T &&
                   C->fixed_bits,
               "%s: a block-diagonal coarse space needs agg_seg, seg_of_agg, blk_ofs, xhat, T and "
               "fixed_bits", who);
    FE_REQUIRE(C->max_block >= 1 && C->max_block <= kSegMax && C->blk_total >= 0,
               "%s: coarse blocks of 1..%d rows expected", who, kSegMax);
  } else {
    FE_REQUIRE(C->n_agg * C->nc <= 16384, "%s: coarse dimension %lld above 16384", who,
               (long long)(C->n_agg * C->nc));
  }
  return FE_OK;
}

}  // namespace

// Launch the three coarse-correction kernels on `stream` (used by the PCG and by fe_coarse_apply).
// Ainv: the inverse in the iteration precision; for a block-diagonal space always the FP64 blocks.
// state: the segment's {done, need_ru} words (or null); mode as in coarse_active.
int fe_coarse_correction(hipStream_t stream, const fe_coarse* C, int fp32, const void* Ainv,
                         const void* r, void* u, void* rc, void* zc, const int32_t* state,
                         int state_stride, int mode) {
  const int64_t nct = C->n_agg * C->nc, ld = fe_coarse_ld(nct);
  const dim3 gr((unsigned)C->n_agg), gg((unsigned)ceil_div(nct, 4)),
      gp((unsigned)ceil_div(C->n_nodes * C->dof, 256));
  const bool rm = fp32 ? (C->xhat32 && C->T32 && C->fixed_bits) : (C->xhat && C->T && C->fixed_bits);
  const dim3 gpn((unsigned)ceil_div(C->n_nodes, 256));
  const int64_t avg_agg = C->n_nodes / std::max<int64_t>(1, C->n_agg);  // nodes per aggregate
#define FE_CC(V, D, NC)                                                                            \
  {                                                                                                \
    const V* P = fp32 ? (const V*)C->P32 : (const V*)C->P;                                         \
    const V* xh = fp32 ? (const V*)C->xhat32 : (const V*)C->xhat;                                  \
    const V* Tm = fp32 ? (const V*)C->T32 : (const V*)C->T;                                        \
    if (rm && avg_agg >= 600)                                                                      \
      hipLaunchKernelGGL((k_coarse_restrict_rm<V, D, NC, 1024>), gr, dim3(1024), 0, stream, *C, xh,   \
                         Tm, (const V*)r, (V*)rc, state, state_stride, mode);                      \
    else if (rm && avg_agg >= 200)                                                                 \
      hipLaunchKernelGGL((k_coarse_restrict_rm<V, D, NC, 256>), gr, dim3(256), 0, stream, *C, xh, Tm, \
                         (const V*)r, (V*)rc, state, state_stride, mode);                          \
    else if (rm)                                                                                   \
      hipLaunchKernelGGL((k_coarse_restrict_rm<V, D, NC, 128>), gr, dim3(128), 0, stream, *C, xh, Tm, \
                         (const V*)r, (V*)rc, state, state_stride, mode);                          \
    else                                                                                           \
      hipLaunchKernelGGL((k_coarse_restrict<V, D, NC>), gr, dim3(kRestrictThreads), 0, stream, *C, P, \
                         (const V*)r, (V*)rc, state, mode);                                        \
    if (C->n_seg)                                                                                  \
      hipLaunchKernelGGL((k_coarse_gemv_seg<V>), dim3((unsigned)C->n_seg), dim3(kSegMax), 0, stream,  \
                         *C, (const double*)Ainv, (const V*)rc, (V*)zc, state, state_stride, mode); \
    else                                                                                           \
      hipLaunchKernelGGL((k_coarse_gemv<V>), gg, dim3(256), 0, stream, nct, ld, (const V*)Ainv,     \
                         (const V*)rc, (V*)zc, state, mode);                                       \
    if (rm)                                                                                        \
      hipLaunchKernelGGL((k_coarse_prolong_rm<V, D, NC>), gpn, dim3(256), 0, stream, *C, xh, Tm,      \
                         (const V*)zc, (V*)u, state, state_stride, mode);                          \
    else                                                                                           \
      hipLaunchKernelGGL((k_coarse_prolong<V, D, NC>), gp, dim3(256), 0, stream, *C, P, (const V*)zc, \
                         (V*)u, state, mode);                                                      \
  }
  if (fp32) {
    if (C->dof == 1) FE_CC(float, 1, 1) else if (C->dof == 2) FE_CC(float, 2, 3) else FE_CC(float, 3, 6)
  } else {
    if (C->dof == 1) FE_CC(double, 1, 1) else if (C->dof == 2) FE_CC(double, 2, 3) else FE_CC(double, 3, 6)
  }
#undef FE_CC
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int64_t fe_coarse_ld(int64_t n_coarse) { return (n_coarse + kNB - 1) / kNB * kNB; }

extern "C" int fe_coarse_operator(fe_handle* h, const fe_bsr* A, const fe_coarse* C, int invert,
                                  double* Ac, float* Ac32, double* scratch) {
  FE_ENTER(h);
  int rc = check_bsr(A, "fe_coarse_operator");
  if (rc) return rc;
  rc = check_coarse(C, "fe_coarse_operator");
  if (rc) return rc;
  FE_REQUIRE(A->precision == 0 && A->layout == FE_LAYOUT_PLANES,
             "fe_coarse_operator: A must be FP64 in FE_LAYOUT_PLANES");
  FE_REQUIRE(A->dof == C->dof && A->n_brows == C->n_nodes, "fe_coarse_operator: A and C disagree");
  FE_REQUIRE(A->row_align == 0, "fe_coarse_operator: the pattern of the coarse pair map expected");
  FE_REQUIRE(Ac && (!invert || scratch || C->n_seg), "fe_coarse_operator: null output");
  const int64_t nct = C->n_agg * C->nc, ld = fe_coarse_ld(nct);
  hipStream_t st = h->stream;
  FE_CHECK_HIP(hipMemsetAsync(Ac, 0, (size_t)(C->n_seg ? C->blk_total : ld * ld) * sizeof(double), st));
  const double* v = (const double*)A->vals;
  const bool rm = C->xhat && C->T && C->fixed_bits;
#define FE_AC(D, NC)                                                                               \
  {                                                                                                \
    if (rm)                                                                                        \
      hipLaunchKernelGGL((k_coarse_ac_rm<D, NC>), dim3((unsigned)C->n_pairs), dim3(kAcThreads), 0, st, \
                         *C, A->nnzb, A->bcol, v, ld, Ac);                                         \
    else                                                                                           \
      hipLaunchKernelGGL((k_coarse_ac<D, NC>), dim3((unsigned)C->n_pairs), dim3(kAcThreads), 0, st,   \
                         *C, A->nnzb, A->bcol, v, ld, Ac);                                         \
  }
  if (C->dof == 1) FE_AC(1, 1) else if (C->dof == 2) FE_AC(2, 3) else FE_AC(3, 6)
#undef FE_AC
  if (C->n_seg) {
    // block-diagonal: one workgroup per segment inverts its block in LDS
    FE_CHECK_LAUNCH();
    if (!invert) return FE_OK;
    const size_t lds = (size_t)C->max_block * C->max_block * sizeof(double);
    if (lds > 64 * 1024)
      FE_CHECK_HIP(hipFuncSetAttribute((const void*)k_coarse_inv_seg,
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
    hipLaunchKernelGGL(k_coarse_inv_seg, dim3((unsigned)C->n_seg), dim3(kSegThreads), lds, st, *C, Ac,
                       Ac32);
    FE_CHECK_LAUNCH();
    return FE_OK;
  }
  if (ld > nct)
    hipLaunchKernelGGL(k_identity_pad, dim3((unsigned)ceil_div(ld - nct, 64)), dim3(64), 0, st, nct,
                       ld, Ac);
  FE_CHECK_LAUNCH();
  if (!invert) return FE_OK;
  // blocked Gauss-Jordan, the matrix alternating between Ac and the scratch copy; the two pivot
  // inverses alternate between the two tiles after it
  const int nb = (int)(ld / kNB);
  double* Sb[2] = {scratch + ld * ld, scratch + ld * ld + kNB * kNB};
  double* cur = Ac;
  double* nxt = scratch;
  hipLaunchKernelGGL(k_gj_inv, dim3(1), dim3(kGjThreads), 0, st, ld, Ac, Sb[0]);
  for (int k = 0; k < nb; ++k) {
    hipLaunchKernelGGL(k_gj_row, dim3((unsigned)nb), dim3(kGjThreads), 0, st, ld, k, cur, nxt, Sb[k & 1]);
    hipLaunchKernelGGL(k_gj_update, dim3((unsigned)(nb * nb)), dim3(kGjThreads), 0, st, ld, k, cur, nxt,
                       Sb[k & 1], Sb[(k + 1) & 1]);
    std::swap(cur, nxt);
  }
  if (cur != Ac) FE_CHECK_HIP(hipMemcpyAsync(Ac, cur, (size_t)ld * ld * sizeof(double),
                                             hipMemcpyDeviceToDevice, st));
  hipLaunchKernelGGL(k_sym_cast, dim3((unsigned)ceil_div(nct * nct, 256)), dim3(256), 0, st, nct, ld,
                     Ac, Ac32);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

// ---------------------------------------------------------------------------------------------------
// allocation-free coarse correction
// ---------------------------------------------------------------------------------------------------
// The two coarse vectors (r_c = P^T r and z_c = A_c^-1 r_c) of fe_coarse_apply_noalloc live in this
// file-scope device array instead of a caller workspace.  Every GPU holds its own copy until the
// process ends; its device address is looked up once per device (fe_coarse_scratch_setup, called by
// fe_create with the device current) and only that address is handed to kernels.  Nothing in it needs
// clearing: the restriction writes every r_c entry and the coarse solve every z_c entry.
constexpr int64_t kNoallocScratchBytes = (int64_t)1 << 20;
constexpr int kMaxScratchDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoallocScratchBytes];
static void* g_scratch_ptr[kMaxScratchDevices];
static std::mutex g_scratch_mutex;

static_assert(2 * (int64_t)FE_COARSE_NOALLOC_MAX * 8 <= kNoallocScratchBytes,
              "FE_COARSE_NOALLOC_MAX FP64 coarse DOF must fit twice in the scratch array");

int fe_coarse_scratch_setup(int device) {
  if (device < 0 || device >= kMaxScratchDevices)
    return fe_set_error(FE_ERR_UNSUPPORTED, "coarse scratch: device %d above %d", device,
                        kMaxScratchDevices - 1);
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (g_scratch_ptr[device]) return FE_OK;
  int cur = -1;
  FE_CHECK_HIP(hipGetDevice(&cur));
  if (cur != device) return fe_set_error(FE_ERR_ARG, "coarse scratch: device %d is not current", device);
  void* ptr = nullptr;
  FE_CHECK_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
  if (!ptr) return fe_set_error(FE_ERR_NOMEM, "coarse scratch: null symbol address on device %d", device);
  g_scratch_ptr[device] = ptr;
  return FE_OK;
}

extern "C" int fe_coarse_apply_noalloc(fe_handle* h, const fe_coarse* C, int precision,
                                       const void* Ainv, const void* r, void* u) {
  FE_ENTER(h);
  int rc = check_coarse(C, "fe_coarse_apply_noalloc");
  if (rc) return rc;
  FE_REQUIRE(precision == 0 || precision == 1, "fe_coarse_apply_noalloc: precision must be 0 or 1");
  FE_REQUIRE(precision == 0 || C->P32, "fe_coarse_apply_noalloc: FP32 needs P32");
  FE_REQUIRE(Ainv && r && u, "fe_coarse_apply_noalloc: null argument");
  const int64_t nct = C->n_agg * C->nc;
  FE_REQUIRE(nct <= FE_COARSE_NOALLOC_MAX,
             "fe_coarse_apply_noalloc: %lld coarse DOF above the allocation-free maximum %d "
             "(use fe_coarse_apply with a workspace)", (long long)nct, FE_COARSE_NOALLOC_MAX);
  void* base = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    base = h->device >= 0 && h->device < kMaxScratchDevices ? g_scratch_ptr[h->device] : nullptr;
  }
  if (!base) {
    rc = fe_coarse_scratch_setup(h->device);  // first call on a device the handle did not set up
    if (rc) return rc;
    std::lock_guard<std::mutex> lock(g_scratch_mutex);
    base = g_scratch_ptr[h->device];
  }
  const size_t el = precision ? sizeof(float) : sizeof(double);
  char* w = (char*)base;
  void* rcv = w;
  void* zcv = w + ((nct * el + 255) / 256) * 256;
  return fe_coarse_correction(h->stream, C, precision, Ainv, r, u, rcv, zcv, nullptr, 0, 2);
}

extern "C" int fe_coarse_apply(fe_handle* h, const fe_coarse* C, int precision, const void* Ainv,
                               const void* r, void* u, void* work) {
  FE_ENTER(h);
  int rc = check_coarse(C, "fe_coarse_apply");
  if (rc) return rc;
  FE_REQUIRE(precision == 0 || precision == 1, "fe_coarse_apply: precision must be 0 or 1");
  FE_REQUIRE(precision == 0 || C->P32, "fe_coarse_apply: FP32 needs P32");
  FE_REQUIRE(Ainv && r && u && work, "fe_coarse_apply: null argument");
  const int64_t nct = C->n_agg * C->nc;
  const size_t el = precision ? sizeof(float) : sizeof(double);
  char* w = (char*)work;
  void* rcv = w;
  void* zcv = w + ((nct * el + 255) / 256) * 256;
  return fe_coarse_correction(h->stream, C, precision, Ainv, r, u, rcv, zcv, nullptr, 0, 2);
}
