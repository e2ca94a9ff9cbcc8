// Device helpers shared by the sparse kernels (SpMV, PCG, learned preconditioner): value indexing of
// the block layouts, wavefront / workgroup reductions and the node-block row product.
#pragma once

#include "fe_internal.h"

namespace fe_sparse {

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

// 16-byte vectors: W = 16 / sizeof(T) consecutive entries of one value plane, and their columns.
template <typename T>
struct Vec16;
template <>
struct Vec16<float> {
  typedef float V __attribute__((ext_vector_type(4)));
  typedef int32_t I __attribute__((ext_vector_type(4)));
  static constexpr int W = 4;
};
template <>
struct Vec16<double> {
  typedef double V __attribute__((ext_vector_type(2)));
  typedef int32_t I __attribute__((ext_vector_type(2)));
  static constexpr int W = 2;
};

// Row product on a padded pattern (row_align = 4: every row starts at a multiple of 4 blocks and
// holds a multiple of 4; padding blocks are zero with the row as column).  G lanes per row; a lane
// takes W consecutive blocks per step, one 16-byte load per value plane and one for the columns.
// The summation order depends on the row alone (not on where the row sits in the arrays).
// active false: an empty row (the lanes still take part in the group's shuffles).
// The row's block range [s, e) as read from bptr (by the caller, e.g. ahead of a barrier).
template <typename T, int D, int G>
__device__ inline void bsr_row_vec_range(const int32_t* __restrict__ bcol, const T* __restrict__ vals,
                                         int64_t nnzb, const T* __restrict__ x, int32_t s, int32_t e,
                                         int lane, T (&acc)[D]) {
  using V = typename Vec16<T>::V;
  using I = typename Vec16<T>::I;
  constexpr int W = Vec16<T>::W;
#pragma unroll
  for (int d = 0; d < D; ++d) acc[d] = T(0);
  for (int32_t q = s + lane * W; q < e; q += G * W) {
    const I cols = *reinterpret_cast<const I*>(bcol + q);
    V a[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) a[k] = *reinterpret_cast<const V*>(vals + (int64_t)k * nnzb + q);
#pragma unroll
    for (int w = 0; w < W; ++w) {
      const int64_t j = cols[w];
      T xj[D];
#pragma unroll
      for (int d = 0; d < D; ++d) xj[d] = x[j * D + d];
#pragma unroll
      for (int di = 0; di < D; ++di)
#pragma unroll
        for (int dj = 0; dj < D; ++dj) acc[di] += a[di * D + dj][w] * xj[dj];
    }
  }
#pragma unroll
  for (int off = G / 2; off > 0; off >>= 1)
#pragma unroll
    for (int d = 0; d < D; ++d) acc[d] += __shfl_xor(acc[d], off, G);
}

template <typename T, int D, int G>
__device__ inline void bsr_row_vec(const int32_t* __restrict__ bptr, const int32_t* __restrict__ bcol,
                                   const T* __restrict__ vals, int64_t nnzb, const T* __restrict__ x,
                                   int64_t i, int lane, T (&acc)[D], bool active = true) {
  const int32_t s = active ? bptr[i] : 0, e = active ? bptr[i + 1] : 0;
  bsr_row_vec_range<T, D, G>(bcol, vals, nnzb, x, s, e, lane, acc);
}

// Whether A is on a padded pattern the vector row product may walk.
inline bool row_aligned(const fe_bsr* A) { return A->layout == FE_LAYOUT_PLANES && A->row_align == 4; }

template <typename T, int D>
__device__ inline T pick(const T (&acc)[D], int lane) {
  T out = acc[0];
#pragma unroll
  for (int d = 1; d < D; ++d)
    if (lane == d) out = acc[d];
  return out;
}

template <int D>
constexpr int lanes_for() {
  return D == 3 ? 16 : 8;
}

inline int lanes_for_dof(int d) { return d == 3 ? 16 : 8; }

// Argument check shared by the sparse entry points.
inline int check_bsr(const fe_bsr* A, const char* who) {
  if (!A) return fe_set_error(FE_ERR_ARG, "%s: null matrix", who);
  if (A->dof < 1 || A->dof > 3) return fe_set_error(FE_ERR_ARG, "%s: dof must be 1..3", who);
  if (A->precision != 0 && A->precision != 1)
    return fe_set_error(FE_ERR_ARG, "%s: precision must be 0 (FP64) or 1 (FP32)", who);
  if (A->layout != FE_LAYOUT_BSR && A->layout != FE_LAYOUT_PLANES)
    return fe_set_error(FE_ERR_ARG, "%s: layout must be FE_LAYOUT_BSR or FE_LAYOUT_PLANES", who);
  if (A->n_brows < 0 || A->nnzb < 0) return fe_set_error(FE_ERR_ARG, "%s: negative sizes", who);
  if (A->row_align != 0) {
    if (A->row_align != 4 || A->layout != FE_LAYOUT_PLANES || A->nnzb % 4 != 0)
      return fe_set_error(FE_ERR_ARG, "%s: row_align must be 0, or 4 with FE_LAYOUT_PLANES and nnzb a "
                          "multiple of 4", who);
    if (((uintptr_t)A->bcol | (uintptr_t)A->vals) % 16 != 0)
      return fe_set_error(FE_ERR_ARG, "%s: row_align = 4 needs 16-byte aligned bcol and vals", who);
  }
  if (A->n_brows > 0 && (!A->bptr || !A->bcol || !A->vals))
    return fe_set_error(FE_ERR_ARG, "%s: null matrix array", who);
  return FE_OK;
}


}  // namespace fe_sparse
