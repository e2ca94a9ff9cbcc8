// Sparse matrix-vector products, value conversion between precisions / layouts and the
// (block-)Jacobi preconditioner setup.  The PCG lives in fe_pcg.hip.
#include "fe_sparse.h"

namespace {

using namespace fe_sparse;

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

// The same product on a padded pattern (16-byte loads of W blocks, G = lanes_for<D>() / W lanes).
template <typename T, int D, int G>
__global__ __launch_bounds__(kBlock) void k_spmv_vec(int64_t nb, int64_t nnzb,
                                                     const int32_t* __restrict__ bptr,
                                                     const int32_t* __restrict__ bcol,
                                                     const T* __restrict__ vals,
                                                     const T* __restrict__ x, T* __restrict__ y) {
  const int64_t gid = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t i = gid / G;
  const int lane = (int)(gid % G);
  if (i >= nb) return;
  T acc[D];
  bsr_row_vec<T, D, G>(bptr, bcol, vals, nnzb, x, i, lane, acc);
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

// dst block q = src block map[q] (or zero where map[q] < 0), in destination order.
template <typename S, typename T, int D, int LS, int LT>
__global__ void k_bsr_repack(int64_t nnzb_dst, int64_t nnzb_src, const int32_t* __restrict__ map,
                             const S* __restrict__ src, T* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nnzb_dst * D * D) return;
  int64_t q;
  int k;
  if (LT == FE_LAYOUT_PLANES) {
    k = (int)(t / nnzb_dst);
    q = t - (int64_t)k * nnzb_dst;
  } else {
    q = t / (D * D);
    k = (int)(t - q * D * D);
  }
  const int32_t qs = map[q];
  dst[vix<D, LT>(q, k, nnzb_dst)] = qs >= 0 ? (T)src[vix<D, LS>(qs, k, nnzb_src)] : T(0);
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

template <typename T, int D, int LAYOUT>
void launch_spmv(fe_handle* h, const fe_bsr* A, const void* x, void* y) {
  if (row_aligned(A)) {
    constexpr int GV = lanes_for<D>() / Vec16<T>::W;
    hipLaunchKernelGGL((k_spmv_vec<T, D, GV>), dim3((unsigned)ceil_div(A->n_brows * GV, kBlock)),
                       dim3(kBlock), 0, h->stream, A->n_brows, A->nnzb, A->bptr, A->bcol,
                       (const T*)A->vals, (const T*)x, (T*)y);
    return;
  }
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

template <typename S, typename T, int D>
void launch_repack(fe_handle* h, const fe_bsr* A, const fe_bsr* B, const int32_t* map) {
  const int64_t n = B->nnzb * D * D;
  dim3 grid((unsigned)ceil_div(n, 256)), block(256);
  const S* src = (const S*)A->vals;
  T* dst = (T*)B->vals;
#define FE_REPACK(LS, LT)                                                                           \
  hipLaunchKernelGGL((k_bsr_repack<S, T, D, LS, LT>), grid, block, 0, h->stream, B->nnzb, A->nnzb, map, \
                     src, dst)
  if (A->layout == FE_LAYOUT_PLANES) {
    if (B->layout == FE_LAYOUT_PLANES) FE_REPACK(FE_LAYOUT_PLANES, FE_LAYOUT_PLANES);
    else FE_REPACK(FE_LAYOUT_PLANES, FE_LAYOUT_BSR);
  } else {
    if (B->layout == FE_LAYOUT_PLANES) FE_REPACK(FE_LAYOUT_BSR, FE_LAYOUT_PLANES);
    else FE_REPACK(FE_LAYOUT_BSR, FE_LAYOUT_BSR);
  }
#undef FE_REPACK
}

template <typename S, typename T>
void launch_repack_d(fe_handle* h, const fe_bsr* A, const fe_bsr* B, const int32_t* map) {
  if (A->dof == 1) launch_repack<S, T, 1>(h, A, B, map);
  else if (A->dof == 2) launch_repack<S, T, 2>(h, A, B, map);
  else launch_repack<S, T, 3>(h, A, B, map);
}

template <typename S, typename T>
void launch_convert_d(fe_handle* h, const fe_bsr* A, const fe_bsr* B) {
  if (A->dof == 1) launch_convert<S, T, 1>(h, A, B);
  else if (A->dof == 2) launch_convert<S, T, 2>(h, A, B);
  else launch_convert<S, T, 3>(h, A, B);
}

// --------------------------------------------------------------------------------------------------
// folding: dst block q = sum over k in [seg[q], seg[q+1]) of coef[k] * src block src_block[k]
// (one thread per value, the sum in the map's order)
// --------------------------------------------------------------------------------------------------
template <int D, int LAYOUT>
__global__ void k_bsr_fold(int64_t nnzb_dst, int64_t nnzb_src, const int32_t* __restrict__ seg,
                           const int32_t* __restrict__ src_block, const double* __restrict__ coef,
                           const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= nnzb_dst * D * D) return;
  int64_t q;
  int k;
  if (LAYOUT == FE_LAYOUT_PLANES) {
    k = (int)(t / nnzb_dst);
    q = t - (int64_t)k * nnzb_dst;
  } else {
    q = t / (D * D);
    k = (int)(t - q * D * D);
  }
  double v = 0.0;
  for (int32_t c = seg[q]; c < seg[q + 1]; ++c) v += coef[c] * src[vix<D, LAYOUT>(src_block[c], k, nnzb_src)];
  dst[vix<D, LAYOUT>(q, k, nnzb_dst)] = v;
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

extern "C" int fe_bsr_repack(fe_handle* h, const fe_bsr* src, const fe_bsr* dst,
                             const int32_t* src_block) {
  FE_ENTER(h);
  int rc = check_bsr(src, "fe_bsr_repack(src)");
  if (rc) return rc;
  rc = check_bsr(dst, "fe_bsr_repack(dst)");
  if (rc) return rc;
  FE_REQUIRE(src->dof == dst->dof && src->n_brows == dst->n_brows,
             "fe_bsr_repack: src and dst must have the same rows and dof");
  FE_REQUIRE(src->vals != dst->vals, "fe_bsr_repack: must be out-of-place");
  if (dst->nnzb == 0) return FE_OK;
  FE_REQUIRE(src_block, "fe_bsr_repack: null block map");
  if (src->precision == 0 && dst->precision == 0) launch_repack_d<double, double>(h, src, dst, src_block);
  else if (src->precision == 0) launch_repack_d<double, float>(h, src, dst, src_block);
  else if (dst->precision == 0) launch_repack_d<float, double>(h, src, dst, src_block);
  else launch_repack_d<float, float>(h, src, dst, src_block);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_bsr_fold(fe_handle* h, const fe_bsr* src, const fe_bsr* dst, const int32_t* seg_ptr,
                           const int32_t* src_block, const double* coef) {
  FE_ENTER(h);
  int rc = check_bsr(src, "fe_bsr_fold(src)");
  if (rc) return rc;
  rc = check_bsr(dst, "fe_bsr_fold(dst)");
  if (rc) return rc;
  FE_REQUIRE(src->dof == dst->dof && src->n_brows == dst->n_brows && src->layout == dst->layout &&
                 src->precision == 0 && dst->precision == 0 && src->row_align == 0 && dst->row_align == 0,
             "fe_bsr_fold: FP64 matrices of the same rows, dof and layout expected (unpadded)");
  FE_REQUIRE(src->vals != dst->vals, "fe_bsr_fold: must be out-of-place");
  if (dst->nnzb == 0) return FE_OK;
  FE_REQUIRE(seg_ptr && src_block && coef, "fe_bsr_fold: null fold map");
  const int D = dst->dof;
  const int64_t total = dst->nnzb * D * D;
  const dim3 grid((unsigned)ceil_div(total, 256)), block(256);
  const double* a = (const double*)src->vals;
  double* b = (double*)dst->vals;
#define FE_FOLD(DD, L)                                                                              \
  hipLaunchKernelGGL((k_bsr_fold<DD, L>), grid, block, 0, h->stream, dst->nnzb, src->nnzb, seg_ptr, \
                     src_block, coef, a, b)
  if (src->layout == FE_LAYOUT_PLANES) {
    if (D == 1) FE_FOLD(1, FE_LAYOUT_PLANES); else if (D == 2) FE_FOLD(2, FE_LAYOUT_PLANES); else FE_FOLD(3, FE_LAYOUT_PLANES);
  } else {
    if (D == 1) FE_FOLD(1, FE_LAYOUT_BSR); else if (D == 2) FE_FOLD(2, FE_LAYOUT_BSR); else FE_FOLD(3, FE_LAYOUT_BSR);
  }
#undef FE_FOLD
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_precond_setup(fe_handle* h, const fe_bsr* A, int kind, double* out) {
  FE_ENTER(h);
  int rc = check_bsr(A, "fe_precond_setup");
  if (rc) return rc;
  FE_REQUIRE(out, "fe_precond_setup: null output");
  FE_REQUIRE(A->precision == 0, "fe_precond_setup: FP64 matrix expected");
  FE_REQUIRE(A->row_align == 0, "fe_precond_setup: the matrix's own pattern expected (row_align 0)");
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

