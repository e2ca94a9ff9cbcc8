// Global assembly kernels.  Both are atomic-free and bitwise deterministic: every output entry is
// owned by one thread that sums its contributions in ascending element order.
#include "fe_element_math.h"
#include "fe_internal.h"

namespace {

// ---------------------------------------------------------------------------------------------------
// Two-phase assembly: element matrices already in HBM (entry-major), contributions pre-sorted by
// destination (gather) with segment pointers.  A workgroup owns 256 consecutive CSR entries; its
// contribution range is contiguous, so it is staged into LDS with one independent load per lane
// (all gathers in flight at once), then each thread reduces its own segment from LDS.
// ---------------------------------------------------------------------------------------------------
constexpr int kGatherBlock = 256;
constexpr int kGatherStage = 4096;  // contributions staged per workgroup pass (32 KB of LDS)

__global__ __launch_bounds__(kGatherBlock) void k_assemble_gather(
    const double* __restrict__ ke, const int32_t* __restrict__ gather,
    const int32_t* __restrict__ seg_ptr, int64_t nnz, double* __restrict__ vals) {
  __shared__ double stage[kGatherStage];
  const int64_t p0 = (int64_t)blockIdx.x * kGatherBlock;
  const int64_t p = p0 + threadIdx.x;
  const int64_t pend = min(p0 + kGatherBlock, nnz);
  const int64_t c0 = seg_ptr[p0];
  const int64_t c1 = seg_ptr[pend];
  int64_t s = 0, e = 0;
  if (p < nnz) {
    s = seg_ptr[p];
    e = seg_ptr[p + 1];
  }
  double sum = 0.0;
  // Stage in passes of kGatherStage contributions; a segment may straddle passes.
  for (int64_t base = c0; base < c1; base += kGatherStage) {
    const int64_t n = min((int64_t)kGatherStage, c1 - base);
    for (int64_t t = threadIdx.x; t < n; t += kGatherBlock) stage[t] = ke[gather[base + t]];
    __syncthreads();
    const int64_t lo = max(s, base), hi = min(e, base + n);
    for (int64_t c = lo; c < hi; ++c) sum += stage[c - base];
    __syncthreads();
  }
  if (p < nnz) vals[p] = sum;
}

// ---------------------------------------------------------------------------------------------------
// Fused assembly.  A workgroup owns the node range [blk[w], blk[w+1]).  Phase 1: one thread per
// node-element incidence (node i = local vertex a of element e) computes row block a of the element
// matrix (npe blocks of d x d) straight from coordinates into LDS.  Phase 2: one thread per output
// value of the range (contiguous in HBM, so stores are coalesced) sums the LDS values listed by the
// block-level contribution map, in ascending element order.
// ---------------------------------------------------------------------------------------------------
constexpr int kFusedBlock = 256;
constexpr int kFusedMaxNodes = 256;

template <int KIND>
struct KindTraits;
template <>
struct KindTraits<FE_KIND_TRI_ELASTIC> {
  static constexpr int NPE = 3, D = 2;
};
template <>
struct KindTraits<FE_KIND_TET_ELASTIC> {
  static constexpr int NPE = 4, D = 3;
};
template <>
struct KindTraits<FE_KIND_TRI_MAGNETIC> {
  static constexpr int NPE = 3, D = 1;
};

// Row block a of the element matrix of element e; out[(b*D + di)*D + dj] for b < NPE.
template <int KIND, int WHAT>
__device__ inline void element_row_block(const fe_mesh& m, int64_t e, int a, double* out) {
  constexpr int NPE = KindTraits<KIND>::NPE, D = KindTraits<KIND>::D;
  const int64_t ne = m.n_elem;
  double px[NPE], py[NPE], pz[NPE];
#pragma unroll
  for (int b = 0; b < NPE; ++b) {
    const int n = m.conn[b * ne + e];
    px[b] = m.x[n];
    py[b] = m.y[n];
    if (NPE == 4) pz[b] = m.z[n];
  }
  if constexpr (KIND == FE_KIND_TET_ELASTIC) {
    TetGeom g;
    tet_geom(px, py, pz, g);
    if (WHAT == FE_MASS) {
      const double f = m.rho[e] * g.vol / 20;
#pragma unroll
      for (int b = 0; b < NPE; ++b)
#pragma unroll
        for (int di = 0; di < D; ++di)
#pragma unroll
          for (int dj = 0; dj < D; ++dj)
            out[(b * D + di) * D + dj] = (di == dj) ? ((a == b) ? 2 * f : f) : 0.0;
    } else {
      double diag, off, shear;
      tet_coeffs(m.mat0[e], m.mat1[e], diag, off, shear);
      double ga[3];
#pragma unroll
      for (int k = 0; k < 3; ++k) ga[k] = (a == 0) ? g.g[0][k] : (a == 1) ? g.g[1][k]
                                        : (a == 2) ? g.g[2][k] : g.g[3][k];
#pragma unroll
      for (int b = 0; b < NPE; ++b)
#pragma unroll
        for (int di = 0; di < D; ++di)
#pragma unroll
          for (int dj = 0; dj < D; ++dj)
            out[(b * D + di) * D + dj] = g.vol * bdb_entry<3>(ga, g.g[b], di, dj, diag, off, shear);
    }
  } else if constexpr (KIND == FE_KIND_TRI_ELASTIC) {
    TriGeom g;
    tri_elastic_geom(px, py, g);
    if (WHAT == FE_MASS) {
      const double f = m.rho[e] * g.area * m.mat2[e] / 12;
#pragma unroll
      for (int b = 0; b < NPE; ++b)
#pragma unroll
        for (int di = 0; di < D; ++di)
#pragma unroll
          for (int dj = 0; dj < D; ++dj)
            out[(b * D + di) * D + dj] = (di == dj) ? ((a == b) ? 2 * f : f) : 0.0;
    } else {
      double diag, off, shear;
      // plane mode is passed in mat slot-independent field of the mesh
      tri_elastic_coeffs(m.mat0[e], m.mat1[e], m.plane_mode, diag, off, shear);
      const double scale = m.mat2[e] * g.area;
      const double ga[2] = {(a == 0) ? g.gx[0] : (a == 1) ? g.gx[1] : g.gx[2],
                            (a == 0) ? g.gy[0] : (a == 1) ? g.gy[1] : g.gy[2]};
#pragma unroll
      for (int b = 0; b < NPE; ++b) {
        const double gb[2] = {g.gx[b], g.gy[b]};
#pragma unroll
        for (int di = 0; di < D; ++di)
#pragma unroll
          for (int dj = 0; dj < D; ++dj)
            out[(b * D + di) * D + dj] = scale * bdb_entry<2>(ga, gb, di, dj, diag, off, shear);
      }
    }
  } else {
    TriGeom g;
    tri_magnetic_geom(px, py, g);
    const double inv_mu = 1 / m.mat0[e];
    const double gax = (a == 0) ? g.gx[0] : (a == 1) ? g.gx[1] : g.gx[2];
    const double gay = (a == 0) ? g.gy[0] : (a == 1) ? g.gy[1] : g.gy[2];
#pragma unroll
    for (int b = 0; b < NPE; ++b) out[b] = inv_mu * (gax * g.gx[b] + gay * g.gy[b]) * g.area;
  }
}

template <int KIND, int WHAT, int LAYOUT>
__global__ __launch_bounds__(kFusedBlock) void k_assemble_fused(fe_mesh m, fe_graph g,
                                                                double* __restrict__ vals) {
  constexpr int NPE = KindTraits<KIND>::NPE, D = KindTraits<KIND>::D;
  constexpr int BS = NPE * D * D;
  extern __shared__ double S[];  // [BS][stride]
  __shared__ int32_t s_bptr[kFusedMaxNodes + 1];
  __shared__ int32_t s_iptr[kFusedMaxNodes + 1];
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  const int32_t n0 = g.blk[w], n1 = g.blk[w + 1];
  const int32_t nn = n1 - n0;
  for (int t = threadIdx.x; t <= nn; t += kFusedBlock) {
    s_bptr[t] = g.bptr[n0 + t];
    s_iptr[t] = g.inc_ptr[n0 + t];
  }
  __syncthreads();
  const int32_t I0 = s_iptr[0];
  const int32_t ninc = s_iptr[nn] - I0;
  const int stride = g.max_inc;
  // Phase 1: element-matrix row blocks of every incidence of the range.
  for (int t = threadIdx.x; t < ninc; t += kFusedBlock) {
    const int32_t packed = g.inc[I0 + t];
    const int64_t e = packed / NPE;
    const int a = packed - (int)e * NPE;
    double blk[BS];
    element_row_block<KIND, WHAT>(m, e, a, blk);
#pragma unroll
    for (int k = 0; k < BS; ++k) S[k * stride + t] = blk[k];
  }
  __syncthreads();
  // Phase 2: one thread per output value, ascending-element sums.
  const int64_t o0 = (int64_t)s_bptr[0] * D * D;
  const int64_t o1 = (int64_t)s_bptr[nn] * D * D;
  for (int64_t o = o0 + threadIdx.x; o < o1; o += kFusedBlock) {
    // node of o: largest j with s_bptr[j]*D*D <= o
    int lo = 0, hi = nn - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if ((int64_t)s_bptr[mid] * D * D <= o)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int j = lo;
    const int32_t deg = s_bptr[j + 1] - s_bptr[j];
    const int32_t l = (int32_t)(o - (int64_t)s_bptr[j] * D * D);
    int q, di, dj;
    if (LAYOUT == FE_LAYOUT_CSR) {
      di = l / (deg * D);
      const int rem = l - di * deg * D;
      q = rem / D;
      dj = rem - q * D;
    } else {
      q = l / (D * D);
      const int r2 = l - q * D * D;
      di = r2 / D;
      dj = r2 - di * D;
    }
    const int32_t bq = s_bptr[j] + q;
    const int32_t c0 = g.bseg_ptr[bq], c1 = g.bseg_ptr[bq + 1];
    const int32_t base = s_iptr[j] - I0;
    double sum = 0.0;
    for (int32_t c = c0; c < c1; ++c) {
      const int v = g.bseg[c];
      const int kl = v >> 2, b = v & 3;
      sum += S[((b * D + di) * D + dj) * stride + base + kl];
    }
    vals[o] = sum;
  }
}

template <int KIND>
int launch_fused(fe_handle* h, const fe_mesh* m, const fe_graph* g, int what, int layout,
                 double* vals) {
  constexpr int NPE = KindTraits<KIND>::NPE, D = KindTraits<KIND>::D;
  const size_t lds = (size_t)NPE * D * D * g->max_inc * sizeof(double);
  if (lds > 150 * 1024)
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: schedule needs %zu bytes of LDS", lds);
  dim3 grid((unsigned)g->n_blk), block(kFusedBlock);
#define FE_LAUNCH_FUSED(W, L)                                                              \
  do {                                                                                     \
    auto kern = k_assemble_fused<KIND, W, L>;                                              \
    if (lds > 64 * 1024)                                                                   \
      FE_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                  \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
    hipLaunchKernelGGL(kern, grid, block, lds, h->stream, *m, *g, vals);                   \
  } while (0)
  if (what == FE_STIFFNESS) {
    if (layout == FE_LAYOUT_CSR)
      FE_LAUNCH_FUSED(FE_STIFFNESS, FE_LAYOUT_CSR);
    else
      FE_LAUNCH_FUSED(FE_STIFFNESS, FE_LAYOUT_BSR);
  } else {
    if constexpr (KIND == FE_KIND_TRI_MAGNETIC) {
      return fe_set_error(FE_ERR_UNSUPPORTED, "fe_assemble_fused: magnetic elements have no mass matrix");
    } else {
      if (layout == FE_LAYOUT_CSR)
        FE_LAUNCH_FUSED(FE_MASS, FE_LAYOUT_CSR);
      else
        FE_LAUNCH_FUSED(FE_MASS, FE_LAYOUT_BSR);
    }
  }
#undef FE_LAUNCH_FUSED
  FE_CHECK_LAUNCH();
  return FE_OK;
}

// ---------------------------------------------------------------------------------------------------
// Layout permutation within each node chunk: CSR [di][q][dj] <-> BSR [q][di][dj].
// ---------------------------------------------------------------------------------------------------
template <int D>
__global__ void k_relayout(const int32_t* __restrict__ bptr, int64_t nn, int to_bsr,
                           const double* __restrict__ src, double* __restrict__ dst) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  const int64_t b0 = bptr[i], deg = bptr[i + 1] - b0;
  const int64_t base = b0 * D * D;
  for (int64_t q = 0; q < deg; ++q)
    for (int di = 0; di < D; ++di)
      for (int dj = 0; dj < D; ++dj) {
        const int64_t csr = base + (int64_t)di * deg * D + q * D + dj;
        const int64_t bsr = base + q * D * D + di * D + dj;
        if (to_bsr)
          dst[bsr] = src[csr];
        else
          dst[csr] = src[bsr];
      }
}

// ---------------------------------------------------------------------------------------------------
// Dirichlet elimination on BSR values: entries in a fixed row or fixed column become identity.
// ---------------------------------------------------------------------------------------------------
template <int D>
__global__ void k_dirichlet_bsr(int64_t nn, const int32_t* __restrict__ bptr,
                                const int32_t* __restrict__ bcol, const uint8_t* __restrict__ fixed,
                                double* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
    const int64_t j = bcol[q];
    for (int di = 0; di < D; ++di)
      for (int dj = 0; dj < D; ++dj) {
        const int64_t r = i * D + di, c = j * D + dj;
        if (fixed[r] || fixed[c]) vals[(int64_t)q * D * D + di * D + dj] = (r == c) ? 1.0 : 0.0;
      }
  }
}

__global__ void k_dirichlet_rhs(int64_t n, const uint8_t* __restrict__ fixed,
                                const double* __restrict__ gv, const double* __restrict__ f,
                                const double* __restrict__ kg, double* __restrict__ b) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  b[i] = fixed[i] ? gv[i] : f[i] - kg[i];
}

}  // namespace

extern "C" int fe_assemble_gather(fe_handle* h, const double* ke, const int32_t* gather,
                                  const int32_t* seg_ptr, int64_t nnz, double* vals) {
  FE_ENTER(h);
  FE_REQUIRE(nnz >= 0, "fe_assemble_gather: negative nnz");
  if (nnz == 0) return FE_OK;
  FE_REQUIRE(ke && gather && seg_ptr && vals, "fe_assemble_gather: null pointer argument");
  hipLaunchKernelGGL(k_assemble_gather, dim3((unsigned)ceil_div(nnz, kGatherBlock)),
                     dim3(kGatherBlock), 0, h->stream, ke, gather, seg_ptr, nnz, vals);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_assemble_fused(fe_handle* h, const fe_mesh* m, const fe_graph* g, int what,
                                 int layout, double* vals) {
  FE_ENTER(h);
  FE_REQUIRE(m && g && vals, "fe_assemble_fused: null argument");
  FE_REQUIRE(what == FE_STIFFNESS || what == FE_MASS, "fe_assemble_fused: bad 'what'");
  FE_REQUIRE(layout == FE_LAYOUT_CSR || layout == FE_LAYOUT_BSR, "fe_assemble_fused: bad layout");
  FE_REQUIRE(g->n_blk >= 0 && g->max_inc >= 0, "fe_assemble_fused: bad schedule");
  if (g->n_blk == 0 || g->nnzb == 0) return FE_OK;
  FE_REQUIRE(m->conn && m->x && m->y && g->bptr && g->inc_ptr && g->inc && g->bseg_ptr && g->bseg &&
                 g->blk,
             "fe_assemble_fused: null mesh or graph array");
  FE_REQUIRE(m->n_nodes == g->n_nodes, "fe_assemble_fused: mesh and graph node counts differ");
  switch (m->kind) {
    case FE_KIND_TRI_ELASTIC:
      FE_REQUIRE(m->plane_mode == FE_PLANE_STRAIN || m->plane_mode == FE_PLANE_STRESS,
                 "fe_assemble_fused: bad plane_mode");
      FE_REQUIRE(what == FE_MASS ? (m->rho && m->mat2) : (m->mat0 && m->mat1 && m->mat2),
                 "fe_assemble_fused: missing material arrays");
      return launch_fused<FE_KIND_TRI_ELASTIC>(h, m, g, what, layout, vals);
    case FE_KIND_TET_ELASTIC:
      FE_REQUIRE(m->z, "fe_assemble_fused: tets need z coordinates");
      FE_REQUIRE(what == FE_MASS ? (m->rho != nullptr) : (m->mat0 && m->mat1),
                 "fe_assemble_fused: missing material arrays");
      return launch_fused<FE_KIND_TET_ELASTIC>(h, m, g, what, layout, vals);
    case FE_KIND_TRI_MAGNETIC:
      FE_REQUIRE(m->mat0, "fe_assemble_fused: missing mu_total array");
      return launch_fused<FE_KIND_TRI_MAGNETIC>(h, m, g, what, layout, vals);
    default:
      return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: unknown element kind %d", m->kind);
  }
}

extern "C" int fe_relayout(fe_handle* h, const int32_t* bptr, int64_t n_nodes, int dof,
                           int to_layout, const double* src, double* dst) {
  FE_ENTER(h);
  FE_REQUIRE(n_nodes >= 0 && dof >= 1 && dof <= 3, "fe_relayout: bad sizes");
  FE_REQUIRE(to_layout == FE_LAYOUT_CSR || to_layout == FE_LAYOUT_BSR, "fe_relayout: bad layout");
  if (n_nodes == 0) return FE_OK;
  FE_REQUIRE(bptr && src && dst && src != dst, "fe_relayout: bad pointers (must be out-of-place)");
  const int to_bsr = to_layout == FE_LAYOUT_BSR;
  dim3 grid((unsigned)ceil_div(n_nodes, 256)), block(256);
  if (dof == 1)
    hipLaunchKernelGGL(k_relayout<1>, grid, block, 0, h->stream, bptr, n_nodes, to_bsr, src, dst);
  else if (dof == 2)
    hipLaunchKernelGGL(k_relayout<2>, grid, block, 0, h->stream, bptr, n_nodes, to_bsr, src, dst);
  else
    hipLaunchKernelGGL(k_relayout<3>, grid, block, 0, h->stream, bptr, n_nodes, to_bsr, src, dst);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_apply_dirichlet_bsr(fe_handle* h, const fe_bsr* A, const uint8_t* fixed,
                                      double* vals) {
  FE_ENTER(h);
  FE_REQUIRE(A && fixed && vals, "fe_apply_dirichlet_bsr: null argument");
  FE_REQUIRE(A->precision == 0, "fe_apply_dirichlet_bsr: FP64 matrix expected");
  if (A->n_brows == 0) return FE_OK;
  dim3 grid((unsigned)ceil_div(A->n_brows, 256)), block(256);
  if (A->dof == 1)
    hipLaunchKernelGGL(k_dirichlet_bsr<1>, grid, block, 0, h->stream, A->n_brows, A->bptr, A->bcol,
                       fixed, vals);
  else if (A->dof == 2)
    hipLaunchKernelGGL(k_dirichlet_bsr<2>, grid, block, 0, h->stream, A->n_brows, A->bptr, A->bcol,
                       fixed, vals);
  else if (A->dof == 3)
    hipLaunchKernelGGL(k_dirichlet_bsr<3>, grid, block, 0, h->stream, A->n_brows, A->bptr, A->bcol,
                       fixed, vals);
  else
    return fe_set_error(FE_ERR_ARG, "fe_apply_dirichlet_bsr: dof must be 1..3");
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_dirichlet_rhs(fe_handle* h, int64_t n, const uint8_t* fixed, const double* g,
                                const double* f, const double* kg, double* b) {
  FE_ENTER(h);
  FE_REQUIRE(n >= 0, "fe_dirichlet_rhs: negative n");
  if (n == 0) return FE_OK;
  FE_REQUIRE(fixed && g && f && kg && b, "fe_dirichlet_rhs: null argument");
  hipLaunchKernelGGL(k_dirichlet_rhs, dim3((unsigned)ceil_div(n, 256)), dim3(256), 0, h->stream, n,
                     fixed, g, f, kg, b);
  FE_CHECK_LAUNCH();
  return FE_OK;
}
