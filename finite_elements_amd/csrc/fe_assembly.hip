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
// matrix (npe blocks of d x d) straight from coordinates into LDS.  Phase 2: one thread per block
// entry of the range sums the LDS values listed by the block-level contribution map, in ascending
// element order, and stores its d x d values (coalesced across threads).
// ---------------------------------------------------------------------------------------------------
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

// Contribution-map entries staged in LDS per workgroup (larger ranges read HBM directly).
constexpr int kStageBq = 512;
constexpr int kStageSeg = 1024;

template <int KIND, int WHAT, int LAYOUT, int NT>
__global__ __launch_bounds__(NT) void k_assemble_fused(fe_mesh m, fe_graph g,
                                                       double* __restrict__ vals) {
  constexpr int NPE = KindTraits<KIND>::NPE, D = KindTraits<KIND>::D;
  constexpr int BS = NPE * D * D;
  extern __shared__ double S[];  // [BS][stride]
  __shared__ int32_t s_bptr[kFusedMaxNodes + 1];
  __shared__ int32_t s_iptr[kFusedMaxNodes + 1];
  __shared__ int32_t s_bsp[kStageBq + 1];
  __shared__ uint16_t s_bseg[kStageSeg];
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  const int32_t n0 = g.blk[w], n1 = g.blk[w + 1];
  const int32_t nn = n1 - n0;
  for (int t = threadIdx.x; t <= nn; t += NT) {
    s_bptr[t] = g.bptr[n0 + t];
    s_iptr[t] = g.inc_ptr[n0 + t];
  }
  __syncthreads();
  const int32_t I0 = s_iptr[0];
  const int32_t ninc = s_iptr[nn] - I0;
  const int32_t B0 = s_bptr[0], B1 = s_bptr[nn];
  const int32_t nbq = B1 - B0;
  const int stride = g.max_inc;
  // Stage the block-level contribution map of the range (independent of phase 1's loads).
  const bool staged_bq = nbq + 1 <= kStageBq + 1;
  if (staged_bq)
    for (int t = threadIdx.x; t <= nbq; t += NT) s_bsp[t] = g.bseg_ptr[B0 + t];
  const int32_t c_lo = g.bseg_ptr[B0];
  const int32_t nseg = g.bseg_ptr[B1] - c_lo;
  const bool staged_seg = nseg <= kStageSeg;
  if (staged_seg)
    for (int t = threadIdx.x; t < nseg; t += NT) s_bseg[t] = g.bseg[c_lo + t];
  // Phase 1: element-matrix row blocks of every incidence of the range.
  for (int t = threadIdx.x; t < ninc; t += NT) {
    const int32_t packed = g.inc[I0 + t];
    const int64_t e = packed / NPE;
    const int a = packed - (int)e * NPE;
    double blk[BS];
    element_row_block<KIND, WHAT>(m, e, a, blk);
#pragma unroll
    for (int k = 0; k < BS; ++k) S[k * stride + t] = blk[k];
  }
  __syncthreads();
  // Phase 2: one thread per block entry (node j, neighbour q) of the range sums its d x d values
  // over the contributing incidences in ascending element order.  Consecutive threads own
  // consecutive block entries, so plane and BSR stores are coalesced.
  for (int32_t t = threadIdx.x; t < nbq; t += NT) {
    const int32_t bq = B0 + t;
    int lo = 0, hi = nn - 1;
    while (lo < hi) {
      const int mid = (lo + hi + 1) >> 1;
      if (s_bptr[mid] <= bq)
        lo = mid;
      else
        hi = mid - 1;
    }
    const int j = lo;
    const int32_t c0 = staged_bq ? s_bsp[t] : g.bseg_ptr[bq];
    const int32_t c1 = staged_bq ? s_bsp[t + 1] : g.bseg_ptr[bq + 1];
    const int32_t base = s_iptr[j] - I0;
    double sum[D * D];
#pragma unroll
    for (int k = 0; k < D * D; ++k) sum[k] = 0.0;
    for (int32_t c = c0; c < c1; ++c) {
      const int v = staged_seg ? s_bseg[c - c_lo] : g.bseg[c];
      const int kl = v >> 2, b = v & 3;
      const double* src = S + (b * D * D) * stride + base + kl;
#pragma unroll
      for (int k = 0; k < D * D; ++k) sum[k] += src[k * stride];
    }
    if (LAYOUT == FE_LAYOUT_PLANES) {
#pragma unroll
      for (int k = 0; k < D * D; ++k) vals[(int64_t)k * g.nnzb + bq] = sum[k];
    } else if (LAYOUT == FE_LAYOUT_BSR) {
#pragma unroll
      for (int k = 0; k < D * D; ++k) vals[(int64_t)bq * D * D + k] = sum[k];
    } else {
      const int32_t deg = s_bptr[j + 1] - s_bptr[j];
      const int32_t q = bq - s_bptr[j];
      const int64_t row0 = (int64_t)s_bptr[j] * D * D + (int64_t)q * D;
#pragma unroll
      for (int di = 0; di < D; ++di)
#pragma unroll
        for (int dj = 0; dj < D; ++dj) vals[row0 + (int64_t)di * deg * D + dj] = sum[di * D + dj];
    }
  }
}

template <int KIND, int WHAT, int LAYOUT>
int launch_fused_nt(fe_handle* h, const fe_mesh* m, const fe_graph* g, double* vals, size_t lds) {
  dim3 grid((unsigned)g->n_blk);
  const int nt = g->threads > 0 ? g->threads : 256;
#define FE_LAUNCH_NT(NT)                                                                        \
  do {                                                                                          \
    auto kern = k_assemble_fused<KIND, WHAT, LAYOUT, NT>;                                       \
    if (lds > 64 * 1024)                                                                        \
      FE_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                       \
                                       hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));  \
    hipLaunchKernelGGL(kern, grid, dim3(NT), lds, h->stream, *m, *g, vals);                     \
  } while (0)
  if (nt == 64)
    FE_LAUNCH_NT(64);
  else if (nt == 128)
    FE_LAUNCH_NT(128);
  else if (nt == 256)
    FE_LAUNCH_NT(256);
  else
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: threads must be 64, 128 or 256");
#undef FE_LAUNCH_NT
  FE_CHECK_LAUNCH();
  return FE_OK;
}

template <int KIND>
int launch_fused(fe_handle* h, const fe_mesh* m, const fe_graph* g, int what, int layout,
                 double* vals) {
  constexpr int NPE = KindTraits<KIND>::NPE, D = KindTraits<KIND>::D;
  const size_t lds = (size_t)NPE * D * D * g->max_inc * sizeof(double);
  if (lds > 140 * 1024)
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: schedule needs %zu bytes of LDS", lds);
  if (what == FE_MASS && KIND == FE_KIND_TRI_MAGNETIC)
    return fe_set_error(FE_ERR_UNSUPPORTED, "fe_assemble_fused: magnetic elements have no mass matrix");
  if constexpr (KIND != FE_KIND_TRI_MAGNETIC) {
    if (what == FE_MASS) {
      if (layout == FE_LAYOUT_CSR) return launch_fused_nt<KIND, FE_MASS, FE_LAYOUT_CSR>(h, m, g, vals, lds);
      if (layout == FE_LAYOUT_BSR) return launch_fused_nt<KIND, FE_MASS, FE_LAYOUT_BSR>(h, m, g, vals, lds);
      return launch_fused_nt<KIND, FE_MASS, FE_LAYOUT_PLANES>(h, m, g, vals, lds);
    }
  }
  if (layout == FE_LAYOUT_CSR) return launch_fused_nt<KIND, FE_STIFFNESS, FE_LAYOUT_CSR>(h, m, g, vals, lds);
  if (layout == FE_LAYOUT_BSR) return launch_fused_nt<KIND, FE_STIFFNESS, FE_LAYOUT_BSR>(h, m, g, vals, lds);
  return launch_fused_nt<KIND, FE_STIFFNESS, FE_LAYOUT_PLANES>(h, m, g, vals, lds);
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
template <int D, int LAYOUT>
__global__ void k_dirichlet_bsr(int64_t nn, int64_t nnzb, const int32_t* __restrict__ bptr,
                                const int32_t* __restrict__ bcol, const uint8_t* __restrict__ fixed,
                                double* __restrict__ vals) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nn) return;
  for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
    const int64_t j = bcol[q];
    for (int di = 0; di < D; ++di)
      for (int dj = 0; dj < D; ++dj) {
        const int64_t r = i * D + di, c = j * D + dj;
        const int64_t at = LAYOUT == FE_LAYOUT_PLANES ? (int64_t)(di * D + dj) * nnzb + q
                                                      : (int64_t)q * D * D + di * D + dj;
        if (fixed[r] || fixed[c]) vals[at] = (r == c) ? 1.0 : 0.0;
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
  FE_REQUIRE(layout == FE_LAYOUT_CSR || layout == FE_LAYOUT_BSR || layout == FE_LAYOUT_PLANES,
             "fe_assemble_fused: bad layout");
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
  FE_REQUIRE(A->row_align == 0, "fe_apply_dirichlet_bsr: the matrix's own pattern expected (row_align 0)");
  FE_REQUIRE(A->dof >= 1 && A->dof <= 3, "fe_apply_dirichlet_bsr: dof must be 1..3");
  FE_REQUIRE(A->layout == FE_LAYOUT_BSR || A->layout == FE_LAYOUT_PLANES,
             "fe_apply_dirichlet_bsr: layout must be FE_LAYOUT_BSR or FE_LAYOUT_PLANES");
  if (A->n_brows == 0) return FE_OK;
  dim3 grid((unsigned)ceil_div(A->n_brows, 256)), block(256);
#define FE_DIR_LAUNCH(D)                                                                          \
  if (A->layout == FE_LAYOUT_PLANES)                                                              \
    hipLaunchKernelGGL((k_dirichlet_bsr<D, FE_LAYOUT_PLANES>), grid, block, 0, h->stream,         \
                       A->n_brows, A->nnzb, A->bptr, A->bcol, fixed, vals);                       \
  else                                                                                            \
    hipLaunchKernelGGL((k_dirichlet_bsr<D, FE_LAYOUT_BSR>), grid, block, 0, h->stream, A->n_brows, \
                       A->nnzb, A->bptr, A->bcol, fixed, vals);
  if (A->dof == 1) {
    FE_DIR_LAUNCH(1)
  } else if (A->dof == 2) {
    FE_DIR_LAUNCH(2)
  } else {
    FE_DIR_LAUNCH(3)
  }
#undef FE_DIR_LAUNCH
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
