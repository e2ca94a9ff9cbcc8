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
// distinct element incident to the range (welem, ascending) evaluates the element's geometry record
// (shape-function gradients, measure, material coefficients) straight from coordinates into LDS, so
// an element shared by several nodes of the range is evaluated once.  Phase 2: one thread per block
// entry (node i, neighbour j) of the range sums block (a, b) of every element holding both nodes,
// evaluated from the records, in ascending element order, and stores its d x d values (coalesced
// across threads).  No element matrix reaches HBM and no atomics are used.
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

// Doubles in the geometry record of one element (material coefficients pre-scaled by the measure;
// an odd record size keeps the lanes' LDS accesses spread over the banks):
//   tet stiffness  g[4][3], vol*off, vol*shear, 0         tet / triangle mass  f
//   tri stiffness  gx[3], gy[3], tA*off, tA*shear, 0      magnetic  gx[3], gy[3], area/mu
// The isotropic D matrices have diag = off + 2 shear (elements.py:418-453, 781-794), so with
// lambda = off and mu = shear block (a, b) of B^T D B is  lambda g_a g_b^T + mu g_b g_a^T +
// mu (g_a . g_b) I.
template <int KIND, int WHAT>
constexpr int record_size() {
  return WHAT == FE_MASS ? 1 : KIND == FE_KIND_TET_ELASTIC ? 15 : KIND == FE_KIND_TRI_ELASTIC ? 9 : 7;
}

// Inputs of one element record (doubles): the vertex coordinates, then the material values it reads
//   tet  x[4] y[4] z[4], then E, nu (stiffness) or rho (mass)
//   tri  x[3] y[3], then E, nu, thickness (stiffness) or rho, thickness (mass); magnetic: mu
template <int KIND, int WHAT>
constexpr int raw_size() {
  return KIND == FE_KIND_TET_ELASTIC   ? (WHAT == FE_MASS ? 13 : 14)
         : KIND == FE_KIND_TRI_ELASTIC ? (WHAT == FE_MASS ? 8 : 9)
                                       : 7;
}

template <int KIND>
constexpr int coord_dims() {
  return KIND == FE_KIND_TET_ELASTIC ? 3 : 2;
}

// Material values of one record (the inputs after the coordinates).
template <int KIND, int WHAT>
constexpr int n_mats() {
  return raw_size<KIND, WHAT>() - KindTraits<KIND>::NPE * coord_dims<KIND>();
}

template <int KIND, int WHAT>
__device__ inline void load_mats(const fe_mesh& m, int64_t e, double (&mt)[n_mats<KIND, WHAT>()]) {
  if constexpr (WHAT == FE_MASS) {
    mt[0] = m.rho[e];
    if constexpr (KIND == FE_KIND_TRI_ELASTIC) mt[1] = m.mat2[e];
  } else {
    mt[0] = m.mat0[e];
    if constexpr (KIND != FE_KIND_TRI_MAGNETIC) mt[1] = m.mat1[e];
    if constexpr (KIND == FE_KIND_TRI_ELASTIC) mt[2] = m.mat2[e];
  }
}

template <int KIND, int WHAT>
__device__ inline void load_raw(const fe_mesh& m, int64_t e, const int (&nid)[KindTraits<KIND>::NPE],
                                double (&R)[raw_size<KIND, WHAT>()]) {
  constexpr int NPE = KindTraits<KIND>::NPE;
#pragma unroll
  for (int b = 0; b < NPE; ++b) {
    const int n = nid[b];
    R[b] = m.x[n];
    R[NPE + b] = m.y[n];
    if (NPE == 4) R[2 * NPE + b] = m.z[n];
  }
  constexpr int M0 = NPE * coord_dims<KIND>();
  double mt[n_mats<KIND, WHAT>()];
  load_mats<KIND, WHAT>(m, e, mt);
#pragma unroll
  for (int i = 0; i < n_mats<KIND, WHAT>(); ++i) R[M0 + i] = mt[i];
}

// The record of an element from its inputs R (registers or LDS), field k stored at S[k].
template <int KIND, int WHAT>
__device__ inline void record_from_raw(const fe_mesh& m, const double* R, double* S) {
  constexpr int NPE = KindTraits<KIND>::NPE;
  double px[NPE], py[NPE], pz[NPE];
#pragma unroll
  for (int b = 0; b < NPE; ++b) {
    px[b] = R[b];
    py[b] = R[NPE + b];
    if (NPE == 4) pz[b] = R[2 * NPE + b];
  }
  if constexpr (KIND == FE_KIND_TET_ELASTIC) {
    TetGeom g;
    tet_geom(px, py, pz, g);
    if (WHAT == FE_MASS) {
      S[0] = R[12] * g.vol / 20;
    } else {
      double diag, off, shear;
      tet_coeffs(R[12], R[13], diag, off, shear);
#pragma unroll
      for (int v = 0; v < 4; ++v)
#pragma unroll
        for (int k = 0; k < 3; ++k) S[v * 3 + k] = g.g[v][k];
      S[12] = g.vol * off;
      S[13] = g.vol * shear;
      S[14] = 0.0;
    }
  } else if constexpr (KIND == FE_KIND_TRI_ELASTIC) {
    TriGeom g;
    tri_elastic_geom(px, py, g);
    if (WHAT == FE_MASS) {
      S[0] = R[6] * g.area * R[7] / 12;
    } else {
      double diag, off, shear;
      tri_elastic_coeffs(R[6], R[7], m.plane_mode, diag, off, shear);
      const double scale = R[8] * g.area;
#pragma unroll
      for (int v = 0; v < 3; ++v) {
        S[v] = g.gx[v];
        S[3 + v] = g.gy[v];
      }
      S[6] = scale * off;
      S[7] = scale * shear;
      S[8] = 0.0;
    }
  } else {
    TriGeom g;
    tri_magnetic_geom(px, py, g);
#pragma unroll
    for (int v = 0; v < 3; ++v) {
      S[v] = g.gx[v];
      S[3 + v] = g.gy[v];
    }
    S[6] = g.area / R[6];
  }
}

// Record of element e, field k stored at S[k].
template <int KIND, int WHAT>
__device__ inline void element_record(const fe_mesh& m, int64_t e, const int (&nid)[KindTraits<KIND>::NPE],
                                     double* S) {
  double R[raw_size<KIND, WHAT>()];
  load_raw<KIND, WHAT>(m, e, nid, R);
  record_from_raw<KIND, WHAT>(m, R, S);
}

// Gradient of local vertex a from a record (gx[.], gy[.] planes for triangles).
template <int KIND, int D>
__device__ inline void record_grad(const double* R, int a, double (&ga)[D]) {
  if constexpr (KIND == FE_KIND_TET_ELASTIC) {
#pragma unroll
    for (int k = 0; k < 3; ++k) ga[k] = R[a * 3 + k];
  } else {
    ga[0] = R[a];
    ga[1] = R[3 + a];
  }
}

// sum += block (a, b), a != b, of the element matrix whose record starts at R:
//   elasticity  sum_ij += (lambda g_a,i) g_b,j + (mu g_b,i) g_a,j + [i == j] mu (g_a . g_b)
//   magnetic    (A/mu) (b_a b_b + c_a c_b);  masses  f I.
template <int KIND, int WHAT, int D>
__device__ inline void acc_ab(const double* R, int a, int b, double (&sum)[D * D]) {
  if constexpr (WHAT == FE_MASS) {
    const double f = (a == b) ? 2 * R[0] : R[0];
#pragma unroll
    for (int d = 0; d < D; ++d) sum[d * D + d] += f;
  } else if constexpr (KIND == FE_KIND_TRI_MAGNETIC) {
    sum[0] += R[6] * (R[a] * R[b] + R[3 + a] * R[3 + b]);
  } else {
    constexpr int C = KIND == FE_KIND_TET_ELASTIC ? 12 : 6;  // lambda, mu
    double ga[D], gb[D];
    record_grad<KIND, D>(R, a, ga);
    record_grad<KIND, D>(R, b, gb);
    const double lam = R[C], mu = R[C + 1];
    double la[D], mb[D];
    double dot = ga[0] * gb[0];
#pragma unroll
    for (int i = 1; i < D; ++i) dot += ga[i] * gb[i];
#pragma unroll
    for (int i = 0; i < D; ++i) {
      la[i] = lam * ga[i];
      mb[i] = mu * gb[i];
    }
    const double mdot = mu * dot;
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = 0; j < D; ++j) {
        double v = la[i] * gb[j] + mb[i] * ga[j];
        if (i == j) v += mdot;
        sum[i * D + j] += v;
      }
  }
}

// sum += the diagonal block (a, a): (lambda + mu) g_a g_a^T + mu |g_a|^2 I (symmetric).
template <int KIND, int WHAT, int D>
__device__ inline void acc_aa(const double* R, int a, double (&sum)[D * D]) {
  if constexpr (WHAT == FE_MASS) {
    const double f2 = 2 * R[0];
#pragma unroll
    for (int d = 0; d < D; ++d) sum[d * D + d] += f2;
  } else if constexpr (KIND == FE_KIND_TRI_MAGNETIC) {
    const double gx = R[a], gy = R[3 + a];
    sum[0] += R[6] * (gx * gx + gy * gy);
  } else {
    constexpr int C = KIND == FE_KIND_TET_ELASTIC ? 12 : 6;
    double ga[D];
    record_grad<KIND, D>(R, a, ga);
    const double lm = R[C] + R[C + 1], mu = R[C + 1];
    double sq = ga[0] * ga[0];
#pragma unroll
    for (int i = 1; i < D; ++i) sq += ga[i] * ga[i];
    const double msq = mu * sq;
    double lg[D];
#pragma unroll
    for (int i = 0; i < D; ++i) lg[i] = lm * ga[i];
#pragma unroll
    for (int i = 0; i < D; ++i)
#pragma unroll
      for (int j = i; j < D; ++j) {
        double v = lg[i] * ga[j];
        if (i == j) v += msq;
        sum[i * D + j] += v;
        if (j != i) sum[j * D + i] += v;
      }
  }
}

// Dynamic LDS of one workgroup: the element records, the block contribution pointers and entries of
// the range (at most npe entries per incidence, and every block has one) and the packed incidence
// codes.  Everything phase 2 reads is in LDS.
template <int KIND, int WHAT>
inline size_t fused_lds_bytes(int max_elem, int max_inc) {
  constexpr int NPE = KindTraits<KIND>::NPE;
  return (size_t)record_size<KIND, WHAT>() * max_elem * sizeof(double) +
         ((size_t)NPE * max_inc + 1) * sizeof(int32_t) + (size_t)NPE * max_inc * sizeof(uint16_t) +
         (size_t)max_inc * sizeof(uint16_t) + 16;
}

// Store the d x d block sum of block entry bq (row node j of the range) in the requested layout.
template <int D, int LAYOUT>
__device__ inline void store_block(double* __restrict__ vals, const fe_graph& g, const int32_t* s_bptr,
                                   int j, int32_t bq, int k, double v) {
  if (LAYOUT == FE_LAYOUT_PLANES) {
    vals[(int64_t)k * g.nnzb + bq] = v;
  } else if (LAYOUT == FE_LAYOUT_BSR) {
    vals[(int64_t)bq * D * D + k] = v;
  } else {
    const int32_t deg = s_bptr[j + 1] - s_bptr[j];
    const int64_t row0 = (int64_t)s_bptr[j] * D * D + (int64_t)(bq - s_bptr[j]) * D;
    vals[row0 + (int64_t)(k / D) * deg * D + k % D] = v;
  }
}

// Phase 2 of the fused assembly for one node range whose records and maps are in LDS, run by NT
// threads of index tid (a multiple of 64 threads starting at a wavefront boundary).
template <int KIND, int WHAT, int LAYOUT, int NT>
__device__ inline void fused_phase2(const fe_graph& g, double* __restrict__ vals, const double* S,
                                    const int32_t* s_bptr, const int32_t* s_iptr, const int32_t* s_diag,
                                    const int32_t* s_bsp, const uint16_t* s_bseg, const uint16_t* s_sa,
                                    int32_t B0, int32_t nbq, int32_t nn, int tid) {
  constexpr int D = KindTraits<KIND>::D;
  constexpr int RS = record_size<KIND, WHAT>();
  constexpr int DL = 4;  // lanes per diagonal block
  // Phase 2, one loop over two kinds of work item (whole wavefronts take one kind):
  //  * items [0, nbq): one thread per off-diagonal block entry (node j, neighbour q), summing block
  //    (a, b) of the elements holding both nodes in ascending element order; consecutive threads own
  //    consecutive block entries, so the stores are coalesced;
  //  * items from nbq rounded up to 64: DL lanes per diagonal block; lane l sums the node's
  //    incidences l, l + DL, ... (ascending element order), then the lanes combine by a fixed
  //    butterfly.
  const int32_t nbq64 = (nbq + 63) & ~63;
  const int32_t items = nbq64 + nn * DL;
  for (int32_t t = tid; t < items; t += NT) {
    if (t < nbq64) {
      if (t >= nbq) continue;
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
      if (bq == s_diag[j]) continue;
      const int32_t c0 = s_bsp[t], c1 = s_bsp[t + 1];
      const int32_t base = s_iptr[j];
      double sum[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) sum[k] = 0.0;
      for (int32_t c = c0; c < c1; ++c) {
        const int v = s_bseg[c];
        const int sa = s_sa[base + (v >> 2)];
        acc_ab<KIND, WHAT, D>(S + (sa >> 2) * RS, sa & 3, v & 3, sum);
      }
#pragma unroll
      for (int k = 0; k < D * D; ++k) store_block<D, LAYOUT>(vals, g, s_bptr, j, bq, k, sum[k]);
    } else {
      const int u = t - nbq64;
      const int j = u / DL, l = u % DL;
      double sum[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) sum[k] = 0.0;
      for (int32_t k = s_iptr[j] + l; k < s_iptr[j + 1]; k += DL) {
        const int sa = s_sa[k];
        acc_aa<KIND, WHAT, D>(S + (sa >> 2) * RS, sa & 3, sum);
      }
#pragma unroll
      for (int off = DL / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < D * D; ++k) sum[k] += __shfl_xor(sum[k], off, DL);
      const int32_t bq = s_diag[j];
      if (bq >= 0)
#pragma unroll
        for (int k = 0; k < D * D; ++k)
          if (k % DL == l) store_block<D, LAYOUT>(vals, g, s_bptr, j, bq, k, sum[k]);
    }
  }
}

template <int KIND, int WHAT, int LAYOUT, int NT>
__global__ __launch_bounds__(NT) void k_assemble_fused(fe_mesh m, fe_graph g,
                                                       double* __restrict__ vals) {
  constexpr int D = KindTraits<KIND>::D;
  constexpr int RS = record_size<KIND, WHAT>();
  constexpr int DL = 4;  // lanes per diagonal block
  constexpr int NPE = KindTraits<KIND>::NPE;
  extern __shared__ double S[];  // [max_elem][RS] records (RS odd: few bank conflicts), then the maps
  __shared__ int32_t s_bptr[kFusedMaxNodes + 1];
  __shared__ int32_t s_iptr[kFusedMaxNodes + 1];
  __shared__ int32_t s_diag[kFusedMaxNodes];
  const int stride = g.max_elem;
  int32_t* s_bsp = reinterpret_cast<int32_t*>(S + (size_t)RS * stride);  // NPE * max_inc + 1
  uint16_t* s_bseg = reinterpret_cast<uint16_t*>(s_bsp + NPE * g.max_inc + 1);  // NPE * max_inc
  uint16_t* s_sa = s_bseg + NPE * g.max_inc;                                     // max_inc
  const int64_t w = xcd_remap(blockIdx.x, gridDim.x);
  int32_t n0, n1, E0, nel, I0, ninc, B0, nbq, c_lo, nseg;
  if (g.wg_hdr) {
    // one record (uniform loads) instead of a chain of dependent index loads
    const int32_t* hd = g.wg_hdr + w * 12;
    n0 = hd[0], n1 = hd[1], E0 = hd[2], nel = hd[3], I0 = hd[4], ninc = hd[5];
    B0 = hd[6], nbq = hd[7], c_lo = hd[8], nseg = hd[9];
  } else {
    n0 = g.blk[w], n1 = g.blk[w + 1];
    E0 = g.welem_ptr[w];
    nel = g.welem_ptr[w + 1] - E0;
    I0 = g.inc_ptr[n0], ninc = g.inc_ptr[n1] - I0;
    B0 = g.bptr[n0];
    nbq = g.bptr[n1] - B0;
    c_lo = g.bseg_ptr[B0];
    nseg = g.bseg_ptr[B0 + nbq] - c_lo;
  }
  const int32_t nn = n1 - n0;
  // a range beyond the schedule's stated maxima would overrun the LDS: such a workgroup writes nothing
  if (nn > kFusedMaxNodes || nel > stride || ninc > g.max_inc || nbq > NPE * g.max_inc ||
      nseg > NPE * g.max_inc)
    return;
  // Stage row pointers, incidence codes and the block-level contribution map of the range.
  for (int t = threadIdx.x; t <= nn; t += NT) {
    s_bptr[t] = g.bptr[n0 + t];
    s_iptr[t] = g.inc_ptr[n0 + t] - I0;
    if (t < nn) s_diag[t] = g.bdiag[n0 + t];
  }
  for (int t = threadIdx.x; t < ninc; t += NT) s_sa[t] = g.inc_sa[I0 + t];
  for (int t = threadIdx.x; t <= nbq; t += NT) s_bsp[t] = g.bseg_ptr[B0 + t] - c_lo;
  for (int t = threadIdx.x; t < nseg; t += NT) s_bseg[t] = g.bseg[c_lo + t];
  // Phase 1: the geometry record of every distinct element of the range.
  for (int t = threadIdx.x; t < nel; t += NT) {
    const int64_t e = g.welem[E0 + t];
    int nid[NPE];
#pragma unroll
    for (int b = 0; b < NPE; ++b)
      nid[b] = g.welem_conn ? g.welem_conn[b * g.n_welem + E0 + t] : m.conn[b * m.n_elem + e];
    element_record<KIND, WHAT>(m, e, nid, S + t * RS);
  }
  __syncthreads();
  fused_phase2<KIND, WHAT, LAYOUT, NT>(g, vals, S, s_bptr, s_iptr, s_diag, s_bsp, s_bseg, s_sa, B0, nbq, nn,
                                       threadIdx.x);
}

// Persistent form of the fused assembly: a workgroup walks several node ranges and, while it
// assembles one, already has the next range's maps, element ids and element nodes in flight (loaded
// into registers, written to LDS at the start of the next round), so the index loads of a range no
// longer sit on its critical path.  Needs the per-workgroup headers and element node lists, a range
// of at most NT - 1 nodes, at most NT incidences and at most NT distinct elements.  The ranges of XCD
// x (workgroups x, x + 8, ...) are the x-th eighth of the schedule, taken in turn by its workgroups.
// Registers of a range's maps as loaded (offsets are made range-relative when written to LDS: a
// subtraction right after a load would wait for it and serialise the loads).
template <int KIND>
struct PipeRegs {
  static constexpr int NPE = KindTraits<KIND>::NPE;
  static constexpr int QB = NPE + 1;  // block-map entries per thread: NPE * max_inc + 1 <= QB * NT
  int32_t bptr, iptr, diag, sa, bsp[QB], bseg[QB], e, nid[NPE];
};

struct RangeHdr {
  int32_t n0, n1, E0, nel, I0, ninc, B0, nbq, c_lo, nseg;
};

__device__ inline RangeHdr load_hdr(const fe_graph& g, int64_t r) {
  const int32_t* hd = g.wg_hdr + r * 12;
  return RangeHdr{hd[0], hd[1], hd[2], hd[3], hd[4], hd[5], hd[6], hd[7], hd[8], hd[9]};
}

// The maps, element ids and element node ids of range h into the registers of thread t of NT.
template <int KIND, int NT>
__device__ inline void load_range(const fe_graph& g, const RangeHdr& h, PipeRegs<KIND>& p, int t) {
  constexpr int NPE = KindTraits<KIND>::NPE;
  const int32_t nn = h.n1 - h.n0;
  if (t <= nn) {
    p.bptr = g.bptr[h.n0 + t];
    p.iptr = g.inc_ptr[h.n0 + t];
  }
  if (t < nn) p.diag = g.bdiag[h.n0 + t];
  if (t < h.ninc) p.sa = g.inc_sa[h.I0 + t];
#pragma unroll
  for (int q = 0; q < PipeRegs<KIND>::QB; ++q) {
    const int idx = t + q * NT;
    if (idx <= h.nbq) p.bsp[q] = g.bseg_ptr[h.B0 + idx];
    if (idx < h.nseg) p.bseg[q] = g.bseg[h.c_lo + idx];
  }
  if (t < h.nel) {
    p.e = g.welem[h.E0 + t];
#pragma unroll
    for (int b = 0; b < NPE; ++b) p.nid[b] = g.welem_conn[b * g.n_welem + h.E0 + t];
  }
}

template <int KIND, int WHAT, int LAYOUT, int NT>
__global__ __launch_bounds__(NT) void k_assemble_pipe(fe_mesh m, fe_graph g, double* __restrict__ vals) {
  constexpr int RS = record_size<KIND, WHAT>();
  constexpr int NPE = KindTraits<KIND>::NPE;
  constexpr int QB = PipeRegs<KIND>::QB;
  extern __shared__ double S[];  // [max_elem][RS] records, then the maps (as k_assemble_fused)
  __shared__ int32_t s_bptr[NT], s_iptr[NT], s_diag[NT];
  int32_t* s_bsp = reinterpret_cast<int32_t*>(S + (size_t)RS * g.max_elem);
  uint16_t* s_bseg = reinterpret_cast<uint16_t*>(s_bsp + NPE * g.max_inc + 1);
  uint16_t* s_sa = s_bseg + NPE * g.max_inc;
  const int x = blockIdx.x % 8, L = gridDim.x / 8;
  const int64_t lo = g.n_blk * x / 8, hi = g.n_blk * (x + 1) / 8;
  int64_t r = lo + blockIdx.x / 8;
  if (r >= hi) return;
  const int t = threadIdx.x;
  RangeHdr h = load_hdr(g, r);
  PipeRegs<KIND> p;
  load_range<KIND, NT>(g, h, p, t);
  for (; r < hi; r += L) {
    const RangeHdr c = h;  // the range assembled in this round
    const int32_t nn = c.n1 - c.n0;
    // a range beyond the limits the launch checked would overrun the LDS: it is left unwritten
    // (uniform over the workgroup, the rounds and barriers go on)
    const bool ok = nn < NT && c.nel <= g.max_elem && c.ninc <= g.max_inc && c.ninc <= NT &&
                    c.nbq <= NPE * g.max_inc && c.nseg <= NPE * g.max_inc;
    const bool more = r + L < hi;
    if (more) h = load_hdr(g, r + L);
    __syncthreads();  // the previous round's phase 2 has finished with the LDS
    if (ok) {
      if (t <= nn) {
        s_bptr[t] = p.bptr;
        s_iptr[t] = p.iptr - c.I0;
      }
      if (t < nn) s_diag[t] = p.diag;
      if (t < c.ninc) s_sa[t] = (uint16_t)p.sa;
#pragma unroll
      for (int q = 0; q < QB; ++q) {
        const int idx = t + q * NT;
        if (idx <= c.nbq) s_bsp[idx] = p.bsp[q] - c.c_lo;
        if (idx < c.nseg) s_bseg[idx] = (uint16_t)p.bseg[q];
      }
      if (t < c.nel) element_record<KIND, WHAT>(m, p.e, p.nid, S + t * RS);
    }
    // the next range's loads fly while this one is assembled
    if (more) load_range<KIND, NT>(g, h, p, t);
    __syncthreads();
    if (ok)
      fused_phase2<KIND, WHAT, LAYOUT, NT>(g, vals, S, s_bptr, s_iptr, s_diag, s_bsp, s_bseg, s_sa,
                                           c.B0, c.nbq, nn, t);
  }
}

// Map image of one range in the wave-specialised form: the arrays phase 2 reads, at fixed 16-byte
// aligned offsets sized by the schedule's maxima, offsets already range-relative.  fe_ws_pack writes
// one image per range; the producers copy a range's image into LDS with 16-byte loads.
struct WsMap {
  int bptr, iptr, diag, bsp, code, row, sa, bytes;
};

__host__ __device__ inline int ws_up16(int v) { return (v + 15) & ~15; }

__host__ __device__ inline WsMap ws_map_layout(const fe_graph& g) {
  WsMap M;
  int o = 0;
  M.bptr = o, o += ws_up16(4 * (g.max_nodes + 1));
  M.iptr = o, o += ws_up16(4 * (g.max_nodes + 1));
  M.diag = o, o += ws_up16(4 * g.max_nodes);
  M.bsp = o, o += ws_up16(4 * (g.max_nbq + 1));
  M.code = o, o += ws_up16(2 * g.max_nseg);
  M.row = o, o += ws_up16(2 * g.max_nbq);
  M.sa = o, o += ws_up16(2 * g.max_inc);
  M.bytes = o;
  return M;
}

// LDS of the wave-specialised form: two slots (the records of one range, then its map image) and two
// staging areas (the node coordinates and the element vertex positions of a range).
struct WsLayout {
  WsMap map;
  int maps, slot;              // byte offset of the map image within a slot, slot size
  int coords, local, stage;    // byte offsets within a staging area, its size
  int bytes;                   // total: two slots, then two staging areas
};

__host__ __device__ inline WsLayout ws_layout(int rs, int dims, const fe_graph& g) {
  WsLayout L;
  L.map = ws_map_layout(g);
  L.maps = ws_up16(rs * 8 * g.max_elem);
  L.slot = L.maps + L.map.bytes;
  int o = 0;
  L.coords = o, o += ws_up16(dims * 8 * g.max_wnode);
  L.local = o, o += ws_up16(8 * g.max_elem);
  L.stage = o;
  L.bytes = 2 * L.slot + 2 * L.stage;
  return L;
}

struct WsHdr {
  int32_t n0, n1, E0, nel, I0, ninc, B0, nbq, c_lo, nseg, N0, nnode;
};

__device__ inline WsHdr ws_hdr(const int32_t* hd) {
  return WsHdr{hd[0], hd[1], hd[2], hd[3], hd[4], hd[5], hd[6], hd[7], hd[8], hd[9], hd[10], hd[11]};
}

// A range fits the map image and the slots (and, with one node and one element per producer thread,
// the producers' registers).
template <int PT>
__device__ inline bool ws_fits(const fe_graph& g, const WsHdr& h) {
  const int32_t nn = h.n1 - h.n0;
  return nn >= 0 && nn <= g.max_nodes && h.nel <= g.max_elem && h.nel <= PT && h.ninc <= g.max_inc &&
         h.nbq <= g.max_nbq && h.nseg <= g.max_nseg && h.nnode <= g.max_wnode && h.nnode <= PT;
}

// One workgroup per range: its map image (the buffer is zeroed beforehand; a range that does not fit
// is left zero and skipped by k_assemble_ws).
__global__ __launch_bounds__(256) void k_ws_pack(fe_graph g, char* __restrict__ out) {
  const WsMap M = ws_map_layout(g);
  const int64_t r = blockIdx.x;
  const WsHdr h = ws_hdr(g.wg_hdr + r * 12);
  if (!ws_fits<1 << 30>(g, h)) return;
  char* b = out + r * M.bytes;
  int32_t* bptr = reinterpret_cast<int32_t*>(b + M.bptr);
  int32_t* iptr = reinterpret_cast<int32_t*>(b + M.iptr);
  int32_t* diag = reinterpret_cast<int32_t*>(b + M.diag);
  int32_t* bsp = reinterpret_cast<int32_t*>(b + M.bsp);
  uint16_t* code = reinterpret_cast<uint16_t*>(b + M.code);
  uint16_t* row = reinterpret_cast<uint16_t*>(b + M.row);
  uint16_t* sa = reinterpret_cast<uint16_t*>(b + M.sa);
  const int32_t nn = h.n1 - h.n0;
  for (int t = threadIdx.x; t <= nn; t += blockDim.x) {
    bptr[t] = g.bptr[h.n0 + t];
    iptr[t] = g.inc_ptr[h.n0 + t] - h.I0;
    if (t < nn) diag[t] = g.bdiag[h.n0 + t];
  }
  for (int t = threadIdx.x; t < h.ninc; t += blockDim.x) sa[t] = g.inc_sa[h.I0 + t];
  for (int t = threadIdx.x; t <= h.nbq; t += blockDim.x) {
    bsp[t] = g.bseg_ptr[h.B0 + t] - h.c_lo;
    if (t < h.nbq) row[t] = g.brow[h.B0 + t];
  }
  for (int t = threadIdx.x; t < h.nseg; t += blockDim.x) code[t] = g.wcode[h.c_lo + t];
}

// Header of range r as one register: lane i < 12 holds word i (a vector load, so that waiting for
// it never holds a barrier; see k_assemble_ws).
__device__ inline int32_t load_hdr_lanes(const fe_graph& g, int64_t r) {
  const int lane = threadIdx.x & 63;
  return lane < 12 ? g.wg_hdr[r * 12 + lane] : 0;
}

__device__ inline WsHdr unpack_hdr(int32_t v) {
  return WsHdr{__builtin_amdgcn_readlane(v, 0), __builtin_amdgcn_readlane(v, 1),
               __builtin_amdgcn_readlane(v, 2), __builtin_amdgcn_readlane(v, 3),
               __builtin_amdgcn_readlane(v, 4), __builtin_amdgcn_readlane(v, 5),
               __builtin_amdgcn_readlane(v, 6), __builtin_amdgcn_readlane(v, 7),
               __builtin_amdgcn_readlane(v, 8), __builtin_amdgcn_readlane(v, 9),
               __builtin_amdgcn_readlane(v, 10), __builtin_amdgcn_readlane(v, 11)};
}

// Phase 2 of the wave-specialised form (the sums of fused_phase2, in the same order): a block's row
// comes from its brow entry and each contribution's record and vertices from its resolved code.
template <int KIND, int WHAT, int LAYOUT, int NT>
__device__ inline void ws_phase2(const fe_graph& g, double* __restrict__ vals, const double* S,
                                 const int32_t* s_bptr, const int32_t* s_iptr, const int32_t* s_diag,
                                 const int32_t* s_bsp, const uint16_t* s_code, const uint16_t* s_row,
                                 const uint16_t* s_sa, int32_t B0, int32_t nbq, int32_t nn, int tid) {
  constexpr int D = KindTraits<KIND>::D;
  constexpr int RS = record_size<KIND, WHAT>();
  constexpr int DL = 4;  // lanes per diagonal block
  const int32_t nbq64 = (nbq + 63) & ~63;
  const int32_t items = nbq64 + nn * DL;
  for (int32_t t = tid; t < items; t += NT) {
    if (t < nbq64) {
      if (t >= nbq) continue;
      const int32_t bq = B0 + t;
      const int rw = s_row[t];
      if (rw & 0x8000) continue;  // the diagonal block
      const int j = rw;
      const int32_t c0 = s_bsp[t], c1 = s_bsp[t + 1];
      double sum[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) sum[k] = 0.0;
      // two contributions per step (their LDS reads overlap), summed in ascending order
      int32_t c = c0;
      for (; c + 1 < c1; c += 2) {
        const int k0 = s_code[c], k1 = s_code[c + 1];
        acc_ab<KIND, WHAT, D>(S + (k0 >> 4) * RS, (k0 >> 2) & 3, k0 & 3, sum);
        acc_ab<KIND, WHAT, D>(S + (k1 >> 4) * RS, (k1 >> 2) & 3, k1 & 3, sum);
      }
      if (c < c1) {
        const int k0 = s_code[c];
        acc_ab<KIND, WHAT, D>(S + (k0 >> 4) * RS, (k0 >> 2) & 3, k0 & 3, sum);
      }
#pragma unroll
      for (int k = 0; k < D * D; ++k) store_block<D, LAYOUT>(vals, g, s_bptr, j, bq, k, sum[k]);
    } else {
      const int u = t - nbq64;
      const int j = u / DL, l = u % DL;
      double sum[D * D];
#pragma unroll
      for (int k = 0; k < D * D; ++k) sum[k] = 0.0;
      for (int32_t k = s_iptr[j] + l; k < s_iptr[j + 1]; k += DL) {
        const int sa = s_sa[k];
        acc_aa<KIND, WHAT, D>(S + (sa >> 2) * RS, sa & 3, sum);
      }
#pragma unroll
      for (int off = DL / 2; off > 0; off >>= 1)
#pragma unroll
        for (int k = 0; k < D * D; ++k) sum[k] += __shfl_xor(sum[k], off, DL);
      const int32_t bq = s_diag[j];
      if (bq >= 0)
#pragma unroll
        for (int k = 0; k < D * D; ++k)
          if (k % DL == l) store_block<D, LAYOUT>(vals, g, s_bptr, j, bq, k, sum[k]);
    }
  }
}

constexpr int kWsImageLoads = 2;  // 16-byte loads of a map image per producer thread

// Wave-specialised persistent form of the fused assembly.  Wavefronts [0, PW) produce and the others
// consume, on two LDS slots: in round a the producers fill slot (a+1)&1 with range a+1 of the
// workgroup while the consumers run phase 2 of range a from slot a&1 and issue its stores; one
// barrier ends a round.
//  * The consumers issue LDS reads and global stores only, so no wait on the vector-memory counter
//    (which counts loads and stores in issue order) ever holds them behind their own stores.
//  * The producers issue no stores.  A range's maps arrive as one contiguous image (fe_ws_pack) in
//    at most kWsImageLoads 16-byte loads per thread; its node coordinates are loaded once per node,
//    coalesced (the node list is ascending), not once per element vertex, and the element records
//    read them from LDS.
//  * The producers' loads run ahead, one stage per round: entering round a they hold range a+1's map
//    image and element materials, range a+2's node coordinates and element ids, range a+3's node ids
//    and the headers of ranges a+1..a+4, all loaded in round a-1 (a wait at the start of a round is
//    for loads issued a round earlier).  Round a writes range a+1's image to its slot and range a+2's
//    coordinates and element vertex positions to a staging area, issues the next stage of every
//    load, and only then computes range a+1's records from the staging area written in round a-1.
//    Headers are vector loads: a scalar load would be waited for at the barrier.
// The ranges of XCD x are the x-th eighth of the schedule, as in k_assemble_pipe.  Writes the same
// bits as the other two forms (the same record and block-sum code on the same values).
template <int KIND, int WHAT, int LAYOUT, int NT, int PW>
__global__ __launch_bounds__(NT) void k_assemble_ws(fe_mesh m, fe_graph g, double* __restrict__ vals) {
  constexpr int RS = record_size<KIND, WHAT>();
  constexpr int RAW = raw_size<KIND, WHAT>();
  constexpr int NPE = KindTraits<KIND>::NPE;
  constexpr int DIM = coord_dims<KIND>();
  constexpr int NM = n_mats<KIND, WHAT>();
  constexpr int PT = PW * 64;  // producer threads
  constexpr int CT = NT - PT;  // consumer threads
  constexpr int QV = kWsImageLoads;
  static_assert(PW > 0 && CT > 0 && CT % 64 == 0, "both roles need whole wavefronts");
  extern __shared__ double S[];     // laid out by ws_layout
  __shared__ int32_t s_meta[2][4];  // per slot: range fits, B0, nbq, node count
  const WsLayout lay = ws_layout(RS, DIM, g);
  const int n4 = lay.map.bytes / 16;  // 16-byte words of a map image
  char* const lds = reinterpret_cast<char*>(S);
  char* const stage0 = lds + 2 * lay.slot;
  const int x = blockIdx.x % 8, L = gridDim.x / 8;
  const int64_t lo = g.n_blk * x / 8, hi = g.n_blk * (x + 1) / 8;
  const int64_t r0 = lo + blockIdx.x / 8;
  if (r0 >= hi) return;                      // uniform over the workgroup
  const int64_t nr = (hi - r0 + L - 1) / L;  // the workgroup's ranges r0, r0 + L, ...
  const int t = threadIdx.x;
  const bool producer = t < PT;  // whole wavefronts
  auto coords_of = [&](int64_t a) { return reinterpret_cast<double*>(stage0 + (a & 1) * lay.stage + lay.coords); };
  auto local_of = [&](int64_t a) { return reinterpret_cast<uint2*>(stage0 + (a & 1) * lay.stage + lay.local); };
  const uint2* wlocal = reinterpret_cast<const uint2*>(g.welem_local);
  const uint4* image = reinterpret_cast<const uint4*>(g.wmaps);
  // producer state entering round a: H2..H5 the headers of ranges a+1..a+4 (moved to H1..H4 at the
  // start of the round, when they have long arrived: a register copy waits for its load); IM the map
  // image and MATn the element materials of range a+1 (moved to MAT likewise); C the coordinates of
  // node t and (ee, el) element t's id and vertex positions of range a+2; W the id of node t of
  // range a+3
  int32_t H1 = 0, H2 = 0, H3 = 0, H4 = 0, H5 = 0;
  uint4 IM0 = make_uint4(0, 0, 0, 0), IM1 = make_uint4(0, 0, 0, 0);
  static_assert(QV == 2, "two image words per producer thread");
  double MAT[NM], MATn[NM], C[DIM];
  int32_t ee = 0, W = 0;
  uint2 el = make_uint2(0, 0);
  if (producer) {
    H2 = load_hdr_lanes(g, r0);
    if (nr > 1) H3 = load_hdr_lanes(g, r0 + L);
    if (nr > 2) H4 = load_hdr_lanes(g, r0 + 2 * L);
    if (nr > 3) H5 = load_hdr_lanes(g, r0 + 3 * L);
    if (t < n4) IM0 = image[r0 * n4 + t];
    if (t + PT < n4) IM1 = image[r0 * n4 + t + PT];
    const WsHdr h1 = unpack_hdr(H2);
    const bool fit1 = ws_fits<PT>(g, h1);
    // range 0: its elements' materials and its node coordinates, staged below
    if (fit1 && t < h1.nel) load_mats<KIND, WHAT>(m, g.welem[h1.E0 + t], MATn);
    double C0[DIM];
    if (fit1 && t < h1.nnode) {
      const int n = g.wnode[h1.N0 + t];
      C0[0] = m.x[n];
      C0[1] = m.y[n];
      if (DIM == 3) C0[DIM - 1] = m.z[n];
    }
    const uint2 el0 = fit1 && t < h1.nel ? wlocal[h1.E0 + t] : make_uint2(0, 0);
    if (nr > 1) {  // range 1: node coordinates and element ids
      const WsHdr h2 = unpack_hdr(H3);
      if (t < h2.nnode) {
        const int n = g.wnode[h2.N0 + t];
        C[0] = m.x[n];
        C[1] = m.y[n];
        if (DIM == 3) C[DIM - 1] = m.z[n];
      }
      if (t < h2.nel) {
        ee = g.welem[h2.E0 + t];
        el = wlocal[h2.E0 + t];
      }
    }
    if (nr > 2) {  // range 2: node ids
      const WsHdr h3 = unpack_hdr(H4);
      if (t < h3.nnode) W = g.wnode[h3.N0 + t];
    }
    if (fit1) {
      if (t < h1.nnode)
#pragma unroll
        for (int d = 0; d < DIM; ++d) coords_of(0)[t * DIM + d] = C0[d];
      if (t < h1.nel) local_of(0)[t] = el0;
    }
  }
  __syncthreads();
  for (int64_t a = -1; a < nr; ++a) {
    if (producer) {
      if (a + 1 < nr) {
        H1 = H2, H2 = H3, H3 = H4, H4 = H5;
#pragma unroll
        for (int i = 0; i < NM; ++i) MAT[i] = MATn[i];
        H5 = a + 5 < nr ? load_hdr_lanes(g, r0 + (a + 5) * L) : 0;
        const int sl = (a + 1) & 1;
        char* sb = lds + sl * lay.slot;
        const WsHdr c = unpack_hdr(H1);
        // a range beyond the slots' or the registers' capacity is left unwritten (the launch checks
        // the schedule's maxima, so this only guards against inconsistent headers)
        const bool ok = ws_fits<PT>(g, c);
        if (t == 0) {
          s_meta[sl][0] = ok;
          s_meta[sl][1] = c.B0;
          s_meta[sl][2] = c.nbq;
          s_meta[sl][3] = c.n1 - c.n0;
        }
        if (ok) {
          uint4* s_img = reinterpret_cast<uint4*>(sb + lay.maps);
          if (t < n4) s_img[t] = IM0;
          if (t + PT < n4) s_img[t + PT] = IM1;
        }
        if (a + 2 < nr) {
          const WsHdr h2 = unpack_hdr(H2);
          const bool fit2 = ws_fits<PT>(g, h2);
          // stage range a+2's node coordinates and element vertex positions for the next round
          if (fit2) {
            if (t < h2.nnode)
#pragma unroll
              for (int d = 0; d < DIM; ++d) coords_of(a + 2)[t * DIM + d] = C[d];
            if (t < h2.nel) local_of(a + 2)[t] = el;
          }
          // the next stage of every load, issued before this range's records are computed
          const int64_t r2 = r0 + (a + 2) * L;
          if (t < n4) IM0 = image[r2 * n4 + t];
          if (t + PT < n4) IM1 = image[r2 * n4 + t + PT];
          if (fit2 && t < h2.nel) load_mats<KIND, WHAT>(m, ee, MATn);
          if (a + 3 < nr) {
            const WsHdr h3 = unpack_hdr(H3);
            if (t < h3.nnode) {
              C[0] = m.x[W];
              C[1] = m.y[W];
              if (DIM == 3) C[DIM - 1] = m.z[W];
            }
            if (t < h3.nel) {
              ee = g.welem[h3.E0 + t];
              el = wlocal[h3.E0 + t];
            }
            if (a + 4 < nr) {
              const WsHdr h4 = unpack_hdr(H4);
              if (t < h4.nnode) W = g.wnode[h4.N0 + t];
            }
          }
        }
        // range a+1's records from the coordinates and vertex positions staged in round a-1
        if (ok && t < c.nel) {
          const double* cs = coords_of(a + 1);
          const uint2 lp = local_of(a + 1)[t];
          double R[RAW];
#pragma unroll
          for (int b = 0; b < NPE; ++b) {
            const uint32_t w2 = b < 2 ? lp.x : lp.y;
            const int loc = (int)((b & 1) ? (w2 >> 16) : (w2 & 0xffff));
#pragma unroll
            for (int d = 0; d < DIM; ++d) R[d * NPE + b] = cs[loc * DIM + d];
          }
#pragma unroll
          for (int i = 0; i < NM; ++i) R[NPE * DIM + i] = MAT[i];
          record_from_raw<KIND, WHAT>(m, R, reinterpret_cast<double*>(sb) + t * RS);
        }
      }
    } else if (a >= 0) {
      const int sl = a & 1;
      const char* sb = lds + sl * lay.slot;
      const char* mb = sb + lay.maps;
      if (s_meta[sl][0])
        ws_phase2<KIND, WHAT, LAYOUT, CT>(
            g, vals, reinterpret_cast<const double*>(sb), reinterpret_cast<const int32_t*>(mb + lay.map.bptr),
            reinterpret_cast<const int32_t*>(mb + lay.map.iptr), reinterpret_cast<const int32_t*>(mb + lay.map.diag),
            reinterpret_cast<const int32_t*>(mb + lay.map.bsp), reinterpret_cast<const uint16_t*>(mb + lay.map.code),
            reinterpret_cast<const uint16_t*>(mb + lay.map.row), reinterpret_cast<const uint16_t*>(mb + lay.map.sa),
            s_meta[sl][1], s_meta[sl][2], s_meta[sl][3], t - PT);
    }
    __syncthreads();
  }
}

template <int KIND, int WHAT, int LAYOUT, int NT, int PW>
int launch_ws(fe_handle* h, const fe_mesh* m, const fe_graph* g, double* vals) {
  constexpr int PT = PW * 64;
  if (!(g->wg_hdr && g->wnode && g->welem_local && g->wmaps && g->max_nodes > 0 && g->max_nbq > 0 &&
        g->max_nseg > 0 && g->max_wnode > 0 && g->n_blk >= 8 && g->max_elem <= PT &&
        g->max_wnode <= PT && ws_map_layout(*g).bytes <= 16 * kWsImageLoads * PT))
    return fe_set_error(FE_ERR_ARG,
                        "fe_assemble_fused: the wave-specialised form needs headers, node lists, map "
                        "images, max_nbq/max_nseg/max_wnode, at least 8 ranges, max_elem and max_wnode "
                        "<= %d and a map image of at most %d bytes", PT, 16 * kWsImageLoads * PT);
  const WsLayout lay = ws_layout(record_size<KIND, WHAT>(), coord_dims<KIND>(), *g);
  const size_t lds = (size_t)lay.bytes;
  if (lds > 160 * 1024)
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: the slots of this schedule need %zu bytes of LDS", lds);
  auto kern = k_assemble_ws<KIND, WHAT, LAYOUT, NT, PW>;
  if (lds > 64 * 1024)
    FE_CHECK_HIP(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  int per_cu = 0;
  FE_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds));
  int64_t nwg = (int64_t)std::max(per_cu, 1) * std::max(h->num_cus, 8);
  nwg = std::min<int64_t>(nwg, g->n_blk) / 8 * 8;
  hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(NT), lds, h->stream, *m, *g, vals);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

}  // namespace

extern "C" int fe_ws_map_bytes(const fe_graph* g, int64_t* bytes) {
  if (!g || !bytes) return fe_set_error(FE_ERR_ARG, "fe_ws_map_bytes: null argument");
  *bytes = ws_map_layout(*g).bytes;
  return FE_OK;
}

extern "C" int fe_ws_pack(fe_handle* h, const fe_graph* g, void* out) {
  if (!h || !g || !out) return fe_set_error(FE_ERR_ARG, "fe_ws_pack: null argument");
  if (!g->wg_hdr || !g->wcode || !g->brow || g->max_nodes <= 0 || g->max_nbq <= 0 || g->max_nseg <= 0)
    return fe_set_error(FE_ERR_ARG, "fe_ws_pack: needs headers, resolved codes, block rows and the maxima");
  if (g->n_blk > 0) {
    hipLaunchKernelGGL(k_ws_pack, dim3((unsigned)g->n_blk), dim3(256), 0, h->stream, *g,
                       static_cast<char*>(out));
    FE_CHECK_LAUNCH();
  }
  return FE_OK;
}

namespace {

template <int KIND, int WHAT, int LAYOUT>
int launch_fused_nt(fe_handle* h, const fe_mesh* m, const fe_graph* g, double* vals, size_t lds) {
  dim3 grid((unsigned)g->n_blk);
  const int nt = g->threads > 0 ? g->threads : 256;
  if (g->producers > 0) {
    if (nt == 1024 && g->producers == 4) return launch_ws<KIND, WHAT, LAYOUT, 1024, 4>(h, m, g, vals);
    if (nt == 1024 && g->producers == 6) return launch_ws<KIND, WHAT, LAYOUT, 1024, 6>(h, m, g, vals);
    if (nt == 1024 && g->producers == 8) return launch_ws<KIND, WHAT, LAYOUT, 1024, 8>(h, m, g, vals);
    if (nt == 512 && g->producers == 2) return launch_ws<KIND, WHAT, LAYOUT, 512, 2>(h, m, g, vals);
    if (nt == 512 && g->producers == 4) return launch_ws<KIND, WHAT, LAYOUT, 512, 4>(h, m, g, vals);
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: producers %d with %d threads is not built "
                        "(1024 threads: 4, 6 or 8; 512 threads: 2 or 4)", g->producers, nt);
  }
  // the persistent form when the schedule fits it (see k_assemble_pipe)
  const bool pipe = g->wg_hdr && g->welem_conn && g->max_nodes > 0 && g->max_nodes < nt &&
                    g->max_inc <= nt && g->max_elem <= nt && g->n_blk >= 8;
#define FE_LAUNCH_NT(NT)                                                                        \
  do {                                                                                          \
    if (pipe) {                                                                                 \
      auto kern = k_assemble_pipe<KIND, WHAT, LAYOUT, NT>;                                      \
      if (lds > 64 * 1024)                                                                      \
        FE_CHECK_HIP(hipFuncSetAttribute((const void*)kern,                                     \
                                         hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds)); \
      int per_cu = 0;                                                                           \
      FE_CHECK_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, NT, lds));       \
      int64_t nwg = (int64_t)std::max(per_cu, 1) * std::max(h->num_cus, 8);                     \
      nwg = std::min<int64_t>(nwg, g->n_blk) / 8 * 8;                                            \
      hipLaunchKernelGGL(kern, dim3((unsigned)nwg), dim3(NT), lds, h->stream, *m, *g, vals);    \
      break;                                                                                    \
    }                                                                                           \
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
  else if (nt == 512)
    FE_LAUNCH_NT(512);
  else if (nt == 1024)
    FE_LAUNCH_NT(1024);
  else
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: threads must be 64, 128, 256, 512 or 1024");
#undef FE_LAUNCH_NT
  FE_CHECK_LAUNCH();
  return FE_OK;
}

template <int KIND>
int launch_fused(fe_handle* h, const fe_mesh* m, const fe_graph* g, int what, int layout,
                 double* vals) {
  if (what == FE_MASS && KIND == FE_KIND_TRI_MAGNETIC)
    return fe_set_error(FE_ERR_UNSUPPORTED, "fe_assemble_fused: magnetic elements have no mass matrix");
  const size_t lds = what == FE_MASS ? fused_lds_bytes<KIND, FE_MASS>(g->max_elem, g->max_inc)
                                     : fused_lds_bytes<KIND, FE_STIFFNESS>(g->max_elem, g->max_inc);
  if (g->producers == 0 && lds > 144 * 1024)  // (the wave-specialised form sizes its own slots)
    return fe_set_error(FE_ERR_ARG, "fe_assemble_fused: schedule needs %zu bytes of LDS", lds);
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
  FE_REQUIRE(g->n_blk >= 0 && g->max_inc >= 0 && g->max_elem >= 0 && g->max_elem <= 16383,
             "fe_assemble_fused: bad schedule");
  if (g->n_blk == 0 || g->nnzb == 0) return FE_OK;
  FE_REQUIRE(m->conn && m->x && m->y && g->bptr && g->inc_ptr && g->bseg_ptr && g->bseg && g->blk &&
                 g->welem_ptr && g->welem && g->inc_sa && g->bdiag,
             "fe_assemble_fused: null mesh or graph array (element schedule from fe_schedule_elements)");
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
