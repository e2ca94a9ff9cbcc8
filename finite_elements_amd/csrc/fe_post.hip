// Per-element fields of a solution (the step after the solve): one thread per element, SoA
// coordinates and connectivity, outputs component-major out[k * n_elem + e] (coalesced stores).
//
//   elasticity  strain = B u_e (Voigt: tri [exx, eyy, gxy], tet [exx, eyy, ezz, gxy, gyz, gzx]),
//               stress = D strain, energy = 1/2 u_e^T Ke u_e = 1/2 (t A | V) strain . stress
//               (finite_elements/results.py:689-830, elements.py:275-292)
//   magnetic    B = (dA/dy, -dA/dx) = (sum c_i A_i, -sum b_i A_i)  (finite_elements/results.py:121-152)
#include "fe_element_math.h"
#include "fe_internal.h"

namespace {

constexpr int kBlock = 256;

__global__ __launch_bounds__(kBlock) void k_fields_tri_elastic(fe_mesh m, const double* __restrict__ u,
                                                               double* __restrict__ strain,
                                                               double* __restrict__ stress,
                                                               double* __restrict__ energy) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t ne = m.n_elem;
  if (e >= ne) return;
  double px[3], py[3], ux[3], uy[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t n = m.conn[a * ne + e];
    px[a] = m.x[n];
    py[a] = m.y[n];
    ux[a] = u[2 * n];
    uy[a] = u[2 * n + 1];
  }
  TriGeom g;
  tri_elastic_geom(px, py, g);
  double eps[3] = {0.0, 0.0, 0.0};
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    eps[0] += g.gx[a] * ux[a];
    eps[1] += g.gy[a] * uy[a];
    eps[2] += g.gy[a] * ux[a] + g.gx[a] * uy[a];
  }
  double diag, off, shear;
  tri_elastic_coeffs(m.mat0[e], m.mat1[e], m.plane_mode, diag, off, shear);
  const double sig[3] = {diag * eps[0] + off * eps[1], off * eps[0] + diag * eps[1], shear * eps[2]};
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    if (strain) strain[k * ne + e] = eps[k];
    if (stress) stress[k * ne + e] = sig[k];
  }
  if (energy)
    energy[e] = 0.5 * m.mat2[e] * g.area * (eps[0] * sig[0] + eps[1] * sig[1] + eps[2] * sig[2]);
}

__global__ __launch_bounds__(kBlock) void k_fields_tet_elastic(fe_mesh m, const double* __restrict__ u,
                                                               double* __restrict__ strain,
                                                               double* __restrict__ stress,
                                                               double* __restrict__ energy) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t ne = m.n_elem;
  if (e >= ne) return;
  double px[4], py[4], pz[4], uv[4][3];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int64_t n = m.conn[a * ne + e];
    px[a] = m.x[n];
    py[a] = m.y[n];
    pz[a] = m.z[n];
#pragma unroll
    for (int c = 0; c < 3; ++c) uv[a][c] = u[3 * n + c];
  }
  TetGeom g;
  tet_geom(px, py, pz, g);
  // displacement gradient H_ij = du_i/dx_j
  double H[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) {
      double s = 0.0;
#pragma unroll
      for (int a = 0; a < 4; ++a) s += g.g[a][j] * uv[a][i];
      H[i][j] = s;
    }
  const double eps[6] = {H[0][0], H[1][1], H[2][2], H[0][1] + H[1][0], H[1][2] + H[2][1],
                         H[2][0] + H[0][2]};
  double diag, off, shear;
  tet_coeffs(m.mat0[e], m.mat1[e], diag, off, shear);
  const double sig[6] = {diag * eps[0] + off * (eps[1] + eps[2]), diag * eps[1] + off * (eps[0] + eps[2]),
                         diag * eps[2] + off * (eps[0] + eps[1]), shear * eps[3], shear * eps[4],
                         shear * eps[5]};
  double w = 0.0;
#pragma unroll
  for (int k = 0; k < 6; ++k) {
    w += eps[k] * sig[k];
    if (strain) strain[k * ne + e] = eps[k];
    if (stress) stress[k * ne + e] = sig[k];
  }
  if (energy) energy[e] = 0.5 * g.vol * w;
}

__global__ __launch_bounds__(kBlock) void k_fields_magnetic(fe_mesh m, const double* __restrict__ a_pot,
                                                            double* __restrict__ bfield) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  const int64_t ne = m.n_elem;
  if (e >= ne) return;
  double px[3], py[3], av[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int64_t n = m.conn[a * ne + e];
    px[a] = m.x[n];
    py[a] = m.y[n];
    av[a] = a_pot[n];
  }
  TriGeom g;
  tri_magnetic_geom(px, py, g);
  double bx = 0.0, by = 0.0;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    bx += g.gy[a] * av[a];
    by -= g.gx[a] * av[a];
  }
  bfield[e] = bx;
  bfield[ne + e] = by;
}

}  // namespace

extern "C" int fe_element_fields(fe_handle* h, const fe_mesh* m, const double* u, double* strain,
                                 double* stress, double* energy) {
  FE_ENTER(h);
  FE_REQUIRE(m && u, "fe_element_fields: null argument");
  FE_REQUIRE(m->n_elem >= 0, "fe_element_fields: negative element count");
  if (m->n_elem == 0) return FE_OK;
  FE_REQUIRE(m->conn && m->x && m->y, "fe_element_fields: null mesh array");
  const dim3 grid((unsigned)ceil_div(m->n_elem, kBlock)), block(kBlock);
  switch (m->kind) {
    case FE_KIND_TRI_ELASTIC:
      FE_REQUIRE(m->mat0 && m->mat1 && m->mat2, "fe_element_fields: missing material arrays");
      FE_REQUIRE(m->plane_mode == FE_PLANE_STRAIN || m->plane_mode == FE_PLANE_STRESS,
                 "fe_element_fields: bad plane_mode");
      FE_REQUIRE(strain || stress || energy, "fe_element_fields: no output requested");
      hipLaunchKernelGGL(k_fields_tri_elastic, grid, block, 0, h->stream, *m, u, strain, stress, energy);
      break;
    case FE_KIND_TET_ELASTIC:
      FE_REQUIRE(m->z && m->mat0 && m->mat1, "fe_element_fields: missing tet arrays");
      FE_REQUIRE(strain || stress || energy, "fe_element_fields: no output requested");
      hipLaunchKernelGGL(k_fields_tet_elastic, grid, block, 0, h->stream, *m, u, strain, stress, energy);
      break;
    case FE_KIND_TRI_MAGNETIC:
      FE_REQUIRE(strain && !stress && !energy,
                 "fe_element_fields: magnetic elements give B in 'strain' only");
      hipLaunchKernelGGL(k_fields_magnetic, grid, block, 0, h->stream, *m, u, strain);
      break;
    default:
      return fe_set_error(FE_ERR_ARG, "fe_element_fields: unknown element kind %d", m->kind);
  }
  FE_CHECK_LAUNCH();
  return FE_OK;
}
