// Element geometry and material coefficients shared by the element kernels and the fused assembly.
//
// All three stiffness forms are closed forms of the reference's per-element products:
//   tri elasticity  Ke = t*A * B^T D B  (finite_elements/elements.py:395-453, 466-511)
//   tet elasticity  Ke = V * B^T D B    (finite_elements/elements.py:719-796, 809-828)
//   magnetic        Ke_ij = (1/mu)*(b_i b_j + c_i c_j)*A  (finite_elements/elements.py:93-118)
// With g_a the gradient of shape function a, the d x d block (a, b) of B^T D B is
//   lambda * g_a g_b^T + mu * g_b g_a^T + mu * (g_a . g_b) * I
// where, in the reference's D matrices, lambda is the off-diagonal normal entry and mu the shear entry
// (the diagonal normal entry is lambda + 2 mu).  The kernels evaluate exactly those D entries.
#pragma once

#include <hip/hip_runtime.h>

#include "fe_hip.h"

// Triangle elasticity: gradients from the reference's B matrix (elements.py:403-414): with
// det = (x1-x3)(y2-y3) - (y1-y3)(x2-x3), dN/dx = [y2-y3, y3-y1, y1-y2]/det and
// dN/dy = [x3-x2, x1-x3, x2-x1]/det.  Area is the unsigned |(p2-p1) x (p3-p1)| / 2.
struct TriGeom {
  double gx[3], gy[3];
  double area;
};

__device__ __host__ static inline void tri_elastic_geom(const double px[3], const double py[3],
                                                        TriGeom& t) {
  const double det = (px[0] - px[2]) * (py[1] - py[2]) - (py[0] - py[2]) * (px[1] - px[2]);
  const double inv = 1.0 / det;
  t.gx[0] = inv * (py[1] - py[2]);
  t.gx[1] = inv * (py[2] - py[0]);
  t.gx[2] = inv * (py[0] - py[1]);
  t.gy[0] = inv * (px[2] - px[1]);
  t.gy[1] = inv * (px[0] - px[2]);
  t.gy[2] = inv * (px[1] - px[0]);
  const double cr = (px[1] - px[0]) * (py[2] - py[0]) - (px[2] - px[0]) * (py[1] - py[0]);
  t.area = fabs(cr) / 2.0;
}

// Magnetic triangle: gradients are the (b_i, c_i) of N_i = a_i + b_i x + c_i y, the unique linear
// interpolants (columns of the inverse of [[1, x_j, y_j]]), with the signed double area
// (x2-x1)(y3-y1) - (x3-x1)(y2-y1) as the common denominator.
__device__ __host__ static inline void tri_magnetic_geom(const double px[3], const double py[3],
                                                         TriGeom& t) {
  const double det = (px[1] - px[0]) * (py[2] - py[0]) - (px[2] - px[0]) * (py[1] - py[0]);
#pragma unroll
  for (int i = 0; i < 3; ++i) {
    const int j = (i + 1) % 3, k = (i + 2) % 3;
    t.gx[i] = (py[j] - py[k]) / det;
    t.gy[i] = (px[k] - px[j]) / det;
  }
  t.area = fabs(det) / 2.0;
}

// Plane elasticity D entries (elements.py:418-453): diag = c, off = a, shear = b.
__device__ __host__ static inline void tri_elastic_coeffs(double E, double nu, int plane_mode,
                                                          double& diag, double& off, double& shear) {
  double a;
  if (plane_mode == FE_PLANE_STRAIN)
    a = (E * nu) / ((1 + nu) * (1 - 2 * nu));
  else
    a = (E * nu) / (1 - nu * nu);
  const double b = E / (2 * (1 + nu));
  diag = a + 2 * b;
  off = a;
  shear = b;
}

// Tet: gradients of the linear shape functions from the edge vectors e_k = p_k - p_0:
// g_1 = (e2 x e3)/det, g_2 = (e3 x e1)/det, g_3 = (e1 x e2)/det, g_0 = -(g_1 + g_2 + g_3),
// det = e1 . (e2 x e3) = 6 V (signed); V is unsigned.
struct TetGeom {
  double g[4][3];
  double vol;
};

__device__ __host__ static inline void tet_geom(const double px[4], const double py[4],
                                                const double pz[4], TetGeom& t) {
  const double e1[3] = {px[1] - px[0], py[1] - py[0], pz[1] - pz[0]};
  const double e2[3] = {px[2] - px[0], py[2] - py[0], pz[2] - pz[0]};
  const double e3[3] = {px[3] - px[0], py[3] - py[0], pz[3] - pz[0]};
  const double c23[3] = {e2[1] * e3[2] - e2[2] * e3[1], e2[2] * e3[0] - e2[0] * e3[2],
                         e2[0] * e3[1] - e2[1] * e3[0]};
  const double c31[3] = {e3[1] * e1[2] - e3[2] * e1[1], e3[2] * e1[0] - e3[0] * e1[2],
                         e3[0] * e1[1] - e3[1] * e1[0]};
  const double c12[3] = {e1[1] * e2[2] - e1[2] * e2[1], e1[2] * e2[0] - e1[0] * e2[2],
                         e1[0] * e2[1] - e1[1] * e2[0]};
  const double det = e1[0] * c23[0] + e1[1] * c23[1] + e1[2] * c23[2];
  const double inv = 1.0 / det;
#pragma unroll
  for (int k = 0; k < 3; ++k) {
    t.g[1][k] = c23[k] * inv;
    t.g[2][k] = c31[k] * inv;
    t.g[3][k] = c12[k] * inv;
    t.g[0][k] = -(t.g[1][k] + t.g[2][k] + t.g[3][k]);
  }
  t.vol = fabs(det) / 6.0;
}

// 3D isotropic D entries (elements.py:781-794): coeff = E/((1+nu)(1-2nu)),
// diag = coeff*(1-nu), off = coeff*nu, shear = coeff*((1-2nu)/2).
__device__ __host__ static inline void tet_coeffs(double E, double nu, double& diag, double& off,
                                                  double& shear) {
  const double coeff = E / ((1 + nu) * (1 - 2 * nu));
  diag = coeff * (1 - nu);
  off = coeff * nu;
  shear = coeff * ((1 - 2 * nu) / 2);
}

// Entry (di, dj) of the d x d block (a, b) of B^T D B from gradients ga, gb.
template <int D>
__device__ __host__ static inline double bdb_entry(const double* ga, const double* gb, int di, int dj,
                                                   double diag, double off, double shear) {
  if (di == dj) {
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < D; ++k)
      if (k != di) s += ga[k] * gb[k];
    return diag * (ga[di] * gb[di]) + shear * s;
  }
  return off * (ga[di] * gb[dj]) + shear * (ga[dj] * gb[di]);
}
