// Element-matrix kernels: one thread per element, SoA coordinates and connectivity (coalesced loads),
// entry-major output ke[k * n_elem + e] (coalesced stores).
#include "fe_element_math.h"
#include "fe_internal.h"

namespace {

constexpr int kBlock = 256;

template <int SYM>
__global__ __launch_bounds__(kBlock) void k_ke_tri_elastic(
    const double* __restrict__ x, const double* __restrict__ y, const int32_t* __restrict__ conn,
    const double* __restrict__ E, const double* __restrict__ nu, const double* __restrict__ th,
    int plane_mode, int64_t ne, double* __restrict__ ke) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (e >= ne) return;
  double px[3], py[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = conn[a * ne + e];
    px[a] = x[n];
    py[a] = y[n];
  }
  TriGeom g;
  tri_elastic_geom(px, py, g);
  double diag, off, shear;
  tri_elastic_coeffs(E[e], nu[e], plane_mode, diag, off, shear);
  const double scale = th[e] * g.area;
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const double ga[2] = {g.gx[a], g.gy[a]};
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (SYM && b < a) continue;
      const double gb[2] = {g.gx[b], g.gy[b]};
#pragma unroll
      for (int di = 0; di < 2; ++di)
#pragma unroll
        for (int dj = 0; dj < 2; ++dj) {
          const int r = 2 * a + di, c = 2 * b + dj;
          if (SYM && c < r) continue;
          const int k = SYM ? sym_index(r, c, 6) : r * 6 + c;
          ke[k * ne + e] = scale * bdb_entry<2>(ga, gb, di, dj, diag, off, shear);
        }
    }
  }
}

template <int SYM>
__global__ __launch_bounds__(kBlock) void k_ke_tet_elastic(
    const double* __restrict__ x, const double* __restrict__ y, const double* __restrict__ z,
    const int32_t* __restrict__ conn, const double* __restrict__ E, const double* __restrict__ nu,
    int64_t ne, double* __restrict__ ke) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (e >= ne) return;
  double px[4], py[4], pz[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int n = conn[a * ne + e];
    px[a] = x[n];
    py[a] = y[n];
    pz[a] = z[n];
  }
  TetGeom g;
  tet_geom(px, py, pz, g);
  double diag, off, shear;
  tet_coeffs(E[e], nu[e], diag, off, shear);
#pragma unroll
  for (int a = 0; a < 4; ++a)
#pragma unroll
    for (int b = 0; b < 4; ++b) {
      if (SYM && b < a) continue;
#pragma unroll
      for (int di = 0; di < 3; ++di)
#pragma unroll
        for (int dj = 0; dj < 3; ++dj) {
          const int r = 3 * a + di, c = 3 * b + dj;
          if (SYM && c < r) continue;
          const int k = SYM ? sym_index(r, c, 12) : r * 12 + c;
          ke[k * ne + e] = g.vol * bdb_entry<3>(g.g[a], g.g[b], di, dj, diag, off, shear);
        }
    }
}

template <int SYM>
__global__ __launch_bounds__(kBlock) void k_ke_tri_magnetic(
    const double* __restrict__ x, const double* __restrict__ y, const int32_t* __restrict__ conn,
    const double* __restrict__ mu, int64_t ne, double* __restrict__ ke) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (e >= ne) return;
  double px[3], py[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = conn[a * ne + e];
    px[a] = x[n];
    py[a] = y[n];
  }
  TriGeom g;
  tri_magnetic_geom(px, py, g);
  const double inv_mu = 1 / mu[e];
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if (SYM && b < a) continue;
      const int k = SYM ? sym_index(a, b, 3) : a * 3 + b;
      ke[k * ne + e] = inv_mu * (g.gx[a] * g.gx[b] + g.gy[a] * g.gy[b]) * g.area;
    }
}

// Consistent mass: factor * (1 + delta_ab) on matching components, zero across components
// (elements.py:513-536 with factor rho*A*t/12; 830-857 with factor rho*V/20).
template <int SYM, int NPE, int D>
__device__ inline void store_mass(double factor, int64_t ne, int64_t e, double* __restrict__ me) {
  constexpr int M = NPE * D;
#pragma unroll
  for (int r = 0; r < M; ++r)
#pragma unroll
    for (int c = 0; c < M; ++c) {
      if (SYM && c < r) continue;
      const int k = SYM ? sym_index(r, c, M) : r * M + c;
      double v = 0.0;
      if (r % D == c % D) v = (r == c) ? 2 * factor : factor;
      me[k * ne + e] = v;
    }
}

template <int SYM>
__global__ __launch_bounds__(kBlock) void k_me_tri(const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const int32_t* __restrict__ conn,
                                                   const double* __restrict__ rho,
                                                   const double* __restrict__ th, int64_t ne,
                                                   double* __restrict__ me) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (e >= ne) return;
  double px[3], py[3];
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    const int n = conn[a * ne + e];
    px[a] = x[n];
    py[a] = y[n];
  }
  TriGeom g;
  tri_elastic_geom(px, py, g);
  store_mass<SYM, 3, 2>(rho[e] * g.area * th[e] / 12, ne, e, me);
}

template <int SYM>
__global__ __launch_bounds__(kBlock) void k_me_tet(const double* __restrict__ x,
                                                   const double* __restrict__ y,
                                                   const double* __restrict__ z,
                                                   const int32_t* __restrict__ conn,
                                                   const double* __restrict__ rho, int64_t ne,
                                                   double* __restrict__ me) {
  const int64_t e = blockIdx.x * (int64_t)kBlock + threadIdx.x;
  if (e >= ne) return;
  double px[4], py[4], pz[4];
#pragma unroll
  for (int a = 0; a < 4; ++a) {
    const int n = conn[a * ne + e];
    px[a] = x[n];
    py[a] = y[n];
    pz[a] = z[n];
  }
  TetGeom g;
  tet_geom(px, py, pz, g);
  store_mass<SYM, 4, 3>(rho[e] * g.vol / 20, ne, e, me);
}

inline dim3 grid_for(int64_t n) { return dim3((unsigned)ceil_div(n, kBlock)); }

}  // namespace

#define FE_CHECK_ELEM_ARGS(name, ne, ...)                                                   \
  FE_ENTER(h);                                                                              \
  FE_REQUIRE((ne) >= 0, name ": negative n_elem");                                          \
  FE_REQUIRE((ne) < (int64_t(1) << 31), name ": n_elem too large"); \
  if ((ne) == 0) return FE_OK;                                                              \
  FE_REQUIRE(__VA_ARGS__, name ": null pointer argument");

extern "C" int fe_ke_tri_elastic(fe_handle* h, const double* x, const double* y,
                                 const int32_t* conn, const double* young, const double* nu,
                                 const double* thickness, int plane_mode, int64_t n_elem, int sym,
                                 double* ke) {
  FE_CHECK_ELEM_ARGS("fe_ke_tri_elastic", n_elem, x && y && conn && young && nu && thickness && ke);
  FE_REQUIRE(plane_mode == FE_PLANE_STRAIN || plane_mode == FE_PLANE_STRESS,
             "fe_ke_tri_elastic: plane_mode must be FE_PLANE_STRAIN or FE_PLANE_STRESS");
  if (sym)
    hipLaunchKernelGGL(k_ke_tri_elastic<1>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, conn,
                       young, nu, thickness, plane_mode, n_elem, ke);
  else
    hipLaunchKernelGGL(k_ke_tri_elastic<0>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, conn,
                       young, nu, thickness, plane_mode, n_elem, ke);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_ke_tet_elastic(fe_handle* h, const double* x, const double* y, const double* z,
                                 const int32_t* conn, const double* young, const double* nu,
                                 int64_t n_elem, int sym, double* ke) {
  FE_CHECK_ELEM_ARGS("fe_ke_tet_elastic", n_elem, x && y && z && conn && young && nu && ke);
  if (sym)
    hipLaunchKernelGGL(k_ke_tet_elastic<1>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, z,
                       conn, young, nu, n_elem, ke);
  else
    hipLaunchKernelGGL(k_ke_tet_elastic<0>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, z,
                       conn, young, nu, n_elem, ke);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_ke_tri_magnetic(fe_handle* h, const double* x, const double* y,
                                  const int32_t* conn, const double* mu_total, int64_t n_elem,
                                  int sym, double* ke) {
  FE_CHECK_ELEM_ARGS("fe_ke_tri_magnetic", n_elem, x && y && conn && mu_total && ke);
  if (sym)
    hipLaunchKernelGGL(k_ke_tri_magnetic<1>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y,
                       conn, mu_total, n_elem, ke);
  else
    hipLaunchKernelGGL(k_ke_tri_magnetic<0>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y,
                       conn, mu_total, n_elem, ke);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_me_tri(fe_handle* h, const double* x, const double* y, const int32_t* conn,
                         const double* rho, const double* thickness, int64_t n_elem, int sym,
                         double* me) {
  FE_CHECK_ELEM_ARGS("fe_me_tri", n_elem, x && y && conn && rho && thickness && me);
  if (sym)
    hipLaunchKernelGGL(k_me_tri<1>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, conn, rho,
                       thickness, n_elem, me);
  else
    hipLaunchKernelGGL(k_me_tri<0>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, conn, rho,
                       thickness, n_elem, me);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_me_tet(fe_handle* h, const double* x, const double* y, const double* z,
                         const int32_t* conn, const double* rho, int64_t n_elem, int sym,
                         double* me) {
  FE_CHECK_ELEM_ARGS("fe_me_tet", n_elem, x && y && z && conn && rho && me);
  if (sym)
    hipLaunchKernelGGL(k_me_tet<1>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, z, conn, rho,
                       n_elem, me);
  else
    hipLaunchKernelGGL(k_me_tet<0>, grid_for(n_elem), dim3(kBlock), 0, h->stream, x, y, z, conn, rho,
                       n_elem, me);
  FE_CHECK_LAUNCH();
  return FE_OK;
}
