// Learned block-Jacobi preconditioner: a per-node graph network over the node graph of K.
//
//   features  f_i (8): local stiffness contrast and shape of the diagonal block, fixed-DOF fraction,
//                      degree, neighbourhood heterogeneity (all scale-free)
//   layer l   h^{l+1}_i = relu(W1_l h^l_i + W2_l mean_{j in N(i), j != i} h^l_j + b_l)   (3 layers, 32)
//   head      o_i = Wo h^3_i + bo  (dof values), s_i = exp(clamp(o_i, -4, 4))
//   block     M_i = diag(s_i) K_ii^{-1} diag(s_i)       (SPD by congruence, so PCG stays valid)
//
// Training loss for a residual/error pair (r, e) of a plain PCG run (A e = r), with z = M r:
//   kind 0: L = 1 - (r.z)^2 / ((z.Az)(r.e))
//     the fraction of the error's energy left after one optimally damped preconditioned Richardson
//     step (scale-free in M).  dL/dz = -2 (r.z)/((z.Az)(r.e)) * (r - (r.z)/(z.Az) Az).
//   kind 1: L = |e - z|_A^2 / |e|_A^2 = 1 - 2 (r.z)/(r.e) + (z.Az)/(r.e)
//     the same step undamped: it also sees the top of the spectrum of MA (a step that overshoots
//     rough errors costs energy).  dL/dz = 2 (Az - r)/(r.e).
// Weight gradients are reduced per workgroup in registers/LDS and summed in a fixed order by a second
// kernel: no atomics, bitwise reproducible.
#include "fe_internal.h"

namespace {

constexpr int GW = 32;   // hidden width
constexpr int GF = 8;    // input features
constexpr int kBlock = 256;
constexpr int kNodesPerBlock = kBlock / GW;
constexpr int kRedBlocks = 512;  // fixed grid of the weight-gradient reductions

// parameter offsets (floats)
__host__ __device__ constexpr int layer_fin(int l) { return l == 0 ? GF : GW; }
__host__ __device__ constexpr int layer_size(int l) { return 2 * GW * layer_fin(l) + GW; }
__host__ __device__ constexpr int layer_off(int l) {
  return l == 0 ? 0 : (l == 1 ? layer_size(0) : layer_size(0) + layer_size(1));
}
__host__ __device__ constexpr int head_off() { return layer_size(0) + 2 * layer_size(1); }
__host__ __device__ constexpr int n_params(int d) { return head_off() + d * GW + d; }

template <int D, int LAYOUT>
__device__ inline int64_t gvix(int64_t q, int k, int64_t nnzb) {
  return LAYOUT == FE_LAYOUT_PLANES ? (int64_t)k * nnzb + q : q * D * D + k;
}

// ---- features ---------------------------------------------------------------------------------
template <int D, int LAYOUT>
__global__ void k_gnn_logtr(int64_t nb, int64_t nnzb, const int32_t* __restrict__ bptr,
                            const int32_t* __restrict__ bcol, const double* __restrict__ vals,
                            const uint8_t* __restrict__ fixed, float* __restrict__ feat,
                            double* __restrict__ logtr) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  int64_t qd = -1;
  for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q)
    if (bcol[q] == i) qd = q;
  double blk[D * D];
#pragma unroll
  for (int k = 0; k < D * D; ++k) blk[k] = qd >= 0 ? vals[gvix<D, LAYOUT>(qd, k, nnzb)] : 0.0;
  double tr = 0.0, off = 0.0;
#pragma unroll
  for (int d = 0; d < D; ++d) tr += fabs(blk[d * D + d]);
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b)
      if (a != b) off += blk[a * D + b] * blk[a * D + b];
  const double safe = tr > 0 ? tr : 1.0;
  logtr[i] = log(safe);
  float* f = feat + i * GF;
  int nfix = 0;
#pragma unroll
  for (int d = 0; d < D; ++d) nfix += fixed[i * D + d] ? 1 : 0;
#pragma unroll
  for (int d = 0; d < 3; ++d) f[1 + d] = d < D ? (float)(fabs(blk[d * D + d]) / safe) : 0.f;
  f[4] = (float)(sqrt(off) / safe);
  f[5] = (float)nfix / (float)D;
  f[6] = (float)(bptr[i + 1] - bptr[i]) / 16.f;
}

// Neighbourhood contrast: own log-trace minus the neighbours' mean, and the neighbours' standard
// deviation (accumulated in double on values centred at the node's own log-trace).
__global__ void k_gnn_contrast(int64_t nb, const int32_t* __restrict__ bptr,
                               const int32_t* __restrict__ bcol, const double* __restrict__ logtr,
                               float* __restrict__ feat) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double li = logtr[i];
  double s = 0.0, s2 = 0.0;
  int n = 0;
  for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
    const int32_t j = bcol[q];
    if (j == i) continue;
    const double v = logtr[j] - li;
    s += v;
    s2 += v * v;
    ++n;
  }
  const double m = n ? s / n : 0.0;
  const double var = n ? fmax(s2 / n - m * m, 0.0) : 0.0;
  feat[i * GF + 0] = (float)(n ? -m : 0.0);
  feat[i * GF + 7] = (float)sqrt(var);
}

// ---- forward ------------------------------------------------------------------------------------
// 8 nodes per workgroup, 32 lanes per node: lane k aggregates input channel k over the neighbours,
// then lane c evaluates output channel c.
template <int FIN>
__global__ __launch_bounds__(kBlock) void k_gnn_layer(int64_t nb, const int32_t* __restrict__ bptr,
                                                      const int32_t* __restrict__ bcol,
                                                      const float* __restrict__ hin,
                                                      const float* __restrict__ P,
                                                      float* __restrict__ hout,
                                                      float* __restrict__ agg_out) {
  __shared__ float sW1[GW * FIN], sW2[GW * FIN], sb[GW];
  __shared__ float sh[kNodesPerBlock][FIN], sa[kNodesPerBlock][FIN];
  for (int t = threadIdx.x; t < GW * FIN; t += kBlock) {
    sW1[t] = P[t];
    sW2[t] = P[GW * FIN + t];
  }
  if (threadIdx.x < GW) sb[threadIdx.x] = P[2 * GW * FIN + threadIdx.x];
  const int slot = threadIdx.x / GW, lane = threadIdx.x % GW;
  const int64_t i = (int64_t)blockIdx.x * kNodesPerBlock + slot;
  if (i < nb && lane < FIN) {
    float a = 0.f;
    int n = 0;
    for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
      const int32_t j = bcol[q];
      if (j == i) continue;
      a += hin[(int64_t)j * FIN + lane];
      ++n;
    }
    a = n ? a / (float)n : 0.f;
    sh[slot][lane] = hin[i * FIN + lane];
    sa[slot][lane] = a;
    agg_out[i * FIN + lane] = a;
  }
  __syncthreads();
  if (i >= nb) return;
  float acc = sb[lane];
#pragma unroll
  for (int k = 0; k < FIN; ++k) acc += sW1[lane * FIN + k] * sh[slot][k] + sW2[lane * FIN + k] * sa[slot][k];
  hout[i * GW + lane] = fmaxf(acc, 0.f);
}

// s = exp(clamp(o, -4, 4)); a non-finite o gives NaN (fminf/fmaxf alone would map NaN to -4 and
// hide a broken network behind a valid scaling): the PCG then breaks down and its caller falls back.
__device__ inline double head_scale(float o) {
  return o == o ? exp((double)fminf(fmaxf(o, -4.f), 4.f)) : __builtin_nan("");
}

// Head + learned blocks: o = Wo h + bo, s = exp(clamp(o)), M_i = diag(s) Binv_i diag(s).
template <int D>
__global__ void k_gnn_head(int64_t nb, const float* __restrict__ h, const float* __restrict__ P,
                           const double* __restrict__ binv, float* __restrict__ o,
                           double* __restrict__ M) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const float* Wo = P + head_off();
  const float* bo = Wo + D * GW;
  double s[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    float acc = bo[k];
#pragma unroll 8
    for (int c = 0; c < GW; ++c) acc += Wo[k * GW + c] * h[i * GW + c];
    o[i * D + k] = acc;
    s[k] = head_scale(acc);
  }
#pragma unroll
  for (int a = 0; a < D; ++a)
#pragma unroll
    for (int b = 0; b < D; ++b) M[i * D * D + a * D + b] = s[a] * binv[i * D * D + a * D + b] * s[b];
}

// ---- loss pieces ----------------------------------------------------------------------------------
// z = M r (+ zc, a correction independent of the network, e.g. the coarse part of a two-level
// preconditioner), and y = Binv (s * r) (kept for the gradient).
template <int D>
__global__ void k_gnn_apply(int64_t nb, const double* __restrict__ M, const double* __restrict__ binv,
                            const float* __restrict__ o, const double* __restrict__ r,
                            const double* __restrict__ zc, double* __restrict__ z,
                            double* __restrict__ y) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  double s[D], u[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    s[k] = head_scale(o[i * D + k]);
    u[k] = s[k] * r[i * D + k];
  }
#pragma unroll
  for (int a = 0; a < D; ++a) {
    double zy = 0.0, zz = 0.0;
#pragma unroll
    for (int b = 0; b < D; ++b) {
      zy += binv[i * D * D + a * D + b] * u[b];
      zz += M[i * D * D + a * D + b] * r[i * D + b];
    }
    y[i * D + a] = zy;
    z[i * D + a] = zz + (zc ? zc[i * D + a] : 0.0);
  }
}

__device__ inline double wsum(double v) {
#pragma unroll
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Per-workgroup partial dots (r.z, z.Az, r.e) over a fixed grid.
__global__ __launch_bounds__(kBlock) void k_gnn_dots(int64_t n, const double* __restrict__ r,
                                                     const double* __restrict__ z,
                                                     const double* __restrict__ az,
                                                     const double* __restrict__ e,
                                                     double* __restrict__ partials) {
  __shared__ double red[3][kBlock / 64];
  double a = 0, b = 0, c = 0;
  for (int64_t i = blockIdx.x * (int64_t)kBlock + threadIdx.x; i < n; i += (int64_t)gridDim.x * kBlock) {
    a += r[i] * z[i];
    b += z[i] * az[i];
    c += r[i] * e[i];
  }
  a = wsum(a);
  b = wsum(b);
  c = wsum(c);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    red[0][w] = a;
    red[1][w] = b;
    red[2][w] = c;
  }
  __syncthreads();
  if (threadIdx.x < 3) {
    double s = 0;
    for (int k = 0; k < kBlock / 64; ++k) s += red[threadIdx.x][k];
    partials[threadIdx.x * gridDim.x + blockIdx.x] = s;
  }
}

// Sum the dot partials in fixed order; scal = {r.z, z.Az, r.e, loss}; loss accumulated into *loss_acc.
__global__ void k_gnn_dots_final(int nblocks, const double* __restrict__ partials,
                                 double* __restrict__ scal, double* __restrict__ loss_acc, int kind) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  double v[3] = {0, 0, 0};
  for (int k = 0; k < 3; ++k)
    for (int b = 0; b < nblocks; ++b) v[k] += partials[k * nblocks + b];
  scal[0] = v[0];
  scal[1] = v[1];
  scal[2] = v[2];
  const double loss = !(v[1] > 0 && v[2] > 0) ? 0.0
                      : kind == 0          ? 1.0 - v[0] * v[0] / (v[1] * v[2])
                                           : 1.0 - 2.0 * v[0] / v[2] + v[1] / v[2];
  scal[3] = loss;
  *loss_acc += loss;
}

// dL/do for one pair, accumulated: with g = dL/dz = c_r r + c_a Az,
// dL/do_ik = s_ik [g_ik y_ik + r_ik (Binv_i (s_i * g_i))_k]   (zero where o is clamped).
template <int D>
__global__ void k_gnn_dlo(int64_t nb, const double* __restrict__ scal, const double* __restrict__ r,
                          const double* __restrict__ az, const double* __restrict__ y,
                          const double* __restrict__ binv, const float* __restrict__ o,
                          float* __restrict__ dlo, int kind) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= nb) return;
  const double rz = scal[0], zaz = scal[1], re = scal[2];
  if (!(zaz > 0 && re > 0)) return;
  const double cr = kind == 0 ? -2.0 * rz / (zaz * re) : -2.0 / re;
  const double ca = kind == 0 ? 2.0 * rz * rz / (zaz * zaz * re) : 2.0 / re;
  double s[D], g[D], sg[D];
#pragma unroll
  for (int k = 0; k < D; ++k) {
    s[k] = head_scale(o[i * D + k]);
    g[k] = cr * r[i * D + k] + ca * az[i * D + k];
    sg[k] = s[k] * g[k];
  }
#pragma unroll
  for (int k = 0; k < D; ++k) {
    double bsg = 0.0;
#pragma unroll
    for (int b = 0; b < D; ++b) bsg += binv[i * D * D + k * D + b] * sg[b];
    const float ok = o[i * D + k];
    const double grad = (ok > -4.f && ok < 4.f) ? s[k] * (g[k] * y[i * D + k] + r[i * D + k] * bsg) : 0.0;
    dlo[i * D + k] += (float)grad;
  }
}

// ---- backward -------------------------------------------------------------------------------------
// Head: dWo = sum_i dlo_i h_i^T, dbo = sum_i dlo_i (partials per workgroup), dh_i = Wo^T dlo_i.
template <int D>
__global__ __launch_bounds__(kBlock) void k_gnn_head_bwd(int64_t nb, const float* __restrict__ h,
                                                         const float* __restrict__ P,
                                                         const float* __restrict__ dlo,
                                                         float* __restrict__ dh,
                                                         float* __restrict__ partials) {
  // thread t of the workgroup owns gradient entry t of (Wo, bo) (D*GW + D <= 99 entries)
  constexpr int NG = D * GW + D;
  const float* Wo = P + head_off();
  float acc = 0.f;
  for (int64_t i = blockIdx.x; i < nb; i += gridDim.x) {
    if (threadIdx.x < NG) {
      const int k = threadIdx.x < D * GW ? threadIdx.x / GW : threadIdx.x - D * GW;
      const float gk = dlo[i * D + k];
      acc += threadIdx.x < D * GW ? gk * h[i * GW + threadIdx.x % GW] : gk;
    }
    if (threadIdx.x < GW) {
      float v = 0.f;
#pragma unroll
      for (int k = 0; k < D; ++k) v += Wo[k * GW + threadIdx.x] * dlo[i * D + k];
      dh[i * GW + threadIdx.x] = v;
    }
  }
  if (threadIdx.x < NG) partials[(int64_t)blockIdx.x * NG + threadIdx.x] = acc;
}

// Layer: delta_i = dh_out_i * (h_out_i > 0); dW1 += delta h_in^T, dW2 += delta agg^T, db += delta.
// Grid-stride over nodes; each thread owns one gradient entry and accumulates over the workgroup's
// nodes (fixed assignment -> deterministic).
template <int FIN>
__global__ __launch_bounds__(kBlock) void k_gnn_layer_wgrad(int64_t nb, const float* __restrict__ hin,
                                                            const float* __restrict__ agg,
                                                            const float* __restrict__ hout,
                                                            const float* __restrict__ dhout,
                                                            float* __restrict__ partials) {
  constexpr int NG = 2 * GW * FIN + GW;
  float acc[(NG + kBlock - 1) / kBlock];
#pragma unroll
  for (int u = 0; u < (NG + kBlock - 1) / kBlock; ++u) acc[u] = 0.f;
  for (int64_t i = blockIdx.x; i < nb; i += gridDim.x) {
#pragma unroll
    for (int u = 0; u < (NG + kBlock - 1) / kBlock; ++u) {
      const int t = threadIdx.x + u * kBlock;
      if (t >= NG) break;
      int c, k, which;
      if (t < GW * FIN) {
        which = 0;
        c = t / FIN;
        k = t % FIN;
      } else if (t < 2 * GW * FIN) {
        which = 1;
        c = (t - GW * FIN) / FIN;
        k = (t - GW * FIN) % FIN;
      } else {
        which = 2;
        c = t - 2 * GW * FIN;
        k = 0;
      }
      const float hv = hout[i * GW + c];
      const float delta = hv > 0.f ? dhout[i * GW + c] : 0.f;
      const float x = which == 0 ? hin[i * FIN + k] : (which == 1 ? agg[i * FIN + k] : 1.f);
      acc[u] += delta * x;
    }
  }
#pragma unroll
  for (int u = 0; u < (NG + kBlock - 1) / kBlock; ++u) {
    const int t = threadIdx.x + u * kBlock;
    if (t < NG) partials[(int64_t)blockIdx.x * NG + t] = acc[u];
  }
}

// dh_in_i = W1^T delta_i + sum_{j in N(i), j != i} W2^T delta_j / (deg_j - 1)
template <int FIN>
__global__ __launch_bounds__(kBlock) void k_gnn_layer_dx(int64_t nb, const int32_t* __restrict__ bptr,
                                                         const int32_t* __restrict__ bcol,
                                                         const float* __restrict__ P,
                                                         const float* __restrict__ hout,
                                                         const float* __restrict__ dhout,
                                                         float* __restrict__ dhin) {
  __shared__ float sW1[GW * FIN], sW2[GW * FIN];
  __shared__ float sd[kNodesPerBlock][GW], sn[kNodesPerBlock][GW];
  for (int t = threadIdx.x; t < GW * FIN; t += kBlock) {
    sW1[t] = P[t];
    sW2[t] = P[GW * FIN + t];
  }
  const int slot = threadIdx.x / GW, lane = threadIdx.x % GW;
  const int64_t i = (int64_t)blockIdx.x * kNodesPerBlock + slot;
  if (i < nb) {
    const float hv = hout[i * GW + lane];
    sd[slot][lane] = hv > 0.f ? dhout[i * GW + lane] : 0.f;
    float nsum = 0.f;
    for (int32_t q = bptr[i]; q < bptr[i + 1]; ++q) {
      const int32_t j = bcol[q];
      if (j == i) continue;
      const int32_t degm1 = bptr[j + 1] - bptr[j] - 1;
      const float hj = hout[(int64_t)j * GW + lane];
      const float dj = hj > 0.f ? dhout[(int64_t)j * GW + lane] : 0.f;
      nsum += degm1 > 0 ? dj / (float)degm1 : 0.f;
    }
    sn[slot][lane] = nsum;
  }
  __syncthreads();
  if (i >= nb || lane >= FIN) return;
  float acc = 0.f;
#pragma unroll 8
  for (int c = 0; c < GW; ++c) acc += sW1[c * FIN + lane] * sd[slot][c] + sW2[c * FIN + lane] * sn[slot][c];
  dhin[i * FIN + lane] = acc;
}

// Fixed-order sum of per-workgroup partials into the gradient vector.
__global__ void k_gnn_reduce(int nparts, int ng, const float* __restrict__ partials,
                             float* __restrict__ grad) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ng) return;
  float s = 0.f;
  for (int b = 0; b < nparts; ++b) s += partials[(int64_t)b * ng + t];
  grad[t] += s;
}

__global__ void k_gnn_adam(int n, float* __restrict__ P, const float* __restrict__ grad,
                           float* __restrict__ m, float* __restrict__ v, float lr, float b1, float b2,
                           float eps, float c1, float c2, float scale) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const float g = grad[t] * scale;
  m[t] = b1 * m[t] + (1.f - b1) * g;
  v[t] = b2 * v[t] + (1.f - b2) * g * g;
  P[t] -= lr * (m[t] / c1) / (sqrtf(v[t] / c2) + eps);
}

inline dim3 g1(int64_t n, int b = 256) { return dim3((unsigned)ceil_div(n, b)); }

template <int D>
int features_d(fe_handle* h, const fe_bsr* A, const uint8_t* fixed, float* feat, double* scratch) {
  if (A->layout == FE_LAYOUT_PLANES)
    hipLaunchKernelGGL((k_gnn_logtr<D, FE_LAYOUT_PLANES>), g1(A->n_brows), dim3(256), 0, h->stream,
                       A->n_brows, A->nnzb, A->bptr, A->bcol, (const double*)A->vals, fixed, feat, scratch);
  else
    hipLaunchKernelGGL((k_gnn_logtr<D, FE_LAYOUT_BSR>), g1(A->n_brows), dim3(256), 0, h->stream,
                       A->n_brows, A->nnzb, A->bptr, A->bcol, (const double*)A->vals, fixed, feat, scratch);
  hipLaunchKernelGGL(k_gnn_contrast, g1(A->n_brows), dim3(256), 0, h->stream, A->n_brows, A->bptr,
                     A->bcol, scratch, feat);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

}  // namespace

extern "C" int fe_gnn_num_params(int dof) {
  if (dof < 1 || dof > 3) return fe_set_error(FE_ERR_ARG, "fe_gnn_num_params: dof must be 1..3");
  return n_params(dof);
}

extern "C" int fe_gnn_features(fe_handle* h, const fe_bsr* A, const uint8_t* fixed, float* feat,
                               double* scratch) {
  FE_ENTER(h);
  FE_REQUIRE(A && fixed && feat && scratch, "fe_gnn_features: null argument");
  FE_REQUIRE(A->precision == 0, "fe_gnn_features: FP64 matrix expected");
  FE_REQUIRE(A->row_align == 0, "fe_gnn_features: the matrix's own pattern expected (row_align 0)");
  if (A->n_brows == 0) return FE_OK;
  switch (A->dof) {
    case 1: return features_d<1>(h, A, fixed, feat, scratch);
    case 2: return features_d<2>(h, A, fixed, feat, scratch);
    case 3: return features_d<3>(h, A, fixed, feat, scratch);
    default: return fe_set_error(FE_ERR_ARG, "fe_gnn_features: dof must be 1..3");
  }
}

extern "C" int fe_gnn_forward(fe_handle* h, int64_t n_nodes, int dof, const int32_t* bptr,
                              const int32_t* bcol, const float* params, const float* feat,
                              float* acts, float* aggs, const double* binv, float* o, double* M) {
  FE_ENTER(h);
  FE_REQUIRE(dof >= 1 && dof <= 3, "fe_gnn_forward: dof must be 1..3");
  FE_REQUIRE(n_nodes >= 0, "fe_gnn_forward: negative node count");
  if (n_nodes == 0) return FE_OK;
  FE_REQUIRE(bptr && bcol && params && feat && acts && aggs && binv && o && M,
             "fe_gnn_forward: null argument");
  const dim3 gl((unsigned)ceil_div(n_nodes, kNodesPerBlock));
  float* h1 = acts;
  float* h2 = acts + n_nodes * GW;
  float* h3 = acts + 2 * n_nodes * GW;
  hipLaunchKernelGGL(k_gnn_layer<GF>, gl, dim3(kBlock), 0, h->stream, n_nodes, bptr, bcol, feat,
                     params + layer_off(0), h1, aggs);
  hipLaunchKernelGGL(k_gnn_layer<GW>, gl, dim3(kBlock), 0, h->stream, n_nodes, bptr, bcol, h1,
                     params + layer_off(1), h2, aggs + n_nodes * GF);
  hipLaunchKernelGGL(k_gnn_layer<GW>, gl, dim3(kBlock), 0, h->stream, n_nodes, bptr, bcol, h2,
                     params + layer_off(2), h3, aggs + n_nodes * (GF + GW));
  if (dof == 1)
    hipLaunchKernelGGL(k_gnn_head<1>, g1(n_nodes), dim3(256), 0, h->stream, n_nodes, h3, params, binv, o, M);
  else if (dof == 2)
    hipLaunchKernelGGL(k_gnn_head<2>, g1(n_nodes), dim3(256), 0, h->stream, n_nodes, h3, params, binv, o, M);
  else
    hipLaunchKernelGGL(k_gnn_head<3>, g1(n_nodes), dim3(256), 0, h->stream, n_nodes, h3, params, binv, o, M);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_gnn_loss_grad(fe_handle* h, const fe_bsr* A, const double* binv, const double* M,
                                const float* o, const double* r, const double* e, const double* zc,
                                double* work, double* loss_acc, float* dlo, int kind) {
  FE_ENTER(h);
  FE_REQUIRE(kind == 0 || kind == 1, "fe_gnn_loss_grad: kind must be 0 (damped) or 1 (undamped)");
  FE_REQUIRE(A && binv && M && o && r && e && work && loss_acc && dlo, "fe_gnn_loss_grad: null argument");
  FE_REQUIRE(A->row_align == 0, "fe_gnn_loss_grad: the matrix's own pattern expected (row_align 0)");
  FE_REQUIRE(A->precision == 0 && A->layout == FE_LAYOUT_PLANES,
             "fe_gnn_loss_grad: FP64 FE_LAYOUT_PLANES matrix expected");
  const int64_t nb = A->n_brows, n = nb * A->dof;
  if (nb == 0) return FE_OK;
  double* z = work;
  double* y = work + n;
  double* az = work + 2 * n;
  double* partials = work + 3 * n;        // 3 * kRedBlocks
  double* scal = partials + 3 * kRedBlocks;  // 4
  const int D = A->dof;
  if (D == 1) hipLaunchKernelGGL(k_gnn_apply<1>, g1(nb), dim3(256), 0, h->stream, nb, M, binv, o, r, zc, z, y);
  else if (D == 2) hipLaunchKernelGGL(k_gnn_apply<2>, g1(nb), dim3(256), 0, h->stream, nb, M, binv, o, r, zc, z, y);
  else hipLaunchKernelGGL(k_gnn_apply<3>, g1(nb), dim3(256), 0, h->stream, nb, M, binv, o, r, zc, z, y);
  FE_CHECK_LAUNCH();
  int rc = fe_spmv(h, A, z, az);
  if (rc) return rc;
  hipLaunchKernelGGL(k_gnn_dots, dim3(kRedBlocks), dim3(kBlock), 0, h->stream, n, r, z, az, e, partials);
  hipLaunchKernelGGL(k_gnn_dots_final, dim3(1), dim3(64), 0, h->stream, kRedBlocks, partials, scal, loss_acc,
                     kind);
  if (D == 1) hipLaunchKernelGGL(k_gnn_dlo<1>, g1(nb), dim3(256), 0, h->stream, nb, scal, r, az, y, binv, o, dlo, kind);
  else if (D == 2) hipLaunchKernelGGL(k_gnn_dlo<2>, g1(nb), dim3(256), 0, h->stream, nb, scal, r, az, y, binv, o, dlo, kind);
  else hipLaunchKernelGGL(k_gnn_dlo<3>, g1(nb), dim3(256), 0, h->stream, nb, scal, r, az, y, binv, o, dlo, kind);
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_gnn_loss_workspace(int64_t n_dofs, int64_t* doubles) {
  if (!doubles || n_dofs < 0) return fe_set_error(FE_ERR_ARG, "fe_gnn_loss_workspace: bad argument");
  *doubles = 3 * n_dofs + 3 * kRedBlocks + 4;
  return FE_OK;
}

extern "C" int fe_gnn_backward_workspace(int64_t n_nodes, int64_t* floats) {
  if (!floats || n_nodes < 0) return fe_set_error(FE_ERR_ARG, "fe_gnn_backward_workspace: bad argument");
  // two node-feature buffers of width 32 + per-workgroup gradient partials of the largest layer
  *floats = 2 * n_nodes * GW + (int64_t)kRedBlocks * layer_size(1);
  return FE_OK;
}

extern "C" int fe_gnn_backward(fe_handle* h, int64_t n_nodes, int dof, const int32_t* bptr,
                               const int32_t* bcol, const float* params, const float* feat,
                               const float* acts, const float* aggs, const float* dlo, float* work,
                               float* grad) {
  FE_ENTER(h);
  FE_REQUIRE(dof >= 1 && dof <= 3, "fe_gnn_backward: dof must be 1..3");
  if (n_nodes == 0) return FE_OK;
  FE_REQUIRE(bptr && bcol && params && feat && acts && aggs && dlo && work && grad,
             "fe_gnn_backward: null argument");
  const float* h1 = acts;
  const float* h2 = acts + n_nodes * GW;
  const float* h3 = acts + 2 * n_nodes * GW;
  float* dA = work;
  float* dB = work + n_nodes * GW;
  float* partials = work + 2 * n_nodes * GW;
  const int nred = (int)std::min<int64_t>(kRedBlocks, n_nodes);
  // head
  const int ngh = dof * GW + dof;
  if (dof == 1) hipLaunchKernelGGL(k_gnn_head_bwd<1>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, h3, params, dlo, dA, partials);
  else if (dof == 2) hipLaunchKernelGGL(k_gnn_head_bwd<2>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, h3, params, dlo, dA, partials);
  else hipLaunchKernelGGL(k_gnn_head_bwd<3>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, h3, params, dlo, dA, partials);
  hipLaunchKernelGGL(k_gnn_reduce, g1(ngh), dim3(256), 0, h->stream, nred, ngh, partials, grad + head_off());
  // layer 2 (h2 -> h3), layer 1 (h1 -> h2), layer 0 (feat -> h1)
  const dim3 gl((unsigned)ceil_div(n_nodes, kNodesPerBlock));
  const float* aggs0 = aggs;
  const float* aggs1 = aggs + n_nodes * GF;
  const float* aggs2 = aggs + n_nodes * (GF + GW);
  hipLaunchKernelGGL(k_gnn_layer_wgrad<GW>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, h2, aggs2, h3, dA, partials);
  hipLaunchKernelGGL(k_gnn_reduce, g1(layer_size(2)), dim3(256), 0, h->stream, nred, layer_size(2), partials, grad + layer_off(2));
  hipLaunchKernelGGL(k_gnn_layer_dx<GW>, gl, dim3(kBlock), 0, h->stream, n_nodes, bptr, bcol, params + layer_off(2), h3, dA, dB);
  hipLaunchKernelGGL(k_gnn_layer_wgrad<GW>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, h1, aggs1, h2, dB, partials);
  hipLaunchKernelGGL(k_gnn_reduce, g1(layer_size(1)), dim3(256), 0, h->stream, nred, layer_size(1), partials, grad + layer_off(1));
  hipLaunchKernelGGL(k_gnn_layer_dx<GW>, gl, dim3(kBlock), 0, h->stream, n_nodes, bptr, bcol, params + layer_off(1), h2, dB, dA);
  hipLaunchKernelGGL(k_gnn_layer_wgrad<GF>, dim3(nred), dim3(kBlock), 0, h->stream, n_nodes, feat, aggs0, h1, dA, partials);
  hipLaunchKernelGGL(k_gnn_reduce, g1(layer_size(0)), dim3(256), 0, h->stream, nred, layer_size(0), partials, grad + layer_off(0));
  FE_CHECK_LAUNCH();
  return FE_OK;
}

extern "C" int fe_gnn_adam(fe_handle* h, int n, float* params, const float* grad, float* m, float* v,
                           float lr, float beta1, float beta2, float eps, int step, float grad_scale) {
  FE_ENTER(h);
  FE_REQUIRE(n > 0 && params && grad && m && v && step >= 1, "fe_gnn_adam: bad argument");
  const float c1 = 1.f - powf(beta1, (float)step), c2 = 1.f - powf(beta2, (float)step);
  hipLaunchKernelGGL(k_gnn_adam, g1(n), dim3(256), 0, h->stream, n, params, grad, m, v, lr, beta1, beta2,
                     eps, c1, c2, grad_scale);
  FE_CHECK_LAUNCH();
  return FE_OK;
}
