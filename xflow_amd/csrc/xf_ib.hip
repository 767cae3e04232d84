// Immersed-boundary interpolation and spreading on the device
// (TImmersedBoundaryProblem::Interpolate / SpreadForce / Index / Coord /
// OptimalDirichlet, XProblem.cpp:8550-8791, 8931-9046).
//
// Interpolation: one thread per Lagrangian node adds its 7^3 window in the
// reference's i->j->k order, so each node's value is the reference's sum.
// Spreading is a deterministic gather: every (node, window cell) product is
// keyed by its cell, a stable radix sort groups the keys while keeping node
// order inside each cell, and one thread per cell adds its products in node
// order - the order of the reference's serial scatter.  Products with a zero
// weight (r >= 2h, 3 of the 7 cells per axis) are skipped: they add +-0.
#include <hipcub/hipcub.hpp>

#include <cmath>

#include "xf_internal.h"

constexpr int RANGE = 3;  // XProblem.cpp:28

struct IbGeom {
  double h[3];  // Separation[0] of each separator
  int E[3];     // EndIndex of each separator
  int l, m;
};

__device__ inline int ib_index(double coord, int type, int axis, double h, int end_index) {
  int index = (int)floor((coord + (type != axis ? 0.5 * h : 0)) / h);  // XProblem.cpp:8954
  if (index > (end_index - 1) - RANGE) index = (end_index - 1) - RANGE;
  if (index < RANGE) index = RANGE;
  return index;
}

__device__ inline double ib_coord(int index, int type, int axis, double h) {
  return index * h - (type != axis ? 0.5 * h : 0);  // XProblem.cpp:9005
}

__device__ inline double ib_delta(double r, double h) {  // XProblem.cpp:9026-9031
  if (r < 2 * h) return 1.0 / (4 * h) * (1 + cos(M_PI * r / (2 * h)));
  return 0;
}

__global__ void k_ib_interp(IbGeom G, int axis, int nn, const double *__restrict__ xyz, const double *__restrict__ vel,
                            double *__restrict__ U) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nn) return;
  const double X = xyz[3 * q], Y = xyz[3 * q + 1], Z = xyz[3 * q + 2];
  const int b0 = ib_index(X, 1, axis, G.h[0], G.E[0]);
  const int b1 = ib_index(Y, 2, axis, G.h[1], G.E[1]);
  const int b2 = ib_index(Z, 3, axis, G.h[2], G.E[2]);
  double wy[7], wz[7];
  for (int t = 0; t < 7; t++) {
    wy[t] = ib_delta(fabs(Y - ib_coord(b1 - RANGE + t, 2, axis, G.h[1])), G.h[1]);
    wz[t] = ib_delta(fabs(Z - ib_coord(b2 - RANGE + t, 3, axis, G.h[2])), G.h[2]);
  }
  const double ha = G.h[axis - 1];
  const double cb = ha * ha * ha;  // CB(separator->Separation[0]), XProblem.cpp:8782
  double acc = 0;
  for (int i = b0 - RANGE; i <= b0 + RANGE; i++) {
    const double wx = ib_delta(fabs(X - ib_coord(i, 1, axis, G.h[0])), G.h[0]);
    for (int j = 0; j < 7; j++)
      for (int k = 0; k < 7; k++) {
        const int64_t c = ((int64_t)i * G.l + (b1 - RANGE + j)) * G.m + (b2 - RANGE + k);
        acc += (vel[c] * wx * wy[j] * wz[k]) * cb;
      }
  }
  U[q] = acc;
}

// Emits 343 (cell, value) products per node; zero-weight products get the
// key 0xffffffff and sort to the end.
__global__ void k_ib_products(IbGeom G, int axis, int nn, const double *__restrict__ xyz,
                              const double *__restrict__ F, uint32_t *__restrict__ keys, double *__restrict__ vals) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= (int64_t)nn * 343) return;
  const int q = (int)(t / 343), w = (int)(t % 343);
  const int di = w / 49, dj = (w / 7) % 7, dk = w % 7;
  const double X = xyz[3 * q], Y = xyz[3 * q + 1], Z = xyz[3 * q + 2];
  const int i = ib_index(X, 1, axis, G.h[0], G.E[0]) - RANGE + di;
  const int j = ib_index(Y, 2, axis, G.h[1], G.E[1]) - RANGE + dj;
  const int k = ib_index(Z, 3, axis, G.h[2], G.E[2]) - RANGE + dk;
  const double wx = ib_delta(fabs(X - ib_coord(i, 1, axis, G.h[0])), G.h[0]);
  const double wy = ib_delta(fabs(Y - ib_coord(j, 2, axis, G.h[1])), G.h[1]);
  const double wz = ib_delta(fabs(Z - ib_coord(k, 3, axis, G.h[2])), G.h[2]);
  const double v = (F[q] * wx * wy * wz);  // XProblem.cpp:8596
  const bool zero = wx == 0 || wy == 0 || wz == 0;
  keys[t] = zero ? 0xffffffffu : (uint32_t)(((int64_t)i * G.l + j) * G.m + k);
  vals[t] = v;
}

__global__ void k_ib_segsum(int64_t np, const uint32_t *__restrict__ keys, const double *__restrict__ vals,
                            double *__restrict__ force) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (t >= np) return;
  const uint32_t k = keys[t];
  if (k == 0xffffffffu || (t > 0 && keys[t - 1] == k)) return;
  double acc = 0;  // force |= 0.0 (XProblem.cpp:8552)
  for (int64_t u = t; u < np && keys[u] == k; u++) acc += vals[u];
  force[k] = acc;
}

__global__ void k_ib_move(int nn, double *__restrict__ xyz, double *__restrict__ prev, const double *__restrict__ U,
                          int axis, double dt) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nn) return;
  const int a = axis - 1;
  if (prev) prev[3 * q + a] = xyz[3 * q + a];
  xyz[3 * q + a] += U[q] * dt;  // XProblem.cpp:8863-8880
}

static IbGeom geom(xf_grid g) {
  IbGeom G;
  G.h[0] = g->hx[0];
  G.h[1] = g->hy[0];
  G.h[2] = g->hz[0];
  G.E[0] = g->n - 1;
  G.E[1] = g->l - 1;
  G.E[2] = g->m - 1;
  G.l = g->l;
  G.m = g->m;
  return G;
}

static void ib_check(xf_grid g, int axis, int nn) {
  XF_REQUIRE(axis >= 1 && axis <= 3, XF_EINVAL, "Incorrect type of axis");
  XF_REQUIRE(nn >= 0, XF_EINVAL, "negative node count");
  XF_REQUIRE(g->n >= 2 * RANGE + 2 && g->l >= 2 * RANGE + 2 && g->m >= 2 * RANGE + 2, XF_EINVAL,
             "immersed boundary: the grid must have at least 8 nodes per axis (7-point window)");
  XF_REQUIRE(g->ncell < 0xffffffffll, XF_EINVAL, "immersed boundary: grid too large for 32-bit cell keys");
}

extern "C" int xf_ib_interp(xf_ctx c, xf_grid g, int axis, int nn, const double *xyz, const double *vel, double *U) {
  return guarded(c, [&] {
    XF_REQUIRE(g && (nn == 0 || (xyz && vel && U)), XF_EINVAL, "xf_ib_interp: null argument");
    ib_check(g, axis, nn);
    if (nn == 0) return;
    hipLaunchKernelGGL(k_ib_interp, dim3((nn + 63) / 64), dim3(64), 0, c->stream, geom(g), axis, nn, xyz, vel, U);
    check_kernel("k_ib_interp");
  });
}

extern "C" int xf_ib_spread(xf_ctx c, xf_grid g, int axis, int nn, const double *xyz, const double *F,
                            double *force) {
  return guarded(c, [&] {
    XF_REQUIRE(g && force && (nn == 0 || (xyz && F)), XF_EINVAL, "xf_ib_spread: null argument");
    ib_check(g, axis, nn);
    XF_HIP(hipMemsetAsync(force, 0, sizeof(double) * g->ncell, c->stream));
    if (nn == 0) return;
    const int64_t np = (int64_t)nn * 343;
    uint32_t *k0 = (uint32_t *)ws_get(c, WS_IB_K0, sizeof(uint32_t) * np);
    uint32_t *k1 = (uint32_t *)ws_get(c, WS_IB_K1, sizeof(uint32_t) * np);
    double *v0 = (double *)ws_get(c, WS_IB_V0, sizeof(double) * np);
    double *v1 = (double *)ws_get(c, WS_IB_V1, sizeof(double) * np);
    hipLaunchKernelGGL(k_ib_products, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, geom(g), axis, nn,
                       xyz, F, k0, v0);
    check_kernel("k_ib_products");
    size_t tb = 0;
    XF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, (int)np, 0, 32, c->stream));
    void *tmp = ws_get(c, WS_IB_TMP, tb);
    XF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)np, 0, 32, c->stream));
    hipLaunchKernelGGL(k_ib_segsum, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, np, k1, v1, force);
    check_kernel("k_ib_segsum");
  });
}

extern "C" int xf_ib_update_position(xf_ctx c, int axis, int nn, double *xyz, double *xyz_prev, const double *U,
                                     double dt) {
  return guarded(c, [&] {
    XF_REQUIRE(axis >= 1 && axis <= 3, XF_EINVAL, "Incorrect type of axis");
    XF_REQUIRE(nn == 0 || (xyz && U), XF_EINVAL, "xf_ib_update_position: null argument");
    if (nn == 0) return;
    hipLaunchKernelGGL(k_ib_move, dim3((nn + 127) / 128), dim3(128), 0, c->stream, nn, xyz, xyz_prev, U, axis, dt);
    check_kernel("k_ib_move");
  });
}
