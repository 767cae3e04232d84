// Immersed-boundary interpolation and spreading on the device
// (TImmersedBoundaryProblem::Interpolate / SpreadForce / Index / Coord /
// OptimalDirichlet, XProblem.cpp:8550-8791, 8931-9046).
//
// Interpolation: one thread per Lagrangian node adds its window in the
// reference's i->j->k order, so each node's value is the reference's sum.
// Only the cells with a non-zero weight are visited (r < 2h: 4 of the 7 per
// axis, 5 where rounding puts a cell at 2h just inside; always consecutive);
// the others add +-0 to the sum and leave it unchanged.  Spreading is a deterministic gather: every (node, non-zero
// window cell) product is keyed by its cell, a stable radix sort over the
// bits a cell index needs groups the keys while keeping node order inside
// each cell, and one thread per cell adds its products in node order - the
// order of the reference's serial scatter.
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

// The non-zero weights of one axis: the window's first index with r < 2h,
// the count and the weights.
constexpr int kIbPer = 5;  // non-zero weights per axis kept (4 of 7, 5 when rounding lets one more in)
struct AxisWeights {
  int first, count;
  double w[kIbPer];
};

__device__ inline AxisWeights axis_weights(double X, int type, int axis, double h, int end_index) {
  const int b = ib_index(X, type, axis, h, end_index);
  AxisWeights a{b - RANGE, 0, {0, 0, 0, 0, 0}};
  bool seen = false;
  for (int t = 0; t < 7; t++) {
    const int i = b - RANGE + t;
    const double wt = ib_delta(fabs(X - ib_coord(i, type, axis, h)), h);
    if (wt == 0) {
      if (seen) break;
      continue;
    }
    if (!seen) {
      a.first = i;
      seen = true;
    }
    if (a.count < kIbPer) a.w[a.count++] = wt;
  }
  return a;
}

__global__ void k_ib_interp(IbGeom G, int axis, int nn, const double *__restrict__ xyz, const double *__restrict__ vel,
                            double *__restrict__ U) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nn) return;
  const AxisWeights ax = axis_weights(xyz[3 * q], 1, axis, G.h[0], G.E[0]);
  const AxisWeights ay = axis_weights(xyz[3 * q + 1], 2, axis, G.h[1], G.E[1]);
  const AxisWeights az = axis_weights(xyz[3 * q + 2], 3, axis, G.h[2], G.E[2]);
  const double ha = G.h[axis - 1];
  const double cb = ha * ha * ha;  // CB(separator->Separation[0]), XProblem.cpp:8782
  double acc = 0;
  for (int a = 0; a < ax.count; a++)
    for (int b = 0; b < ay.count; b++)
      for (int e = 0; e < az.count; e++) {
        const int64_t c = ((int64_t)(ax.first + a) * G.l + (ay.first + b)) * G.m + (az.first + e);
        acc += (vel[c] * ax.w[a] * ay.w[b] * az.w[e]) * cb;  // XProblem.cpp:8742-8783
      }
  U[q] = acc;
}

// 125 (cell, value) slots per node: the products of its non-zero window
// (at most 5 x 5 x 5) in i->j->k order; unused slots get the key 0xffffffff,
// which sorts after every cell.
__global__ void k_ib_products(IbGeom G, int axis, int nn, const double *__restrict__ xyz,
                              const double *__restrict__ F, uint32_t *__restrict__ keys, double *__restrict__ vals) {
  const int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  constexpr int kSlots = kIbPer * kIbPer * kIbPer;
  if (t >= (int64_t)nn * kSlots) return;
  const int q = (int)(t / kSlots), w = (int)(t % kSlots);
  const int di = w / (kIbPer * kIbPer), dj = (w / kIbPer) % kIbPer, dk = w % kIbPer;
  const AxisWeights ax = axis_weights(xyz[3 * q], 1, axis, G.h[0], G.E[0]);
  const AxisWeights ay = axis_weights(xyz[3 * q + 1], 2, axis, G.h[1], G.E[1]);
  const AxisWeights az = axis_weights(xyz[3 * q + 2], 3, axis, G.h[2], G.E[2]);
  if (di >= ax.count || dj >= ay.count || dk >= az.count) {
    keys[t] = 0xffffffffu;
    vals[t] = 0;
    return;
  }
  vals[t] = (F[q] * ax.w[di] * ay.w[dj] * az.w[dk]);  // XProblem.cpp:8596
  keys[t] = (uint32_t)(((int64_t)(ax.first + di) * G.l + (ay.first + dj)) * G.m + (az.first + dk));
}

__global__ void k_ib_zero(int64_t n, double *__restrict__ f) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < n; c += (int64_t)gridDim.x * blockDim.x) f[c] = 0;
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
    // a fill kernel rather than a memset: callers may capture this call in a graph
    hipLaunchKernelGGL(k_ib_zero, dim3((unsigned)std::min<int64_t>((g->ncell + 255) / 256, 8192)), dim3(256), 0,
                       c->stream, g->ncell, force);
    if (nn == 0) return;
    const int64_t np = (int64_t)nn * kIbPer * kIbPer * kIbPer;
    int bits = 1;  // cell keys < 2^bits - 1, so that the unused slots' key sorts after them
    while (((int64_t)1 << bits) <= g->ncell) bits++;
    uint32_t *k0 = (uint32_t *)ws_get(c, WS_IB_K0, sizeof(uint32_t) * np);
    uint32_t *k1 = (uint32_t *)ws_get(c, WS_IB_K1, sizeof(uint32_t) * np);
    double *v0 = (double *)ws_get(c, WS_IB_V0, sizeof(double) * np);
    double *v1 = (double *)ws_get(c, WS_IB_V1, sizeof(double) * np);
    hipLaunchKernelGGL(k_ib_products, dim3((unsigned)((np + 255) / 256)), dim3(256), 0, c->stream, geom(g), axis, nn,
                       xyz, F, k0, v0);
    check_kernel("k_ib_products");
    size_t tb = 0;
    XF_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, k0, k1, v0, v1, (int)np, 0, bits, c->stream));
    void *tmp = ws_get(c, WS_IB_TMP, tb);
    XF_HIP(hipcub::DeviceRadixSort::SortPairs(tmp, tb, k0, k1, v0, v1, (int)np, 0, bits, c->stream));
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
