// Immersed-boundary interpolation and spreading on the device
// (TImmersedBoundaryProblem::Interpolate / SpreadForce / Index / Coord /
// OptimalDirichlet, XProblem.cpp:8550-8791, 8931-9046).
//
// Interpolation: one wave per Lagrangian node; its lanes form the window's
// products, one lane adds them in the reference's i->j->k order, so each
// node's value is the reference's sum.
// Only the cells with a non-zero weight are visited (r < 2h: 4 of the 7 per
// axis, 5 where rounding puts a cell at 2h just inside; always consecutive);
// the others add +-0 to the sum and leave it unchanged.  Spreading is a
// deterministic gather: one pass forms every node's window (first index,
// count and weights per axis) and the box all windows cover; then each block
// owns a 1 x 4 x 64 tile of cells, writes zeros if the tile lies outside
// that box, and otherwise walks the nodes in chunks of 256, keeps (in node
// order) those whose window meets the tile, and lets every cell add the
// products F * wx * wy * wz of the kept nodes that cover it, in node order -
// the order of the reference's serial scatter, so each cell gets the same
// sum.
#include <climits>
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

// One block (one wave) per node: three lanes form the axis windows, every
// lane forms the products of its window cells, one lane adds them in the
// window's i->j->k order.  Slots outside the window hold +0, which leaves the
// sum unchanged (it starts at +0 and so is never -0).
__global__ __launch_bounds__(64) void k_ib_interp(IbGeom G, int axis, int nn, const double *__restrict__ xyz,
                                                  const double *__restrict__ vel, double *__restrict__ U) {
  constexpr int kSlots = kIbPer * kIbPer * kIbPer;
  __shared__ AxisWeights aw[3];
  __shared__ double prod[kSlots];
  const int q = blockIdx.x, t = threadIdx.x;
  if (q >= nn) return;
  if (t < 3) aw[t] = axis_weights(xyz[3 * q + t], t + 1, axis, G.h[t], G.E[t]);
  __syncthreads();
  const double ha = G.h[axis - 1];
  const double cb = ha * ha * ha;  // CB(separator->Separation[0]), XProblem.cpp:8782
  for (int w = t; w < kSlots; w += 64) {
    const int a = w / (kIbPer * kIbPer), b = (w / kIbPer) % kIbPer, e = w % kIbPer;
    double pr = 0;
    if (a < aw[0].count && b < aw[1].count && e < aw[2].count) {
      const int64_t c = ((int64_t)(aw[0].first + a) * G.l + (aw[1].first + b)) * G.m + (aw[2].first + e);
      pr = (vel[c] * aw[0].w[a] * aw[1].w[b] * aw[2].w[e]) * cb;  // XProblem.cpp:8742-8783
    }
    prod[w] = pr;
  }
  __syncthreads();
  if (t != 0) return;
  double acc = 0;
  for (int w = 0; w < kSlots; w++) acc += prod[w];
  U[q] = acc;
}

// One node's window for spreading: per axis the first index with a
// non-zero weight, the count and the weights, and the node's value.
struct IbWin {
  int first[3], count[3];
  double w[3][kIbPer];
  double F;
};

// Windows of the nodes, and per block the box [lo, hi] (per axis) that the
// block's windows cover.
__global__ __launch_bounds__(256) void k_ib_windows(IbGeom G, int axis, int nn, const double *__restrict__ xyz,
                                                    const double *__restrict__ F, IbWin *__restrict__ win,
                                                    int *__restrict__ nbox, int *__restrict__ part) {
  __shared__ int red[6][256];
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  int lo[3] = {INT_MAX, INT_MAX, INT_MAX}, hi[3] = {INT_MIN, INT_MIN, INT_MIN};
  if (q < nn) {
    IbWin wq;
    for (int d = 0; d < 3; d++) {
      const AxisWeights a = axis_weights(xyz[3 * q + d], d + 1, axis, G.h[d], G.E[d]);
      wq.first[d] = a.first;
      wq.count[d] = a.count;
      for (int v = 0; v < kIbPer; v++) wq.w[d][v] = a.w[v];
    }
    wq.F = F[q];
    win[q] = wq;
    if (wq.count[0] > 0 && wq.count[1] > 0 && wq.count[2] > 0)
      for (int d = 0; d < 3; d++) {
        lo[d] = wq.first[d];
        hi[d] = wq.first[d] + wq.count[d] - 1;
      }
    for (int d = 0; d < 3; d++) {  // an empty window gets an empty box
      nbox[6 * q + d] = lo[d];
      nbox[6 * q + 3 + d] = hi[d];
    }
  }
  for (int d = 0; d < 3; d++) {
    red[d][threadIdx.x] = lo[d];
    red[3 + d][threadIdx.x] = hi[d];
  }
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h)
      for (int d = 0; d < 3; d++) {
        red[d][threadIdx.x] = min(red[d][threadIdx.x], red[d][threadIdx.x + h]);
        red[3 + d][threadIdx.x] = max(red[3 + d][threadIdx.x], red[3 + d][threadIdx.x + h]);
      }
    __syncthreads();
  }
  if (threadIdx.x < 6) part[blockIdx.x * 6 + threadIdx.x] = red[threadIdx.x][0];
}

// force on every cell of the grid (zero where no window reaches); one block
// per 1 x 4 x 64 tile (i, j, k), one thread per cell.  Node chunks (the
// blocks of k_ib_windows) whose box misses the tile are skipped whole; in
// the others each node's box is tested before its window is read.
constexpr int kGj = 4, kGk = 64;
__global__ __launch_bounds__(256) void k_ib_gather(int n, int l, int m, int nn, int nparts,
                                                   const IbWin *__restrict__ win, const int *__restrict__ nbox,
                                                   const int *__restrict__ part, double *__restrict__ force) {
  constexpr int kChunksLds = 64;  // chunk boxes kept in LDS (the rest read from memory)
  __shared__ IbWin kept[256];
  __shared__ int box[6], wcount[4], cbox[6 * kChunksLds];
  const int mt = (m + kGk - 1) / kGk, lt = (l + kGj - 1) / kGj;
  const int64_t tl = blockIdx.x;
  const int k0 = (int)(tl % mt) * kGk, j0 = (int)((tl / mt) % lt) * kGj, i = (int)(tl / ((int64_t)mt * lt));
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const int j = j0 + (tid >> 6), k = k0 + (tid & 63);
  const int tlo[3] = {i, j0, k0}, thi[3] = {i, min(j0 + kGj, l) - 1, min(k0 + kGk, m) - 1};
  auto meets = [&](const int *bx) {  // box (lo[3], hi[3]) meets the tile
    bool r = true;
    for (int d = 0; d < 3; d++) r = r && bx[d] <= thi[d] && bx[3 + d] >= tlo[d];
    return r;
  };
  // the box of all windows, folded from the chunks' boxes
  if (tid < 6) box[tid] = tid < 3 ? INT_MAX : INT_MIN;
  __syncthreads();
  for (int t = tid; t < nparts * 6; t += 256) {
    const int d = t % 6, v = part[t];
    if (t < 6 * kChunksLds) cbox[t] = v;
    if (d < 3) atomicMin(&box[d], v);
    else atomicMax(&box[d], v);
  }
  __syncthreads();
  double acc = 0;  // force |= 0.0 (XProblem.cpp:8552)
  if (meets(box)) {
    for (int ch = 0; ch < nparts; ch++) {
      int cb[6];
      for (int d = 0; d < 6; d++) cb[d] = ch < kChunksLds ? cbox[6 * ch + d] : part[6 * ch + d];
      if (!meets(cb)) continue;  // the same for every thread of the block
      const int q = ch * 256 + tid;
      bool hit = false;
      if (q < nn) {
        int nb[6];
        for (int d = 0; d < 6; d++) nb[d] = nbox[6 * q + d];
        hit = meets(nb);
      }
      const unsigned long long mask = __ballot(hit);
      if (lane == 0) wcount[wave] = __popcll(mask);
      __syncthreads();
      int pos = __popcll(mask & ((1ull << lane) - 1)), cnt = 0;
      for (int w = 0; w < 4; w++) {
        if (w < wave) pos += wcount[w];
        cnt += wcount[w];
      }
      if (hit) kept[pos] = win[q];
      __syncthreads();
      for (int e = 0; e < cnt; e++) {  // XProblem.cpp:8583-8599, in node order
        const IbWin &w = kept[e];
        const unsigned di = (unsigned)(i - w.first[0]), dj = (unsigned)(j - w.first[1]),
                       dk = (unsigned)(k - w.first[2]);
        if (di < (unsigned)w.count[0] && dj < (unsigned)w.count[1] && dk < (unsigned)w.count[2])
          acc += (w.F * w.w[0][di] * w.w[1][dj] * w.w[2][dk]);  // XProblem.cpp:8596
      }
      __syncthreads();
    }
  }
  if (j < l && k < m) force[((int64_t)i * l + j) * m + k] = acc;
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
  XF_REQUIRE(g->ncell < 0xffffffffll, XF_EINVAL, "immersed boundary: grid too large (2^32 - 1 cells or more)");
}

extern "C" int xf_ib_interp(xf_ctx c, xf_grid g, int axis, int nn, const double *xyz, const double *vel, double *U) {
  return guarded(c, [&] {
    XF_REQUIRE(g && (nn == 0 || (xyz && vel && U)), XF_EINVAL, "xf_ib_interp: null argument");
    ib_check(g, axis, nn);
    if (nn == 0) return;
    hipLaunchKernelGGL(k_ib_interp, dim3(nn), dim3(64), 0, c->stream, geom(g), axis, nn, xyz, vel, U);
    check_kernel("k_ib_interp");
  });
}

extern "C" int xf_ib_spread(xf_ctx c, xf_grid g, int axis, int nn, const double *xyz, const double *F,
                            double *force) {
  return guarded(c, [&] {
    XF_REQUIRE(g && force && (nn == 0 || (xyz && F)), XF_EINVAL, "xf_ib_spread: null argument");
    ib_check(g, axis, nn);
    const int nparts = (nn + 255) / 256;
    IbWin *win = (IbWin *)ws_get(c, WS_IB_WIN, sizeof(IbWin) * (size_t)std::max(nn, 1));
    int *nbox = (int *)ws_get(c, WS_IB_BOX, sizeof(int) * 6 * (size_t)std::max(nn, 1));
    int *part = (int *)ws_get(c, WS_IB_PART, sizeof(int) * 6 * (size_t)std::max(nparts, 1));
    if (nn > 0) {
      hipLaunchKernelGGL(k_ib_windows, dim3(nparts), dim3(256), 0, c->stream, geom(g), axis, nn, xyz, F, win, nbox,
                         part);
      check_kernel("k_ib_windows");
    }
    const int64_t tiles = (int64_t)g->n * ((g->l + kGj - 1) / kGj) * ((g->m + kGk - 1) / kGk);
    hipLaunchKernelGGL(k_ib_gather, dim3((unsigned)tiles), dim3(256), 0, c->stream, g->n, g->l, g->m, nn, nparts,
                       win, nbox, part, force);
    check_kernel("k_ib_gather");
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
