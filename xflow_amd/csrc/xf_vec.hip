// Reference-row operator and masked vector kernels (parity path).
//
// The operator rows repeat T3dLaplaceOperator::operator() (XOperator.cpp:938-1129)
// and T3dVariableDivergentedOperator::operator() (XOperator.cpp:1901-2233) with
// the reference's operand order; the library is built with -ffp-contract=off
// and f64 division on gfx950 is correctly rounded, so each row is bit-identical
// to the reference's.  Reductions are deterministic (fixed block partition,
// fixed-order combination of partials) but sum in a different order than the
// reference's serial i->j->k loop.
#include <cmath>

#include "xf_rows.h"

__global__ void k_apply_exact(GridView g, const double *__restrict__ x, const double *__restrict__ b,
                              double sign_b, double *__restrict__ y) {
  int64_t nc = (int64_t)g.n * g.l * g.m;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (!g.space[c]) continue;
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    double v = row_value(g, x, c, i, j, k);
    if (sign_b > 0) v = v + b[c];
    else if (sign_b < 0) v = v - b[c];
    y[c] = v;
  }
}

// sum over unknowns of (A x)_c * w_c; result into *out (device).
__global__ void k_apply_dot_exact(GridView g, const double *__restrict__ x, const double *__restrict__ w,
                                  double *partials, unsigned *counter, double *out) {
  __shared__ double lds[16];
  int64_t nc = (int64_t)g.n * g.l * g.m;
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (!g.space[c]) continue;
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    acc += row_value(g, x, c, i, j, k) * w[c];
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

// res may alias s1 or s2 (in-place updates such as result |= result - t*r).
__global__ void k_lincomb(int64_t nc, const uint8_t *__restrict__ space, double p1, const double *s1,
                          double p2, const double *s2, double *res) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (space && !space[c]) continue;  // null space: every entry (Rn vectors)
    double v;
    if (p1 != 0 && p2 != 0) v = p1 * s1[c] + p2 * s2[c];
    else if (p1 != 0) v = p1 * s1[c];
    else if (p2 != 0) v = p2 * s2[c];
    else v = 0;
    res[c] = v;
  }
}

__global__ void k_copy_masked(int64_t nc, const uint8_t *__restrict__ space, const double *s, double *res) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (space[c]) res[c] = s[c];
}

__global__ void k_dot(int64_t nc, const uint8_t *__restrict__ space, const double *__restrict__ a,
                      const double *__restrict__ b, double *partials, unsigned *counter, double *out) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (!space || space[c]) acc += a[c] * b[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

// max |x| with the smallest Rn index among the maxima (null space and rn: an
// Rn vector, entry c has index c + 1) (the reference keeps the
// first strict maximum in i->j->k order, which is the smallest Rn index).
__global__ void k_absmax(int64_t nc, const uint8_t *__restrict__ space, const int32_t *__restrict__ rn,
                         const double *__restrict__ x, double *pv, int32_t *pi, unsigned *counter, double *out,
                         int32_t *out_idx) {
  __shared__ double lv[kBlock];
  __shared__ int32_t li[kBlock];
  __shared__ bool am_last;
  double best = 0;
  int32_t bi = 0x7fffffff;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (space && !space[c]) continue;
    const int32_t rc = rn ? rn[c] : (int32_t)(c + 1);
    double v = fabs(x[c]);
    if (v > best || (v == best && v > 0 && rc < bi)) { best = v; bi = rc; }
  }
  lv[threadIdx.x] = best;
  li[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double v2 = lv[threadIdx.x + s];
      int32_t i2 = li[threadIdx.x + s];
      if (v2 > lv[threadIdx.x] || (v2 == lv[threadIdx.x] && i2 < li[threadIdx.x])) {
        lv[threadIdx.x] = v2;
        li[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = lv[0];
    pi[blockIdx.x] = li[0];
    __threadfence();
    unsigned prev = atomicAdd(counter, 1u);
    am_last = (prev == gridDim.x - 1);
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  // the last block folds the partials: a strided share per thread, then the
  // block (ties to the lowest Rn index, as in the serial scan)
  double b = 0;
  int32_t idx = 0x7fffffff;
  for (unsigned q = threadIdx.x; q < gridDim.x; q += blockDim.x) {
    double v = __hip_atomic_load(&pv[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int32_t iq = __hip_atomic_load(&pi[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (v > b || (v == b && iq < idx)) { b = v; idx = iq; }
  }
  __syncthreads();
  lv[threadIdx.x] = b;
  li[threadIdx.x] = idx;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double v2 = lv[threadIdx.x + s];
      int32_t i2 = li[threadIdx.x + s];
      if (v2 > lv[threadIdx.x] || (v2 == lv[threadIdx.x] && i2 < li[threadIdx.x])) {
        lv[threadIdx.x] = v2;
        li[threadIdx.x] = i2;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  *out = lv[0];
  *out_idx = (lv[0] > 0) ? li[0] : 1;
  *counter = 0;
}

static int reduce_blocks(int64_t nc) {
  int nb = grid_blocks(nc);
  return nb > 2048 ? 2048 : nb;
}
static int stream_blocks(int64_t nc) {
  int nb = grid_blocks(nc);
  return nb > 65536 ? 65536 : nb;
}

void launch_apply_exact(xf_ctx ctx, xf_grid g, const double *x, const double *b, double sign_b, double *y) {
  XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
             "operator(): not implemented for a row of this mask (the reference throws here)");
  hipLaunchKernelGGL(k_apply_exact, dim3(stream_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, view(g), x, b,
                     sign_b, y);
  check_kernel("k_apply_exact");
}

double run_apply_dot_exact(xf_ctx ctx, xf_grid g, const double *x, const double *w, int slot, double *dev_out) {
  XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
             "operator(): not implemented for a row of this mask (the reference throws here)");
  double *out = dev_out ? dev_out : ctx->d_scalar + slot;
  hipLaunchKernelGGL(k_apply_dot_exact, dim3(reduce_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, view(g), x, w,
                     ctx->partials + (size_t)slot * kMaxPartials, ctx->counters + slot, out);
  check_kernel("k_apply_dot_exact");
  if (dev_out) return 0;
  double h;
  XF_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  XF_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

void launch_lincomb(xf_ctx ctx, xf_grid g, double p1, const double *s1, double p2, const double *s2, double *res) {
  hipLaunchKernelGGL(k_lincomb, dim3(stream_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, g->ncell, g->d_space,
                     p1, s1, p2, s2, res);
  check_kernel("k_lincomb");
}

double run_dot(xf_ctx ctx, xf_grid g, const double *a, const double *b, int slot, double *dev_out) {
  double *out = dev_out ? dev_out : ctx->d_scalar + slot;
  hipLaunchKernelGGL(k_dot, dim3(reduce_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, g->ncell, g->d_space, a,
                     b, ctx->partials + (size_t)slot * kMaxPartials, ctx->counters + slot, out);
  check_kernel("k_dot");
  if (dev_out) return 0;
  double h;
  XF_HIP(hipMemcpyAsync(&h, out, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  XF_HIP(hipStreamSynchronize(ctx->stream));
  return h;
}

// --------------------------------------------------------------- C ABI
extern "C" int xf_apply(xf_ctx c, xf_grid g, const double *x, const double *b, double sign_b, double *y) {
  return guarded(c, [&] {
    XF_REQUIRE(g && x && y, XF_EINVAL, "xf_apply: null argument");
    XF_REQUIRE(sign_b == 0 || b, XF_EINVAL, "xf_apply: b is null");
    launch_apply_exact(c, g, x, b, sign_b, y);
  });
}

extern "C" int xf_apply_dot(xf_ctx c, xf_grid g, const double *x, double *out) {
  return guarded(c, [&] {
    XF_REQUIRE(g && x && out, XF_EINVAL, "xf_apply_dot: null argument");
    *out = run_apply_dot_exact(c, g, x, x, 0, nullptr);
  });
}

extern "C" int xf_lincomb(xf_ctx c, xf_grid g, double p1, const double *s1, double p2, const double *s2,
                          double *res) {
  return guarded(c, [&] {
    XF_REQUIRE(g && res, XF_EINVAL, "xf_lincomb: null argument");
    XF_REQUIRE((p1 == 0 || s1) && (p2 == 0 || s2), XF_EINVAL,
               "Incorrect combination: p<>0 but s=NULL.");  // XSpace.cpp:2522-2525
    launch_lincomb(c, g, p1, s1, p2, s2, res);
  });
}

extern "C" int xf_copy(xf_ctx c, xf_grid g, const double *s, double *res) {
  return guarded(c, [&] {
    XF_REQUIRE(g && s && res, XF_EINVAL, "xf_copy: null argument");
    hipLaunchKernelGGL(k_copy_masked, dim3(stream_blocks(g->ncell)), dim3(kBlock), 0, c->stream, g->ncell,
                       g->d_space, s, res);
    check_kernel("k_copy_masked");
  });
}

extern "C" int xf_dot(xf_ctx c, xf_grid g, const double *a, const double *b, double *out) {
  return guarded(c, [&] {
    XF_REQUIRE(g && a && b && out, XF_EINVAL, "xf_dot: null argument");
    *out = run_dot(c, g, a, b, 0, nullptr);
  });
}

extern "C" int xf_absmax(xf_ctx c, xf_grid g, const double *x, double *val, int32_t *rn_index) {
  return guarded(c, [&] {
    XF_REQUIRE(g && x && val, XF_EINVAL, "xf_absmax: null argument");
    int nb = reduce_blocks(g->ncell);
    if (nb > kBlock * 8) nb = kBlock * 8;
    double *pv = c->partials;
    int32_t *pi = (int32_t *)(c->partials + kMaxPartials);
    hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(kBlock), 0, c->stream, g->ncell, g->d_space, g->d_rn, x, pv, pi,
                       c->counters, c->d_scalar, c->d_iscalar);
    check_kernel("k_absmax");
    double v;
    int32_t idx;
    XF_HIP(hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipMemcpyAsync(&idx, c->d_iscalar, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    *val = v;
    if (rn_index) *rn_index = idx;
  });
}

// Rn vectors (TRnRelease1dSpace over an all-Actual T1dNumberMask, the vectors
// of the packed operator): XSpace.cpp 1-D linear combination, dot, abs-max
extern "C" int xf_rn_lincomb(xf_ctx c, int64_t n, double p1, const double *s1, double p2, const double *s2,
                             double *res) {
  return guarded(c, [&] {
    XF_REQUIRE(n >= 0 && res, XF_EINVAL, "xf_rn_lincomb: null argument");
    XF_REQUIRE((p1 == 0 || s1) && (p2 == 0 || s2), XF_EINVAL, "Incorrect combination: p<>0 but s=NULL.");
    if (n == 0) return;
    hipLaunchKernelGGL(k_lincomb, dim3(stream_blocks(n)), dim3(kBlock), 0, c->stream, n, (const uint8_t *)nullptr,
                       p1, s1, p2, s2, res);
    check_kernel("k_lincomb");
  });
}

extern "C" int xf_rn_dot(xf_ctx c, int64_t n, const double *a, const double *b, double *out) {
  return guarded(c, [&] {
    XF_REQUIRE(n > 0 && a && b && out, XF_EINVAL, "xf_rn_dot: null argument");
    hipLaunchKernelGGL(k_dot, dim3(reduce_blocks(n)), dim3(kBlock), 0, c->stream, n, (const uint8_t *)nullptr, a, b,
                       c->partials, c->counters, c->d_scalar);
    check_kernel("k_dot");
    XF_HIP(hipMemcpyAsync(out, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  });
}

extern "C" int xf_rn_absmax(xf_ctx c, int64_t n, const double *x, double *val, int32_t *rn_index) {
  return guarded(c, [&] {
    XF_REQUIRE(n > 0 && x && val, XF_EINVAL, "xf_rn_absmax: null argument");
    XF_REQUIRE(n < ((int64_t)1 << 31), XF_EINVAL, "xf_rn_absmax: more than 2^31 - 1 entries");
    int nb = reduce_blocks(n);
    if (nb > kBlock * 8) nb = kBlock * 8;
    double *pv = c->partials;
    int32_t *pi = (int32_t *)(c->partials + kMaxPartials);
    hipLaunchKernelGGL(k_absmax, dim3(nb), dim3(kBlock), 0, c->stream, n, (const uint8_t *)nullptr,
                       (const int32_t *)nullptr, x, pv, pi, c->counters, c->d_scalar, c->d_iscalar);
    check_kernel("k_absmax");
    double v;
    int32_t idx;
    XF_HIP(hipMemcpyAsync(&v, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipMemcpyAsync(&idx, c->d_iscalar, sizeof(int32_t), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    *val = v;
    if (rn_index) *rn_index = idx;
  });
}

// ------------------------------------------------ done-gated solver variants
__global__ void k_apply_gated(GridView g, const double *__restrict__ x, const double *__restrict__ b, double sign_b,
                              double *__restrict__ y, const DevState *st) {
  if (st->done) return;
  int64_t nc = (int64_t)g.n * g.l * g.m;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (!g.space[c]) continue;
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    double v = row_value(g, x, c, i, j, k);
    if (sign_b > 0) v = v + b[c];
    else if (sign_b < 0) v = v - b[c];
    y[c] = v;
  }
}

__global__ void k_apply_dot_gated(GridView g, const double *__restrict__ x, const double *__restrict__ w,
                                  double *partials, unsigned *counter, double *out, const DevState *st) {
  __shared__ double lds[16];
  if (st->done) return;
  int64_t nc = (int64_t)g.n * g.l * g.m;
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (!g.space[c]) continue;
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    acc += row_value(g, x, c, i, j, k) * w[c];
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

void launch_apply_gated(xf_ctx ctx, xf_grid g, const double *x, const double *b, double sign_b, double *y) {
  XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
             "operator(): not implemented for a row of this mask (the reference throws here)");
  hipLaunchKernelGGL(k_apply_gated, dim3(stream_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, view(g), x, b,
                     sign_b, y, ctx->state);
  check_kernel("k_apply_gated");
}

void launch_apply_dot_gated(xf_ctx ctx, xf_grid g, const double *x, const double *w, double *out, int rslot) {
  XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
             "operator(): not implemented for a row of this mask (the reference throws here)");
  hipLaunchKernelGGL(k_apply_dot_gated, dim3(reduce_blocks(g->ncell)), dim3(kBlock), 0, ctx->stream, view(g), x, w,
                     ctx->partials + (size_t)rslot * kMaxPartials, ctx->counters + rslot, out, ctx->state);
  check_kernel("k_apply_dot_gated");
}
