// Throughput BiCGStab (XF_BICGSTAB_T): the recurrence of
// TSLAE::SolveByBiConjugateGradientsStab (XSLAE.cpp:667-767) with its
// stopping rule (max |r| <= tol, absolute, tested before each iteration) on
// the reference's operator rows, in five fused passes per iteration instead
// of the parity solver's two dozen:
//   KP  loop test, beta = (rho'/rho)(alpha/omega), p = r + beta (p - omega v)
//   KV  v = A p                                   and the partials of (r^, v)
//   KS  alpha = rho' / (r^, v), s = r - alpha v
//   KT  t = A s                                   and the partials of (t, s), (t, t)
//   KX  omega = (t, s) / (t, t), x += alpha p + omega s, r = s - omega t
//                                                 and the partials of (r^, r), max |r|
// (Forming s at the stencil points inside KT instead cost 819 us per pass at
// 256^3 against 94 + 216 us for KS and KT: the exact rows read their field
// several times per point.)
// Each vector statement keeps the reference's rounding sequence (XSLAE.cpp:
// 723-745).  Reductions are per-block partials that the next pass folds in a
// fixed order in every block (no grid-wide atomics), so all blocks take the
// same scalars and the same loop decision; the loop state lives on the
// device and every pass returns at once after convergence, so the host only
// polls every few iterations.  The sums' order differs from the parity
// solver's, so iteration counts may differ by a few.
#include <algorithm>
#include <cmath>

#include "xf_internal.h"
#include "xf_rows.h"

namespace {

constexpr int kBtBlocks = 1024;  // blocks of every pass (partials per reduction)
enum { BT_PN = 0, BT_AN, BT_WN, BT_PN1, BT_BN, BT_NORM, BT_EPS };
enum { PS_RR = 0, PS_MAX, PS_RV, PS_TS, PS_TT };  // partial slots

struct BtArgs {
  GridView g;
  const double *b;
  double *x, *r, *rh, *p, *v, *s, *t;
  double *part;  // [5][kBtBlocks]
  DevState *st;
  int nb;
  int maxit;
};

__device__ inline double *slot(const BtArgs &A, int k) { return A.part + (size_t)k * kBtBlocks; }

// the block's fold of the nb partials of slot k (every block gets the same bits)
template <bool MAX>
__device__ double fold(const BtArgs &A, int k, double *lds) {
  const double *pp = slot(A, k);
  double v = 0.0;
  for (int q = threadIdx.x; q < A.nb; q += blockDim.x) {
    const double w = pp[q];
    if (MAX) {
      if (w > v || w != w) v = w;
    } else {
      v += w;
    }
  }
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const double o = lds[threadIdx.x + h];
      if (MAX) {
        if (o > lds[threadIdx.x] || o != o) lds[threadIdx.x] = o;
      } else {
        lds[threadIdx.x] += o;
      }
    }
    __syncthreads();
  }
  const double res = lds[0];
  __syncthreads();
  return res;
}

// the block's sum (MAX: max) of per-thread values into partial slot k
template <bool MAX>
__device__ void put(const BtArgs &A, int k, double v, double *lds) {
  lds[threadIdx.x] = v;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
      const double o = lds[threadIdx.x + h];
      if (MAX) {
        if (o > lds[threadIdx.x] || o != o) lds[threadIdx.x] = o;
      } else {
        lds[threadIdx.x] += o;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) slot(A, k)[blockIdx.x] = lds[0];
  __syncthreads();
}

// body(c, i, j, k) for every cell once, grid-stride over the flat index;
// (i, j, k) by 32-bit division where the grid has fewer than 2^32 cells.
template <class F>
__device__ inline void for_cells(const GridView &g, const F &body) {
  const int64_t nc = (int64_t)g.n * g.l * g.m;
  const int64_t step = (int64_t)gridDim.x * blockDim.x;
  if (nc < ((int64_t)1 << 32)) {
    const uint32_t m = (uint32_t)g.m, l = (uint32_t)g.l;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += step) {
      const uint32_t cc = (uint32_t)c, row = cc / m, i = row / l;
      body(c, (int)i, (int)(row - i * l), (int)(cc - row * m));
    }
  } else {
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += step) {
      int i, j, k;
      decompose(c, g.l, g.m, i, j, k);
      body(c, i, j, k);
    }
  }
}

// r = b - A x, r^ = r, p = v = 0, s = t = 0 (XSLAE.cpp:681-703); partials of
// (r^, r) and max |r|
__global__ __launch_bounds__(kBlock) void k_bt_init(BtArgs A) {
  __shared__ double lds[kBlock];
  double rr = 0, mx = 0;
  for_cells(A.g, [&](int64_t c, int i, int j, int k) {
    double rv = 0;
    if (A.g.space[c]) {
      rv = A.b[c] - row_value(A.g, A.x, c, i, j, k);
      rr += rv * rv;
      const double a = fabs(rv);
      if (a > mx || a != a) mx = a;
    }
    A.r[c] = rv;
    A.rh[c] = rv;
    A.p[c] = 0;  // the vectors are 0 off the unknowns, where the rows read them too
    A.v[c] = 0;
    A.s[c] = 0;
    A.t[c] = 0;
  });
  put<false>(A, PS_RR, rr, lds);
  put<true>(A, PS_MAX, mx, lds);
}

// the loop test (XSLAE.cpp:716) and p = r + beta (p - omega v) (XSLAE.cpp:721-727)
__global__ __launch_bounds__(kBlock) void k_bt_p(BtArgs A) {
  __shared__ double lds[kBlock];
  DevState *st = A.st;
  if (st->done) return;
  const double norm = fold<true>(A, PS_MAX, lds);
  const double pn1 = fold<false>(A, PS_RR, lds);
  // st->iter is advanced by k_bt_v, so no block of this pass sees it change
  const int stop = !(norm > st->s[BT_EPS]) ? 1 : (st->iter >= A.maxit ? 2 : 0);
  if (stop) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st->s[BT_NORM] = norm;
      st->done = 1;
      if (stop == 2) st->status = XF_EMAXIT;
      else if (norm != norm) st->status = XF_ENAN;
    }
    return;
  }
  const double bn = (pn1 / st->s[BT_PN]) * (st->s[BT_AN] / st->s[BT_WN]);
  const double wn = st->s[BT_WN];
  for_cells(A.g, [&](int64_t c, int, int, int) {
    if (!A.g.space[c]) return;
    double q = wn * A.v[c];
    q = A.p[c] - q;
    q = bn * q;
    A.p[c] = A.r[c] + q;
  });
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->s[BT_NORM] = norm;
    st->s[BT_PN1] = pn1;
    st->s[BT_BN] = bn;
  }
}

// v = A p (XSLAE.cpp:729); partials of (r^, v)
__global__ __launch_bounds__(kBlock) void k_bt_v(BtArgs A) {
  __shared__ double lds[kBlock];
  if (A.st->done) return;
  double acc = 0;
  for_cells(A.g, [&](int64_t c, int i, int j, int k) {
    if (!A.g.space[c]) return;
    const double vv = row_value(A.g, A.p, c, i, j, k);
    A.v[c] = vv;
    acc += A.rh[c] * vv;
  });
  put<false>(A, PS_RV, acc, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) A.st->iter++;  // XSLAE.cpp:719
}

// alpha = rho' / (r^, v) (XSLAE.cpp:730) and s = r - alpha v (732-733)
__global__ __launch_bounds__(kBlock) void k_bt_s(BtArgs A) {
  __shared__ double lds[kBlock];
  DevState *st = A.st;
  if (st->done) return;
  const double an = st->s[BT_PN1] / fold<false>(A, PS_RV, lds);
  for_cells(A.g, [&](int64_t c, int, int, int) {
    if (!A.g.space[c]) return;
    const double av = an * A.v[c];
    A.s[c] = A.r[c] - av;
  });
  if (blockIdx.x == 0 && threadIdx.x == 0) st->s[BT_AN] = an;
}

// t = A s (XSLAE.cpp:735); partials of (t, s) and (t, t)
__global__ __launch_bounds__(kBlock) void k_bt_t(BtArgs A) {
  __shared__ double lds[kBlock];
  if (A.st->done) return;
  double ts = 0, tt = 0;
  for_cells(A.g, [&](int64_t c, int i, int j, int k) {
    if (!A.g.space[c]) return;
    const double tv = row_value(A.g, A.s, c, i, j, k);
    A.t[c] = tv;
    ts += tv * A.s[c];
    tt += tv * tv;
  });
  put<false>(A, PS_TS, ts, lds);
  put<false>(A, PS_TT, tt, lds);
}

// omega = (t, s) / (t, t) (XSLAE.cpp:736), x += alpha p + omega s (738-741),
// r = s - omega t (743-744); partials of (r^, r) and max |r| (749)
__global__ __launch_bounds__(kBlock) void k_bt_x(BtArgs A) {
  __shared__ double lds[kBlock];
  DevState *st = A.st;
  if (st->done) return;
  const double wn = fold<false>(A, PS_TS, lds) / fold<false>(A, PS_TT, lds);
  const double an = st->s[BT_AN];
  double rr = 0, mx = 0;
  for_cells(A.g, [&](int64_t c, int, int, int) {
    if (!A.g.space[c]) return;
    const double sv = A.s[c];
    const double tmp1 = an * A.p[c];
    const double tmp2 = tmp1 + wn * sv;
    A.x[c] = A.x[c] + tmp2;
    const double wt = wn * A.t[c];
    const double rv = sv - wt;
    A.r[c] = rv;
    rr += A.rh[c] * rv;
    const double a = fabs(rv);
    if (a > mx || a != a) mx = a;
  });
  put<false>(A, PS_RR, rr, lds);
  put<true>(A, PS_MAX, mx, lds);
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->s[BT_WN] = wn;
    st->s[BT_PN] = st->s[BT_PN1];  // pn = pn1 (XSLAE.cpp:746)
  }
}

}  // namespace

void solve_bicgstab_t(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                      xf_solve_stats &s) {
  XF_REQUIRE((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) == 0, XF_EINVAL,
             "operator(): not implemented for a row of this mask (the reference throws here)");
  XF_REQUIRE(o.tol_mode == XF_TOL_ABS, XF_EINVAL, "XF_BICGSTAB_T: the reference's absolute max-norm tolerance");
  const int64_t nc = g->ncell;
  double *w = (double *)ws_get(c, WS_BT_VECS, sizeof(double) * nc * 6);
  double *part = (double *)ws_get(c, WS_BT_PART, sizeof(double) * 5 * kBtBlocks);
  BtArgs A;
  A.g = view(g);
  A.b = b;
  A.x = x;
  A.r = w;
  A.rh = w + nc;
  A.p = w + 2 * nc;
  A.v = w + 3 * nc;
  A.s = w + 4 * nc;
  A.t = w + 5 * nc;
  A.part = part;
  A.st = c->state;
  A.nb = (int)std::max<int64_t>(1, std::min<int64_t>(kBtBlocks, (nc + kBlock - 1) / kBlock));
  A.maxit = o.maxit;
  DevState h{};
  h.s[BT_PN] = 1.0;
  h.s[BT_AN] = 1.0;
  h.s[BT_WN] = 1.0;
  h.s[BT_BN] = 1.0;
  h.s[BT_EPS] = o.tol;
  *c->h_state = h;
  XF_HIP(hipMemcpyAsync(c->state, c->h_state, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  // the partial slots of the first loop test are those of k_bt_init; the
  // other slots are written before they are read
  hipLaunchKernelGGL(k_bt_init, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
  check_kernel("k_bt_init");
  const int every = o.check_every > 0 ? o.check_every : 16;
  for (;;) {
    for (int q = 0; q < every; q++) {
      hipLaunchKernelGGL(k_bt_p, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      hipLaunchKernelGGL(k_bt_v, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      hipLaunchKernelGGL(k_bt_s, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      hipLaunchKernelGGL(k_bt_t, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      hipLaunchKernelGGL(k_bt_x, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
    }
    check_kernel("BiCGStab iteration");
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (c->h_state->done) break;
  }
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[BT_NORM];
}
