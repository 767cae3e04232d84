// Throughput conjugate gradients for the masked 7-point pressure operator.
//
// Hestenes-Stiefel CG with two fused kernels per iteration on zero-padded
// vectors (entries off the unknowns are 0 in r, p, q):
//   K1  p' = r + beta p (on the fly, double-buffered), q = A p', (p', q)
//       -> alpha = (r, r) / (p', q)                  [2.5-D marching stencil]
//   K2  x += alpha p', r -= alpha q, (r, r)
//       -> beta, convergence flag                    [streaming]
// Algorithmic traffic per cell and iteration: K1 reads r, p, the flag byte and
// writes p', q; K2 reads x, p', r, q and writes x, r: 33 + 48 = 81 B in FP64,
// 41 B in FP32 (x stays FP64 in the mixed-precision outer loop).
//
// The stencil is the reference's Actual-row formula (XOperator.cpp:971-1019)
// with the step ratios folded into per-node coefficients
//   c-[i] = 1/(h[i-1] (h[i-1]+h[i])/2),  c+[i] = 1/(h[i] (h[i-1]+h[i])/2)
// so that A f = -sum_axes (c+ (f[+1]-f) - c- (f - f[-1])) with a face term
// dropped where the neighbour is Fictive.  For power-of-two steps this is
// bit-identical to the reference's divisions; otherwise it differs in
// rounding only.  It applies to grids whose unknown rows are all Actual
// (Neumann-by-mask and Dirichlet-by-Defined-point masks); other masks use the
// reference-semantics solvers.
//
// Reductions are deterministic: a fixed block partition, block sums in wave
// order, and the last block adds the partials in block order.
#include <cmath>

#include "xf_internal.h"

constexpr int TK = 64;   // k extent of a tile (one wavefront)
constexpr int TJ = 4;    // j extent of a tile (waves per block)
constexpr int IC = 32;   // i planes marched per block
constexpr int kStreamBlocks = 4096;

struct CgWork {
  int64_t n = 0;
  double *r = nullptr, *p[2] = {nullptr, nullptr}, *q = nullptr, *tmp = nullptr;
  float *r32 = nullptr, *p32[2] = {nullptr, nullptr}, *q32 = nullptr, *e32 = nullptr;
};

void free_cg_work(CgWork *w) {
  if (!w) return;
  hipFree(w->r); hipFree(w->p[0]); hipFree(w->p[1]); hipFree(w->q); hipFree(w->tmp);
  hipFree(w->r32); hipFree(w->p32[0]); hipFree(w->p32[1]); hipFree(w->q32); hipFree(w->e32);
  delete w;
}

static CgWork *work(xf_ctx c, xf_grid g, bool fp32) {
  if (!g->work) {
    g->work = new CgWork();
    g->work->n = g->ncell;
  }
  CgWork *w = g->work;
  size_t n = (size_t)g->ncell;
  auto get = [&](auto **p, size_t bytes) {
    if (!*p) {
      hipError_t e = hipMalloc((void **)p, bytes);
      XF_REQUIRE(e == hipSuccess, XF_ENOMEM, "solver workspace: out of device memory");
      XF_HIP(hipMemsetAsync(*p, 0, bytes, c->stream));
    }
  };
  get(&w->r, n * 8); get(&w->p[0], n * 8); get(&w->p[1], n * 8); get(&w->q, n * 8);
  if (fp32) {
    get(&w->r32, n * 4); get(&w->p32[0], n * 4); get(&w->p32[1], n * 4); get(&w->q32, n * 4); get(&w->e32, n * 4);
  }
  return w;
}

// State slots of the throughput solvers.
enum { HS_RR = 0, HS_PQ, HS_ALPHA, HS_BETA, HS_STOP2, HS_RR0, HS_SUM, HS_CNT, HS_AUX };

struct Tiles {
  int kt, jt, it, total;
  int blocks;  // 8 * ceil(total / 8): the XCD-aware mapping is a bijection on [0, blocks)
};
static Tiles tiles(xf_grid g) {
  Tiles t;
  t.kt = (g->m - 2 + TK - 1) / TK;
  t.jt = (g->l - 2 + TJ - 1) / TJ;
  t.it = (g->n - 2 + IC - 1) / IC;
  t.total = t.kt * t.jt * t.it;
  t.blocks = 8 * ((t.total + 7) / 8);
  return t;
}

// K1: optional p' = r + beta p_old, q = A p', deterministic (p', q).
// One wavefront per (j) row of 64 k-cells; each block marches IC planes in i,
// keeping p'(i-1), p'(i), p'(i+1) of its own column in registers; the four
// in-plane neighbours come from the L1/L2-resident rows of the same plane.
// Block index -> tile mapping is XCD-aware: the hardware deals workgroups to
// the 8 XCDs round-robin, so consecutive tiles (which share halo rows) are
// given to workgroups on the same XCD.
template <typename T, bool FUSE>
__global__ __launch_bounds__(256) void k_stencil_cg(int n, int l, int m, Tiles tl, const uint8_t *__restrict__ flags,
                                                     const T *__restrict__ cx, const T *__restrict__ cy,
                                                     const T *__restrict__ cz, const T *__restrict__ r,
                                                     const T *__restrict__ pold, T *__restrict__ pnew,
                                                     T *__restrict__ q, DevState *st, double *partials,
                                                     unsigned *counter) {
  __shared__ double lds[16];
  if (st->done) return;
  const int per = (tl.total + 7) / 8;
  const int w = blockIdx.x;
  const int t = (w & 7) * per + (w >> 3);
  double acc = 0;
  if (t < tl.total) {
    const int ktile = t % tl.kt;
    const int jtile = (t / tl.kt) % tl.jt;
    const int itile = t / (tl.kt * tl.jt);
    const int k = 1 + ktile * TK + (threadIdx.x & 63);
    const int j = 1 + jtile * TJ + (threadIdx.x >> 6);
    const int i0 = 1 + itile * IC;
    const int i1 = min(i0 + IC, n - 1);
    if (k <= m - 2 && j <= l - 2) {
      const int64_t sI = (int64_t)l * m, sJ = m;
      const T beta = FUSE ? (T)st->s[HS_BETA] : T(0);
      auto P = [&](int64_t c) -> T {
        if (FUSE) return r[c] + beta * pold[c];
        return pold[c];
      };
      const T czm = cz[2 * k], czp = cz[2 * k + 1];
      const T cym = cy[2 * j], cyp = cy[2 * j + 1];
      int64_t c = ((int64_t)i0 * l + j) * m + k;
      T pm = P(c - sI), pc = P(c);
      for (int i = i0; i < i1; i++, c += sI) {
        const T pp = P(c + sI);
        const uint8_t fl = flags[c];
        T v = 0;
        if (fl & FL_ACTUAL) {
          const T cxm = cx[2 * i], cxp = cx[2 * i + 1];
          const T pjm = P(c - sJ), pjp = P(c + sJ), pkm = P(c - 1), pkp = P(c + 1);
          T tx = ((fl & FL_IP) ? cxp * (pp - pc) : T(0)) - ((fl & FL_IM) ? cxm * (pc - pm) : T(0));
          T ty = ((fl & FL_JP) ? cyp * (pjp - pc) : T(0)) - ((fl & FL_JM) ? cym * (pc - pjm) : T(0));
          T tz = ((fl & FL_KP) ? czp * (pkp - pc) : T(0)) - ((fl & FL_KM) ? czm * (pc - pkm) : T(0));
          v = -(tx + ty + tz);
          acc += (double)pc * (double)v;
        }
        q[c] = v;
        if (FUSE) pnew[c] = pc;
        pm = pc;
        pc = pp;
      }
    }
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) {
    st->s[HS_PQ] = tot;
    if (tot == 0 || tot != tot) {
      st->done = 1;
      st->status = (tot != tot) ? XF_ENAN : st->status;
    } else {
      st->s[HS_ALPHA] = st->s[HS_RR] / tot;
    }
  }
}

// K2: x += alpha p, r -= alpha q, (r, r) -> beta and the stopping test.
template <typename T, typename X>
__global__ __launch_bounds__(256) void k_cg_update(int64_t nc, X *__restrict__ x, const T *__restrict__ p,
                                                    T *__restrict__ r, const T *__restrict__ q, DevState *st,
                                                    double *partials, unsigned *counter, int maxit) {
  __shared__ double lds[16];
  if (st->done) return;
  const double alpha = st->s[HS_ALPHA];
  const T a = (T)alpha;
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const T pc = p[c];
    const T rn = r[c] - a * q[c];
    r[c] = rn;
    x[c] = x[c] + (X)a * (X)pc;
    acc += (double)rn * (double)rn;
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) {
    const double rr = st->s[HS_RR];
    st->s[HS_BETA] = tot / rr;
    st->s[HS_RR] = tot;
    st->iter += 1;
    if (tot != tot) {
      st->done = 1;
      st->status = XF_ENAN;
    } else if (tot <= st->s[HS_STOP2]) {
      st->done = 1;
    } else if (st->iter >= maxit) {
      st->done = 1;
      st->status = XF_EMAXIT;
    }
  }
}

// r = b - A x over the interior with the fast rows (x read everywhere, so
// boundary values of Defined neighbours enter as in the reference), written
// as type R; (r, r) and sum(r) over unknowns into slots.
template <typename R>
__global__ __launch_bounds__(256) void k_residual_fast(int n, int l, int m, Tiles tl,
                                                        const uint8_t *__restrict__ flags,
                                                        const double *__restrict__ cx, const double *__restrict__ cy,
                                                        const double *__restrict__ cz, const double *__restrict__ b,
                                                        const double *__restrict__ x, R *__restrict__ rout,
                                                        DevState *st, double *partials, unsigned *counter,
                                                        int slot_rr) {
  __shared__ double lds[16];
  const int per = (tl.total + 7) / 8;
  const int w = blockIdx.x;
  const int t = (w & 7) * per + (w >> 3);
  double acc = 0;
  if (t < tl.total) {
    const int ktile = t % tl.kt;
    const int jtile = (t / tl.kt) % tl.jt;
    const int itile = t / (tl.kt * tl.jt);
    const int k = 1 + ktile * TK + (threadIdx.x & 63);
    const int j = 1 + jtile * TJ + (threadIdx.x >> 6);
    const int i0 = 1 + itile * IC;
    const int i1 = min(i0 + IC, n - 1);
    if (k <= m - 2 && j <= l - 2) {
      const int64_t sI = (int64_t)l * m, sJ = m;
      const double czm = cz[2 * k], czp = cz[2 * k + 1];
      const double cym = cy[2 * j], cyp = cy[2 * j + 1];
      int64_t c = ((int64_t)i0 * l + j) * m + k;
      double pm = x[c - sI], pc = x[c];
      for (int i = i0; i < i1; i++, c += sI) {
        const double pp = x[c + sI];
        const uint8_t fl = flags[c];
        double v = 0;
        if (fl & FL_ACTUAL) {
          const double cxm = cx[2 * i], cxp = cx[2 * i + 1];
          const double pjm = x[c - sJ], pjp = x[c + sJ], pkm = x[c - 1], pkp = x[c + 1];
          double tx = ((fl & FL_IP) ? cxp * (pp - pc) : 0.0) - ((fl & FL_IM) ? cxm * (pc - pm) : 0.0);
          double ty = ((fl & FL_JP) ? cyp * (pjp - pc) : 0.0) - ((fl & FL_JM) ? cym * (pc - pjm) : 0.0);
          double tz = ((fl & FL_KP) ? czp * (pkp - pc) : 0.0) - ((fl & FL_KM) ? czm * (pc - pkm) : 0.0);
          v = b[c] + (tx + ty + tz);  // b - A x
          acc += v * v;
        }
        rout[c] = (R)v;
        pm = pc;
        pc = pp;
      }
    }
  }
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) st->s[slot_rr] = tot;
}

// sum and count over unknowns (for the mean projection)
template <typename T>
__global__ void k_sum_unknowns(int64_t nc, const uint8_t *__restrict__ flags, const T *__restrict__ v,
                               DevState *st, double *partials, unsigned *counter, int slot) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) acc += (double)v[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) st->s[slot] = tot;
}

// v -= s[slot]/count at unknowns
template <typename T>
__global__ void k_sub_mean(int64_t nc, const uint8_t *__restrict__ flags, T *__restrict__ v, const DevState *st,
                           int slot, double count) {
  const double mean = st->s[slot] / count;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) v[c] = (T)((double)v[c] - mean);
}

__global__ void k_hs_start(DevState *st, double tol, int rel, int slot_rr) {
  const double rr0 = st->s[slot_rr];
  st->s[HS_RR] = rr0;
  st->s[HS_RR0] = rr0;
  st->s[HS_BETA] = 0;
  st->s[HS_STOP2] = rel ? tol * tol * rr0 : tol * tol;
  st->iter = 0;
  st->status = 0;
  st->done = (rr0 <= st->s[HS_STOP2]) ? 1 : 0;
  if (rr0 != rr0) {
    st->done = 1;
    st->status = XF_ENAN;
  }
}

// restart plain CG inside FCG: rr = (r, r), beta = 0, keep the stopping target
__global__ void k_restart_beta(DevState *st) {
  st->s[HS_RR] = st->s[20 + 2];  // FC_RR2
  st->s[HS_BETA] = 0;
  st->done = 0;
}

// x += (X) e at unknowns
__global__ void k_add_correction(int64_t nc, double *__restrict__ x, const float *__restrict__ e) {
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    x[c] += (double)e[c];
}

// sum of squares over unknowns
__global__ void k_resid_sq(int64_t nc, const uint8_t *__restrict__ space, const double *__restrict__ r,
                           double *partials, unsigned *counter, double *out) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (space[c]) acc += r[c] * r[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

// ------------------------------------------------------------------ driver
static void require_fast(xf_grid g) {
  XF_REQUIRE(!g->use_rho && g->special_rows == 0 && g->bad_rows_lap == 0, XF_EINVAL,
             "throughput CG needs a symmetric operator: every unknown row Actual and no density "
             "(use XF_BICGSTAB or XF_CG3 for other masks)");
}

static int stream_nb(int64_t nc) {
  int64_t nb = (nc + kBlock - 1) / kBlock;
  return (int)(nb > kStreamBlocks ? kStreamBlocks : (nb < 1 ? 1 : nb));
}
static int reduce_nb(int64_t nc) {
  int nb = stream_nb(nc);
  return nb > 2048 ? 2048 : nb;
}

// Runs HS-CG iterations on A e = r (r, p zero-padded; x is the FP64 or FP32
// accumulator).  State must hold HS_RR/HS_STOP2 (k_hs_start).
template <typename T, typename X>
static void hs_iterations(xf_ctx c, xf_grid g, X *x, T *r, T *p0, T *p1, T *q, const T *cx, const T *cy,
                          const T *cz, int maxit, int every) {
  Tiles tl = tiles(g);
  T *pb[2] = {p0, p1};
  const bool f64 = sizeof(T) == 8;
  const char *n1 = f64 ? "cg_stencil_fp64" : "cg_stencil_fp32";
  const char *n2 = f64 ? "cg_update_fp64" : "cg_update_fp32";
  std::vector<ProfMark> marks;
  int it = 0;
  for (;;) {
    const int before = c->h_state->iter;
    for (int s = 0; s < every; s++, it++) {
      T *pold = pb[it & 1], *pnew = pb[(it + 1) & 1];
      int e1 = c->prof ? prof_begin(c) : -1;
      hipLaunchKernelGGL((k_stencil_cg<T, true>), dim3(tl.blocks), dim3(256), 0, c->stream, g->n, g->l, g->m, tl,
                         g->d_flags, cx, cy, cz, r, pold, pnew, q, c->state, c->partials, c->counters);
      if (c->prof) {
        prof_end(c, e1);
        marks.push_back({n1, e1});
      }
      int e2 = c->prof ? prof_begin(c) : -1;
      hipLaunchKernelGGL((k_cg_update<T, X>), dim3(stream_nb(g->ncell)), dim3(256), 0, c->stream, g->ncell, x,
                         pnew, r, q, c->state, c->partials + kMaxPartials, c->counters + 1, maxit);
      if (c->prof) {
        prof_end(c, e2);
        marks.push_back({n2, e2});
      }
    }
    check_kernel("hs_iterations");
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) {
      long long ran = c->h_state->iter - before;
      // the stencil of the converging iteration ran too; a stencil that hit a
      // zero (p, Ap) stopped the loop before its update
      prof_collect(c, marks, {{n1, ran}, {n2, ran}});
    }
    if (c->h_state->done) break;
  }
}

static void upload_state_zero(xf_ctx c) {
  XF_HIP(hipMemsetAsync(c->state, 0, sizeof(DevState), c->stream));
}

// Plain FP64 HS-CG from x0 on b - A x0.
void solve_hs_cg(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                 xf_solve_stats &s) {
  require_fast(g);
  CgWork *w = work(c, g, false);
  const int64_t nc = g->ncell;
  Tiles tl = tiles(g);
  upload_state_zero(c);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  XF_HIP(hipMemsetAsync(w->p[0], 0, sizeof(double) * nc, c->stream));
  XF_HIP(hipMemsetAsync(w->p[1], 0, sizeof(double) * nc, c->stream));
  hipLaunchKernelGGL((k_residual_fast<double>), dim3(tl.blocks), dim3(256), 0, c->stream, g->n, g->l, g->m, tl,
                     g->d_flags, g->d_cx, g->d_cy, g->d_cz, b, x, w->r, c->state, c->partials, c->counters,
                     (int)HS_AUX);
  if (o.project_mean && g->nunk > 0) {
    hipLaunchKernelGGL((k_sum_unknowns<double>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r,
                       c->state, c->partials, c->counters, (int)HS_SUM);
    hipLaunchKernelGGL((k_sub_mean<double>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r,
                       c->state, (int)HS_SUM, (double)g->nunk);
    hipLaunchKernelGGL(k_resid_sq, dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, g->d_space, w->r, c->partials,
                       c->counters, c->state->s + HS_AUX);
  }
  check_kernel("hs start");
  hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, o.tol, o.tol_mode == XF_TOL_REL ? 1 : 0,
                     (int)HS_AUX);
  c->h_state->iter = 0;
  hs_iterations<double, double>(c, g, x, w->r, w->p[0], w->p[1], w->q, g->d_cx, g->d_cy, g->d_cz, o.maxit,
                                o.check_every > 0 ? o.check_every : 16);
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[HS_RR0] > 0 ? sqrt(c->h_state->s[HS_RR] / c->h_state->s[HS_RR0]) : 0;
}

template <typename T>
__global__ void k_sq_unknowns(int64_t nc, const uint8_t *__restrict__ flags, const T *__restrict__ v, DevState *st,
                              double *partials, unsigned *counter, int slot) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) acc += (double)v[c] * (double)v[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) st->s[slot] = tot;
}

// Mixed precision: FP32 Hestenes-Stiefel CG sweeps on the correction
// equation A e = r, with the residual r = b - A x recomputed in FP64 from the
// FP64 iterate after every sweep (iterative refinement), FP64 dot-product
// accumulation everywhere and the correction added to x in FP64.
void solve_mp_cg(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                 xf_solve_stats &s) {
  require_fast(g);
  CgWork *w = work(c, g, true);
  const int64_t nc = g->ncell;
  Tiles tl = tiles(g);
  const int every = o.check_every > 0 ? o.check_every : 16;
  const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  upload_state_zero(c);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  std::vector<ProfMark> marks;
  auto residual = [&]() {
    int e = c->prof ? prof_begin(c) : -1;
    hipLaunchKernelGGL((k_residual_fast<float>), dim3(tl.blocks), dim3(256), 0, c->stream, g->n, g->l, g->m, tl,
                       g->d_flags, g->d_cx, g->d_cy, g->d_cz, b, x, w->r32, c->state, c->partials, c->counters,
                       (int)HS_AUX);
    if (c->prof) {
      prof_end(c, e);
      marks.push_back({"mp_residual_fp64", e});
    }
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) prof_collect(c, marks, {{"mp_residual_fp64", 1}, {"mp_correct", 1}});
    return c->h_state->s[HS_AUX];
  };
  double rr = residual();
  const double rr0 = rr;
  const double stop2 = o.tol_mode == XF_TOL_REL ? o.tol * o.tol * rr0 : o.tol * o.tol;
  int outer = 0, inner_total = 0, status = XF_OK;
  const int max_outer = 60;
  while (rr > stop2) {
    if (rr != rr) { status = XF_ENAN; break; }
    if (outer >= max_outer || inner_total >= o.maxit) { status = XF_EMAXIT; break; }
    outer++;
    if (o.project_mean && g->nunk > 0) {
      hipLaunchKernelGGL((k_sum_unknowns<float>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags,
                         w->r32, c->state, c->partials, c->counters, (int)HS_SUM);
      hipLaunchKernelGGL((k_sub_mean<float>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r32,
                         c->state, (int)HS_SUM, (double)g->nunk);
    }
    hipLaunchKernelGGL((k_sq_unknowns<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r32,
                       c->state, c->partials, c->counters, (int)HS_AUX);
    XF_HIP(hipMemsetAsync(w->e32, 0, sizeof(float) * nc, c->stream));
    XF_HIP(hipMemsetAsync(w->p32[0], 0, sizeof(float) * nc, c->stream));
    hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, inner_tol, 1, (int)HS_AUX);
    check_kernel("mp inner start");
    c->h_state->iter = 0;
    hs_iterations<float, float>(c, g, w->e32, w->r32, w->p32[0], w->p32[1], w->q32, g->d_cxf, g->d_cyf, g->d_czf,
                                o.maxit - inner_total, every);
    inner_total += c->h_state->iter;
    if (c->h_state->status == XF_ENAN) { status = XF_ENAN; break; }
    int e = c->prof ? prof_begin(c) : -1;
    hipLaunchKernelGGL(k_add_correction, dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, x, w->e32);
    if (c->prof) {
      prof_end(c, e);
      marks.push_back({"mp_correct", e});
    }
    rr = residual();
  }
  s.iters = outer;
  s.inner_iters = inner_total;
  s.status = status;
  s.resid = rr0 > 0 ? sqrt(rr / rr0) : 0;
}

// --------------------------------------------------- flexible CG + network
enum { FC_PR = 20, FC_ZQ, FC_RR2, FC_SUMSQ };

template <typename T>
__global__ void k_dot_t(int64_t nc, const T *__restrict__ a, const T *__restrict__ b, DevState *st, double *partials,
                        unsigned *counter, int slot) {
  __shared__ double lds[16];
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    acc += (double)a[c] * (double)b[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) st->s[slot] = tot;
}

// p = z + beta p, beta = -(z, q_old)/(p_old, q_old) (first=1: p = z)
__global__ void k_fcg_dir(int64_t nc, const float *__restrict__ z, float *__restrict__ p, const DevState *st,
                          int first) {
  const double pq = st->s[HS_PQ];
  const float beta = first || pq == 0 ? 0.0f : (float)(-st->s[FC_ZQ] / pq);
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    p[c] = first ? z[c] : z[c] + beta * p[c];
}

// alpha = (p, r)/(p, q)
__global__ void k_fcg_alpha(DevState *st) {
  const double pq = st->s[HS_PQ];
  st->s[HS_ALPHA] = pq != 0 ? st->s[FC_PR] / pq : 0.0;
}

// Flexible CG (A-orthogonalisation against the previous direction, Notay's
// FCG(1)) in FP32 with the network as preconditioner, inside the same FP64
// iterative refinement as the mixed-precision solver.  If the inner residual
// does not improve for stall_window iterations the inner solve continues with
// plain FP32 CG (fallback).
void solve_fcg_nn(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                  xf_solve_stats &s) {
  require_fast(g);
  XF_REQUIRE(o.precond, XF_EINVAL, "XF_FCG_NN needs opts.precond (an xf_unet)");
  struct xf_unet_s *net = (struct xf_unet_s *)o.precond;
  CgWork *w = work(c, g, true);
  const int64_t nc = g->ncell;
  Tiles tl = tiles(g);
  const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  const int stall = o.stall_window > 0 ? o.stall_window : 20;
  float *z = nullptr;
  {
    hipError_t e = hipMalloc(&z, sizeof(float) * nc);
    XF_REQUIRE(e == hipSuccess, XF_ENOMEM, "FCG: out of device memory");
    XF_HIP(hipMemsetAsync(z, 0, sizeof(float) * nc, c->stream));
  }
  struct Free {
    float *p;
    ~Free() { hipFree(p); }
  } fz{z};
  upload_state_zero(c);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  auto poll = [&]() {
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  };
  auto residual = [&]() {
    hipLaunchKernelGGL((k_residual_fast<float>), dim3(tl.blocks), dim3(256), 0, c->stream, g->n, g->l, g->m, tl,
                       g->d_flags, g->d_cx, g->d_cy, g->d_cz, b, x, w->r32, c->state, c->partials, c->counters,
                       (int)HS_AUX);
    poll();
    return c->h_state->s[HS_AUX];
  };
  double rr = residual();
  const double rr0 = rr;
  const double stop2 = o.tol_mode == XF_TOL_REL ? o.tol * o.tol * rr0 : o.tol * o.tol;
  int outer = 0, inner_total = 0, status = XF_OK, fallback = 0;
  float *p = w->p32[0];
  while (rr > stop2) {
    if (rr != rr) { status = XF_ENAN; break; }
    if (outer >= 60 || inner_total >= o.maxit) { status = XF_EMAXIT; break; }
    outer++;
    if (o.project_mean && g->nunk > 0) {
      hipLaunchKernelGGL((k_sum_unknowns<float>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags,
                         w->r32, c->state, c->partials, c->counters, (int)HS_SUM);
      hipLaunchKernelGGL((k_sub_mean<float>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r32,
                         c->state, (int)HS_SUM, (double)g->nunk);
    }
    hipLaunchKernelGGL((k_sq_unknowns<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags, w->r32,
                       c->state, c->partials, c->counters, (int)HS_AUX);
    XF_HIP(hipMemsetAsync(w->e32, 0, sizeof(float) * nc, c->stream));
    XF_HIP(hipMemsetAsync(p, 0, sizeof(float) * nc, c->stream));
    hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, inner_tol, 1, (int)HS_AUX);
    poll();
    double best = c->h_state->s[HS_RR];
    int since_best = 0, it = 0;
    bool use_net = !fallback;
    while (!c->h_state->done) {
      if (!use_net) {
        // plain CG from the current (e, r): restart with p = r
        XF_HIP(hipMemsetAsync(w->p32[0], 0, sizeof(float) * nc, c->stream));
        XF_HIP(hipMemsetAsync(w->p32[1], 0, sizeof(float) * nc, c->stream));
        hipLaunchKernelGGL((k_sq_unknowns<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags,
                           w->r32, c->state, c->partials, c->counters, (int)FC_RR2);
        hipLaunchKernelGGL(k_restart_beta, dim3(1), dim3(1), 0, c->stream, c->state);
        const int before = c->h_state->iter;
        hs_iterations<float, float>(c, g, w->e32, w->r32, w->p32[0], w->p32[1], w->q32, g->d_cxf, g->d_cyf,
                                    g->d_czf, o.maxit - inner_total, 16);
        it += c->h_state->iter - before;
        break;
      }
      // z = M(r)
      hipLaunchKernelGGL((k_sq_unknowns<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, g->d_flags,
                         w->r32, c->state, c->partials, c->counters, (int)FC_SUMSQ);
      unet_apply_grid(c, net, g, w->r32, z, c->state->s + FC_SUMSQ);
      if (it > 0)
        hipLaunchKernelGGL((k_dot_t<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, z, w->q32, c->state,
                           c->partials, c->counters, (int)FC_ZQ);
      hipLaunchKernelGGL(k_fcg_dir, dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, z, p, c->state, it == 0 ? 1 : 0);
      // q = A p, (p, q); (p, r); alpha
      hipLaunchKernelGGL((k_stencil_cg<float, false>), dim3(tl.blocks), dim3(256), 0, c->stream, g->n, g->l, g->m, tl,
                         g->d_flags, g->d_cxf, g->d_cyf, g->d_czf, w->r32, p, p, w->q32, c->state, c->partials,
                         c->counters);
      hipLaunchKernelGGL((k_dot_t<float>), dim3(reduce_nb(nc)), dim3(256), 0, c->stream, nc, p, w->r32, c->state,
                         c->partials + kMaxPartials, c->counters + 1, (int)FC_PR);
      hipLaunchKernelGGL(k_fcg_alpha, dim3(1), dim3(1), 0, c->stream, c->state);
      hipLaunchKernelGGL((k_cg_update<float, float>), dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, w->e32, p,
                         w->r32, w->q32, c->state, c->partials + kMaxPartials, c->counters + 1,
                         o.maxit - inner_total);
      check_kernel("fcg iteration");
      poll();
      it++;
      const double rrn = c->h_state->s[HS_RR];
      if (rrn < 0.99 * best) {
        best = rrn;
        since_best = 0;
      } else if (++since_best >= stall) {
        use_net = false;
        fallback = 1;
      }
    }
    inner_total += it;
    if (c->h_state->status == XF_ENAN) { status = XF_ENAN; break; }
    hipLaunchKernelGGL(k_add_correction, dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, x, w->e32);
    rr = residual();
  }
  s.iters = outer;
  s.inner_iters = inner_total;
  s.status = status;
  s.fallback = fallback;
  s.resid = rr0 > 0 ? sqrt(rr / rr0) : 0;
}

// ||b - A x|| / ||b - A x0|| with the reference rows.
double true_residual(xf_ctx c, xf_grid g, const double *b, const double *x, const double *x0) {
  if ((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) != 0) return NAN;
  double *tmp = nullptr;
  XF_HIP(hipMalloc(&tmp, sizeof(double) * g->ncell));
  double v[2];
  const double *xs[2] = {x, x0};
  for (int t = 0; t < 2; t++) {
    launch_apply_exact(c, g, xs[t], b, -1.0, tmp);
    hipLaunchKernelGGL(k_resid_sq, dim3(reduce_nb(g->ncell)), dim3(256), 0, c->stream, g->ncell, g->d_space, tmp,
                       c->partials, c->counters, c->d_scalar);
    XF_HIP(hipMemcpyAsync(&v[t], c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  }
  hipFree(tmp);
  return v[1] > 0 ? sqrt(v[0] / v[1]) : sqrt(v[0]);
}
