// Reference-semantics Krylov solvers on the device: three-term CG, BiCG,
// minimal residuals and BiCGStab, statement for statement as in XSLAE.cpp,
// with every scalar kept on the device.  The host only enqueues kernels and
// polls the convergence flag every few iterations; kernels queued past
// convergence see `done` and return at once, so the iteration count is the
// reference's.  Dot products are deterministic device reductions (a different
// summation order than the reference's serial loop).
#include <algorithm>
#include <cmath>

#include "xf_internal.h"

// ------------------------------------------------------------ gated kernels
// res may alias s1 or s2.
__global__ void k_lincomb_dev(int64_t nc, const uint8_t *__restrict__ space, const double *p1p, const double *s1,
                              const double *p2p, const double *s2, double *res, const DevState *st) {
  if (st->done) return;
  const double p1 = p1p ? *p1p : 0.0, p2 = p2p ? *p2p : 0.0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (space && !space[c]) continue;
    double v;
    if (p1 != 0 && p2 != 0) v = p1 * s1[c] + p2 * s2[c];
    else if (p1 != 0) v = p1 * s1[c];
    else if (p2 != 0) v = p2 * s2[c];
    else v = 0;
    __builtin_nontemporal_store(v, res + c);  // read again only by a later launch
  }
}

__global__ void k_copy_dev(int64_t nc, const uint8_t *__restrict__ space, const double *s, double *res,
                           const DevState *st) {
  if (st->done) return;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (!space || space[c]) __builtin_nontemporal_store(s[c], res + c);
}

__global__ void k_dot_dev(int64_t nc, const uint8_t *__restrict__ space, const double *__restrict__ a,
                          const double *__restrict__ b, double *partials, unsigned *counter, double *out,
                          const DevState *st) {
  __shared__ double lds[16];
  if (st && st->done) return;  // null state: not gated
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (!space || space[c]) acc += a[c] * b[c];
  double bs = block_sum(acc, lds);
  double tot;
  if (grid_sum_last(bs, partials, counter, &tot, lds) && threadIdx.x == 0) *out = tot;
}

__global__ void k_absmax_dev(int64_t nc, const uint8_t *__restrict__ space, const double *__restrict__ x,
                             double *partials, unsigned *counter, double *out, const DevState *st) {
  __shared__ double lv[kBlock];
  __shared__ bool am_last;
  if (st->done) return;
  double best = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (!space || space[c]) {
      double v = fabs(x[c]);
      if (v > best || v != v) best = v;  // NaN propagates
    }
  lv[threadIdx.x] = best;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double o = lv[threadIdx.x + s];
      if (o > lv[threadIdx.x] || o != o) lv[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    partials[blockIdx.x] = lv[0];
    __threadfence();
    am_last = atomicAdd(counter, 1u) == gridDim.x - 1;
  }
  __syncthreads();
  if (!am_last) return;
  __threadfence();
  // the last block folds the partials: a strided share per thread, then the block
  double b = 0;
  for (unsigned q = threadIdx.x; q < gridDim.x; q += blockDim.x) {
    double o = __hip_atomic_load(&partials[q], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (o > b || o != o) b = o;
  }
  __syncthreads();
  lv[threadIdx.x] = b;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double o = lv[threadIdx.x + s];
      if (o > lv[threadIdx.x] || o != o) lv[threadIdx.x] = o;
    }
    __syncthreads();
  }
  if (threadIdx.x != 0) return;
  *out = lv[0];
  *counter = 0;
}

// ------------------------------------------------------------ scalar steps
enum { ONE = 28, MONE = 29 };  // constant slots
enum { CG3_RR = 0, CG3_RR1, CG3_TAU, CG3_ALFA, CG3_DEN, CG3_P1, CG3_P2, CG3_P3, CG3_EPS, CG3_EPSCUR, CG3_MTAU };
enum { BI_A1 = 0, BI_A2, BI_ALPHA, BI_MALPHA, BI_B1, BI_B2, BI_BETA, BI_RR, BI_EPS, BI_EPSCUR };
enum { MR_D1 = 0, MR_D2, MR_ALFA, MR_RR, MR_EPS, MR_EPSCUR };
enum { BS_PN = 0, BS_AN, BS_BN, BS_WN, BS_PN1, BS_NORM, BS_EPS, BS_DEN, BS_TT, BS_TS };

// Loop top of the 2-norm solvers: `while (epsCur > eps)` with the square
// root of the latest (r, r) (XSLAE.cpp:109-114, 197-210, 375-380); a NaN
// residual stops with XF_ENAN instead of counting as converged.
__global__ void k_loop_top(DevState *st, int rr_slot, int eps_slot, int epscur_slot, int maxit) {
  if (st->done) return;
  double e = sqrt(st->s[rr_slot]);
  st->s[epscur_slot] = e;
  if (!(e > st->s[eps_slot])) {
    st->done = 1;
    if (e != e) st->status = XF_ENAN;
    return;
  }
  if (st->iter >= maxit) {
    st->done = 1;
    st->status = XF_EMAXIT;
    return;
  }
  st->iter++;
}

// CG3 start-up (XSLAE.cpp:69-90): tau_0 = (r,r)/(A r, r).
__global__ void k_cg3_start(DevState *st) {
  double rr1 = st->s[CG3_RR1];
  st->s[CG3_EPSCUR] = sqrt(rr1);
  if (rr1 == 0) {
    st->done = 1;
    return;
  }
  st->s[CG3_TAU] = rr1 / st->s[CG3_DEN];
  st->s[CG3_MTAU] = -st->s[CG3_TAU];
}
// XSLAE.cpp:94-103
__global__ void k_cg3_second(DevState *st) {
  if (st->done) return;
  double rr = st->s[CG3_RR];
  st->s[CG3_EPSCUR] = sqrt(rr);
  if (rr == 0) {
    st->done = 1;
    return;
  }
  st->s[CG3_ALFA] = 1;
}
// XSLAE.cpp:119-131
__global__ void k_cg3_coef(DevState *st) {
  if (st->done) return;
  double rr = st->s[CG3_RR], rr1 = st->s[CG3_RR1], tauk = st->s[CG3_TAU], alfak = st->s[CG3_ALFA];
  double tauk1 = rr / st->s[CG3_DEN];
  double alfak1 = 1 - (1 / alfak) * (tauk1 / tauk) * (rr / rr1);
  alfak1 = 1 / alfak1;
  st->s[CG3_P1] = alfak1;
  st->s[CG3_P2] = 1 - alfak1;
  st->s[CG3_P3] = -(tauk1 * alfak1);
  st->s[CG3_RR1] = rr;
  st->s[CG3_TAU] = tauk1;
  st->s[CG3_ALFA] = alfak1;
}
__global__ void k_bicg_alpha(DevState *st) {  // XSLAE.cpp:233
  if (st->done) return;
  st->s[BI_ALPHA] = st->s[BI_A1] / st->s[BI_A2];
  st->s[BI_MALPHA] = -st->s[BI_ALPHA];
}
__global__ void k_bicg_beta(DevState *st) {  // XSLAE.cpp:274-276
  if (st->done) return;
  if (st->s[BI_B1] == 0 || st->s[BI_B2] == 0) {
    st->done = 1;
    st->status = XF_EBREAKDOWN;
    return;
  }
  st->s[BI_BETA] = st->s[BI_B1] / st->s[BI_B2];
}
__global__ void k_mr_alfa(DevState *st) {  // XSLAE.cpp:385
  if (st->done) return;
  st->s[MR_ALFA] = (-1) * st->s[MR_D1] / st->s[MR_D2];
}
__global__ void k_bs_top(DevState *st, int maxit) {  // XSLAE.cpp:716-719, max-norm test
  if (st->done) return;
  double nrm = st->s[BS_NORM];
  if (!(nrm > st->s[BS_EPS])) {
    st->done = 1;
    if (nrm != nrm) st->status = XF_ENAN;
    return;
  }
  if (st->iter >= maxit) {
    st->done = 1;
    st->status = XF_EMAXIT;
    return;
  }
  st->iter++;
}
__global__ void k_bs_bn(DevState *st) {  // XSLAE.cpp:721
  if (st->done) return;
  st->s[BS_BN] = (st->s[BS_PN1] / st->s[BS_PN]) * (st->s[BS_AN] / st->s[BS_WN]);
}
__global__ void k_bs_an(DevState *st) {  // XSLAE.cpp:730
  if (st->done) return;
  st->s[BS_AN] = st->s[BS_PN1] / st->s[BS_DEN];
}
__global__ void k_bs_wn(DevState *st) {  // XSLAE.cpp:736, 746
  if (st->done) return;
  st->s[BS_WN] = st->s[BS_TS] / st->s[BS_TT];
  st->s[BS_PN] = st->s[BS_PN1];
}

// ----------------------------------------------------------------- driver
namespace {

struct Exec {
  xf_ctx c;
  xf_grid g;
  xf_pack_s *pk = nullptr;  // set: Rn vectors of the packed operator (no mask)
  int pt = 0;               // packed operator taken transposed (its operator~ state)
  int64_t n;                // vector length
  const uint8_t *space;     // mask of the vectors (null: every entry)
  int nbs, nbr;
  Exec(xf_ctx c_, xf_grid g_) : c(c_), g(g_), n(g_->ncell), space(g_->d_space) { blocks(); }
  Exec(xf_ctx c_, xf_pack_s *p, int transpose) : c(c_), g(p->g), pk(p), pt(transpose), n(p->nunk), space(nullptr) {
    blocks();
  }
  void blocks() {
    int64_t nb = (n + kBlock - 1) / kBlock;
    nbs = (int)(nb > 65536 ? 65536 : (nb < 1 ? 1 : nb));
    nbr = (int)(nb > 2048 ? 2048 : (nb < 1 ? 1 : nb));
  }
  double *S(int slot) const { return c->state->s + slot; }
  // res = s[p1]*s1 + s[p2]*s2 (slot < 0: coefficient 0)
  void lin(int p1, const double *s1, int p2, const double *s2, double *res) {
    hipLaunchKernelGGL(k_lincomb_dev, dim3(nbs), dim3(kBlock), 0, c->stream, n, space,
                       p1 >= 0 ? S(p1) : nullptr, s1, p2 >= 0 ? S(p2) : nullptr, s2, res, c->state);
    check_kernel("k_lincomb_dev");
  }
  void copy(const double *s, double *res) {
    hipLaunchKernelGGL(k_copy_dev, dim3(nbs), dim3(kBlock), 0, c->stream, n, space, s, res, c->state);
    check_kernel("k_copy_dev");
  }
  void dot(const double *a, const double *b, int slot) {
    hipLaunchKernelGGL(k_dot_dev, dim3(nbr), dim3(kBlock), 0, c->stream, n, space, a, b, c->partials,
                       c->counters, S(slot), c->state);
    check_kernel("k_dot_dev");
  }
  void absmax(const double *x, int slot) {
    hipLaunchKernelGGL(k_absmax_dev, dim3(nbr), dim3(kBlock), 0, c->stream, n, space, x, c->partials,
                       c->counters, S(slot), c->state);
    check_kernel("k_absmax_dev");
  }
  void apply(const double *x, const double *b, double sign, double *y) {
    if (pk) launch_pack_apply(c, pk, pt, x, b, sign, y, true);
    else launch_apply_gated(c, g, x, b, sign, y);
  }
  // the transposed operator (operator~): the packed A^T, or A on a grid (the
  // symmetric masks the matrix-free path is used on)
  void applyT(const double *x, double *y) {
    if (pk) launch_pack_apply(c, pk, 1 - pt, x, nullptr, 0, y, true);
    else launch_apply_gated(c, g, x, nullptr, 0, y);
  }
  void apply_dot(const double *x, const double *w, int slot) {
    if (pk) launch_pack_apply_dot(c, pk, pt, x, w, S(slot));
    else launch_apply_dot_gated(c, g, x, w, S(slot), 0);
  }
  // result |= v after the loop (not gated by the convergence flag)
  void final_copy(const double *v, double *res) {
    if (pk) XF_HIP(hipMemcpyAsync(res, v, sizeof(double) * n, hipMemcpyDeviceToDevice, c->stream));
    else launch_lincomb(c, g, 1.0, v, 0.0, nullptr, res);
  }
  void one(void (*k)(DevState *)) {
    hipLaunchKernelGGL(k, dim3(1), dim3(1), 0, c->stream, c->state);
    check_kernel("scalar step");
  }
  void top(int rr, int eps, int epscur, int maxit) {
    hipLaunchKernelGGL(k_loop_top, dim3(1), dim3(1), 0, c->stream, c->state, rr, eps, epscur, maxit);
    check_kernel("k_loop_top");
  }
  // copies the device state to the pinned mirror; returns done
  bool poll() {
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    return c->h_state->done != 0;
  }
  void init(std::initializer_list<std::pair<int, double>> slots) {
    DevState h{};
    h.s[ONE] = 1.0;
    h.s[MONE] = -1.0;
    for (auto &p : slots) h.s[p.first] = p.second;
    *c->h_state = h;
    XF_HIP(hipMemcpyAsync(c->state, c->h_state, sizeof(DevState), hipMemcpyHostToDevice, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  }
};

struct Buf {
  double *p = nullptr;
  Buf(int64_t n, const double *init, hipStream_t s) {
    XF_HIP(hipMalloc(&p, sizeof(double) * (n ? n : 1)));
    if (init) XF_HIP(hipMemcpyAsync(p, init, sizeof(double) * n, hipMemcpyDeviceToDevice, s));
    else XF_HIP(hipMemsetAsync(p, 0, sizeof(double) * n, s));
  }
  ~Buf() { hipFree(p); }
  Buf(const Buf &) = delete;
  Buf &operator=(const Buf &) = delete;
};

int poll_every(const xf_solve_opts &o) { return o.check_every > 0 ? o.check_every : 32; }

}  // namespace

// Three-term CG, XSLAE.cpp:56-157.  x receives TSLAE::Result.
static void solve_cg3(Exec &E, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                      xf_solve_stats &s) {
  xf_ctx c = E.c;
  int64_t nc = E.n;
  E.init({{CG3_EPS, o.tol}});
  Buf xk_1(nc, x0, c->stream), rk_1(nc, nullptr, c->stream), xk(nc, x0, c->stream), rk(nc, nullptr, c->stream);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));  // fResult = x.Copy()
  E.lin(-1, nullptr, -1, nullptr, x);                                                         // *fResult |= 0
  E.apply(xk_1.p, b, -1, rk_1.p);                                                             // rk_1 |= A(xk_1) - b
  E.dot(rk_1.p, rk_1.p, CG3_RR1);
  E.copy(xk_1.p, x);
  E.apply_dot(rk_1.p, rk_1.p, CG3_DEN);
  E.one(k_cg3_start);
  E.lin(ONE, xk_1.p, CG3_MTAU, rk_1.p, xk.p);  // xk |= xk_1 - tauk*rk_1
  E.apply(xk.p, b, -1, rk.p);
  E.dot(rk.p, rk.p, CG3_RR);
  E.copy(xk.p, x);
  E.one(k_cg3_second);
  const int every = poll_every(o);
  while (!E.poll()) {
    for (int t = 0; t < every; t++) {
      E.top(CG3_RR, CG3_EPS, CG3_EPSCUR, o.maxit);
      E.apply_dot(rk.p, rk.p, CG3_DEN);
      E.one(k_cg3_coef);
      E.lin(CG3_P1, xk.p, CG3_P2, xk_1.p, x);  // result |= alfak1*xk + (1-alfak1)*xk_1
      E.lin(ONE, x, CG3_P3, rk.p, x);          // result |= result - (tauk1*alfak1)*rk
      E.copy(xk.p, xk_1.p);
      E.copy(rk.p, rk_1.p);
      E.copy(x, xk.p);
      E.apply(xk.p, b, -1, rk.p);
      E.dot(rk.p, rk.p, CG3_RR);
    }
  }
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[CG3_EPSCUR];
}

// BiCG, XSLAE.cpp:159-348 (A^T taken as A: defined for symmetric operators).
static void solve_bicg(Exec &E, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                       xf_solve_stats &s) {
  xf_ctx c = E.c;
  int64_t nc = E.n;
  E.init({{BI_EPS, o.tol}});
  Buf xk(nc, x0, c->stream), xk_new(nc, x0, c->stream);
  Buf rk(nc, nullptr, c->stream), rt(nc, nullptr, c->stream), rk_new(nc, nullptr, c->stream),
      rt_new(nc, nullptr, c->stream), pk(nc, nullptr, c->stream), pt(nc, nullptr, c->stream),
      pk_new(nc, nullptr, c->stream), pt_new(nc, nullptr, c->stream);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  E.lin(-1, nullptr, -1, nullptr, x);
  E.lin(-1, nullptr, -1, nullptr, xk.p);  // xk |= 0
  E.apply(xk.p, b, -1, rk.p);
  E.lin(MONE, rk.p, -1, nullptr, rt.p);   // rt |= (-1)*rk
  E.copy(rt.p, rk.p);
  E.copy(rk.p, pk.p);
  E.copy(rt.p, pt.p);
  E.dot(rk.p, rk.p, BI_RR);
  E.copy(xk.p, x);
  const int every = poll_every(o);
  while (!E.poll()) {
    for (int t = 0; t < every; t++) {
      E.top(BI_RR, BI_EPS, BI_EPSCUR, o.maxit);
      E.dot(rk.p, rt.p, BI_A1);
      E.apply_dot(pk.p, pt.p, BI_A2);  // (A(pk), pt)
      E.one(k_bicg_alpha);
      E.lin(ONE, xk.p, BI_ALPHA, pk.p, xk_new.p);
      E.apply(pk.p, nullptr, 0, rk_new.p);
      E.lin(BI_MALPHA, rk_new.p, ONE, rk.p, rk_new.p);
      E.applyT(pt.p, rt_new.p);  // A_ = A.Copy(), ~A_
      E.lin(BI_MALPHA, rt_new.p, ONE, rt.p, rt_new.p);
      E.dot(rk_new.p, rt_new.p, BI_B1);
      E.dot(rk.p, rt.p, BI_B2);
      E.one(k_bicg_beta);
      E.lin(ONE, rk_new.p, BI_BETA, pk.p, pk_new.p);
      E.lin(ONE, rt_new.p, BI_BETA, pt.p, pt_new.p);
      E.copy(xk_new.p, xk.p);
      E.copy(rk_new.p, rk.p);
      E.copy(rt_new.p, rt.p);
      E.copy(pk_new.p, pk.p);
      E.copy(pt_new.p, pt.p);
      E.dot(rk.p, rk.p, BI_RR);
    }
  }
  // result |= xk (after the loop, also after a breakdown)
  XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
  XF_HIP(hipStreamSynchronize(c->stream));
  E.final_copy(xk.p, x);
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[BI_EPSCUR];
}

// Minimal residuals, XSLAE.cpp:350-414.
static void solve_mr(Exec &E, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                     xf_solve_stats &s) {
  xf_ctx c = E.c;
  int64_t nc = E.n;
  E.init({{MR_EPS, o.tol}});
  Buf xk(nc, x0, c->stream), rk(nc, nullptr, c->stream), ark(nc, nullptr, c->stream);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  E.lin(-1, nullptr, -1, nullptr, x);
  E.apply(xk.p, b, -1, rk.p);
  E.dot(rk.p, rk.p, MR_RR);
  E.copy(xk.p, x);
  const int every = poll_every(o) * 4;
  while (!E.poll()) {
    for (int t = 0; t < every; t++) {
      E.top(MR_RR, MR_EPS, MR_EPSCUR, o.maxit);
      E.apply(rk.p, nullptr, 0, ark.p);
      E.dot(ark.p, rk.p, MR_D1);
      E.dot(ark.p, ark.p, MR_D2);
      E.one(k_mr_alfa);
      E.lin(ONE, xk.p, MR_ALFA, rk.p, xk.p);
      E.apply(xk.p, b, -1, rk.p);
      E.dot(rk.p, rk.p, MR_RR);
    }
  }
  E.final_copy(xk.p, x);
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[MR_EPSCUR];
}

// BiCGStab, XSLAE.cpp:667-767 (max-norm stopping test).
static void solve_bicgstab(Exec &E, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                           xf_solve_stats &s) {
  xf_ctx c = E.c;
  int64_t nc = E.n;
  E.init({{BS_EPS, o.tol}, {BS_PN, 1.0}, {BS_AN, 1.0}, {BS_BN, 1.0}, {BS_WN, 1.0}});
  Buf Un(nc, x0, c->stream), tmp1(nc, nullptr, c->stream), tmp2(nc, nullptr, c->stream), Rn(nc, nullptr, c->stream),
      sRn(nc, nullptr, c->stream), Pn(nc, nullptr, c->stream), Pn1(nc, nullptr, c->stream), Vn(nc, nullptr, c->stream),
      Sn(nc, nullptr, c->stream), Tn(nc, nullptr, c->stream);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  E.lin(-1, nullptr, -1, nullptr, x);
  E.apply(Un.p, nullptr, 0, Rn.p);
  E.lin(ONE, b, MONE, Rn.p, Rn.p);  // Rn |= b - Rn
  E.copy(Rn.p, sRn.p);
  E.absmax(Rn.p, BS_NORM);
  const int every = poll_every(o);
  while (!E.poll()) {
    for (int t = 0; t < every; t++) {
      hipLaunchKernelGGL(k_bs_top, dim3(1), dim3(1), 0, c->stream, c->state, o.maxit);
      E.dot(sRn.p, Rn.p, BS_PN1);
      E.one(k_bs_bn);
      E.lin(BS_WN, Vn.p, -1, nullptr, Pn1.p);
      E.lin(ONE, Pn.p, MONE, Pn1.p, Pn1.p);
      E.lin(BS_BN, Pn1.p, -1, nullptr, Pn1.p);
      E.lin(ONE, Rn.p, ONE, Pn1.p, Pn1.p);
      E.apply(Pn1.p, nullptr, 0, Vn.p);
      E.dot(sRn.p, Vn.p, BS_DEN);
      E.one(k_bs_an);
      E.lin(BS_AN, Vn.p, -1, nullptr, Sn.p);
      E.lin(ONE, Rn.p, MONE, Sn.p, Sn.p);
      E.apply(Sn.p, nullptr, 0, Tn.p);
      E.dot(Tn.p, Sn.p, BS_TS);
      E.dot(Tn.p, Tn.p, BS_TT);
      E.one(k_bs_wn);
      E.lin(BS_AN, Pn1.p, -1, nullptr, tmp1.p);
      E.lin(ONE, tmp1.p, BS_WN, Sn.p, tmp2.p);
      E.lin(ONE, Un.p, ONE, tmp2.p, Un.p);
      E.lin(BS_WN, Tn.p, -1, nullptr, Rn.p);
      E.lin(ONE, Sn.p, MONE, Rn.p, Rn.p);
      E.copy(Pn1.p, Pn.p);
      E.absmax(Rn.p, BS_NORM);
    }
  }
  E.final_copy(Un.p, x);
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[BS_NORM];
}

// ------------------------------------------------------------------ C ABI
extern "C" void xf_solve_opts_default(int method, xf_solve_opts *o) {
  if (!o) return;
  xf_solve_opts d{};
  d.method = method;
  d.tol_mode = XF_TOL_ABS;
  d.maxit = 1000000;
  d.check_every = 0;
  d.project_mean = 0;
  d.inner_tol = 1e-4;
  d.stall_window = 20;
  d.precond = nullptr;
  d.verify = 1;
  switch (method) {
    case XF_CG3: d.tol = 1e-5; break;       // XSLAE.cpp:64
    case XF_BICG: d.tol = 1e-8; break;      // XSLAE.cpp:166
    case XF_MR: d.tol = 1e-7; break;        // XSLAE.cpp:358
    case XF_BICGSTAB: d.tol = 1e-5; break;  // XSLAE.cpp:669
    case XF_BICGSTAB_T:
      d.tol = 1e-5;  // XSLAE.cpp:669
      d.check_every = 16;
      break;
    default:
      d.tol = 1e-8;
      d.tol_mode = XF_TOL_REL;
      d.check_every = 16;
  }
  *o = d;
}

extern "C" int xf_solve(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts *opts,
                        xf_solve_stats *stats) {
  xf_solve_stats s{};
  int st = guarded(c, [&] {
    XF_REQUIRE(g && b && x0 && x && opts, XF_EINVAL, "xf_solve: null argument");
    XF_REQUIRE(x != x0 && x != b, XF_EINVAL, "xf_solve: x must not alias b or x0");
    xf_solve_opts o = *opts;
    XF_REQUIRE(o.maxit >= 0, XF_EINVAL, "xf_solve: negative iteration cap");
    XF_HIP(hipEventRecord(c->ev0, c->stream));
    switch (o.method) {
      case XF_CG3:
      case XF_BICG:
      case XF_MR:
      case XF_BICGSTAB: {
        XF_REQUIRE(o.tol_mode == XF_TOL_ABS, XF_EINVAL, "xf_solve: the reference solvers use absolute tolerances");
        Exec E(c, g);
        if (o.method == XF_CG3) solve_cg3(E, b, x0, x, o, s);
        else if (o.method == XF_BICG) solve_bicg(E, b, x0, x, o, s);
        else if (o.method == XF_MR) solve_mr(E, b, x0, x, o, s);
        else solve_bicgstab(E, b, x0, x, o, s);
        break;
      }
      case XF_CG: solve_hs_cg(c, g, b, x0, x, o, s); break;
      case XF_BICGSTAB_T: solve_bicgstab_t(c, g, b, x0, x, o, s); break;
      case XF_CG_MP: solve_mp_cg(c, g, b, x0, x, o, s); break;
      case XF_FCG_NN:
      case XF_FCG_MGNN: solve_fcg_nn(c, g, b, x0, x, o, s); break;
      default: XF_REQUIRE(false, XF_EINVAL, "xf_solve: unknown method");
    }
    XF_HIP(hipEventRecord(c->ev1, c->stream));
    XF_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    XF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    s.ms = ms;
    s.resid_true = o.verify ? true_residual(c, g, b, x, x0, nullptr) : -1.0;
  });
  if (st != XF_OK) s.status = st;
  if (stats) *stats = s;
  return st != XF_OK ? st : s.status;
}

extern "C" int xf_solve_noalloc(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x,
                                const xf_solve_opts *opts, xf_solve_stats *stats) {
  xf_solve_stats s{};
  int st = guarded(c, [&] {
    XF_REQUIRE(g && b && x0 && x && opts, XF_EINVAL, "xf_solve_noalloc: null argument");
    XF_REQUIRE(x != x0 && x != b, XF_EINVAL, "xf_solve_noalloc: x must not alias b or x0");
    xf_solve_opts o = *opts;
    XF_REQUIRE(o.maxit >= 0, XF_EINVAL, "xf_solve_noalloc: negative iteration cap");
    XF_HIP(hipEventRecord(c->ev0, c->stream));
    double *tmp = nullptr;
    solve_hs_cg_noalloc(c, g, b, x0, x, o, s, &tmp);
    XF_HIP(hipEventRecord(c->ev1, c->stream));
    XF_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    XF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    s.ms = ms;
    s.resid_true = o.verify ? true_residual(c, g, b, x, x0, tmp) : -1.0;
  });
  if (st != XF_OK) s.status = st;
  if (stats) *stats = s;
  return st != XF_OK ? st : s.status;
}

// ||b - A x|| / ||b - A x0|| on Rn vectors of the packed operator
static double pack_true_residual(xf_ctx c, xf_pack_s *p, int transpose, const double *b, const double *x,
                                 const double *x0) {
  double *tmp = nullptr;
  XF_HIP(hipMalloc(&tmp, sizeof(double) * p->nunk));
  double v[2];
  const double *xs[2] = {x, x0};
  const int nb = (int)std::min<int64_t>(2048, (p->nunk + kBlock - 1) / kBlock);
  for (int t = 0; t < 2; t++) {
    launch_pack_apply(c, p, transpose, xs[t], b, -1.0, tmp, false);
    hipLaunchKernelGGL(k_dot_dev, dim3(nb), dim3(kBlock), 0, c->stream, p->nunk, (const uint8_t *)nullptr, tmp, tmp,
                       c->partials, c->counters, c->d_scalar, (const DevState *)nullptr);
    XF_HIP(hipMemcpyAsync(&v[t], c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  }
  hipFree(tmp);
  return v[1] > 0 ? sqrt(v[0] / v[1]) : sqrt(v[0]);
}

extern "C" int xf_pack_solve(xf_ctx c, xf_pack p, int transpose, const double *b, const double *x0, double *x,
                             const xf_solve_opts *opts, xf_solve_stats *stats) {
  xf_solve_stats s{};
  int st = guarded(c, [&] {
    XF_REQUIRE(p && b && x0 && x && opts, XF_EINVAL, "xf_pack_solve: null argument");
    XF_REQUIRE(x != x0 && x != b, XF_EINVAL, "xf_pack_solve: x must not alias b or x0");
    xf_solve_opts o = *opts;
    XF_REQUIRE(o.maxit >= 0, XF_EINVAL, "xf_pack_solve: negative iteration cap");
    XF_REQUIRE(o.tol_mode == XF_TOL_ABS, XF_EINVAL, "xf_pack_solve: the reference solvers use absolute tolerances");
    XF_HIP(hipEventRecord(c->ev0, c->stream));
    XF_REQUIRE(transpose == 0 || transpose == 1, XF_EINVAL, "xf_pack_solve: transpose must be 0 or 1");
    Exec E(c, p, transpose);
    switch (o.method) {
      case XF_CG3: solve_cg3(E, b, x0, x, o, s); break;
      case XF_BICG: solve_bicg(E, b, x0, x, o, s); break;
      case XF_MR: solve_mr(E, b, x0, x, o, s); break;
      case XF_BICGSTAB: solve_bicgstab(E, b, x0, x, o, s); break;
      default: XF_REQUIRE(false, XF_EINVAL, "xf_pack_solve: the packed operator runs the reference solvers only");
    }
    XF_HIP(hipEventRecord(c->ev1, c->stream));
    XF_HIP(hipEventSynchronize(c->ev1));
    float ms = 0;
    XF_HIP(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    s.ms = ms;
    s.resid_true = o.verify ? pack_true_residual(c, p, transpose, b, x, x0) : -1.0;
  });
  if (st != XF_OK) s.status = st;
  if (stats) *stats = s;
  return st != XF_OK ? st : s.status;
}
