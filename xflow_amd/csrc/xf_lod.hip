// LOD transport step and batched Thomas solves on the device.
//
// THydrodynamicsProblem::SolveTransferEquation__ (XProblem.cpp:2784-3475) runs,
// for each global direction, one implicit sweep along that direction (a
// tridiagonal solve per "sector" of Actual points between two border points)
// followed by two border-only passes along the other directions.  Lines are
// independent within a sweep or pass (a sweep reads only the previous time
// level and the stages of earlier directions; a border pass reads sector
// interiors that no pass writes), so each line is one GPU thread; sweeps and
// passes are separate launches, in the reference's order.  The Thomas
// elimination (XSLAE.cpp:416-445) keeps its A[m], B[m] in a line-interleaved
// scratch array (element m of line t at m*nlines + t) so neighbouring threads
// touch neighbouring addresses.  Arithmetic follows the reference's operand
// order (built with -ffp-contract=off), so results are bit-identical.
#include <algorithm>

#include "xf_internal.h"

struct LodArgs {
  int n, l, m;
  const uint8_t *codes;
  const double *h[3];        // steps per axis
  const double *vel[3];      // u, v, w
  const double *f_prev, *rp, *kappa;
  double *stage[3];          // f_1/3, f_2/3, f (output)
  double tau;
  int scheme;
  double *scrA, *scrB;       // [maxlen][nlines]
  int *err;
};

__device__ inline void line_geometry(const LodArgs &a, int axis, int t, int64_t &base, int64_t &st, int &len,
                                     int &nlines) {
  const int64_t S[3] = {(int64_t)a.l * a.m, (int64_t)a.m, 1};
  const int E[3] = {a.n, a.l, a.m};
  // lines along `axis`, indexed by the other two axes (slower one first)
  int ax1 = axis == 0 ? 1 : 0;
  int ax2 = axis == 2 ? 1 : 2;
  nlines = E[ax1] * E[ax2];
  int d1 = t / E[ax2], d2 = t % E[ax2];
  base = (int64_t)d1 * S[ax1] + (int64_t)d2 * S[ax2];
  st = S[axis];
  len = E[axis];
}

// Explicit half-weight convection/diffusion terms of direction e for cell c,
// from field fd (a stage or the previous level) and f_prev (XProblem.cpp:3246-3292).
__device__ inline void side_terms(const LodArgs &a, int e, int64_t c, int ii, const double *fd, double kap,
                                  double &FC, double &FD) {
  const double *h = a.h[e];
  const int64_t s = e == 0 ? (int64_t)a.l * a.m : e == 1 ? (int64_t)a.m : 1;
  const double h_1 = h[ii - 1], hh = h[ii], hx2 = h[ii - 1] + h[ii];
  const double hxhxh = h[ii - 1] * h[ii] * (h[ii - 1] + h[ii]) / 2;
  const double vel = a.vel[e][c];
  double fc = 0, fdd = 0;
  for (int item = 1; item <= 2; item++) {
    const double *f_ = item == 1 ? fd : a.f_prev;
    if (a.scheme == 0) fc = fc + 0.5 * (-vel) * (f_[c + s] - f_[c - s]) / hx2;
    else if (vel < 0) fc = fc + 0.5 * (-vel) * (f_[c + s] - f_[c]) / hh;
    else fc = fc + 0.5 * (-vel) * (f_[c] - f_[c - s]) / hh;
    fdd = fdd + 0.5 * (kap) * ((hh * f_[c - s] - hx2 * f_[c] + h_1 * f_[c + s]) / hxhxh);
  }
  FC = fc;
  FD = fdd;
}

// Row `s` of the sector system: a, b, c, f (XProblem.cpp:3058-3222, 3224-3419).
struct Row {
  double a, b, c, f;
};

__device__ inline Row sector_row(const LodArgs &A, int axis, int dir, int64_t base, int64_t sa, int beg, int end,
                                 int mbeg, int mend, int fl, int fr, int sd, int s) {
  Row r{0, 0, 0, 0};
  const double *h = A.h[axis];
  if (fl && s == 1) {
    const int64_t cb = base + (int64_t)beg * sa;
    if (mbeg == PT_PREDEFINED) r = Row{1, 0.5, 0.5, A.rp[cb]};
    else r = Row{1, -1 / h[beg], 1 / h[beg], A.rp[cb]};
    return r;
  }
  if (fr && s == sd - 1) {
    const int64_t ce = base + (int64_t)end * sa;
    if (mend == PT_PREDEFINED) r = Row{0.5, 0.5, 1, A.rp[ce]};
    else r = Row{-1 / h[end - 1], 1 / h[end - 1], 1, A.rp[ce]};
    return r;
  }
  const int da = beg + s - fl;
  const int64_t c = base + (int64_t)da * sa;
  const int64_t S[3] = {(int64_t)A.l * A.m, (int64_t)A.m, 1};
  const int idx3[3] = {(int)(c / S[0]), (int)((c / S[1]) % A.l), (int)(c % A.m)};
  const int ii = idx3[axis];
  const double h_1 = h[ii - 1], hh = h[ii], hx2 = h[ii - 1] + h[ii];
  const double hxhxh = h[ii - 1] * h[ii] * (h[ii - 1] + h[ii]) / 2;
  const double kap = A.kappa[c];
  const double vt = A.vel[axis][c];
  const double tau = A.tau;
  if (A.scheme == 0) {
    r.a = 1 * (0.5 * (-vt / hx2)) + 1 * (0.5 * (-(kap)*hh / hxhxh));
    r.b = 1 / tau + 1 * (0.5 * (kap)*hx2 / hxhxh);
    r.c = 1 * (0.5 * (vt / hx2)) + 1 * (0.5 * (-(kap)*h_1 / hxhxh));
  } else if (vt < 0) {
    r.a = 1 * (0.5 * (-(kap)*hh / hxhxh));
    r.b = 1 / tau + 1 * (0.5 * (-vt / hh)) + 1 * (0.5 * (kap)*hx2 / hxhxh);
    r.c = 1 * (0.5 * (vt / hh)) + 1 * (0.5 * (-(kap)*h_1 / hxhxh));
  } else {
    r.a = 1 * (0.5 * (-vt / hh)) + 1 * (0.5 * (-(kap)*hh / hxhxh));
    r.b = 1 / tau + 1 * (0.5 * (vt / hh)) + 1 * (0.5 * (kap)*hx2 / hxhxh);
    r.c = 1 * (0.5 * (-(kap)*h_1 / hxhxh));
  }
  const double *fp = A.f_prev;
  double F = (1 / tau) * fp[c] + A.rp[c];
  if (A.scheme == 0) F = F + 1 * (0.5 * (-vt) * (fp[c + sa] - fp[c - sa]) / hx2);
  else if (vt < 0) F = F + 1 * (0.5 * (-vt) * (fp[c + sa] - fp[c]) / hh);
  else F = F + 1 * (0.5 * (-vt) * (fp[c] - fp[c - sa]) / hh);
  F = F + 1 * (0.5 * (kap) * ((hh * fp[c - sa] - hx2 * fp[c] + h_1 * fp[c + sa]) / hxhxh));
  for (int e = 0; e < 3; e++) {
    if (e == axis) continue;
    const int de = e + 1;
    const double *fd = de > dir ? fp : A.stage[de - 1];
    double FC, FD;
    side_terms(A, e, c, idx3[e], fd, kap, FC, FD);
    F = F + 1 * FC + 1 * FD;
  }
  r.f = F;
  return r;
}

// One sweep (implicit solve along `axis`, writes stage[dg-1]) or one border
// pass (rewrites sector ends of stage[dg-1]) over every line along `axis`.
__global__ void k_lod_lines(LodArgs A, int axis, int dg, int border_only) {
  int64_t base, sa;
  int len, nlines;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  line_geometry(A, axis, 0, base, sa, len, nlines);
  if (t >= nlines) return;
  line_geometry(A, axis, t, base, sa, len, nlines);
  const uint8_t *mk = A.codes;
  const double *h = A.h[axis];
  double *fa = A.stage[dg - 1];
  const int dir = axis + 1;
  int beg = -1, mbeg = 0;
  for (int q = 0; q < len; q++) {
    const int mq = mk[base + (int64_t)q * sa];
    if (beg < 0) {
      if (mq == PT_ACTUAL) {
        if (q == 0) { atomicOr(A.err, 1); return; }  // XProblem.cpp:2932-2935
        beg = q - 1;
        mbeg = mk[base + (int64_t)beg * sa];
      }
      continue;
    }
    if (mq == PT_ACTUAL) continue;
    const int end = q, mend = mq;
    // XProblem.cpp:2952-2960
    if (mbeg == PT_ACTUAL || mbeg == PT_FICTIVE || mend == PT_ACTUAL || mend == PT_FICTIVE || end - beg < 2) {
      atomicOr(A.err, 2);
      return;
    }
    const int64_t cb = base + (int64_t)beg * sa, ce = base + (int64_t)end * sa;
    if (border_only) {  // XProblem.cpp:2965-3033
      if (mbeg == PT_DEFINED || mbeg == PT_EQUATION) fa[cb] = A.rp[cb];
      else if (mbeg == PT_PREDEFINED) fa[cb] = 2 * A.rp[cb] - fa[cb + sa];
      else if (mbeg == PT_NORMAL || mbeg == PT_PRENORMAL) fa[cb] = -(h[beg]) * A.rp[cb] + fa[cb + sa];
      else { atomicOr(A.err, 4); return; }
      if (mend == PT_DEFINED || mend == PT_EQUATION) fa[ce] = A.rp[ce];
      else if (mend == PT_PREDEFINED) fa[ce] = 2 * A.rp[ce] - fa[ce - sa];
      else if (mend == PT_NORMAL || mend == PT_PRENORMAL) fa[ce] = (h[end - 1]) * A.rp[ce] + fa[ce - sa];
      else { atomicOr(A.err, 4); return; }
      beg = -1;
      continue;
    }
    const bool okb = mbeg == PT_DEFINED || mbeg == PT_EQUATION || mbeg == PT_PREDEFINED || mbeg == PT_NORMAL ||
                     mbeg == PT_PRENORMAL;
    const bool oke = mend == PT_DEFINED || mend == PT_EQUATION || mend == PT_PREDEFINED || mend == PT_NORMAL ||
                     mend == PT_PRENORMAL;
    if (!okb || !oke) { atomicOr(A.err, 4); return; }
    const int fl = (mbeg != PT_DEFINED && mbeg != PT_EQUATION);
    const int fr = (mend != PT_DEFINED && mend != PT_EQUATION);
    const int sd = end - beg + fl + fr;  // dim - 1
    const double u0 = fl ? 0.0 : A.rp[cb];
    const double uM = fr ? 0.0 : A.rp[ce];
    // forward elimination (XSLAE.cpp:428-436)
    double Am1 = 0, Bm1 = u0;
    for (int s = 1; s < sd; s++) {
      const Row rw = sector_row(A, axis, dir, base, sa, beg, end, mbeg, mend, fl, fr, sd, s);
      const double Am = (-rw.c) / (rw.a * Am1 + rw.b);
      const double Bm = (rw.f - rw.a * Bm1) / (rw.a * Am1 + rw.b);
      A.scrA[(int64_t)s * nlines + t] = Am;
      A.scrB[(int64_t)s * nlines + t] = Bm;
      Am1 = Am;
      Bm1 = Bm;
    }
    // back substitution (XSLAE.cpp:438-441) and assignment (XProblem.cpp:3439-3449)
    double u = uM;
    if (!fr) fa[base + (int64_t)(beg + sd - fl) * sa] = u;
    for (int s = sd - 1; s > 0; s--) {
      u = A.scrA[(int64_t)s * nlines + t] * u + A.scrB[(int64_t)s * nlines + t];
      if (s >= fl && s <= sd - fr) fa[base + (int64_t)(beg + s - fl) * sa] = u;
    }
    if (!fl) fa[cb] = u0;
    beg = -1;
  }
}

namespace {
struct LodWork {
  int64_t ncell = 0, scr = 0;
  double *st0 = nullptr, *st1 = nullptr, *scrA = nullptr, *scrB = nullptr;
  int *err = nullptr;
  ~LodWork() {
    hipFree(st0); hipFree(st1); hipFree(scrA); hipFree(scrB); hipFree(err);
  }
};
}  // namespace

static void run_lod(xf_ctx c, xf_grid g, double *f, const double *f_prev, const double *u, const double *v,
                    const double *w, const double *rp, const double *kappa, double tau, int scheme) {
  const int64_t nc = g->ncell;
  const int E[3] = {g->n, g->l, g->m};
  int64_t maxlines = 0;
  int maxlen = 0;
  for (int a = 0; a < 3; a++) {
    int64_t nl = nc / E[a];
    maxlines = std::max(maxlines, nl);
    maxlen = std::max(maxlen, E[a]);
  }
  LodWork wk;
  XF_HIP(hipMalloc(&wk.st0, sizeof(double) * nc));
  XF_HIP(hipMalloc(&wk.st1, sizeof(double) * nc));
  XF_HIP(hipMalloc(&wk.scrA, sizeof(double) * maxlines * (maxlen + 2)));
  XF_HIP(hipMalloc(&wk.scrB, sizeof(double) * maxlines * (maxlen + 2)));
  XF_HIP(hipMalloc(&wk.err, sizeof(int)));
  XF_HIP(hipMemsetAsync(wk.err, 0, sizeof(int), c->stream));
  // f_1/3 and f_2/3 start as copies of the previous level (XProblem.cpp:2796-2797)
  XF_HIP(hipMemcpyAsync(wk.st0, f_prev, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  XF_HIP(hipMemcpyAsync(wk.st1, f_prev, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  LodArgs A;
  A.n = g->n; A.l = g->l; A.m = g->m;
  A.codes = g->d_codes;
  A.h[0] = g->d_hx; A.h[1] = g->d_hy; A.h[2] = g->d_hz;
  A.vel[0] = u; A.vel[1] = v; A.vel[2] = w;
  A.f_prev = f_prev; A.rp = rp; A.kappa = kappa;
  A.stage[0] = wk.st0; A.stage[1] = wk.st1; A.stage[2] = f;
  A.tau = tau;
  A.scheme = scheme;
  A.scrA = wk.scrA; A.scrB = wk.scrB;
  A.err = wk.err;
  for (int dg = 1; dg < 4; dg++) {
    for (int dfk = 1; dfk < 4; dfk++) {  // XProblem.cpp:2849-2864
      int dir;
      if (dfk == 1) dir = dg;
      else if (dg == 1) dir = dfk;
      else if (dg == 2) dir = dfk == 2 ? 1 : 3;
      else dir = dfk - 1;
      const int axis = dir - 1;
      const int nl = (int)(nc / E[axis]);
      hipLaunchKernelGGL(k_lod_lines, dim3((nl + 127) / 128), dim3(128), 0, c->stream, A, axis, dg,
                         dir != dg ? 1 : 0);
      check_kernel("k_lod_lines");
    }
  }
  int err = 0;
  XF_HIP(hipMemcpyAsync(&err, wk.err, sizeof(int), hipMemcpyDeviceToHost, c->stream));
  XF_HIP(hipStreamSynchronize(c->stream));
  XF_REQUIRE(err == 0, XF_EINVAL,
             err & 1 ? "SolveTransferEquation: Initial border not found."
                     : "SolveTransferEquation: Incorrect sector border or size.");
}

extern "C" int xf_lod_transport(xf_ctx c, xf_grid g, double *f, const double *f_prev, const double *u,
                                const double *v, const double *w, const double *rp, const double *kappa, double tau,
                                int scheme) {
  return guarded(c, [&] {
    XF_REQUIRE(g && f && f_prev && u && v && w && rp && kappa, XF_EINVAL, "xf_lod_transport: null argument");
    XF_REQUIRE(scheme == 0 || scheme == 1, XF_EINVAL, "Unknown scheme type.");
    XF_REQUIRE(tau > 0, XF_EINVAL, "xf_lod_transport: time step must be positive");
    run_lod(c, g, f, f_prev, u, v, w, rp, kappa, tau, scheme);
  });
}

// Independent systems a[m] u[m-1] + b[m] u[m] + c[m] u[m+1] = f[m], m = 1..M-1,
// with u[0] and u[M] given; system s occupies [s*(M+1), (s+1)*(M+1)).
__global__ void k_thomas_batched(int nsys, int M, double *u, const double *a, const double *b, const double *cc,
                                 const double *f, double *scrA, double *scrB) {
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsys) return;
  const int64_t o = (int64_t)s * (M + 1);
  double Am1 = 0, Bm1 = u[o];
  for (int q = 1; q < M; q++) {
    const double Am = (-cc[o + q]) / (a[o + q] * Am1 + b[o + q]);
    const double Bm = (f[o + q] - a[o + q] * Bm1) / (a[o + q] * Am1 + b[o + q]);
    scrA[(int64_t)q * nsys + s] = Am;
    scrB[(int64_t)q * nsys + s] = Bm;
    Am1 = Am;
    Bm1 = Bm;
  }
  double x = u[o + M];
  for (int q = M - 1; q > 0; q--) {
    x = scrA[(int64_t)q * nsys + s] * x + scrB[(int64_t)q * nsys + s];
    u[o + q] = x;
  }
}

extern "C" int xf_thomas_batched(xf_ctx c, int nsys, int M, double *u, const double *a, const double *b,
                                 const double *cc, const double *f) {
  return guarded(c, [&] {
    XF_REQUIRE(nsys >= 0 && M >= 1 && u && a && b && cc && f, XF_EINVAL, "xf_thomas_batched: bad argument");
    if (nsys == 0) return;
    double *sa = nullptr, *sb = nullptr;
    XF_HIP(hipMalloc(&sa, sizeof(double) * nsys * (M + 1)));
    XF_HIP(hipMalloc(&sb, sizeof(double) * nsys * (M + 1)));
    hipLaunchKernelGGL(k_thomas_batched, dim3((nsys + 127) / 128), dim3(128), 0, c->stream, nsys, M, u, a, b, cc, f,
                       sa, sb);
    check_kernel("k_thomas_batched");
    XF_HIP(hipStreamSynchronize(c->stream));
    hipFree(sa);
    hipFree(sb);
  });
}
