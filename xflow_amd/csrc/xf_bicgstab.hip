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
// several times per point.)  KV and KT march: a block of four waves owns four
// rows of 64 cells and walks them along i, the field and the density at
// planes i - 1, i, i + 1 in registers, the k-neighbours by lane shuffles, the
// Fictive tests from the flag byte's face bits; an Actual row is the
// reference's expression in its operand order (the same bits as the flat
// rows); the other unknown rows (PreDefined, PreNormal) go through the
// generic row in a small list-driven pass beside them, which keeps the
// generic row's registers out of the march.  Grids with unknowns on the
// array faces (which the marching tiles do not visit) use the flat passes.
// Each vector statement keeps the reference's rounding sequence (XSLAE.cpp:
// 723-745).  The streaming passes store non-temporally (their outputs are
// read only by later passes; 0.70 -> 0.66 ms per iteration at 256^3).  Reductions are per-block partials that the next pass folds in a
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

constexpr int kBtBlocks = 2048 + 64;  // blocks of every pass at most (partials per reduction)
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

// the block's sum (MAX: max) of per-thread values into partial slot k, entry
// first + blockIdx.x
template <bool MAX>
__device__ void put(const BtArgs &A, int k, double v, double *lds, int first = 0) {
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
  if (threadIdx.x == 0) slot(A, k)[first + blockIdx.x] = lds[0];
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
    __builtin_nontemporal_store(A.r[c] + q, A.p + c);
  });
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->s[BT_NORM] = norm;
    st->s[BT_PN1] = pn1;
    st->s[BT_BN] = bn;
  }
}

// ---------------------------------------------------------- marching rows
constexpr int kBtRows = 4;  // rows (waves) per block of the marching passes

struct BtTiles {
  int kt, jt, it, ic, total;
};

// The Actual row of T3dLaplaceOperator / T3dVariableDivergentedOperator
// (XOperator.cpp:985-1031, 1990-2025) from values already in registers: f and
// rho at the cell and its six neighbours, the face bits of the flag byte
// (set where the neighbour is not Fictive), the steps around the cell.
template <bool RHO>
__device__ inline double bt_actual_row(uint8_t fl, double fc, double fim, double fip, double fjm, double fjp,
                                       double fkm, double fkp, double rc, double rim, double rip, double rjm,
                                       double rjp, double rkm, double rkp, double hxm, double hxc, double hym,
                                       double hyc, double hzm, double hzc) {
  const double hx_2X = (hxm + hxc) / 2;
  const double hx_2Y = (hym + hyc) / 2;
  const double hx_2Z = (hzm + hzc) / 2;
  if (!RHO) {
    double f_i_1 = (fc - fim) / (hxm);
    double f_i = (fip - fc) / (hxc);
    double f_j_1 = (fc - fjm) / (hym);
    double f_j = (fjp - fc) / (hyc);
    double f_k_1 = (fc - fkm) / (hzm);
    double f_k = (fkp - fc) / (hzc);
    if (!(fl & FL_IM)) f_i_1 = 0;
    if (!(fl & FL_IP)) f_i = 0;
    if (!(fl & FL_JM)) f_j_1 = 0;
    if (!(fl & FL_JP)) f_j = 0;
    if (!(fl & FL_KM)) f_k_1 = 0;
    if (!(fl & FL_KP)) f_k = 0;
    return -((f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z);
  }
  double f_i_1 = rim * (fc - fim) / (hxm);
  double f_i = rip * (fip - fc) / (hxc);
  double f_j_1 = rjm * (fc - fjm) / (hym);
  double f_j = rjp * (fjp - fc) / (hyc);
  double f_k_1 = rkm * (fc - fkm) / (hzm);
  double f_k = rkp * (fkp - fc) / (hzc);
  double d_i = (rip - rc) / (hxm);
  double d_j = (rjp - rc) / (hym);
  double d_k = (rkp - rc) / (hzm);
  if (!(fl & FL_IM)) f_i_1 = 0;
  if (!(fl & FL_IP)) { f_i = 0; d_i = 0; }
  if (!(fl & FL_JM)) f_j_1 = 0;
  if (!(fl & FL_JP)) { f_j = 0; d_j = 0; }
  if (!(fl & FL_KM)) f_k_1 = 0;
  if (!(fl & FL_KP)) { f_k = 0; d_k = 0; }
  double lap = (f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z;
  double divd = d_i + d_j + d_k;
  return -(rc * lap - divd * fc);
}

// MODE 0: v = A p with the partials of (r^, v) (XSLAE.cpp:729); MODE 1:
// t = A s with the partials of (t, s) and (t, t) (XSLAE.cpp:735).  Blocks
// are persistent over the tiles (a multiple of 8 of them), consecutive tiles
// on one XCD.
template <bool RHO, int MODE>
__global__ __launch_bounds__(kBlock) void k_bt_rows(BtArgs A, BtTiles tl, const uint8_t *__restrict__ flags) {
  __shared__ double lds[kBlock];
  if (A.st->done) return;
  const GridView &g = A.g;
  const int n = g.n, l = g.l, m = g.m;
  const double *__restrict__ f = MODE == 0 ? A.p : A.s;
  double *__restrict__ out = MODE == 0 ? A.v : A.t;
  const double *__restrict__ rho = g.rho;
  const int64_t sI = (int64_t)l * m, sJ = m;
  const int lane = threadIdx.x & 63;
  const int per = (tl.total + 7) / 8, bx = gridDim.x >> 3, xcd = blockIdx.x & 7;
  double acc0 = 0, acc1 = 0;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    const int ktile = t % tl.kt, jtile = (t / tl.kt) % tl.jt, itile = t / (tl.kt * tl.jt);
    // the last k-tile is shifted back to end at m - 2; its repeated cells are not owned
    const int kcount = 1 + ktile * 64;
    const int kstart = max(1, min(kcount, m - 1 - 64));
    const int kraw = kstart + lane;
    const int k = min(kraw, m - 2);
    const bool kown = kraw >= kcount && kraw <= m - 2;
    const int j = 1 + jtile * kBtRows + (threadIdx.x >> 6);
    if (j > l - 2) continue;  // uniform per wave; no barrier below
    const int i0 = 1 + itile * tl.ic, i1 = min(i0 + tl.ic, n - 1);
    const double hym = g.hy[j - 1], hyc = g.hy[j], hzm = g.hz[k - 1], hzc = g.hz[k];
    int64_t c = ((int64_t)i0 * l + j) * m + k;
    double fm = f[c - sI], fc = f[c];
    double rm = 0, rcur = 0;
    if (RHO) {
      rm = rho[c - sI];
      rcur = rho[c];
    }
    const bool kload_r = lane == 63 || kraw >= m - 2;
    for (int i = i0; i < i1; i++, c += sI) {
      const double fp = f[c + sI], fjm = f[c - sJ], fjp = f[c + sJ];
      double rp = 0, rjm = 0, rjp = 0;
      if (RHO) {
        rp = rho[c + sI];
        rjm = rho[c - sJ];
        rjp = rho[c + sJ];
      }
      const uint8_t fl = flags[c];
      double fkm = __shfl_up(fc, 1, 64), fkp = __shfl_down(fc, 1, 64);
      double rkm = 0, rkp = 0;
      if (RHO) {
        rkm = __shfl_up(rcur, 1, 64);
        rkp = __shfl_down(rcur, 1, 64);
      }
      if (lane == 0) {
        fkm = f[c - 1];
        if (RHO) rkm = rho[c - 1];
      }
      if (kload_r) {
        fkp = f[c + 1];
        if (RHO) rkp = rho[c + 1];
      }
      if (kown && (fl & FL_UNKNOWN)) {
        if (fl & FL_ACTUAL) {
          const double v = bt_actual_row<RHO>(fl, fc, fm, fp, fjm, fjp, fkm, fkp, rcur, rm, rp, rjm, rjp, rkm, rkp,
                                              g.hx[i - 1], g.hx[i], hym, hyc, hzm, hzc);
          out[c] = v;
          if (MODE == 0) {
            acc0 += A.rh[c] * v;
          } else {
            acc0 += v * fc;
            acc1 += v * v;
          }
        }
      }
      fm = fc;
      fc = fp;
      if (RHO) {
        rm = rcur;
        rcur = rp;
      }
    }
  }
  if (MODE == 0) {
    put<false>(A, PS_RV, acc0, lds);
    if (blockIdx.x == 0 && threadIdx.x == 0) A.st->iter++;  // XSLAE.cpp:719
  } else {
    put<false>(A, PS_TS, acc0, lds);
    put<false>(A, PS_TT, acc1, lds);
  }
}

// The unknown rows that are not Actual (PreDefined, PreNormal), from a list
// of their cells, for the marching passes (which evaluate Actual rows only):
// the generic row, and partials into the entries after the marching blocks'.
template <int MODE>
__global__ __launch_bounds__(kBlock) void k_bt_other(BtArgs A, const int64_t *__restrict__ cells, int64_t count,
                                                     int first) {
  __shared__ double lds[kBlock];
  if (A.st->done) return;
  const GridView &g = A.g;
  const double *__restrict__ f = MODE == 0 ? A.p : A.s;
  double *__restrict__ out = MODE == 0 ? A.v : A.t;
  double acc0 = 0, acc1 = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < count; q += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = cells[q];
    int i, j, k;
    decompose(c, g.l, g.m, i, j, k);
    const double v = row_value(g, f, c, i, j, k);
    out[c] = v;
    if (MODE == 0) {
      acc0 += A.rh[c] * v;
    } else {
      acc0 += v * f[c];
      acc1 += v * v;
    }
  }
  if (MODE == 0) {
    put<false>(A, PS_RV, acc0, lds, first);
  } else {
    put<false>(A, PS_TS, acc0, lds, first);
    put<false>(A, PS_TT, acc1, lds, first);
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
    __builtin_nontemporal_store(A.r[c] - av, A.s + c);
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
    __builtin_nontemporal_store(A.x[c] + tmp2, A.x + c);
    const double wt = wn * A.t[c];
    const double rv = sv - wt;
    __builtin_nontemporal_store(rv, A.r + c);
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
  // the marching passes need every unknown inside the array faces
  if (g->ring_unknowns < 0) {
    const std::vector<uint8_t> &cd = host_codes(g);
    const int gn = g->n, gl = g->l, gm = g->m;
    auto unknown = [&](int ii, int jj, int kk) {
      const uint8_t cv = cd[((int64_t)ii * gl + jj) * gm + kk];
      return cv == PT_ACTUAL || cv == PT_PREDEFINED || cv == PT_NORMAL || cv == PT_PRENORMAL || cv == PT_PREEQUATION;
    };
    int cnt = 0;  // the six faces (edges counted more than once: only zero matters)
    for (int a = 0; a < gl; a++)
      for (int b = 0; b < gm; b++) cnt += unknown(0, a, b) + unknown(gn - 1, a, b);
    for (int a = 0; a < gn; a++)
      for (int b = 0; b < gm; b++) cnt += unknown(a, 0, b) + unknown(a, gl - 1, b);
    for (int a = 0; a < gn; a++)
      for (int b = 0; b < gl; b++) cnt += unknown(a, b, 0) + unknown(a, b, gm - 1);
    g->ring_unknowns = cnt;
  }
  const bool march = g->ring_unknowns == 0 && g->n >= 3 && g->l >= 3 && g->m >= 3;
  BtTiles tl{};
  int nb_march = 0, nb_other = 0;
  if (march && g->n_other_rows < 0) {
    // the cells whose rows the marching passes leave to k_bt_other
    const std::vector<uint8_t> &cd = host_codes(g);
    std::vector<int64_t> cells;
    for (int64_t q = 0; q < nc; q++)
      if (cd[q] == PT_PREDEFINED || cd[q] == PT_PRENORMAL || cd[q] == PT_NORMAL || cd[q] == PT_PREEQUATION)
        cells.push_back(q);
    g->n_other_rows = (int64_t)cells.size();
    if (!cells.empty()) {
      XF_HIP(hipMalloc(&g->d_other_rows, sizeof(int64_t) * cells.size()));
      XF_HIP(hipMemcpy(g->d_other_rows, cells.data(), sizeof(int64_t) * cells.size(), hipMemcpyHostToDevice));
    }
  }
  if (march) {
    tl.kt = (g->m - 2 + 63) / 64;
    tl.jt = (g->l - 2 + kBtRows - 1) / kBtRows;
    for (tl.ic = 32;; tl.ic /= 2) {
      tl.it = (g->n - 2 + tl.ic - 1) / tl.ic;
      tl.total = tl.kt * tl.jt * tl.it;
      if (tl.total >= 2048 || tl.ic <= 2) break;
    }
    // a multiple of 8 marching blocks (the XCD mapping), then the blocks of
    // the other rows; together at most one partial slot's worth
    nb_other = g->n_other_rows > 0 ? (int)std::min<int64_t>(64, (g->n_other_rows + kBlock - 1) / kBlock) : 0;
    nb_march = (int)std::min<int64_t>(2048, 8 * ((tl.total + 7) / 8));
    A.nb = nb_march + nb_other;
  } else {
    A.nb = (int)std::max<int64_t>(1, std::min<int64_t>(kBtBlocks, (nc + kBlock - 1) / kBlock));
  }
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
      if (!march) {
        hipLaunchKernelGGL(k_bt_v, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      } else {
        if (A.g.rho) {
          hipLaunchKernelGGL((k_bt_rows<true, 0>), dim3(nb_march), dim3(kBlock), 0, c->stream, A, tl, g->d_flags);
        } else {
          hipLaunchKernelGGL((k_bt_rows<false, 0>), dim3(nb_march), dim3(kBlock), 0, c->stream, A, tl, g->d_flags);
        }
        if (nb_other)
          hipLaunchKernelGGL(k_bt_other<0>, dim3(nb_other), dim3(kBlock), 0, c->stream, A, g->d_other_rows,
                             g->n_other_rows, nb_march);
      }
      hipLaunchKernelGGL(k_bt_s, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      if (!march) {
        hipLaunchKernelGGL(k_bt_t, dim3(A.nb), dim3(kBlock), 0, c->stream, A);
      } else {
        if (A.g.rho) {
          hipLaunchKernelGGL((k_bt_rows<true, 1>), dim3(nb_march), dim3(kBlock), 0, c->stream, A, tl, g->d_flags);
        } else {
          hipLaunchKernelGGL((k_bt_rows<false, 1>), dim3(nb_march), dim3(kBlock), 0, c->stream, A, tl, g->d_flags);
        }
        if (nb_other)
          hipLaunchKernelGGL(k_bt_other<1>, dim3(nb_other), dim3(kBlock), 0, c->stream, A, g->d_other_rows,
                             g->n_other_rows, nb_march);
      }
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
