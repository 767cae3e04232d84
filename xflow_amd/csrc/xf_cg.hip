// Throughput conjugate gradients for the masked 7-point pressure operator.
//
// Plain solves use the single-reduction CG of Chronopoulos and Gear ("s-step
// iterative methods for symmetric linear systems", J. Comput. Appl. Math. 25,
// 1989): both inner products of an iteration come from one pass, so one
// iteration is two kernels with one grid-wide reduction between them:
//   KA  per-tile partials of gamma = (r, r) and delta = (A r, r); no store
//                                                   [2.5-D marching stencil]
//   KB  every block reduces the partials itself (same order in every block,
//       so all blocks hold bit-identical gamma, delta), then
//       beta = gamma/gamma_old, alpha = gamma/(delta - beta gamma/alpha_old),
//       p = r + beta p, r -= alpha A p (A p recomputed from r and the old p at
//       the seven stencil points, never stored), x += alpha p
//                                                   [2.5-D marching stencil]
// The textbook formulation also carries s = A p by the recurrence
// s = A r + beta s; recomputing A p instead removes w = A r and s from memory
// (81 -> 50 B per cell and iteration in FP64).
// No kernel needs a device-scope fence or atomic: a fence per block (the
// "last block finishes the sum" pattern) writes back the XCD's L2 and cost
// more than the data movement itself on this chip.  Scalars that carry over
// between iterations (gamma, alpha) live in two parity slots of the device
// state, written by block 0 of KB into the slot the next launch reads.
// Algorithmic traffic per interior cell and iteration: KA reads r and the
// flag byte (9 B in FP64, 5 B in FP32); KB reads r, p and the flag byte and
// writes p and r, and on odd iterations also reads and writes x with two
// steps at once (33 / 49 B in FP64, mean 41 B; 17 / 25 B in FP32, mean 21 B).
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
// Reductions are deterministic: fixed tile partition, block sums in wave
// order, partials added in block order.
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <mutex>
#include <string>

#include "xf_internal.h"

constexpr int TK = 64;   // k extent of a tile (one wavefront)
constexpr int TJ = 4;    // j extent of a tile (waves per block)
constexpr int IC = 32;   // i planes marched per block
constexpr int kStreamBlocks = 4096;
constexpr int kCgTileBlocks = 2048;   // KA blocks (persistent over tiles): bounds the partials KB reads

struct CgWork {
  int64_t n = 0;
  double *r = nullptr, *p[2] = {nullptr, nullptr}, *q = nullptr, *tmp = nullptr;
  float *r32 = nullptr, *p32[2] = {nullptr, nullptr}, *q32 = nullptr, *e32 = nullptr;
  float *z32 = nullptr;    // flexible CG: the preconditioned residual
  double *part = nullptr;  // KA partials: [kCgTileBlocks][2]
};

void free_cg_work(CgWork *w) {
  if (!w) return;
  hipFree(w->r); hipFree(w->p[0]); hipFree(w->p[1]); hipFree(w->q); hipFree(w->tmp);
  hipFree(w->r32); hipFree(w->p32[0]); hipFree(w->p32[1]); hipFree(w->q32); hipFree(w->e32);
  hipFree(w->z32); hipFree(w->part);
  delete w;
}

static CgWork *work(xf_ctx c, xf_grid g, bool fp32, bool fcg = false) {
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
  get(&w->part, sizeof(double) * 2 * kCgTileBlocks);
  if (fp32) {
    get(&w->r32, n * 4); get(&w->p32[0], n * 4); get(&w->p32[1], n * 4); get(&w->q32, n * 4); get(&w->e32, n * 4);
  }
  if (fcg) get(&w->z32, n * 4);
  return w;
}

// State slots of the throughput solvers.
enum { HS_RR = 0, HS_PQ, HS_ALPHA, HS_BETA, HS_STOP2, HS_RR0, HS_SUM, HS_CNT, HS_AUX, HS_RRW,
       CG_GAMMA = 10, CG_ALPHA = 12 };  // HS_RRW: (r, r) of the working-precision residual  // CG_GAMMA, CG_ALPHA: two parity slots each

struct Tiles {
  int kt, jt, it, total;
  int blocks;  // 8 * ceil(total / 8): the XCD-aware mapping is a bijection on [0, blocks)
  int ic = IC; // i planes per tile
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

// The marching stencils below: one wavefront per (j) row of 64 k-cells (or
// 64*CPL with vectors); each block marches IC planes in i, keeping the three
// planes i-1, i, i+1 of its own column in registers; the four in-plane
// neighbours come from the L1/L2-resident rows of the same plane.  Block index
// -> tile mapping is XCD-aware: the hardware deals workgroups to the 8 XCDs
// round-robin, so consecutive tiles (which share halo rows) are given to
// workgroups on the same XCD.

// Block sum valid in every thread.  Butterfly shuffles leave the same value in
// all lanes (a + b == b + a exactly), and the wave sums are added in wave
// order, so every block that sums the same data gets the same bits.
__device__ inline double block_sum_all(double v, double *lds) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6, nw = blockDim.x >> 6;
  __syncthreads();
  if (lane == 0) lds[wid] = v;
  __syncthreads();
  double r = 0;
  for (int q = 0; q < nw; q++) r += lds[q];
  return r;
}

// ------------------------------------------------ plain CG: two kernels per iteration
// The stored vectors are r and p (each ping-ponged between two buffers, so a
// kernel never overwrites what another block still reads as a neighbour) and
// x.  s = A p is not stored: KB recomputes it from r and p_old at the seven
// stencil points (p = r + beta p_old evaluated there with the same operations,
// so the neighbours' p are bit-identical to the ones their owners store).
//   KA  k_cgcg_dots  per-block partials of gamma = (r, r), delta = (A r, r);
//                    reads r and the flag byte, stores nothing      [9 / 5 B]
//   KB  k_cgcg_step  every block reduces the partials (same order in every
//                    block: bit-identical scalars), then
//                    p = r + beta p_old, r_out = r - alpha A p, and x += alpha p
//                    on odd iterations (two steps at once)          [41 / 21 B mean]
// In exact arithmetic (A p, p) = delta - beta gamma/alpha_old, so alpha below
// is the Hestenes-Stiefel step and the method is CG.
//
// Tiles: one wave covers 64*CPL consecutive k-cells of one (i, j) row (16-byte
// vectors: CPL = 2 doubles or 4 floats), TJ rows per block, tl.ic planes marched
// in i with the planes i-1, i, i+1 of the column in registers.  The last k-tile
// of a row is shifted back to end at m-2 so every vector lies inside the row;
// cells it shares with the tile before are computed twice and stored and
// counted only by their owner.  Rows shorter than 64 cells (CPL = 1) clamp the
// lanes past the row end onto its last cell and store nothing from them.
struct CgTile {
  int64_t c;     // array index of the lane's first cell on plane i0
  int i0, i1, j;
  int k0;        // first cell of the lane in its row
  int kraw;      // nominal first cell of the lane (k0 before clamping)
  bool own[4];
  bool allown;
};

template <int CPL>
__device__ inline bool cg_tile(const Tiles &tl, int t, int n, int l, int m, CgTile &w) {
  const int lane = threadIdx.x & 63;
  const int kspan = 64 * CPL;
  const int ktile = t % tl.kt;
  const int jtile = (t / tl.kt) % tl.jt;
  const int itile = t / (tl.kt * tl.jt);
  const int kcount = 1 + ktile * kspan;  // cells from here on are this tile's own
  const int kstart = max(1, min(kcount, m - 1 - kspan));
  w.kraw = kstart + lane * CPL;
  const int k0 = min(w.kraw, m - 1 - CPL);
  w.j = 1 + jtile * TJ + (threadIdx.x >> 6);
  w.i0 = 1 + itile * tl.ic;
  w.i1 = min(w.i0 + tl.ic, n - 1);
  w.allown = true;
#pragma unroll
  for (int e = 0; e < CPL; e++) {
    w.own[e] = w.kraw + e >= kcount && w.kraw + e <= m - 2;
    w.allown = w.allown && w.own[e];
  }
  w.k0 = k0;
  w.c = ((int64_t)w.i0 * l + w.j) * m + k0;
  return w.j <= l - 2;
}

template <typename T, int CPL>
struct CgVec {
  T v[CPL];
};

// A f at the CPL cells of a lane from the centre plane fc, its i-neighbours
// fm / fp, j-neighbours fjm / fjp and k-neighbours left / right (the reference's
// Actual-row formula with folded coefficients; 0 on rows that are not Actual).
template <typename T, int CPL>
__device__ inline T cg_stencil_cell(uint8_t f, int e, const CgVec<T, CPL> &fc, const CgVec<T, CPL> &fm,
                                    const CgVec<T, CPL> &fp, T fjm, T fjp, T left, T right, T cxm, T cxp, T cym,
                                    T cyp, T czm, T czp) {
  const T x0 = fc.v[e];
  const T km = e == 0 ? left : fc.v[e > 0 ? e - 1 : 0];
  const T kp = e == CPL - 1 ? right : fc.v[e < CPL - 1 ? e + 1 : 0];
  T tx = ((f & FL_IP) ? cxp * (fp.v[e] - x0) : T(0)) - ((f & FL_IM) ? cxm * (x0 - fm.v[e]) : T(0));
  T ty = ((f & FL_JP) ? cyp * (fjp - x0) : T(0)) - ((f & FL_JM) ? cym * (x0 - fjm) : T(0));
  T tz = ((f & FL_KP) ? czp * (kp - x0) : T(0)) - ((f & FL_KM) ? czm * (x0 - km) : T(0));
  return (f & FL_ACTUAL) ? -(tx + ty + tz) : T(0);
}

// KA: per-block partials (gamma, delta) = ((r, r), (A r, r)) over Actual cells.
// Blocks are persistent over tiles (bounded partial count), consecutive tiles
// on one XCD.
template <typename T, int CPL>
__global__ __launch_bounds__(256) void k_cgcg_dots(int n, int l, int m, Tiles tl, const uint8_t *__restrict__ flags,
                                                    const T *__restrict__ cx, const T *__restrict__ cy,
                                                    const T *__restrict__ cz, const T *__restrict__ r,
                                                    const DevState *st, double *__restrict__ part) {
  __shared__ double lds[8];
  if (st->done) return;
  using V = CgVec<T, CPL>;
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  double gam = 0, del = 0;
  const int64_t sI = (int64_t)l * m, sJ = m;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    CgTile w;
    if (!cg_tile<CPL>(tl, t, n, l, m, w)) continue;
    const int k0 = w.k0;
    T czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (k0 + e)];
      czp[e] = cz[2 * (k0 + e) + 1];
    }
    const T cym = cy[2 * w.j], cyp = cy[2 * w.j + 1];
    int64_t c = w.c;
    V rm, rc;
    __builtin_memcpy(&rm, r + c - sI, sizeof(V));
    __builtin_memcpy(&rc, r + c, sizeof(V));
    for (int i = w.i0; i < w.i1; i++, c += sI) {
      V rp, rjm, rjp;
      __builtin_memcpy(&rp, r + c + sI, sizeof(V));
      __builtin_memcpy(&rjm, r + c - sJ, sizeof(V));
      __builtin_memcpy(&rjp, r + c + sJ, sizeof(V));
      uint8_t fl[CPL];
      __builtin_memcpy(fl, flags + c, CPL);
      T left = __shfl_up(rc.v[CPL - 1], 1, 64), right = __shfl_down(rc.v[0], 1, 64);
      if (lane == 0) left = r[c - 1];
      if (lane == 63 || w.kraw + CPL > m - 2) right = r[c + CPL];
      const T cxm = cx[2 * i], cxp = cx[2 * i + 1];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const T v = cg_stencil_cell<T, CPL>(fl[e], e, rc, rm, rp, rjm.v[e], rjp.v[e], left, right, cxm, cxp, cym,
                                            cyp, czm[e], czp[e]);
        if ((fl[e] & FL_ACTUAL) && w.own[e]) {
          gam += (double)rc.v[e] * (double)rc.v[e];
          del += (double)v * (double)rc.v[e];
        }
      }
      rm = rc;
      rc = rp;
    }
  }
  gam = block_sum_all(gam, lds);
  del = block_sum_all(del, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = gam;
    part[2 * blockIdx.x + 1] = del;
  }
}

// The step of update `it` from gamma = (r, r) and delta = (A r, r) of the
// current residual: alpha, beta and the decision to stop (converged, NaN,
// iteration cap, breakdown), the same in every block.  Block 0 records it in
// the state: gamma and alpha go to the parity slot the next launch reads.
// Returns true when the solve stops here (no update).
__device__ inline bool cg_decide(double gam, double del, DevState *st, int it, int maxit, double &alpha,
                                 double &beta) {
  const int par = it & 1;
  beta = 0;
  if (it == 0) {
    alpha = gam / del;
  } else {
    beta = gam / st->s[CG_GAMMA + par];
    alpha = gam / (del - beta * gam / st->s[CG_ALPHA + par]);
  }
  int stop = 0, status = XF_OK;
  if (gam != gam || del != del) {
    stop = 1;
    status = XF_ENAN;
  } else if (gam <= st->s[HS_STOP2]) {
    stop = 1;
  } else if (it >= maxit) {
    stop = 1;
    status = XF_EMAXIT;
  } else if (!(fabs(alpha) < INFINITY) || alpha == 0) {
    stop = 1;
    status = XF_EBREAKDOWN;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) {
    st->s[HS_RR] = gam;
    if (stop) {
      st->iter = it;
      st->status = status;
      st->done = 1;
    } else {
      st->s[CG_GAMMA + (par ^ 1)] = gam;
      st->s[CG_ALPHA + (par ^ 1)] = alpha;
      st->iter = it + 1;
    }
  }
  return stop != 0;
}

// KB: reduce KA's partials, decide (converged / NaN / iteration cap /
// breakdown: the same decision in every block), then the update over the
// interior cells.  `it` is the number of updates applied before this launch;
// the first launch (it == 0) takes p = r without reading p_old.  x is updated
// on odd iterations only, with both steps in the order the eager update would
// add them, (x + a_old p_old) + a p: same bits, 16 B per cell less every other
// iteration; k_cgcg_flush adds a pending even step at the end.
template <typename T, typename X, int CPL>
__global__ __launch_bounds__(256) void k_cgcg_step(int n, int l, int m, Tiles tl, int npart,
                                                    const double *__restrict__ part, const uint8_t *__restrict__ flags,
                                                    const T *__restrict__ cx, const T *__restrict__ cy,
                                                    const T *__restrict__ cz, X *__restrict__ x,
                                                    const T *__restrict__ r, T *__restrict__ rout,
                                                    const T *__restrict__ pin, T *__restrict__ pout, DevState *st,
                                                    int it, int maxit) {
  __shared__ double lds[8];
  if (st->done) return;
  using V = CgVec<T, CPL>;
  using VX = CgVec<X, CPL>;
  double gam = 0, del = 0;
  for (int q = threadIdx.x; q < npart; q += blockDim.x) {
    gam += part[2 * q];
    del += part[2 * q + 1];
  }
  gam = block_sum_all(gam, lds);
  del = block_sum_all(del, lds);
  const int par = it & 1;
  double alpha, beta;
  if (cg_decide(gam, del, st, it, maxit, alpha, beta)) return;
  const T a = (T)alpha, bt = (T)beta;
  const bool xup = it & 1;
  const bool first = it == 0;
  const X aold = xup ? (X)st->s[CG_ALPHA + par] : X(0);
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  const int64_t sI = (int64_t)l * m, sJ = m;
  // p = r + bt p_old at one vector (p_old taken as 0 on the first iteration)
  auto pnew = [&](int64_t at, V &rv, V &po) {
    V q;
    __builtin_memcpy(&rv, r + at, sizeof(V));
    if (first) {
#pragma unroll
      for (int e = 0; e < CPL; e++) po.v[e] = T(0);
    } else {
      __builtin_memcpy(&po, pin + at, sizeof(V));
    }
#pragma unroll
    for (int e = 0; e < CPL; e++) q.v[e] = rv.v[e] + bt * po.v[e];
    return q;
  };
  auto pnew1 = [&](int64_t at) { return first ? r[at] : r[at] + bt * pin[at]; };
  // p at the rows j - 1 and j + 1 comes from the waves that own those rows,
  // through LDS (slot s holds row j0 - 1 + s of the block's rows j0 ..
  // j0 + TJ - 1; waves 0 and TJ - 1 also form the two halo rows), one buffer
  // per plane parity and one barrier per plane.  Rows past the interior keep
  // marching so that every wave reaches every barrier; the first of them (the
  // array's last row) still publishes its p for the row below it.
  __shared__ V ldsq[2][TJ + 2][64];
  const int wv = threadIdx.x >> 6;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    CgTile w;
    const bool act = cg_tile<CPL>(tl, t, n, l, m, w);  // row inside the interior
    const bool ld = w.j <= l - 1;                       // row inside the array
    const int k0 = w.k0;
    T czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (k0 + e)];
      czp[e] = cz[2 * (k0 + e) + 1];
    }
    const T cym = act ? cy[2 * w.j] : T(0), cyp = act ? cy[2 * w.j + 1] : T(0);
    int64_t c = w.c;
    V qm{}, qc{}, rc{}, pc{}, tr, tp;
    if (ld) {
      qm = pnew(c - sI, tr, tp);
      qc = pnew(c, rc, pc);
    }
    for (int i = w.i0; i < w.i1; i++, c += sI) {
      V rp{}, pp{}, rh, ph;
      V qp{};
      if (ld) qp = pnew(c + sI, rp, pp);
      const int par = (i - w.i0) & 1;
      ldsq[par][wv + 1][lane] = qc;
      if (wv == 0) ldsq[par][0][lane] = pnew(c - sJ, rh, ph);
      if (wv == TJ - 1 && act) ldsq[par][TJ + 1][lane] = pnew(c + sJ, rh, ph);
      __syncthreads();
      const V qjm = ldsq[par][wv][lane];
      const V qjp = ldsq[par][wv + 2][lane];
      if (!act) {
        qm = qc;
        qc = qp;
        continue;
      }
      VX xx;
      if (xup) __builtin_memcpy(&xx, x + c, sizeof(VX));
      uint8_t fl[CPL];
      __builtin_memcpy(fl, flags + c, CPL);
      T left = __shfl_up(qc.v[CPL - 1], 1, 64), right = __shfl_down(qc.v[0], 1, 64);
      if (lane == 0) left = pnew1(c - 1);
      if (lane == 63 || w.kraw + CPL > m - 2) right = pnew1(c + CPL);
      const T cxm = cx[2 * i], cxp = cx[2 * i + 1];
      V ro;
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const T sv = cg_stencil_cell<T, CPL>(fl[e], e, qc, qm, qp, qjm.v[e], qjp.v[e], left, right, cxm, cxp, cym,
                                             cyp, czm[e], czp[e]);
        if (xup) xx.v[e] = (xx.v[e] + aold * (X)pc.v[e]) + (X)a * (X)qc.v[e];
        ro.v[e] = rc.v[e] - a * sv;
      }
      if (w.allown) {
        __builtin_memcpy(pout + c, &qc, sizeof(V));
        __builtin_memcpy(rout + c, &ro, sizeof(V));
        if (xup) __builtin_memcpy(x + c, &xx, sizeof(VX));
      } else {
#pragma unroll
        for (int e = 0; e < CPL; e++)
          if (w.own[e]) {
            pout[c + e] = qc.v[e];
            rout[c + e] = ro.v[e];
            if (xup) x[c + e] = xx.v[e];
          }
      }
      qm = qc;
      qc = qp;
      rc = rp;
      pc = pp;
    }
    __syncthreads();  // the next tile's first plane reuses buffer 0
  }
}

// After the last KB: an even number of updates leaves the last one (even
// index, so not yet in x) pending: x += alpha_last p.
template <typename T, typename X>
__global__ __launch_bounds__(256) void k_cgcg_flush(int64_t nc, X *__restrict__ x, const T *__restrict__ p,
                                                     const DevState *st) {
  const int applied = st->iter;
  if (!(applied & 1)) return;
  const X a = (X)st->s[CG_ALPHA + (applied & 1)];
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    x[c] = x[c] + a * (X)p[c];
}

// ------------------------------------------ plain CG: one kernel per iteration
// Where one wave covers a whole row, KA's dot products move into the update
// kernel.  Lane L holds cells OFF + L CPL .. OFF + L CPL + CPL - 1 of the row:
// with OFF = 0 the whole row, ring included (m <= 64 CPL, a multiple of
// CPL); with OFF = 1 its interior (m - 2 <= 64 CPL, a multiple of CPL), the
// two ring cells beside it being zeros in p and r'.  KF applies
// update `it` with KB's arithmetic (the same bits in p, r and x) and then
// forms gamma = (r', r') and delta = (A r', r') of the new residual r' for
// update it + 1, so KA runs once, for update 0.  A r' on a row needs r' on
// the rows beside it, and r' there needs p one row further out: a block of
// W + 2 waves has one row each, rows j0 - 2 .. j0 + W - 1.  All of them form
// p; the inner W also form A p and r' (rows j0 - 1 .. j0 + W - 2), and the
// inner TJ = W - 2 of those store p and r' and count the dots.  In i the
// march runs two planes ahead of its dots (r' at plane s + 1 needs p at plane
// s + 2), starting two planes before the tile (recomputing what the
// neighbouring tile stores, with the same operations and so the same bits).
// The k-neighbours are all in the wave or on the ring (the ring cells are
// never Actual, so with OFF = 0 their own stencil values are never used).
// p and r' go to LDS once per plane each, double-buffered by plane parity,
// one barrier per plane.  The loads of plane s + 3 (r, p_old, the flag
// bytes, x) are issued before the work on plane s + 1 and consumed one plane
// later, so they are in flight across the barrier and the stencils; the flag
// bytes stay packed in one integer until the stencil takes them apart (a use
// right after the load would make the wave wait for every load before it).
// The lane's k coefficients sit in LDS, not in registers.  In FP32, p, r'
// and x are stored non-temporally (88 -> 77 us at 256^3; in FP64 the same
// time with 5 % more write traffic, so FP64 stores are plain; the loads stay
// cached: halo rows are read by two blocks).
//   reads  r, p_old (W + 2 rows per TJ stored), flags (W rows), x (odd updates)
//   writes p, r' (TJ rows), x (odd updates)              [41 B per cell, mean]
struct FTiles {
  int jt, it, ic, total;
  int blocks;  // 8 * ceil(total / 8)
};

// The flag bytes of a lane's CPL cells as one integer.
template <int CPL>
struct CgFlagWord {
  using type = typename std::conditional<CPL == 4, uint32_t, typename std::conditional<CPL == 2, uint16_t, uint8_t>::type>::type;
};
template <typename W>
__device__ inline uint8_t flag_byte(W w, int e) {
  return (uint8_t)(w >> (8 * e));
}

template <typename T, typename X, int CPL, int W, int OFF>
__global__ __launch_bounds__(64 * (W + 2)) void k_cgcg_fused(int n, int l, int m, FTiles tl, int npart,
                                                        const double *__restrict__ part_in,
                                                        double *__restrict__ part_out,
                                                        const uint8_t *__restrict__ flags, const T *__restrict__ cx,
                                                        const T *__restrict__ cy, const T *__restrict__ cz,
                                                        X *__restrict__ x, const T *__restrict__ r,
                                                        T *__restrict__ rout, const T *__restrict__ pin,
                                                        T *__restrict__ pout, DevState *st, int it, int maxit) {
  constexpr int TJ = W - 2;
  using V = CgVec<T, CPL>;
  using VX = CgVec<X, CPL>;
  __shared__ double lds[W + 2];
  __shared__ V lp[2][W + 2][64];  // p(s + 1) on rows j0 - 2 + u
  __shared__ V lr[2][W][64];      // r'(s + 1) on rows j0 - 1 + w
  __shared__ V lcz[2][64];        // the lane's k coefficients (kept out of registers)
  if (st->done) return;
  double gam = 0, del = 0;
  for (int q = threadIdx.x; q < npart; q += blockDim.x) {
    gam += part_in[2 * q];
    del += part_in[2 * q + 1];
  }
  gam = block_sum_all(gam, lds);
  del = block_sum_all(del, lds);
  double alpha, beta;
  if (cg_decide(gam, del, st, it, maxit, alpha, beta)) return;
  const T a = (T)alpha, bt = (T)beta;
  const bool xup = it & 1;
  const bool first = it == 0;
  const X aold = xup ? (X)st->s[CG_ALPHA + (it & 1)] : X(0);
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63, u = threadIdx.x >> 6, w = u - 1;
  const bool dotrow = u >= 1 && u <= W;  // waves 0 and W + 1 only form p on the halo rows
  const int64_t sI = (int64_t)l * m;
  const bool kin = lane * CPL < m - 2 * OFF;
  const int k0 = kin ? OFF + lane * CPL : m - OFF - CPL;
  const bool kfirst = OFF && lane == 0, klast = OFF && lane == (m - 2) / CPL - 1;
  bool kown[CPL];
  bool kall = kin;
#pragma unroll
  for (int e = 0; e < CPL; e++) {
    kown[e] = kin && k0 + e >= 1 && k0 + e <= m - 2;
    kall = kall && kown[e];
  }
  if (u == 0) {
    V zm, zp;
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      zm.v[e] = cz[2 * (k0 + e)];
      zp.v[e] = cz[2 * (k0 + e) + 1];
    }
    lcz[0][lane] = zm;
    lcz[1][lane] = zp;
  }
  __syncthreads();
  auto pl = [&](int i) { return (int64_t)min(max(i, 0), n - 1) * sI; };
  // FP32 stores stream past the caches; FP64 ones are plain (same time, less write traffic)
  auto put_row = [](auto *dst, const void *src, int bytes) {
    if constexpr (sizeof(T) == 4) store_stream(dst, src, bytes);
    else __builtin_memcpy(dst, src, bytes);
  };
  double g2 = 0, d2 = 0;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    const int jtile = t % tl.jt, itile = t / tl.jt;
    const int j0 = 1 + jtile * TJ;
    const int jrow = j0 - 2 + u;
    const int jl = min(max(jrow, 0), l - 1);
    const bool rown = w >= 1 && w <= TJ && jrow <= l - 2;
    const int i0 = 1 + itile * tl.ic, i1 = min(i0 + tl.ic, n - 1);
    const T cym = cy[2 * jl], cyp = cy[2 * jl + 1];
    const int64_t rowc = (int64_t)jl * m + k0;
    // raw loads of one plane: r, p_old (zero on the first update), the
    // flags (rows with dots) and x (stored rows on odd updates)
    // the flags stay packed (one byte per cell) until the stencil takes them
    // apart, so that no use right after their load makes it wait
    using FW = typename CgFlagWord<CPL>::type;
    V nr, np;
    FW nf = 0;
    VX nx;
    auto fetch = [&](int i) {
      const int64_t c = pl(i) + rowc;
      __builtin_memcpy(&nr, r + c, sizeof(V));
      if (first) {
#pragma unroll
        for (int e = 0; e < CPL; e++) np.v[e] = T(0);
      } else {
        __builtin_memcpy(&np, pin + c, sizeof(V));
      }
      if (dotrow) __builtin_memcpy(&nf, flags + c, CPL);
      if (xup && rown && i >= i0 && i < i1) __builtin_memcpy(&nx, x + c, sizeof(VX));
    };
    V pA, pB, rB, qA{}, qB{};
    FW fA = 0, fB = 0;
    fetch(i0 - 2);
#pragma unroll
    for (int e = 0; e < CPL; e++) pA.v[e] = nr.v[e] + bt * np.v[e];
    fetch(i0 - 1);
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      rB.v[e] = nr.v[e];
      pB.v[e] = nr.v[e] + bt * np.v[e];
    }
    fB = nf;
    fetch(i0);
    for (int s = i0 - 2; s < i1; s++) {
      const int par = (s - i0) & 1;
      // 1. plane s + 2 from the registers loaded one plane ago, and x there
      // (two steps at once) before those registers take the next plane
      V rC, pC;
      const FW fC = nf;
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        rC.v[e] = nr.v[e];
        pC.v[e] = nr.v[e] + bt * np.v[e];
      }
      if (xup && rown && s + 2 >= i0 && s + 2 < i1) {
        const int64_t c = pl(s + 2) + rowc;
        VX xx;
#pragma unroll
        for (int e = 0; e < CPL; e++) xx.v[e] = (nx.v[e] + aold * (X)np.v[e]) + (X)a * (X)pC.v[e];
        if (kall) {
          put_row(x + c, &xx, sizeof(VX));
        } else {
#pragma unroll
          for (int e = 0; e < CPL; e++)
            if (kown[e]) x[c + e] = xx.v[e];
        }
      }
      // 2. loads of plane s + 3 stay in flight over this plane's work
      if (s + 3 <= i1 + 1) fetch(s + 3);
      // 4. p(s + 1) of every row to LDS
      lp[par][u][lane] = pB;
      __syncthreads();
      V qC{};
      if (dotrow) {
        // 5. A p and r' at plane s + 1
        {
          const V pjm = lp[par][w][lane], pjp = lp[par][w + 2][lane];
          T left = __shfl_up(pB.v[CPL - 1], 1, 64), right = __shfl_down(pB.v[0], 1, 64);
          if (kfirst) left = T(0);
          if (klast) right = T(0);
          const int ip = min(max(s + 1, 0), n - 1);
          const T cxm = cx[2 * ip], cxp = cx[2 * ip + 1];
          const V czm = lcz[0][lane], czp = lcz[1][lane];
#pragma unroll
          for (int e = 0; e < CPL; e++) {
            const T sv = cg_stencil_cell<T, CPL>(flag_byte(fB, e), e, pB, pA, pC, pjm.v[e], pjp.v[e], left, right, cxm, cxp,
                                                 cym, cyp, czm.v[e], czp.v[e]);
            qC.v[e] = rB.v[e] - a * sv;
          }
        }
        if (rown && kin && s + 1 >= i0 && s + 1 < i1) {
          const int64_t c = pl(s + 1) + rowc;
          put_row(pout + c, &pB, sizeof(V));
          put_row(rout + c, &qC, sizeof(V));
        }
        // 6. r'(s + 1) to LDS; dots at plane s from r'(s) on the rows beside
        lr[par][w][lane] = qC;
        if (s >= i0 && rown) {
          const V qjm = lr[par ^ 1][w - 1][lane], qjp = lr[par ^ 1][w + 1][lane];
          T left = __shfl_up(qB.v[CPL - 1], 1, 64), right = __shfl_down(qB.v[0], 1, 64);
          if (kfirst) left = T(0);
          if (klast) right = T(0);
          const T cxm = cx[2 * s], cxp = cx[2 * s + 1];
          const V czm = lcz[0][lane], czp = lcz[1][lane];
#pragma unroll
          for (int e = 0; e < CPL; e++) {
            const uint8_t fa = flag_byte(fA, e);
            const T v = cg_stencil_cell<T, CPL>(fa, e, qB, qA, qC, qjm.v[e], qjp.v[e], left, right, cxm, cxp,
                                                cym, cyp, czm.v[e], czp.v[e]);
            if ((fa & FL_ACTUAL) && kown[e]) {
              g2 += (double)qB.v[e] * (double)qB.v[e];
              d2 += (double)v * (double)qB.v[e];
            }
          }
        }
      }
      pA = pB;
      pB = pC;
      rB = rC;
      qA = qB;
      qB = qC;
      fA = fB;
      fB = fC;
    }
    __syncthreads();  // the next tile reuses both LDS buffers
  }
  g2 = block_sum_all(g2, lds);
  d2 = block_sum_all(d2, lds);
  if (threadIdx.x == 0) {
    part_out[2 * blockIdx.x] = g2;
    part_out[2 * blockIdx.x + 1] = d2;
  }
}

// ------------------------------------- refinement residual and mean projection
// Fence-free reductions: each kernel writes kRefPart doubles per block into
// the context's partial buffer, and k_ref_finish (one block) adds them in
// block order into state slots.
constexpr int kRefPart = 4;      // partial doubles per block
constexpr int kRefBlocks = 2048; // blocks of the refinement kernels (partials fit the context buffer)

// r = b - A x over the interior with the fast rows (x read everywhere, so
// boundary values of Defined neighbours enter as in the reference), written
// as type R (0 off the Actual rows).  Partials per block: (r, r) in FP64
// before rounding, the sum of the rounded r and (R r, R r), over Actual
// cells.  Blocks are persistent over the tiles, consecutive tiles on one XCD.
template <typename R>
__global__ __launch_bounds__(256) void k_residual_fast(int n, int l, int m, Tiles tl,
                                                        const uint8_t *__restrict__ flags,
                                                        const double *__restrict__ cx, const double *__restrict__ cy,
                                                        const double *__restrict__ cz, const double *__restrict__ b,
                                                        const double *__restrict__ x, R *__restrict__ rout,
                                                        double *__restrict__ part) {
  __shared__ double lds[8];
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  double rr = 0, sum = 0, rr2 = 0;
  const int64_t sI = (int64_t)l * m, sJ = m;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    const int ktile = t % tl.kt;
    const int jtile = (t / tl.kt) % tl.jt;
    const int itile = t / (tl.kt * tl.jt);
    const int k = 1 + ktile * TK + (threadIdx.x & 63);
    const int j = 1 + jtile * TJ + (threadIdx.x >> 6);
    const int i0 = 1 + itile * tl.ic;
    const int i1 = min(i0 + tl.ic, n - 1);
    if (k > m - 2 || j > l - 2) continue;
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
        const double vr = (double)(R)v;
        rr += v * v;
        sum += vr;
        rr2 += vr * vr;
      }
      rout[c] = (R)v;
      pm = pc;
      pc = pp;
    }
  }
  rr = block_sum_all(rr, lds);
  sum = block_sum_all(sum, lds);
  rr2 = block_sum_all(rr2, lds);
  if (threadIdx.x == 0) {
    part[kRefPart * blockIdx.x] = rr;
    part[kRefPart * blockIdx.x + 1] = sum;
    part[kRefPart * blockIdx.x + 2] = rr2;
  }
}

// v -= s[slot]/count at unknowns; partials of (v, v) over unknowns after the shift
template <typename T>
__global__ __launch_bounds__(256) void k_sub_mean_sq(int64_t nc, const uint8_t *__restrict__ flags,
                                                      T *__restrict__ v, const DevState *st, int slot, double count,
                                                      double *__restrict__ part) {
  __shared__ double lds[8];
  const double mean = st->s[slot] / count;
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) {
      const T nv = (T)((double)v[c] - mean);
      v[c] = nv;
      acc += (double)nv * (double)nv;
    }
  acc = block_sum_all(acc, lds);
  if (threadIdx.x == 0) part[kRefPart * blockIdx.x] = acc;
}

// out = (float)(r - s[slot]/count) at unknowns, out = (float)r elsewhere;
// partials of the projected residual's (r', r') in FP64 and of the rounded
// (out, out), over unknowns.  Removing the mean before rounding keeps the
// digits of a residual whose mean is large next to its mean-free part.
__global__ __launch_bounds__(256) void k_sub_mean_cast(int64_t nc, const uint8_t *__restrict__ flags,
                                                       const double *__restrict__ r, float *__restrict__ out,
                                                       const DevState *st, int slot, double count,
                                                       double *__restrict__ part) {
  __shared__ double lds[8];
  const double mean = st->s[slot] / count;
  double a64 = 0, a32 = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    double v = r[c];
    if (flags[c] & FL_UNKNOWN) {
      v -= mean;
      const float f = (float)v;
      a64 += v * v;
      a32 += (double)f * (double)f;
      out[c] = f;
    } else {
      out[c] = (float)v;
    }
  }
  a64 = block_sum_all(a64, lds);
  a32 = block_sum_all(a32, lds);
  if (threadIdx.x == 0) {
    part[kRefPart * blockIdx.x] = a64;
    part[kRefPart * blockIdx.x + 1] = a32;
  }
}

// partials of (v, v) over unknowns
template <typename T>
__global__ __launch_bounds__(256) void k_sq_unknowns(int64_t nc, const uint8_t *__restrict__ flags,
                                                      const T *__restrict__ v, const DevState *st,
                                                      double *__restrict__ part) {
  __shared__ double lds[8];
  if (st && st->done) return;
  double acc = 0;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x)
    if (flags[c] & FL_UNKNOWN) acc += (double)v[c] * (double)v[c];
  acc = block_sum_all(acc, lds);
  if (threadIdx.x == 0) part[kRefPart * blockIdx.x] = acc;
}

// s[slot q] = sum over blocks of partial column q (slot < 0: column unused)
__global__ void k_ref_finish(const double *__restrict__ part, int np, DevState *st, int slot0, int slot1,
                             int slot2) {
  __shared__ double lds[16];
  const int slots[3] = {slot0, slot1, slot2};
  for (int q = 0; q < 3; q++) {
    if (slots[q] < 0) continue;
    double s = 0;
    for (int b = threadIdx.x; b < np; b += blockDim.x) s += part[kRefPart * b + q];
    s = block_sum(s, lds);
    if (threadIdx.x == 0) st->s[slots[q]] = s;
  }
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

// Planes per tile of the two plain-CG kernels: the deepest tile (at most
// kCgMaxPlanes) that still leaves >= kCgMinTiles tiles, so a launch keeps 4096
// waves (the step kernel's 4 waves per SIMD on 256 CUs).  Timing lab at 256^3
// (a solve's dots/step sequence): FP64 32 planes for both, 56 + 180 us
// (16/8 planes: 60 + 195 us); FP32 16 planes, 32 + 90 us.
template <typename T>
struct CgShape;
template <>
struct CgShape<double> {
  static constexpr int max_planes = 32;
};
template <>
struct CgShape<float> {
  static constexpr int max_planes = 16;
};
constexpr int kCgMinTiles = 1024;
constexpr int kCgStepBlocks = 4096;  // KB blocks (persistent over tiles)

static Tiles march_tiles(xf_grid g, int cpl, int max_planes) {
  Tiles t;
  t.kt = (g->m - 2 + 64 * cpl - 1) / (64 * cpl);
  t.jt = (g->l - 2 + TJ - 1) / TJ;
  for (t.ic = max_planes;; t.ic /= 2) {
    t.it = (g->n - 2 + t.ic - 1) / t.ic;
    t.total = t.kt * t.jt * t.it;
    if (t.total >= kCgMinTiles || t.ic <= 2) break;
  }
  t.blocks = 8 * ((t.total + 7) / 8);
  return t;
}

template <typename T, typename X, int CPL>
static void cgcg_run(xf_ctx c, xf_grid g, X *x, T *const R[2], T *const P[2], const T *cx, const T *cy,
                     const T *cz, double *part, int maxit, int every) {
  const Tiles ta = march_tiles(g, CPL, CgShape<T>::max_planes);
  const Tiles &tb = ta;
  const int nbA = std::min(ta.blocks, kCgTileBlocks);
  const int nbB = std::min(tb.blocks, kCgStepBlocks);
  const bool f64 = sizeof(T) == 8;
  const char *n1 = f64 ? "cg_dots_fp64" : "cg_dots_fp32";
  const char *n2 = f64 ? "cg_step_fp64" : "cg_step_fp32";
  // per-kernel events on two of every kProfEvery iterations only (markers
  // between kernels cost a few microseconds each)
  constexpr int kProfEvery = 8;
  std::vector<ProfMark> marks;
  std::vector<int> sampled;  // iteration index of each sampled (KA, KB) pair of this batch
  int it = 0;
  for (;;) {
    for (int q = 0; q < every; q++, it++) {
      const bool pm = c->prof && it % kProfEvery < 2;  // an even and an odd update
      T *rin = R[it & 1], *rout = R[(it + 1) & 1];
      T *pin = P[it & 1], *pout = P[(it + 1) & 1];
      int e1 = pm ? prof_begin(c) : -1;
      hipLaunchKernelGGL((k_cgcg_dots<T, CPL>), dim3(nbA), dim3(256), 0, c->stream, g->n, g->l, g->m, ta,
                         g->d_flags, cx, cy, cz, rin, c->state, part);
      if (pm) {
        prof_end(c, e1);
        marks.push_back({n1, e1});
      }
      int e2 = pm ? prof_begin(c) : -1;
      hipLaunchKernelGGL((k_cgcg_step<T, X, CPL>), dim3(nbB), dim3(256), 0, c->stream, g->n, g->l, g->m, tb,
                         nbA, part, g->d_flags, cx, cy, cz, x, rin, rout, pin, pout, c->state, it, maxit);
      if (pm) {
        prof_end(c, e2);
        marks.push_back({n2, e2});
        sampled.push_back(it);
      }
    }
    check_kernel("cg iterations");
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) {
      // every KA up to the deciding one ran in full; the deciding KB launch
      // only reduced, so only the updates before it are counted
      const long long ran = c->h_state->iter, done = c->h_state->done ? 1 : 0;
      long long na = 0, nb = 0;
      for (int v : sampled) {
        if (v < ran + done) na++;
        if (v < ran) nb++;
      }
      prof_collect(c, marks, {{n1, na}, {n2, nb}});
      sampled.clear();
    }
    if (c->h_state->done) break;
  }
  // update index applied-1 wrote its direction into P[applied & 1]
  const int applied = c->h_state->iter;
  if (applied & 1)
    hipLaunchKernelGGL((k_cgcg_flush<T, X>), dim3(stream_nb(g->ncell)), dim3(256), 0, c->stream, g->ncell, x,
                       P[applied & 1], c->state);
  check_kernel("cg flush");
}

// KF's shape: W + 2 waves per block (TJ = W - 2 rows stored), about one
// tile per resident block.  Timing lab at 256^3 (258^3 array): FP64 at 4
// cells per lane needs 163 VGPRs, three waves per SIMD, so one block of 12
// waves per CU (32 x 8 tiles, 172 us per iteration; 12 rows: 191 us; 8 rows:
// 214 us, both before the halo waves and the prefetch); FP32 14 rows, 16
// waves per block (22 x 11 tiles, 88 us).  A block has at most 16 waves.
template <typename T, int CPL>
constexpr int fused_rows() {
  return CPL == 4 && sizeof(T) == 8 ? 10 : 14;
}
constexpr int kFusedBlocks = 1024;  // KF blocks at most: two partial buffers of this many fit the KA buffer

// The vector width (CPL) and offset (OFF) at which one wave covers a whole
// row: 2 CPL + OFF, or 0 where none does.
static int fused_layout(int m) {
  for (int cpl = 1; cpl <= 4; cpl *= 2) {
    if (m - 2 <= 64 * cpl && (m - 2) % cpl == 0) return 2 * cpl + 1;
    if (m <= 64 * cpl && m % cpl == 0) return 2 * cpl;
  }
  return 0;
}

template <typename T, int CPL>
static int dots_launch(xf_ctx c, xf_grid g, const T *cx, const T *cy, const T *cz, const T *r, double *part,
                       int maxblocks) {
  const Tiles ta = march_tiles(g, CPL, CgShape<T>::max_planes);
  const int nb = std::min(ta.blocks, maxblocks);
  hipLaunchKernelGGL((k_cgcg_dots<T, CPL>), dim3(nb), dim3(256), 0, c->stream, g->n, g->l, g->m, ta, g->d_flags, cx,
                     cy, cz, r, c->state, part);
  return nb;
}

// KA over r with the vector width the plain-CG kernels use for this row length.
template <typename T>
static int dots_once(xf_ctx c, xf_grid g, const T *cx, const T *cy, const T *cz, const T *r, double *part,
                     int maxblocks) {
  if constexpr (sizeof(T) == 4) {
    if (g->m - 2 >= 64 * 4) return dots_launch<T, 4>(c, g, cx, cy, cz, r, part, maxblocks);
  }
  if (g->m - 2 >= 64 * 2) return dots_launch<T, 2>(c, g, cx, cy, cz, r, part, maxblocks);
  return dots_launch<T, 1>(c, g, cx, cy, cz, r, part, maxblocks);
}

template <typename T, typename X, int CPL, int OFF>
static void cgcg_run_fused(xf_ctx c, xf_grid g, X *x, T *const R[2], T *const P[2], const T *cx, const T *cy,
                           const T *cz, double *part, int maxit, int every) {
  constexpr int W = fused_rows<T, CPL>(), TJ = W - 2;
  static const int resident = [] {
    int dev = 0, cus = 0, per = 0;
    XF_HIP(hipGetDevice(&dev));
    XF_HIP(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    XF_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_cgcg_fused<T, X, CPL, W, OFF>, 64 * (W + 2), 0));
    return std::max(1, cus * std::max(per, 1));
  }();
  FTiles tf;
  tf.jt = (g->l - 2 + TJ - 1) / TJ;
  const int want_it = std::max(1, resident / tf.jt);
  tf.ic = (g->n - 2 + want_it - 1) / want_it;
  tf.it = (g->n - 2 + tf.ic - 1) / tf.ic;
  tf.total = tf.jt * tf.it;
  tf.blocks = 8 * ((tf.total + 7) / 8);
  const int nbF = std::min(tf.blocks, kFusedBlocks);
  double *const pb[2] = {part, part + 2 * kFusedBlocks};
  const bool f64 = sizeof(T) == 8;
  const char *n1 = f64 ? "cg_dots_fp64" : "cg_dots_fp32";
  const char *n2 = f64 ? "cg_fused_fp64" : "cg_fused_fp32";
  constexpr int kProfEvery = 8;
  std::vector<ProfMark> marks;
  std::vector<int> sampled;
  int e0 = c->prof ? prof_begin(c) : -1;
  const int nbA = dots_once<T>(c, g, cx, cy, cz, R[0], pb[0], kFusedBlocks);
  if (c->prof) {
    prof_end(c, e0);
    marks.push_back({n1, e0});
  }
  long long na = 1;
  int it = 0;
  // Without profiling the host checks each batch's state one batch late:
  // the next batch is already queued when it waits, so the GPU does not idle
  // for the round trip (launches after convergence return at once).
  const bool ahead = !c->prof;
  DevState *const hs[2] = {c->h_state, c->h_state2};
  int batch = 0, pending = -1, final_slot = 0;
  for (;;) {
    for (int q = 0; q < every; q++, it++) {
      const bool pm = c->prof && it % kProfEvery < 2;
      int e1 = pm ? prof_begin(c) : -1;
      hipLaunchKernelGGL((k_cgcg_fused<T, X, CPL, W, OFF>), dim3(nbF), dim3(64 * (W + 2)), 0, c->stream, g->n, g->l, g->m, tf,
                         it == 0 ? nbA : nbF, pb[it & 1], pb[(it + 1) & 1], g->d_flags, cx, cy, cz, x, R[it & 1],
                         R[(it + 1) & 1], P[it & 1], P[(it + 1) & 1], c->state, it, maxit);
      if (pm) {
        prof_end(c, e1);
        marks.push_back({n2, e1});
        sampled.push_back(it);
      }
    }
    check_kernel("cg iterations");
    if (ahead) {
      const int slot = batch & 1;
      XF_HIP(hipMemcpyAsync(hs[slot], c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
      XF_HIP(hipEventRecord(c->evp[slot], c->stream));
      batch++;
      if (pending >= 0) {
        XF_HIP(hipEventSynchronize(c->evp[pending]));
        if (hs[pending]->done) {
          final_slot = pending;
          break;
        }
      }
      pending = slot;
      continue;
    }
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (c->prof) {
      // the deciding launch only reduced: the updates before it are counted
      const long long ran = c->h_state->iter;
      long long nf = 0;
      for (int v : sampled)
        if (v < ran) nf++;
      prof_collect(c, marks, {{n1, na}, {n2, nf}});
      na = 0;
      sampled.clear();
    }
    if (c->h_state->done) break;
  }
  if (ahead) {  // the batch queued after the deciding one has returned; its copy is the same state
    XF_HIP(hipStreamSynchronize(c->stream));
    if (final_slot != 0) *c->h_state = *hs[final_slot];
  }
  const int applied = c->h_state->iter;
  if (applied & 1)
    hipLaunchKernelGGL((k_cgcg_flush<T, X>), dim3(stream_nb(g->ncell)), dim3(256), 0, c->stream, g->ncell, x,
                       P[applied & 1], c->state);
  check_kernel("cg flush");
}

// Runs single-reduction CG iterations on A e = r from p = r (r zero-padded,
// in r0; x the FP64 or FP32 accumulator; r1, p0, p1 work vectors, zero on
// the array's outer ring; part the dot kernel's partials).  State
// must hold HS_STOP2 and done = 0 (k_hs_start).  Polls the device state every
// `every` iterations; h_state->iter counts the updates applied.  The final
// residual is left in r0 or r1 (by the parity of the update count).
template <typename T, typename X>
static void cgcg_iterations(xf_ctx c, xf_grid g, X *x, T *r0, T *r1, T *p0, T *p1, const T *cx, const T *cy,
                            const T *cz, double *part, int maxit, int every) {
  T *const R[2] = {r0, r1};
  T *const P[2] = {p0, p1};
  // one kernel per iteration where a wave holds a whole row
  switch (fused_layout(g->m)) {
    case 2: return cgcg_run_fused<T, X, 1, 0>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    case 3: return cgcg_run_fused<T, X, 1, 1>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    case 4: return cgcg_run_fused<T, X, 2, 0>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    case 5: return cgcg_run_fused<T, X, 2, 1>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    case 8: return cgcg_run_fused<T, X, 4, 0>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    case 9: return cgcg_run_fused<T, X, 4, 1>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
    default: break;
  }
  // 16-byte vectors per lane where a row holds a whole wave of them
  if constexpr (sizeof(T) == 4) {
    if (g->m - 2 >= 64 * 4) return cgcg_run<T, X, 4>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
  }
  if (g->m - 2 >= 64 * 2) return cgcg_run<T, X, 2>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
  cgcg_run<T, X, 1>(c, g, x, R, P, cx, cy, cz, part, maxit, every);
}

// r = b - A x (fast rows) into rout as type R: s[HS_AUX] = (r, r) in FP64,
// s[HS_SUM] = sum of the rounded r, s[HS_RRW] = (R r, R r), all over unknowns.
template <typename R>
static void residual_fast(xf_ctx c, xf_grid g, const double *b, const double *x, R *rout) {
  const Tiles tl = tiles(g);
  const int nb = std::min(tl.blocks, kRefBlocks);
  hipLaunchKernelGGL((k_residual_fast<R>), dim3(nb), dim3(256), 0, c->stream, g->n, g->l, g->m, tl, g->d_flags,
                     g->d_cx, g->d_cy, g->d_cz, b, x, rout, c->partials);
  hipLaunchKernelGGL(k_ref_finish, dim3(1), dim3(256), 0, c->stream, c->partials, nb, c->state, (int)HS_AUX,
                     (int)HS_SUM, (int)HS_RRW);
}

// v -= mean of v over the unknowns (the mean from s[HS_SUM]); s[HS_RRW] = (v, v) after
template <typename T>
static void project_mean(xf_ctx c, xf_grid g, T *v) {
  const int nb = std::min(stream_nb(g->ncell), kRefBlocks);
  hipLaunchKernelGGL((k_sub_mean_sq<T>), dim3(nb), dim3(256), 0, c->stream, g->ncell, g->d_flags, v, c->state,
                     (int)HS_SUM, (double)g->nunk, c->partials);
  hipLaunchKernelGGL(k_ref_finish, dim3(1), dim3(256), 0, c->stream, c->partials, nb, c->state, (int)HS_RRW, -1, -1);
}

// Refinement residual as FP32: r = b - A x formed in FP64 and, with the mean
// projection, its mean over the unknowns removed in FP64 before rounding
// (r64: an FP64 work field).  s[HS_AUX] = the stopping norm (r, r) in FP64 of
// the (projected) residual, s[HS_RRW] = (r32, r32).
static void residual_refine(xf_ctx c, xf_grid g, const double *b, const double *x, bool proj, double *r64,
                            float *r32) {
  if (!proj) {
    residual_fast<float>(c, g, b, x, r32);
    return;
  }
  residual_fast<double>(c, g, b, x, r64);
  const int nb = std::min(stream_nb(g->ncell), kRefBlocks);
  hipLaunchKernelGGL(k_sub_mean_cast, dim3(nb), dim3(256), 0, c->stream, g->ncell, g->d_flags, r64, r32, c->state,
                     (int)HS_SUM, (double)g->nunk, c->partials);
  hipLaunchKernelGGL(k_ref_finish, dim3(1), dim3(256), 0, c->stream, c->partials, nb, c->state, (int)HS_AUX,
                     (int)HS_RRW, -1);
}

// s[slot] = (v, v) over the unknowns
template <typename T>
static void sq_unknowns(xf_ctx c, xf_grid g, const T *v, int slot) {
  const int nb = std::min(stream_nb(g->ncell), kRefBlocks);
  hipLaunchKernelGGL((k_sq_unknowns<T>), dim3(nb), dim3(256), 0, c->stream, g->ncell, g->d_flags, v,
                     (const DevState *)nullptr, c->partials);
  hipLaunchKernelGGL(k_ref_finish, dim3(1), dim3(256), 0, c->stream, c->partials, nb, c->state, slot, -1, -1);
}

static void upload_state_zero(xf_ctx c) {
  XF_HIP(hipMemsetAsync(c->state, 0, sizeof(DevState), c->stream));
}

// Fill and copy kernels for the path that must not record memset or memcpy
// nodes (callers that capture the stream into a graph).
template <typename T>
__global__ __launch_bounds__(256) void k_fill(int64_t n, T *__restrict__ v, T val) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    v[i] = val;
}
__global__ __launch_bounds__(256) void k_copy_f64(int64_t n, double *__restrict__ dst,
                                                  const double *__restrict__ src) {
  for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

// The vectors of one plain FP64 CG solve: r0/r1 and p0/p1 ping-pong pairs
// (zero on the array's outer ring) and the dot kernel's partials.
struct CgVecs {
  double *r0, *r1, *p0, *p1, *part;
};

// Plain FP64 CG from x0 on b - A x0.  kernels_only: state cleared and x0
// copied by kernels instead of memset / memcpy calls.
static void hs_cg_core(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                       xf_solve_stats &s, const CgVecs &v, bool kernels_only) {
  const int64_t nc = g->ncell;
  if (kernels_only) {
    hipLaunchKernelGGL(k_fill<int>, dim3(1), dim3(64), 0, c->stream, (int64_t)(sizeof(DevState) / sizeof(int)),
                       (int *)c->state, 0);
    hipLaunchKernelGGL(k_copy_f64, dim3(stream_nb(nc)), dim3(256), 0, c->stream, nc, x, x0);
  } else {
    upload_state_zero(c);
    XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  }
  residual_fast<double>(c, g, b, x, v.r0);
  if (o.project_mean && g->nunk > 0) project_mean<double>(c, g, v.r0);
  check_kernel("cg start");
  hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, o.tol, o.tol_mode == XF_TOL_REL ? 1 : 0,
                     (int)HS_RRW);
  c->h_state->iter = 0;
  cgcg_iterations<double, double>(c, g, x, v.r0, v.r1, v.p0, v.p1, g->d_cx, g->d_cy, g->d_cz, v.part, o.maxit,
                                  o.check_every > 0 ? o.check_every : 16);
  s.iters = c->h_state->iter;
  s.status = c->h_state->status;
  s.resid = c->h_state->s[HS_RR0] > 0 ? sqrt(c->h_state->s[HS_RR] / c->h_state->s[HS_RR0]) : 0;
}

void solve_hs_cg(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                 xf_solve_stats &s) {
  require_fast(g);
  CgWork *w = work(c, g, false);
  hs_cg_core(c, g, b, x0, x, o, s, CgVecs{w->r, w->q, w->p[0], w->p[1], w->part}, false);
}

// --------------------------------------------- plain FP64 CG without allocation
// The work vectors of xf_solve_noalloc live in one fixed array per device,
// held from the library's load to the end of the process.  Host code never
// uses the array's name as an address: xf_init looks up each device's copy
// with hipGetSymbolAddress (that device current) and keeps the pointer.
// Layout: partials [kCgTileBlocks * 2], then r0, r1, p0, p1 [ncell] each.
constexpr size_t kScratchBytes = size_t(72) << 20;
static __device__ __align__(256) double scratch[kScratchBytes / sizeof(double)];
constexpr size_t kScratchPartDoubles = 2 * kCgTileBlocks;
constexpr int kMaxDevices = 64;
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[kMaxDevices];

// Called by xf_init with `device` current.
void noalloc_setup(int device) {
  XF_REQUIRE(device >= 0 && device < kMaxDevices, XF_EINVAL, "xf_init: device index beyond the library's table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (g_scratch_ptr[device]) return;
  void *p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  XF_REQUIRE(e == hipSuccess && p, XF_EHIP,
             std::string("xf_init: hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
  g_scratch_ptr[device] = (double *)p;
}

int64_t noalloc_max_cells() {
  return (int64_t)((kScratchBytes / sizeof(double) - kScratchPartDoubles) / 4);
}
static_assert((kScratchBytes / sizeof(double) - kScratchPartDoubles) / 4 == XF_NOALLOC_MAX_CELLS,
              "XF_NOALLOC_MAX_CELLS must state the work array's capacity");

void solve_hs_cg_noalloc(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x,
                         const xf_solve_opts &o, xf_solve_stats &s, double **verify_tmp) {
  XF_REQUIRE(o.method == XF_CG, XF_EINVAL, "xf_solve_noalloc: only method XF_CG (plain FP64 CG) is supported");
  require_fast(g);
  XF_REQUIRE(g->ncell <= noalloc_max_cells(), XF_EINVAL,
             "xf_solve_noalloc: grid of " + std::to_string(g->ncell) + " cells exceeds the fixed work array (" +
                 std::to_string(noalloc_max_cells()) + " cells, 72 MiB)");
  double *base;
  {
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    base = (c->device >= 0 && c->device < kMaxDevices) ? g_scratch_ptr[c->device] : nullptr;
  }
  XF_REQUIRE(base, XF_EHIP, "xf_solve_noalloc: the work array of this device was not set up by xf_init");
  const int64_t nc = g->ncell;
  CgVecs v;
  v.part = base;
  v.r0 = base + kScratchPartDoubles;
  v.r1 = v.r0 + nc;
  v.p0 = v.r1 + nc;
  v.p1 = v.p0 + nc;
  // the last call's contents are still there: the rings of r and p must be 0
  hipLaunchKernelGGL(k_fill<double>, dim3(stream_nb(4 * nc)), dim3(256), 0, c->stream, 4 * nc, v.r0, 0.0);
  // per-kernel event sampling creates events: off on this path
  const bool prof = c->prof;
  c->prof = false;
  struct Restore {
    xf_ctx c;
    bool p;
    ~Restore() { c->prof = p; }
  } rs{c, prof};
  hs_cg_core(c, g, b, x0, x, o, s, v, true);
  *verify_tmp = v.r0;  // free after the solve: the after-solve check may use it
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
  const int every = o.check_every > 0 ? o.check_every : 16;
  const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  upload_state_zero(c);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  std::vector<ProfMark> marks;
  // r = b - A x in FP64, rounded to FP32; with project_mean its mean over the
  // unknowns is removed first and the stopping norm is that of the projected
  // residual (the plain CG's rule)
  const bool proj = o.project_mean && g->nunk > 0;
  auto residual = [&]() {
    int e = c->prof ? prof_begin(c) : -1;
    residual_refine(c, g, b, x, proj, w->r, w->r32);
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
    XF_HIP(hipMemsetAsync(w->e32, 0, sizeof(float) * nc, c->stream));
    // the last sweep only needs the factor that is still missing (with margin)
    const double itol = std::min(0.5, std::max(inner_tol, 0.3 * sqrt(stop2 / rr)));
    hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, itol, 1, (int)HS_RRW);
    check_kernel("mp inner start");
    c->h_state->iter = 0;
    cgcg_iterations<float, float>(c, g, w->e32, w->r32, w->q32, w->p32[0], w->p32[1], g->d_cxf, g->d_cyf, g->d_czf,
                                  w->part, o.maxit - inner_total, every);
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

// The flexible-CG step in fence-free kernels: each reduction kernel writes
// one partial (pair) per block, a one-block finisher adds them in block order
// and applies the scalar step (no device-scope fences in the streaming
// passes).

// The step's three passes are marching stencil kernels over the interior (the
// plain-CG tiles), q = A p is never stored: each pass that needs A p (or A
// p_old) forms it from p at the seven stencil points, with the same operations
// as everywhere else, so every value of A p is bit-identical between passes.
// p is ping-ponged (the kernel writing p_new never overwrites a p_old a
// neighbouring block still reads).  FP32 bytes per interior cell: K1 9, K2 17,
// K3 21 (the former four streaming passes with q stored moved 57).
//   K1 k_fcg_zq      partials of (z, A p_old)                    [reads z, p_old]
//   K2 k_fcg_dir_pq  beta = -(z, A p_old)/(p_old, A p_old) (first: 0),
//                    p = z + beta p_old, partials of (p, r), (p, A p)
//                                                  [reads z, p_old, r; writes p]
//   K3 k_fcg_upd     e += alpha p, r -= alpha A p, partials of (r, r)
//                                                  [reads e, p, r; writes e, r]
template <int CPL>
__global__ __launch_bounds__(256) void k_fcg_zq(int n, int l, int m, Tiles tl, const uint8_t *__restrict__ flags,
                                                 const float *__restrict__ cx, const float *__restrict__ cy,
                                                 const float *__restrict__ cz, const float *__restrict__ z,
                                                 const float *__restrict__ p, const DevState *st,
                                                 double *__restrict__ part) {
  __shared__ double lds[8];
  if (st->done) return;
  using V = CgVec<float, CPL>;
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  double acc = 0;
  const int64_t sI = (int64_t)l * m, sJ = m;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    CgTile w;
    if (!cg_tile<CPL>(tl, t, n, l, m, w)) continue;
    float czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (w.k0 + e)];
      czp[e] = cz[2 * (w.k0 + e) + 1];
    }
    const float cym = cy[2 * w.j], cyp = cy[2 * w.j + 1];
    int64_t c = w.c;
    V pm, pc;
    __builtin_memcpy(&pm, p + c - sI, sizeof(V));
    __builtin_memcpy(&pc, p + c, sizeof(V));
    for (int i = w.i0; i < w.i1; i++, c += sI) {
      V pp, pjm, pjp, zc;
      __builtin_memcpy(&pp, p + c + sI, sizeof(V));
      __builtin_memcpy(&pjm, p + c - sJ, sizeof(V));
      __builtin_memcpy(&pjp, p + c + sJ, sizeof(V));
      __builtin_memcpy(&zc, z + c, sizeof(V));
      uint8_t fl[CPL];
      __builtin_memcpy(fl, flags + c, CPL);
      float left = __shfl_up(pc.v[CPL - 1], 1, 64), right = __shfl_down(pc.v[0], 1, 64);
      if (lane == 0) left = p[c - 1];
      if (lane == 63 || w.kraw + CPL > m - 2) right = p[c + CPL];
      const float cxm = cx[2 * i], cxp = cx[2 * i + 1];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const float q = cg_stencil_cell<float, CPL>(fl[e], e, pc, pm, pp, pjm.v[e], pjp.v[e], left, right, cxm, cxp,
                                                    cym, cyp, czm[e], czp[e]);
        if (w.own[e]) acc += (double)zc.v[e] * (double)q;
      }
      pm = pc;
      pc = pp;
    }
  }
  acc = block_sum_all(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc;
    part[2 * blockIdx.x + 1] = 0;
  }
}

template <int CPL>
__global__ __launch_bounds__(256) void k_fcg_dir_pq(int n, int l, int m, Tiles tl, const uint8_t *__restrict__ flags,
                                                     const float *__restrict__ cx, const float *__restrict__ cy,
                                                     const float *__restrict__ cz, const float *__restrict__ z,
                                                     const float *__restrict__ pin, float *__restrict__ pout,
                                                     const float *__restrict__ r, const DevState *st, int first,
                                                     double *__restrict__ part) {
  __shared__ double lds[8];
  if (st->done) return;
  using V = CgVec<float, CPL>;
  const double pq = st->s[HS_PQ];
  const bool nob = first || pq == 0;
  const float beta = nob ? 0.0f : (float)(-st->s[FC_ZQ] / pq);
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  double apr = 0, apq = 0;
  const int64_t sI = (int64_t)l * m, sJ = m;
  // p = z + beta p_old at one vector (p = z when there is no beta)
  auto pnew = [&](int64_t at) {
    V zv, q;
    __builtin_memcpy(&zv, z + at, sizeof(V));
    if (nob) return zv;
    V po;
    __builtin_memcpy(&po, pin + at, sizeof(V));
#pragma unroll
    for (int e = 0; e < CPL; e++) q.v[e] = zv.v[e] + beta * po.v[e];
    return q;
  };
  auto pnew1 = [&](int64_t at) { return nob ? z[at] : z[at] + beta * pin[at]; };
  // p at the rows j +- 1 through LDS, as in k_cgcg_step
  __shared__ V ldsq[2][TJ + 2][64];
  const int wv = threadIdx.x >> 6;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    CgTile w;
    const bool act = cg_tile<CPL>(tl, t, n, l, m, w);  // row inside the interior
    const bool ld = w.j <= l - 1;                       // row inside the array
    float czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (w.k0 + e)];
      czp[e] = cz[2 * (w.k0 + e) + 1];
    }
    const float cym = act ? cy[2 * w.j] : 0.f, cyp = act ? cy[2 * w.j + 1] : 0.f;
    int64_t c = w.c;
    V qm{}, qc{};
    if (ld) {
      qm = pnew(c - sI);
      qc = pnew(c);
    }
    for (int i = w.i0; i < w.i1; i++, c += sI) {
      V qp{};
      if (ld) qp = pnew(c + sI);
      const int par = (i - w.i0) & 1;
      ldsq[par][wv + 1][lane] = qc;
      if (wv == 0) ldsq[par][0][lane] = pnew(c - sJ);
      if (wv == TJ - 1 && act) ldsq[par][TJ + 1][lane] = pnew(c + sJ);
      __syncthreads();
      const V qjm = ldsq[par][wv][lane], qjp = ldsq[par][wv + 2][lane];
      if (!act) {
        qm = qc;
        qc = qp;
        continue;
      }
      V rc;
      __builtin_memcpy(&rc, r + c, sizeof(V));
      uint8_t fl[CPL];
      __builtin_memcpy(fl, flags + c, CPL);
      float left = __shfl_up(qc.v[CPL - 1], 1, 64), right = __shfl_down(qc.v[0], 1, 64);
      if (lane == 0) left = pnew1(c - 1);
      if (lane == 63 || w.kraw + CPL > m - 2) right = pnew1(c + CPL);
      const float cxm = cx[2 * i], cxp = cx[2 * i + 1];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const float q = cg_stencil_cell<float, CPL>(fl[e], e, qc, qm, qp, qjm.v[e], qjp.v[e], left, right, cxm, cxp,
                                                    cym, cyp, czm[e], czp[e]);
        if (w.own[e]) {
          apr += (double)qc.v[e] * (double)rc.v[e];
          apq += (double)q * (double)qc.v[e];
        }
      }
      if (w.allown) {
        __builtin_memcpy(pout + c, &qc, sizeof(V));
      } else {
#pragma unroll
        for (int e = 0; e < CPL; e++)
          if (w.own[e]) pout[c + e] = qc.v[e];
      }
      qm = qc;
      qc = qp;
    }
    __syncthreads();  // the next tile's first plane reuses buffer 0
  }
  apr = block_sum_all(apr, lds);
  apq = block_sum_all(apq, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = apr;
    part[2 * blockIdx.x + 1] = apq;
  }
}

template <int CPL>
__global__ __launch_bounds__(256) void k_fcg_upd(int n, int l, int m, Tiles tl, const uint8_t *__restrict__ flags,
                                                  const float *__restrict__ cx, const float *__restrict__ cy,
                                                  const float *__restrict__ cz, float *__restrict__ ev,
                                                  const float *__restrict__ p, float *__restrict__ r,
                                                  const DevState *st, double *__restrict__ part) {
  __shared__ double lds[8];
  if (st->done) return;
  using V = CgVec<float, CPL>;
  const float a = (float)st->s[HS_ALPHA];
  const int per = (tl.total + 7) / 8;
  const int bx = gridDim.x >> 3;
  const int xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  double acc = 0;
  const int64_t sI = (int64_t)l * m, sJ = m;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    CgTile w;
    if (!cg_tile<CPL>(tl, t, n, l, m, w)) continue;
    float czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (w.k0 + e)];
      czp[e] = cz[2 * (w.k0 + e) + 1];
    }
    const float cym = cy[2 * w.j], cyp = cy[2 * w.j + 1];
    int64_t c = w.c;
    V pm, pc;
    __builtin_memcpy(&pm, p + c - sI, sizeof(V));
    __builtin_memcpy(&pc, p + c, sizeof(V));
    for (int i = w.i0; i < w.i1; i++, c += sI) {
      V pp, pjm, pjp, rc, ec;
      __builtin_memcpy(&pp, p + c + sI, sizeof(V));
      __builtin_memcpy(&pjm, p + c - sJ, sizeof(V));
      __builtin_memcpy(&pjp, p + c + sJ, sizeof(V));
      __builtin_memcpy(&rc, r + c, sizeof(V));
      __builtin_memcpy(&ec, ev + c, sizeof(V));
      uint8_t fl[CPL];
      __builtin_memcpy(fl, flags + c, CPL);
      float left = __shfl_up(pc.v[CPL - 1], 1, 64), right = __shfl_down(pc.v[0], 1, 64);
      if (lane == 0) left = p[c - 1];
      if (lane == 63 || w.kraw + CPL > m - 2) right = p[c + CPL];
      const float cxm = cx[2 * i], cxp = cx[2 * i + 1];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const float q = cg_stencil_cell<float, CPL>(fl[e], e, pc, pm, pp, pjm.v[e], pjp.v[e], left, right, cxm, cxp,
                                                    cym, cyp, czm[e], czp[e]);
        rc.v[e] = rc.v[e] - a * q;
        ec.v[e] = ec.v[e] + a * pc.v[e];
        if (w.own[e]) acc += (double)rc.v[e] * (double)rc.v[e];
      }
      if (w.allown) {
        __builtin_memcpy(r + c, &rc, sizeof(V));
        __builtin_memcpy(ev + c, &ec, sizeof(V));
      } else {
#pragma unroll
        for (int e = 0; e < CPL; e++)
          if (w.own[e]) {
            r[c + e] = rc.v[e];
            ev[c + e] = ec.v[e];
          }
      }
      pm = pc;
      pc = pp;
    }
  }
  acc = block_sum_all(acc, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = acc;
    part[2 * blockIdx.x + 1] = 0;
  }
}

// Finishers (one block): ZQ: s[FC_ZQ] = sum; PRPQ: s[FC_PR] = sum of the first
// partials (p, r), s[HS_PQ] = sum of the second (p, A p), alpha = (p, r)/(p, A p)
// or stop on 0 / NaN; RR: s[HS_RR] = sum, iteration count and the stopping
// test.
enum { FIN_ZQ, FIN_PRPQ, FIN_RR };
template <int MODE>
__global__ void k_fcg_finish(const double *__restrict__ part, int np, DevState *st, int maxit) {
  __shared__ double lds[16];
  if (st->done) return;
  double s0 = 0, s1 = 0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    s0 += part[2 * q];
    if (MODE == FIN_PRPQ) s1 += part[2 * q + 1];
  }
  const double tot = block_sum(s0, lds);
  const double tot1 = MODE == FIN_PRPQ ? block_sum(s1, lds) : 0.0;
  if (threadIdx.x != 0) return;
  if (MODE == FIN_ZQ) {
    st->s[FC_ZQ] = tot;
  } else if (MODE == FIN_PRPQ) {
    st->s[FC_PR] = tot;
    st->s[HS_PQ] = tot1;
    if (tot1 == 0 || tot1 != tot1 || tot != tot) {
      st->done = 1;
      if (tot1 != tot1 || tot != tot) st->status = XF_ENAN;
    } else {
      st->s[HS_ALPHA] = tot / tot1;
    }
  } else {
    st->s[HS_BETA] = tot / st->s[HS_RR];
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

// Launches of the flexible-CG step on a grid: tiles and blocks of the marching
// kernels (16-byte vectors where a row holds a whole wave of them).
struct FcgLaunch {
  int cpl;
  Tiles tl;
  int nb;
};
static FcgLaunch fcg_launch(xf_grid g) {
  FcgLaunch f;
  f.cpl = g->m - 2 >= 64 * 4 ? 4 : (g->m - 2 >= 64 * 2 ? 2 : 1);
  f.tl = march_tiles(g, f.cpl, CgShape<float>::max_planes);
  f.nb = std::min(f.tl.blocks, kCgTileBlocks);
  return f;
}

// K1: partials of (z, A p_old)
static void fcg_zq(xf_ctx c, xf_grid g, const FcgLaunch &f, const float *z, const float *p, double *part) {
#define XF_FCG_ZQ(CPL)                                                                                          \
  hipLaunchKernelGGL((k_fcg_zq<CPL>), dim3(f.nb), dim3(256), 0, c->stream, g->n, g->l, g->m, f.tl, g->d_flags,     \
                     g->d_cxf, g->d_cyf, g->d_czf, z, p, c->state, part)
  if (f.cpl == 4) XF_FCG_ZQ(4);
  else if (f.cpl == 2) XF_FCG_ZQ(2);
  else XF_FCG_ZQ(1);
#undef XF_FCG_ZQ
}

// K2: p_new = z + beta p_old, partials of (p_new, r) and (p_new, A p_new)
static void fcg_dir_pq(xf_ctx c, xf_grid g, const FcgLaunch &f, const float *z, const float *pin, float *pout,
                       const float *r, int first, double *part) {
#define XF_FCG_DIR(CPL)                                                                                         \
  hipLaunchKernelGGL((k_fcg_dir_pq<CPL>), dim3(f.nb), dim3(256), 0, c->stream, g->n, g->l, g->m, f.tl, g->d_flags, \
                     g->d_cxf, g->d_cyf, g->d_czf, z, pin, pout, r, c->state, first, part)
  if (f.cpl == 4) XF_FCG_DIR(4);
  else if (f.cpl == 2) XF_FCG_DIR(2);
  else XF_FCG_DIR(1);
#undef XF_FCG_DIR
}

// K3: e += alpha p, r -= alpha A p, partials of (r, r)
static void fcg_upd(xf_ctx c, xf_grid g, const FcgLaunch &f, float *e, const float *p, float *r, double *part) {
#define XF_FCG_UPD(CPL)                                                                                         \
  hipLaunchKernelGGL((k_fcg_upd<CPL>), dim3(f.nb), dim3(256), 0, c->stream, g->n, g->l, g->m, f.tl, g->d_flags,    \
                     g->d_cxf, g->d_cyf, g->d_czf, e, p, r, c->state, part)
  if (f.cpl == 4) XF_FCG_UPD(4);
  else if (f.cpl == 2) XF_FCG_UPD(2);
  else XF_FCG_UPD(1);
#undef XF_FCG_UPD
}

// Flexible CG (A-orthogonalisation against the previous direction, Notay's
// FCG(1)) in FP32 with the network as preconditioner, inside the same FP64
// iterative refinement as the mixed-precision solver.  If the inner residual
// does not improve for stall_window iterations the inner solve continues with
// plain FP32 CG (fallback).
void solve_fcg_nn(xf_ctx c, xf_grid g, const double *b, const double *x0, double *x, const xf_solve_opts &o,
                  xf_solve_stats &s) {
  require_fast(g);
  const bool mgc = o.method == XF_FCG_MGNN;
  XF_REQUIRE(mgc || o.precond, XF_EINVAL, "XF_FCG_NN needs opts.precond (an xf_unet)");
  struct xf_unet_s *net = (struct xf_unet_s *)o.precond;
  CgWork *w = work(c, g, true, true);
  const int64_t nc = g->ncell;
  const FcgLaunch fl = fcg_launch(g);
  const double inner_tol = o.inner_tol > 0 ? o.inner_tol : 1e-4;
  const int stall = o.stall_window > 0 ? o.stall_window : 20;
  float *z = w->z32;
  upload_state_zero(c);
  XF_HIP(hipMemcpyAsync(x, x0, sizeof(double) * nc, hipMemcpyDeviceToDevice, c->stream));
  auto poll = [&]() {
    XF_HIP(hipMemcpyAsync(c->h_state, c->state, sizeof(DevState), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  };
  // the stopping norm as in solve_mp_cg: the projected residual's with project_mean
  const bool proj = o.project_mean && g->nunk > 0;
  auto residual = [&]() {
    residual_refine(c, g, b, x, proj, w->r, w->r32);
    poll();
    return c->h_state->s[HS_AUX];
  };
  double rr = residual();
  const double rr0 = rr;
  const double stop2 = o.tol_mode == XF_TOL_REL ? o.tol * o.tol * rr0 : o.tol * o.tol;
  int outer = 0, inner_total = 0, status = XF_OK, fallback = 0;
  while (rr > stop2) {
    if (rr != rr) { status = XF_ENAN; break; }
    if (outer >= 60 || inner_total >= o.maxit) { status = XF_EMAXIT; break; }
    outer++;
    XF_HIP(hipMemsetAsync(w->e32, 0, sizeof(float) * nc, c->stream));
    // as in solve_mp_cg: the last sweep only needs the factor still missing
    const double itol = std::min(0.5, std::max(inner_tol, 0.3 * sqrt(stop2 / rr)));
    hipLaunchKernelGGL(k_hs_start, dim3(1), dim3(1), 0, c->stream, c->state, itol, 1, (int)HS_RRW);
    poll();
    double best = c->h_state->s[HS_RR];
    int since_best = 0, it = 0;
    bool use_net = !fallback;
    // one network iteration: z = M(r), p = z + beta p, e += alpha p, r -= alpha A p
    auto launch_iter = [&](int idx) {
      if (mgc) {
        mg_apply(c, g, net, w->r32, z);
      } else {
        sq_unknowns<float>(c, g, w->r32, (int)FC_SUMSQ);
        unet_apply_grid(c, net, g, w->r32, z, c->state->s + FC_SUMSQ);
      }
      float *pold = w->p32[idx & 1], *pnew = w->p32[(idx + 1) & 1];
      if (idx > 0) {
        fcg_zq(c, g, fl, z, pold, w->part);
        hipLaunchKernelGGL(k_fcg_finish<FIN_ZQ>, dim3(1), dim3(256), 0, c->stream, w->part, fl.nb, c->state, 0);
      }
      // p = z + beta p_old with the partials ((p, r), (p, A p)); alpha = (p, r)/(p, A p)
      fcg_dir_pq(c, g, fl, z, pold, pnew, w->r32, idx == 0 ? 1 : 0, w->part);
      hipLaunchKernelGGL(k_fcg_finish<FIN_PRPQ>, dim3(1), dim3(256), 0, c->stream, w->part, fl.nb, c->state, 0);
      fcg_upd(c, g, fl, w->e32, pnew, w->r32, w->part);
      hipLaunchKernelGGL(k_fcg_finish<FIN_RR>, dim3(1), dim3(256), 0, c->stream, w->part, fl.nb, c->state,
                         o.maxit - inner_total);
      check_kernel("fcg iteration");
    };
    // the host reads the state after every iteration (stall test); the
    // cycle and network kernels are not gated on the device's done flag, so
    // no iteration is enqueued before the previous one has been checked
    int issued = 0;
    while (use_net && !c->h_state->done) {
      launch_iter(issued++);
      poll();
      if (c->h_state->done) break;
      const double rrn = c->h_state->s[HS_RR];
      if (rrn < 0.99 * best) {
        best = rrn;
        since_best = 0;
      } else if (++since_best >= stall) {
        use_net = false;
        fallback = 1;
      }
    }
    it = c->h_state->iter;  // network iterations applied (the device counts them)
    if (!use_net && !c->h_state->done) {
      // plain CG from the current (e, r): restart with p = r
      sq_unknowns<float>(c, g, w->r32, (int)FC_RR2);
      hipLaunchKernelGGL(k_restart_beta, dim3(1), dim3(1), 0, c->stream, c->state);
      cgcg_iterations<float, float>(c, g, w->e32, w->r32, w->q32, w->p32[0], w->p32[1], g->d_cxf, g->d_cyf,
                                    g->d_czf, w->part, std::max(1, o.maxit - inner_total - it), 16);
      it += c->h_state->iter;
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

// ||b - A x|| / ||b - A x0|| with the reference rows; work: a field of the
// grid's size to use (nullptr: one is allocated for the call).
double true_residual(xf_ctx c, xf_grid g, const double *b, const double *x, const double *x0, double *work) {
  if ((g->use_rho ? g->bad_rows_div : g->bad_rows_lap) != 0) return NAN;
  double *tmp = work;
  if (!tmp) XF_HIP(hipMalloc(&tmp, sizeof(double) * g->ncell));
  struct Free {
    double *p;
    ~Free() {
      if (p) hipFree(p);
    }
  } ft{work ? nullptr : tmp};
  double v[2];
  const double *xs[2] = {x, x0};
  for (int t = 0; t < 2; t++) {
    launch_apply_exact(c, g, xs[t], b, -1.0, tmp);
    sq_unknowns<double>(c, g, tmp, (int)HS_CNT);
    XF_HIP(hipMemcpyAsync(&v[t], c->state->s + HS_CNT, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
  }
  return v[1] > 0 ? sqrt(v[0] / v[1]) : sqrt(v[0]);
}
