// LOD transport step and batched Thomas solves on the device.
//
// THydrodynamicsProblem::SolveTransferEquation__ (XProblem.cpp:2784-3475) runs,
// for each global direction, one implicit sweep along that direction (a
// tridiagonal solve per "sector" of Actual points between two border points)
// followed by two border-only passes along the other directions.  Here each
// direction is three launches:
//   k_lod_rows   the tridiagonal row (a, b, c, F) of every Actual cell, one
//                thread per cell (the first direction's pass also starts the
//                two stages as copies of the previous level; a row reads only
//                the previous level and the
//                stages of earlier directions); rows of lines along k are
//                written line-interleaved through an LDS transpose, so every
//                line solve reads its rows coalesced;
//   k_lod_solve  one thread per line: the sector walk, the boundary rows and
//                the Thomas elimination (XSLAE.cpp:416-445) in the
//                reference's order, rows in two alternating register blocks
//                of eight, the elimination's (A[m], B[m]) pairs
//                line-interleaved in scratch;
//   k_lod_border the two border-only passes, one thread per cell of a
//                per-grid list of the cells that are neither Actual nor
//                Fictive (a
//                sector end reads only its Actual neighbour, which no pass
//                writes; where a cell ends one sector and starts the next, or
//                borders sectors of both passes, the later rule wins, as in
//                the line walk).
// The sector rules the reference enforces per line (XProblem.cpp:2932-2960)
// depend only on the point types, so they are checked once per grid on the
// host.  Arithmetic follows the reference's operand order (built with
// -ffp-contract=off), so results are bit-identical.
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
  double *scrA;              // (A_m, B_m) pairs, [maxlen][nlines][2]
};

// Up to kLodBatch independent LOD problems advanced by the same launches
// (blockIdx.y, or blockIdx.z for the k-line rows, picks the problem).
constexpr int kLodBatch = 4;
struct LodSet {
  LodArgs a[kLodBatch];
  struct Row4 *rows[kLodBatch];
  const uint8_t *codes_kt[kLodBatch];
  const int32_t *border[kLodBatch];  // cells neither Actual nor Fictive
  int64_t nborder[kLodBatch];
  int count;
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

// Row of an Actual cell c of a sector along `axis` (XProblem.cpp:3198-3222,
// 3224-3419); idx3 = (i, j, k) of c, dir = axis + 1 = the sweep's direction.
struct Row {
  double a, b, c, f;
};
struct alignas(32) Row4 {
  double a, b, c, f;
};

__device__ inline Row cell_row(const LodArgs &A, int axis, int dir, int64_t c, const int idx3[3]) {
  Row r{0, 0, 0, 0};
  const double *h = A.h[axis];
  const int64_t S[3] = {(int64_t)A.l * A.m, (int64_t)A.m, 1};
  const int64_t sa = S[axis];
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

// Rows of every Actual cell for the sweep along AXIS.  AXIS 0 and 1: one
// thread per cell, rows in cell order (the line solves of those axes read
// them coalesced already).  AXIS 2: 16 (k) x 16 (j) tiles of a plane,
// transposed in LDS and written with the row of cell k of line t = i*l + j at
// k * (n*l) + t.
template <int AXIS>
__global__ __launch_bounds__(256) void k_lod_rows(LodSet S) {
  const int p = AXIS < 2 ? (int)blockIdx.y : (int)(blockIdx.z % S.count);
  const LodArgs &A = S.a[p];
  Row4 *__restrict__ rows = S.rows[p];
  const int n = A.n, l = A.l, m = A.m;
  if (AXIS < 2) {
    const int64_t nc = (int64_t)n * l * m;
    for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
      if (AXIS == 0) {  // f_1/3 and f_2/3 start as copies of the previous level (XProblem.cpp:2796-2797)
        const double v = A.f_prev[c];
        A.stage[0][c] = v;
        A.stage[1][c] = v;
      }
      if (A.codes[c] != PT_ACTUAL) continue;
      const int idx3[3] = {(int)(c / ((int64_t)l * m)), (int)((c / m) % l), (int)(c % m)};
      const Row r = cell_row(A, AXIS, AXIS + 1, c, idx3);
      rows[c] = Row4{r.a, r.b, r.c, r.f};
    }
  } else {
    __shared__ Row4 tile[16][16];
    __shared__ uint8_t act[16][16];
    const int kk = threadIdx.x & 15, jj = threadIdx.x >> 4;
    const int k0 = blockIdx.x * 16, j0 = blockIdx.y * 16, i = blockIdx.z / S.count;
    if (i >= n) return;
    const int k = k0 + kk, j = j0 + jj;
    uint8_t on = 0;
    if (k < m && j < l) {
      const int64_t c = ((int64_t)i * l + j) * m + k;
      if (A.codes[c] == PT_ACTUAL) {
        const int idx3[3] = {i, j, k};
        const Row r = cell_row(A, 2, 3, c, idx3);
        tile[kk][jj] = Row4{r.a, r.b, r.c, r.f};
        on = 1;
      }
    }
    act[kk][jj] = on;
    __syncthreads();
    const int kk2 = threadIdx.x >> 4, jj2 = threadIdx.x & 15;
    const int k2 = k0 + kk2, j2 = j0 + jj2;
    if (k2 < m && j2 < l && act[kk2][jj2])
      rows[(int64_t)k2 * ((int64_t)n * l) + (int64_t)i * l + j2] = tile[kk2][jj2];
  }
}

// The sweep along AXIS over every line (one thread per line): sectors of
// Actual cells between border points, the boundary rows of PreDefined /
// Normal / PreNormal ends (XProblem.cpp:3064-3155), forward elimination and
// back substitution (XSLAE.cpp:428-441), written into stage[dg-1]
// (XProblem.cpp:3439-3449).  The host has checked every line's sectors.
//
// Rows arrive in two register blocks of P rows that alternate: while one
// block is eliminated the other block's loads are in flight (no register
// copies between the blocks, which would wait for every load outstanding).
// A block whose rows are all Actual inside an open sector takes the short
// path; any other block (a sector opens or closes in it, or the line ends)
// walks its rows again from memory through the general per-row code, which
// is inlined once per block.  The elimination's (A_m, B_m) pairs go to
// scratch line-interleaved, one 16-byte store per row, and the back
// substitution reads them in two alternating register blocks of Q pairs.
template <int AXIS>
__global__ __launch_bounds__(64) void k_lod_solve(LodSet S, int dg) {
  const LodArgs &A = S.a[blockIdx.y];
  const Row4 *__restrict__ rows = S.rows[blockIdx.y];
  const uint8_t *__restrict__ codes_kt = S.codes_kt[blockIdx.y];
  int64_t base, sa;
  int len, nl;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  line_geometry(A, AXIS, 0, base, sa, len, nl);
  if (t >= nl) return;
  line_geometry(A, AXIS, t, base, sa, len, nl);
  const double *h = A.h[AXIS];
  double *fa = A.stage[dg - 1];
  double2 *__restrict__ sAB = reinterpret_cast<double2 *>(A.scrA);
  // Row q of the line and its point type.  Past the line's end the last
  // point is read again: it is never Actual (the host has checked that every
  // sector closes), so a block that reaches past the end takes the general
  // walk, which stops at the end.  The loads carry no condition, so the
  // waits for one block do not extend to the other block's loads.
  auto load = [&](int q, Row4 &r, int &cd) {
    const int64_t qq = min(q, len - 1);
    if (AXIS == 2) {
      cd = codes_kt[qq * nl + t];
      r = rows[qq * nl + t];
    } else {
      cd = A.codes[base + qq * sa];
      r = rows[base + qq * sa];
    }
  };
  bool in = false;
  int beg = 0, mbeg = 0, fl = 0, s = 0, pc = PT_FICTIVE;
  double Am1 = 0, Bm1 = 0, u0 = 0;
  auto fwd = [&](double ra, double rb, double rc, double rf) {
    const double den = ra * Am1 + rb;
    const double Am = (-rc) / den;
    const double Bm = (rf - ra * Bm1) / den;
    sAB[(int64_t)s * nl + t] = double2{Am, Bm};
    Am1 = Am;
    Bm1 = Bm;
    s++;
  };
  // back substitution of the sector that closes at point `end` of type mend
  auto close = [&](int end, int mend) {
    const int fr = (mend != PT_DEFINED && mend != PT_EQUATION);
    const int sd = end - beg + fl + fr;  // dim - 1
    const int64_t cb = base + (int64_t)beg * sa, ce = base + (int64_t)end * sa;
    if (fr) {
      if (mend == PT_PREDEFINED) fwd(0.5, 0.5, 1, A.rp[ce]);
      else fwd(-1 / h[end - 1], 1 / h[end - 1], 1, A.rp[ce]);
    }
    double u = fr ? 0.0 : A.rp[ce];
    if (!fr) fa[base + (int64_t)(beg + sd - fl) * sa] = u;
    constexpr int Q = 8;
    double2 b0[Q], b1[Q];
    auto ldblk = [&](int top, double2 *x) {  // pairs top, top-1, ... (below 1: row 0, unused)
#pragma unroll
      for (int v = 0; v < Q; v++) x[v] = sAB[(int64_t)max(top - v, 0) * nl + t];
    };
    // Every row of a block stores: values that belong to no cell of the
    // sector (rows below 1, the boundary rows) go to an unused scratch row of
    // this line, so the number of stores in flight is the same on every path.
    double *const spare = reinterpret_cast<double *>(sAB + (int64_t)(len + 1) * nl + t);
    auto useblk = [&](int top, const double2 *x) {
#pragma unroll
      for (int v = 0; v < Q; v++) {
        const int ss = top - v;
        const double un = x[v].x * u + x[v].y;
        u = ss > 0 ? un : u;
        const bool put = ss > 0 && ss >= fl && ss <= sd - fr;
        *(put ? fa + base + (int64_t)(beg + ss - fl) * sa : spare) = u;
      }
    };
    int s2 = sd - 1;
    ldblk(s2, b0);
    while (s2 > 0) {
      ldblk(s2 - Q, b1);
      useblk(s2, b0);
      s2 -= Q;
      if (s2 <= 0) break;
      ldblk(s2 - Q, b0);
      useblk(s2, b1);
      s2 -= Q;
    }
    if (!fl) fa[cb] = u0;
    in = false;
  };
  auto visit = [&](int q, int cq, const Row4 &rw) {
    if (!in) {
      if (cq != PT_ACTUAL) {
        pc = cq;
        return;
      }
      beg = q - 1;
      mbeg = pc;
      fl = (mbeg != PT_DEFINED && mbeg != PT_EQUATION);
      const int64_t cb = base + (int64_t)beg * sa;
      u0 = fl ? 0.0 : A.rp[cb];
      Am1 = 0;
      Bm1 = u0;
      s = 1;
      in = true;
      if (fl) {
        if (mbeg == PT_PREDEFINED) fwd(1, 0.5, 0.5, A.rp[cb]);
        else fwd(1, -1 / h[beg], 1 / h[beg], A.rp[cb]);
      }
    }
    pc = cq;
    if (cq == PT_ACTUAL) fwd(rw.a, rw.b, rw.c, rw.f);
    else close(q, cq);
  };
  constexpr int P = 8;
  // rows q0 .. q0+P-1 held in (r, cd): the short path, or the general walk
  auto block = [&](int q0, const Row4 *r, const int *cd) {
    bool quick = in;
#pragma unroll
    for (int v = 0; v < P; v++) quick = quick && cd[v] == PT_ACTUAL;
    if (quick) {
#pragma unroll
      for (int v = 0; v < P; v++) fwd(r[v].a, r[v].b, r[v].c, r[v].f);
      pc = PT_ACTUAL;
    } else {
      const int qe = min(q0 + P, len);
      for (int q = q0; q < qe; q++) {
        Row4 rw;
        int cq;
        load(q, rw, cq);
        visit(q, cq, rw);
      }
    }
  };
  Row4 r0[P], r1[P];
  int c0[P], c1[P];
#pragma unroll
  for (int v = 0; v < P; v++) load(v, r0[v], c0[v]);
  // (the scheduling barriers keep a block's loads ahead of the other
  // block's elimination instead of letting them sink into it)
  for (int q0 = 0; q0 < len; q0 += 2 * P) {
#pragma unroll
    for (int v = 0; v < P; v++) load(q0 + P + v, r1[v], c1[v]);
    __builtin_amdgcn_sched_barrier(0);
    block(q0, r0, c0);
    if (q0 + P >= len) break;
#pragma unroll
    for (int v = 0; v < P; v++) load(q0 + 2 * P + v, r0[v], c0[v]);
    __builtin_amdgcn_sched_barrier(0);
    block(q0 + P, r1, c1);
  }
}

// The border-only passes of direction group dg along ax1, then ax2
// (XProblem.cpp:2965-3033), one thread per cell of the grid's list of cells
// that are neither Actual nor Fictive (the only ones a pass can write).
__global__ __launch_bounds__(256) void k_lod_border(LodSet set, int dg, int ax1, int ax2) {
  const LodArgs &A = set.a[blockIdx.y];
  const int32_t *__restrict__ list = set.border[blockIdx.y];
  const int64_t cnt = set.nborder[blockIdx.y];
  const int n = A.n, l = A.l, m = A.m;
  const int64_t S[3] = {(int64_t)l * m, (int64_t)m, 1};
  const int E[3] = {n, l, m};
  double *fa = A.stage[dg - 1];
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = list[t];
    const int cc = A.codes[c];
    const int idx3[3] = {(int)(c / S[0]), (int)((c / m) % l), (int)(c % m)};
    bool any = false;
    double v = 0;
    for (int pass = 0; pass < 2; pass++) {
      const int ax = pass ? ax2 : ax1;
      const int64_t sa = S[ax];
      const int q = idx3[ax];
      const double *h = A.h[ax];
      if (q >= 1 && A.codes[c - sa] == PT_ACTUAL) {  // the end of the sector before c
        any = true;
        if (cc == PT_DEFINED || cc == PT_EQUATION) v = A.rp[c];
        else if (cc == PT_PREDEFINED) v = 2 * A.rp[c] - fa[c - sa];
        else v = (h[q - 1]) * A.rp[c] + fa[c - sa];
      }
      if (q + 1 < E[ax] && A.codes[c + sa] == PT_ACTUAL) {  // the start of the sector after c
        any = true;
        if (cc == PT_DEFINED || cc == PT_EQUATION) v = A.rp[c];
        else if (cc == PT_PREDEFINED) v = 2 * A.rp[c] - fa[c + sa];
        else v = -(h[q]) * A.rp[c] + fa[c + sa];
      }
    }
    if (any) fa[c] = v;
  }
}

// The reference's per-line sector rules (XProblem.cpp:2928-2962) over every
// line of every axis: 1 = an Actual point first on its line ("Initial border
// not found"), 2 = a Fictive sector border, 4 = a border type the LOD has no
// rule for, 8 = a sector still open at the line's end (the reference would
// read past the array there).
static int lod_check(xf_grid g) {
  if (g->lod_checked) return g->lod_err;
  const std::vector<uint8_t> &cd = host_codes(g);
  const int E[3] = {g->n, g->l, g->m};
  const int64_t S[3] = {(int64_t)g->l * g->m, (int64_t)g->m, 1};
  auto okb = [](int v) {
    return v == PT_DEFINED || v == PT_EQUATION || v == PT_PREDEFINED || v == PT_NORMAL || v == PT_PRENORMAL;
  };
  int err = 0;
  for (int axis = 0; axis < 3; axis++) {
    const int ax1 = axis == 0 ? 1 : 0, ax2 = axis == 2 ? 1 : 2;
    for (int d1 = 0; d1 < E[ax1]; d1++)
      for (int d2 = 0; d2 < E[ax2]; d2++) {
        const int64_t base = d1 * S[ax1] + d2 * S[ax2];
        bool in = false;
        int mbeg = 0;
        for (int q = 0; q < E[axis]; q++) {
          const int v = cd[base + q * S[axis]];
          if (!in) {
            if (v != PT_ACTUAL) continue;
            if (q == 0) {
              err |= 1;
              break;
            }
            mbeg = cd[base + (q - 1) * S[axis]];
            in = true;
            continue;
          }
          if (v == PT_ACTUAL) continue;
          if (mbeg == PT_FICTIVE || v == PT_FICTIVE) err |= 2;
          else if (!okb(mbeg) || !okb(v)) err |= 4;
          in = false;
        }
        if (in) err |= 8;
      }
  }
  if (!err && !g->d_lod_border) {
    std::vector<int32_t> bl;
    for (int64_t q = 0; q < g->ncell; q++)
      if (cd[(size_t)q] != PT_ACTUAL && cd[(size_t)q] != PT_FICTIVE) bl.push_back((int32_t)q);
    g->n_lod_border = (int64_t)bl.size();
    XF_HIP(hipMalloc(&g->d_lod_border, sizeof(int32_t) * std::max<size_t>(bl.size(), 1)));
    if (!bl.empty())
      XF_HIP(hipMemcpy(g->d_lod_border, bl.data(), sizeof(int32_t) * bl.size(), hipMemcpyHostToDevice));
  }
  if (!err && !g->d_codes_kt) {
    // point types of lines along k in line-interleaved order
    const int64_t nl = (int64_t)g->n * g->l;
    std::vector<uint8_t> kt((size_t)g->ncell);
    for (int64_t t = 0; t < nl; t++)
      for (int k = 0; k < g->m; k++) kt[(size_t)(k * nl + t)] = cd[(size_t)(t * g->m + k)];
    XF_HIP(hipMalloc(&g->d_codes_kt, (size_t)g->ncell));
    XF_HIP(hipMemcpy(g->d_codes_kt, kt.data(), (size_t)g->ncell, hipMemcpyHostToDevice));
  }
  g->lod_checked = true;
  g->lod_err = err;
  return err;
}

// The LOD step of `count` independent problems (xf_lod_transport's
// arguments per problem).
static void run_lod(xf_ctx c, int count, const xf_grid *grids, double *const *f, const double *const *f_prev,
                    const double *const *u, const double *const *v, const double *const *w,
                    const double *const *rp, const double *const *kappa, double tau, int scheme) {
  XF_REQUIRE(count >= 1 && count <= kLodBatch, XF_EINVAL, "xf_lod_transport_batch: 1 to 4 problems");
  int64_t ncs = 0, scs = 0, ncmax = 0, nlmax[3] = {0, 0, 0};
  int nmax = 0, lmax = 0, mmax = 0;
  for (int p = 0; p < count; p++) {
    const xf_grid g = grids[p];
    XF_REQUIRE(g && f[p] && f_prev[p] && u[p] && v[p] && w[p] && rp[p] && kappa[p], XF_EINVAL,
               "xf_lod_transport: null argument");
    const int err = lod_check(g);
    XF_REQUIRE(!(err & 1), XF_EINVAL, "SolveTransferEquation: Initial border not found.");
    XF_REQUIRE(!(err & 8), XF_EINVAL,
               "SolveTransferEquation: a sector of Actual points reaches the end of a grid line (no closing border "
               "point).");
    XF_REQUIRE(err == 0, XF_EINVAL, "SolveTransferEquation: Incorrect sector border or size.");
    const int64_t nc = g->ncell;
    const int E[3] = {g->n, g->l, g->m};
    int64_t maxlines = 0;
    int maxlen = 0;
    for (int a = 0; a < 3; a++) {
      maxlines = std::max(maxlines, nc / E[a]);
      maxlen = std::max(maxlen, E[a]);
      nlmax[a] = std::max(nlmax[a], nc / E[a]);
    }
    ncs += nc;
    scs += maxlines * (maxlen + 2);
    ncmax = std::max(ncmax, nc);
    nmax = std::max(nmax, g->n);
    lmax = std::max(lmax, g->l);
    mmax = std::max(mmax, g->m);
  }
  double *st0 = (double *)ws_get(c, WS_LOD_ST0, sizeof(double) * ncs);
  double *st1 = (double *)ws_get(c, WS_LOD_ST1, sizeof(double) * ncs);
  Row4 *rows = (Row4 *)ws_get(c, WS_LOD_ROWS, sizeof(Row4) * ncs);
  double *scrA = (double *)ws_get(c, WS_LOD_SCRA, 2 * sizeof(double) * scs);  // (A_m, B_m) pairs
  LodSet S{};
  S.count = count;
  int64_t nbmax = 0;
  int64_t oc = 0, os = 0;
  for (int p = 0; p < count; p++) {
    const xf_grid g = grids[p];
    LodArgs &A = S.a[p];
    A.n = g->n; A.l = g->l; A.m = g->m;
    A.codes = g->d_codes;
    A.h[0] = g->d_hx; A.h[1] = g->d_hy; A.h[2] = g->d_hz;
    A.vel[0] = u[p]; A.vel[1] = v[p]; A.vel[2] = w[p];
    A.f_prev = f_prev[p]; A.rp = rp[p]; A.kappa = kappa[p];
    A.stage[0] = st0 + oc; A.stage[1] = st1 + oc; A.stage[2] = f[p];
    A.tau = tau;
    A.scheme = scheme;
    A.scrA = scrA + 2 * os;
    S.rows[p] = rows + oc;
    S.codes_kt[p] = g->d_codes_kt;
    S.border[p] = g->d_lod_border;
    S.nborder[p] = g->n_lod_border;
    nbmax = std::max(nbmax, g->n_lod_border);
    const int E[3] = {g->n, g->l, g->m};
    int64_t maxlines = 0;
    int maxlen = 0;
    for (int a = 0; a < 3; a++) {
      maxlines = std::max(maxlines, g->ncell / E[a]);
      maxlen = std::max(maxlen, E[a]);
    }
    oc += g->ncell;
    os += maxlines * (maxlen + 2);
  }
  const unsigned cb = (unsigned)std::min<int64_t>((ncmax + 255) / 256, 16384);
  const dim3 cells(cb, count);
  const dim3 borders((unsigned)std::max<int64_t>(1, std::min<int64_t>((nbmax + 255) / 256, 4096)), count);
  for (int dg = 1; dg < 4; dg++) {  // XProblem.cpp:2847-2864
    const int axis = dg - 1;
    const dim3 lines((unsigned)((nlmax[axis] + 63) / 64), count);
    if (axis == 0) {
      hipLaunchKernelGGL(k_lod_rows<0>, cells, dim3(256), 0, c->stream, S);
      hipLaunchKernelGGL(k_lod_solve<0>, lines, dim3(64), 0, c->stream, S, dg);
      hipLaunchKernelGGL(k_lod_border, borders, dim3(256), 0, c->stream, S, dg, 1, 2);
    } else if (axis == 1) {
      hipLaunchKernelGGL(k_lod_rows<1>, cells, dim3(256), 0, c->stream, S);
      hipLaunchKernelGGL(k_lod_solve<1>, lines, dim3(64), 0, c->stream, S, dg);
      hipLaunchKernelGGL(k_lod_border, borders, dim3(256), 0, c->stream, S, dg, 0, 2);
    } else {
      hipLaunchKernelGGL(k_lod_rows<2>, dim3((mmax + 15) / 16, (lmax + 15) / 16, nmax * count), dim3(256), 0,
                         c->stream, S);
      hipLaunchKernelGGL(k_lod_solve<2>, lines, dim3(64), 0, c->stream, S, dg);
      hipLaunchKernelGGL(k_lod_border, borders, dim3(256), 0, c->stream, S, dg, 0, 1);
    }
    check_kernel("LOD sweep");
  }
}

extern "C" int xf_lod_transport(xf_ctx c, xf_grid g, double *f, const double *f_prev, const double *u,
                                const double *v, const double *w, const double *rp, const double *kappa, double tau,
                                int scheme) {
  return guarded(c, [&] {
    XF_REQUIRE(g && f && f_prev && u && v && w && rp && kappa, XF_EINVAL, "xf_lod_transport: null argument");
    XF_REQUIRE(scheme == 0 || scheme == 1, XF_EINVAL, "Unknown scheme type.");
    XF_REQUIRE(tau > 0, XF_EINVAL, "xf_lod_transport: time step must be positive");
    run_lod(c, 1, &g, &f, &f_prev, &u, &v, &w, &rp, &kappa, tau, scheme);
  });
}

extern "C" int xf_lod_transport_batch(xf_ctx c, int count, const xf_grid *grids, double *const *f,
                                      const double *const *f_prev, const double *const *u, const double *const *v,
                                      const double *const *w, const double *const *rp, const double *const *kappa,
                                      double tau, int scheme) {
  return guarded(c, [&] {
    XF_REQUIRE(grids && f && f_prev && u && v && w && rp && kappa, XF_EINVAL,
               "xf_lod_transport_batch: null argument");
    XF_REQUIRE(scheme == 0 || scheme == 1, XF_EINVAL, "Unknown scheme type.");
    XF_REQUIRE(tau > 0, XF_EINVAL, "xf_lod_transport: time step must be positive");
    for (int p = 0; p < count; p++)
      for (int q = 0; q < count; q++)
        XF_REQUIRE(p == q || (f[p] != f[q] && f[p] != f_prev[q]), XF_EINVAL,
                   "xf_lod_transport_batch: a problem's output is another problem's field");
    run_lod(c, count, grids, f, f_prev, u, v, w, rp, kappa, tau, scheme);
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
    double *sa = (double *)ws_get(c, WS_THOMAS_A, sizeof(double) * nsys * (M + 1));
    double *sb = (double *)ws_get(c, WS_THOMAS_B, sizeof(double) * nsys * (M + 1));
    hipLaunchKernelGGL(k_thomas_batched, dim3((nsys + 127) / 128), dim3(128), 0, c->stream, nsys, M, u, a, b, cc, f,
                       sa, sb);
    check_kernel("k_thomas_batched");
  });
}
