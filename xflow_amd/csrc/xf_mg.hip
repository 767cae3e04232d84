// Geometric V-cycle preconditioner with the learned network as its
// coarsest-grid solve (XF_FCG_MGNN).
//
// The pressure grid is halved per axis while every interior extent stays
// even.  With the network the cycle stops where the next row (k extent)
// would fall below 64, the network's minimum (64^3 for a 256^3 problem); the
// cycle without a network goes on until the grid, ring included, has at most
// kCoarseCells cells (4^3 for a 256^3 problem) and solves there exactly: a
// dense pseudo-inverse of that grid's operator, formed on the host when the
// hierarchy is built.  Coarse cells are unknowns if
// any of their 8 children is an unknown, "boundary value" cells (a zero
// correction, open face) if any child is a Defined/boundary point, Fictive
// otherwise; coarse steps are the doubled fine steps, so each coarse grid
// carries the same masked 7-point operator rediscretised (its own xf_grid).
// One cycle, all in FP32 with FP64 dot products:
//   level k < coarsest: nu damped-Jacobi sweeps from z = 0 (omega = 6/7),
//     t = r - A z, r_{k+1} = mean of t over the unknown children,
//     V-cycle on level k+1, z += z_{k+1}(parent) on the unknowns,
//     nu damped-Jacobi sweeps;
//   coarsest level: y = net(r) (the network applied as on any grid: input
//     r / rms(r) and the unknown mask, output times rms(r)), scaled by the
//     energy-optimal factor (r, y) / (y, A y), then nu Jacobi sweeps; without
//     a network z = A^+ r on the deepest grid (the classical full-depth
//     cycle: the least-squares solution of least norm, which on a grid with
//     no boundary-value cell is the mean-free one), or nu_c Jacobi sweeps
//     where no grid is small enough.
// The rediscretised coarse equation A_H e_H = mean(t) is consistent for
// smooth errors (both sides approximate -lap e); the flexible CG around the
// cycle absorbs any remaining scale.
#include <algorithm>
#include <cmath>
#include <vector>

#include "xf_internal.h"

constexpr float kOmega = 0.857142857f;  // 6/7
constexpr int kSweeps = 4;        // pre- and post-smoothing sweeps per level
constexpr int kCoarseSweeps = 8;  // coarsest-level sweeps without a network
constexpr int kMgBlocks = 512;    // blocks of the reduction kernels
constexpr int kCoarseCells = 512;   // cells (ring included) of a grid solved densely
constexpr int kBottomCells = 2048;  // cells of the largest level the one-workgroup bottom of the cycle takes
constexpr int kMaxLevels = 12;

struct MgLevel {
  xf_grid g = nullptr;  // the rediscretised grid of this level (owned, except level 0)
  float *r = nullptr, *z = nullptr, *z2 = nullptr, *t = nullptr;
};

struct MgHier {
  int nlev = 0;         // levels of the cycle without a network
  int net_lev = 0;      // levels of the cycle with the network (its coarsest has rows >= 64)
  int ndense = 0;       // unknowns of the deepest grid's dense solve (0: Jacobi sweeps there)
  int bottom = -1;      // first level of the cycle without a network run by k_mg_bottom (-1: none)
  float *minv = nullptr;     // [ndense][ndense] its pseudo-inverse (symmetric)
  int32_t *uidx = nullptr;   // [ncell] cell -> unknown index, -1 elsewhere
  MgLevel lev[kMaxLevels];
  double *part = nullptr;  // [kMgBlocks][2] dot partials
  double *scal = nullptr;  // [4]: sum r^2, (r, y), (y, A y), alpha
};

// ----------------------------------------------------------------- kernels
// One thread per interior cell, k fastest: (64, 4) threads per block over
// (k, j), one i-plane per blockIdx.z.
struct Dims {
  int n, l, m;
};

__device__ inline float diag_of(uint8_t f, float cxm, float cxp, float cym, float cyp, float czm, float czp) {
  float d = 0;
  if (f & FL_IM) d += cxm;
  if (f & FL_IP) d += cxp;
  if (f & FL_JM) d += cym;
  if (f & FL_JP) d += cyp;
  if (f & FL_KM) d += czm;
  if (f & FL_KP) d += czp;
  return d;
}

// The hierarchy's grids have uniform steps per axis, so every interior node
// carries the same coefficients (c-[i] == c+[i], bit for bit) and the damped
// Jacobi weight omega / diag depends on the six face bits only: each block
// tabulates it once in LDS (the same operations as per cell, so the same bits)
// instead of dividing per cell.
__device__ inline void fill_weights(float *lut, const float *__restrict__ cx, const float *__restrict__ cy,
                                    const float *__restrict__ cz) {
  if (threadIdx.x < 64) {
    const float a = cx[2], b = cy[2], e = cz[2];
    const float dg = diag_of((uint8_t)threadIdx.x, a, a, b, b, e, e);
    lut[threadIdx.x] = dg > 0 ? kOmega / dg : 0.f;  // a cell with no open face is left alone
  }
  __syncthreads();
}

// omega D^-1 r at cell q: the first sweep from z = 0 (0 at cells that are not Actual)
__device__ inline float jacobi0(const uint8_t *__restrict__ flags, const float *lut, const float *__restrict__ r,
                                int64_t q) {
  const uint8_t f = flags[q];
  return (f & FL_ACTUAL) ? lut[f & 63] * r[q] : 0.f;
}

// MODE 0: out = omega D^-1 r;  MODE 1: out = z + omega D^-1 (r - A z);
// MODE 2: out = r - A z;  MODE 3: out = A z;  MODE 5: MODE 1 on z = omega D^-1 r
// (MODE 6, vector kernel only: MODE 2 restricted to the coarse grid in the pass)
// (the first two sweeps from z = 0 in one pass: z recomputed at the seven
// stencil points, never stored).  0 at cells that are not Actual.
template <int MODE>
__global__ __launch_bounds__(256) void k_mg_sweep(Dims d, const uint8_t *__restrict__ flags,
                                                   const float *__restrict__ cx, const float *__restrict__ cy,
                                                   const float *__restrict__ cz, const float *__restrict__ r,
                                                   const float *__restrict__ z, float *__restrict__ out) {
  __shared__ float lut[64];
  fill_weights(lut, cx, cy, cz);
  const int k = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int i = 1 + blockIdx.z;
  if (k > d.m - 2 || j > d.l - 2) return;
  const int64_t sI = (int64_t)d.l * d.m, sJ = d.m;
  const int64_t c = ((int64_t)i * d.l + j) * d.m + k;
  const uint8_t f = flags[c];
  float v = 0;
  if (f & FL_ACTUAL) {
    const float cxm = cx[2 * i], cxp = cx[2 * i + 1], cym = cy[2 * j], cyp = cy[2 * j + 1];
    const float czm = cz[2 * k], czp = cz[2 * k + 1];
    const float wd = lut[f & 63];
    if (MODE == 0) {
      v = wd * r[c];
    } else {
      // z at the stencil points (MODE 5: the first sweep's omega D^-1 r)
      auto zat = [&](int64_t q) { return MODE == 5 ? jacobi0(flags, lut, r, q) : z[q]; };
      const float zc = MODE == 5 ? wd * r[c] : z[c];
      float tx = ((f & FL_IP) ? cxp * (zat(c + sI) - zc) : 0.f) - ((f & FL_IM) ? cxm * (zc - zat(c - sI)) : 0.f);
      float ty = ((f & FL_JP) ? cyp * (zat(c + sJ) - zc) : 0.f) - ((f & FL_JM) ? cym * (zc - zat(c - sJ)) : 0.f);
      float tz = ((f & FL_KP) ? czp * (zat(c + 1) - zc) : 0.f) - ((f & FL_KM) ? czm * (zc - zat(c - 1)) : 0.f);
      const float az = -(tx + ty + tz);
      if (MODE == 1 || MODE == 5) v = zc + wd * (r[c] - az);
      else if (MODE == 2) v = r[c] - az;
      else v = az;
    }
  }
  out[c] = v;
}

// The same sweeps on rows of at least 64 CPL cells: CPL (4 or 2) consecutive
// k-cells per lane (16- or 8-byte loads), k-neighbours by lane shuffles, IC i-planes marched
// per block with the planes i-1, i, i+1 in registers; blocks persistent over
// the (k, j, i) tiles, consecutive tiles on one XCD.  The last k-tile of a row
// is shifted back to end at m-2 (cells it shares with the tile before get the
// same value twice).
struct SweepTiles {
  int kt, jt, it, total, ic;  // ic: i-planes marched per tile
};

// MODE 4: MODE 1 on z + zcoarse(parent) at the unknowns (the prolongation of
// the coarse correction folded into the first post-smoothing sweep; the
// corrected z is formed with the same addition as k_mg_prolong4 and not stored).
// MODE 5: MODE 1 on z = omega D^-1 r (the first two sweeps from z = 0 in one
// pass, z formed at the stencil points as MODE 0 forms it and not stored).
template <int MODE, int CPL>
__global__ __launch_bounds__(256) void k_mg_sweep_vec(Dims d, SweepTiles tl, const uint8_t *__restrict__ flags,
                                                       const float *__restrict__ cx, const float *__restrict__ cy,
                                                       const float *__restrict__ cz, const float *__restrict__ r,
                                                       const float *__restrict__ z, float *__restrict__ out,
                                                       const float *__restrict__ zcoarse, Dims cd) {
  struct V {
    float v[CPL];
  };
  __shared__ float lut[64];
  // MODE 6: the odd rows' residuals and unknown bits for the even rows below
  // them, one buffer per plane parity
  __shared__ float xr[2][2][MODE == 6 ? 64 * CPL : 1];
  __shared__ uint8_t xu[2][2][MODE == 6 ? 64 * CPL : 1];
  fill_weights(lut, cx, cy, cz);
  // z (+ the parent's coarse correction at unknowns for MODE 4; omega D^-1 r
  // for MODE 5) of CPL cells from k
  auto zrow = [&](int64_t q, int ii, int jj, int kk) {
    V v;
    if (MODE == 5) {
      V rv;
      uint8_t fq[CPL];
      __builtin_memcpy(&rv, r + q, sizeof(V));
      __builtin_memcpy(fq, flags + q, CPL);
#pragma unroll
      for (int e = 0; e < CPL; e++) v.v[e] = (fq[e] & FL_ACTUAL) ? lut[fq[e] & 63] * rv.v[e] : 0.f;
      return v;
    }
    __builtin_memcpy(&v, z + q, sizeof(V));
    if (MODE == 4) {
      // parents of cells kk..kk+CPL-1 are the CPL/2 + 1 coarse cells from
      // (kk + 1) >> 1 (in bounds for an even m, which the dispatcher checks)
      uint8_t fq[CPL];
      __builtin_memcpy(fq, flags + q, CPL);
      const int p0 = (kk + 1) >> 1;
      const float *pr = zcoarse + ((int64_t)((ii + 1) >> 1) * cd.l + ((jj + 1) >> 1)) * cd.m + p0;
      float pa[CPL / 2 + 1];
#pragma unroll
      for (int e = 0; e <= CPL / 2; e++) pa[e] = pr[e];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const float zc2 = pa[((kk + e + 1) >> 1) - p0];
        v.v[e] = (fq[e] & FL_UNKNOWN) ? v.v[e] + zc2 : v.v[e];
      }
    }
    return v;
  };
  auto zone = [&](int64_t q, int ii, int jj, int kk) {
    if (MODE == 5) return jacobi0(flags, lut, r, q);
    float v = z[q];
    if (MODE == 4 && (flags[q] & FL_UNKNOWN))
      v += zcoarse[((int64_t)((ii + 1) >> 1) * cd.l + ((jj + 1) >> 1)) * cd.m + ((kk + 1) >> 1)];
    return v;
  };
  const int per = (tl.total + 7) / 8, bx = gridDim.x >> 3, xcd = blockIdx.x & 7;
  const int lane = threadIdx.x & 63;
  const int64_t sI = (int64_t)d.l * d.m, sJ = d.m;
  for (int tq = blockIdx.x >> 3; tq < per; tq += bx) {
    const int t = xcd * per + tq;
    if (t >= tl.total) break;
    const int ktile = t % tl.kt, jtile = (t / tl.kt) % tl.jt, itile = t / (tl.kt * tl.jt);
    const int k0 = min(1 + ktile * 64 * CPL, d.m - 1 - 64 * CPL) + lane * CPL;
    const int j = 1 + jtile * 4 + (threadIdx.x >> 6);
    const int i0 = 1 + itile * tl.ic, i1 = min(i0 + tl.ic, d.n - 1);
    // MODE 6 exchanges rows between the waves of a block at a barrier per
    // plane, so rows past the grid keep marching (without memory accesses);
    // the grid's even extents make both rows of a pair active or neither
    const bool act = j <= d.l - 2;
    if (!act && MODE != 6) continue;
    const int jj = act ? j : d.l - 2;  // a valid row for the coefficient reads
    float czm[CPL], czp[CPL];
#pragma unroll
    for (int e = 0; e < CPL; e++) {
      czm[e] = cz[2 * (k0 + e)];
      czp[e] = cz[2 * (k0 + e) + 1];
    }
    const float cym = cy[2 * jj], cyp = cy[2 * jj + 1];
    int64_t c = ((int64_t)i0 * d.l + jj) * d.m + k0;
    V zm{}, zc{};
    if (MODE != 0 && act) {
      zm = zrow(c - sI, i0 - 1, j, k0);
      zc = zrow(c, i0, j, k0);
    }
    float rsum[CPL / 2];
    int rcnt[CPL / 2];
    for (int i = i0; i < i1; i++, c += sI) {
      uint8_t fl[CPL] = {};
      V rr{}, zp{}, zjm{}, zjp{}, o;
      if (act) {
        __builtin_memcpy(fl, flags + c, CPL);
        if (MODE != 3) __builtin_memcpy(&rr, r + c, sizeof(V));
      }
      float left = 0, right = 0;
      if (MODE != 0 && act) {
        zp = zrow(c + sI, i + 1, j, k0);
        zjm = zrow(c - sJ, i, j - 1, k0);
        zjp = zrow(c + sJ, i, j + 1, k0);
        left = __shfl_up(zc.v[CPL - 1], 1, 64);
        right = __shfl_down(zc.v[0], 1, 64);
        if (lane == 0) left = zone(c - 1, i, j, k0 - 1);
        if (lane == 63) right = zone(c + CPL, i, j, k0 + CPL);
      }
      const float cxm = cx[2 * i], cxp = cx[2 * i + 1];
#pragma unroll
      for (int e = 0; e < CPL; e++) {
        const uint8_t f = fl[e];
        const float wd = lut[f & 63];
        float v = 0;
        if (MODE == 0) {
          v = wd * rr.v[e];
        } else {
          const float x0 = zc.v[e];
          const float km = e == 0 ? left : zc.v[e - 1];
          const float kp = e == CPL - 1 ? right : zc.v[e + 1];
          float tx = ((f & FL_IP) ? cxp * (zp.v[e] - x0) : 0.f) - ((f & FL_IM) ? cxm * (x0 - zm.v[e]) : 0.f);
          float ty = ((f & FL_JP) ? cyp * (zjp.v[e] - x0) : 0.f) - ((f & FL_JM) ? cym * (x0 - zjm.v[e]) : 0.f);
          float tz = ((f & FL_KP) ? czp[e] * (kp - x0) : 0.f) - ((f & FL_KM) ? czm[e] * (x0 - km) : 0.f);
          const float az = -(tx + ty + tz);
          if (MODE == 1 || MODE == 4 || MODE == 5) v = x0 + wd * (rr.v[e] - az);
          else if (MODE == 2 || MODE == 6) v = rr.v[e] - az;
          else v = az;
        }
        o.v[e] = (f & FL_ACTUAL) ? v : 0.f;
      }
      if (MODE != 6) {
        __builtin_memcpy(out + c, &o, sizeof(V));
      } else {
        // coarse r = mean of the residual over the unknown children, summed
        // in k_mg_restrict's order: plane, then row, then k; the even row of
        // each pair (wave 0 or 2) adds the odd row's values from LDS.  Cells
        // of a shifted last k-tile repeat the tile before: the same coarse
        // values are written twice.
        const int w = threadIdx.x >> 6, par = (i - i0) & 1;
        if (w & 1) {
#pragma unroll
          for (int e = 0; e < CPL; e++) {
            xr[par][w >> 1][lane * CPL + e] = o.v[e];
            xu[par][w >> 1][lane * CPL + e] = fl[e] & FL_UNKNOWN;
          }
        }
        __syncthreads();
        if (!(w & 1) && act) {
#pragma unroll
          for (int p = 0; p < CPL / 2; p++) {
            float sp = par ? rsum[p] : 0.f;
            int cn = par ? rcnt[p] : 0;
#pragma unroll
            for (int e = 2 * p; e < 2 * p + 2; e++)
              if (fl[e] & FL_UNKNOWN) {
                sp += o.v[e];
                cn++;
              }
#pragma unroll
            for (int e = 2 * p; e < 2 * p + 2; e++)
              if (xu[par][w >> 1][lane * CPL + e]) {
                sp += xr[par][w >> 1][lane * CPL + e];
                cn++;
              }
            rsum[p] = sp;
            rcnt[p] = cn;
            if (par) {
              const int64_t q = ((int64_t)(i >> 1) * cd.l + ((j + 1) >> 1)) * cd.m + ((k0 + 2 * p + 1) >> 1);
              out[q] = cn ? sp / (float)cn : 0.f;
            }
          }
        }
      }
      if (MODE != 0) {
        zm = zc;
        zc = zp;
      }
    }
  }
}

// coarse r = mean of the fine t over the unknown children (one thread per coarse interior cell)
__global__ __launch_bounds__(256) void k_mg_restrict(Dims fd, Dims cd, const uint8_t *__restrict__ fflags,
                                                      const float *__restrict__ t, float *__restrict__ rc) {
  const int K = 1 + blockIdx.x * 64 + (threadIdx.x & 63);
  const int J = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int I = 1 + blockIdx.z;
  if (K > cd.m - 2 || J > cd.l - 2) return;
  float s = 0;
  int cnt = 0;
  for (int a = 0; a < 2; a++)
    for (int b = 0; b < 2; b++)
      for (int e = 0; e < 2; e++) {
        const int64_t q = ((int64_t)(2 * I - 1 + a) * fd.l + (2 * J - 1 + b)) * fd.m + (2 * K - 1 + e);
        if (fflags[q] & FL_UNKNOWN) {
          s += t[q];
          cnt++;
        }
      }
  rc[((int64_t)I * cd.l + J) * cd.m + K] = cnt ? s / (float)cnt : 0.f;
}

// fine z += coarse z(parent) on the fine unknowns, 4 consecutive k-cells per thread
__global__ __launch_bounds__(256) void k_mg_prolong4(Dims fd, Dims cd, const uint8_t *__restrict__ fflags,
                                                     const float *__restrict__ zc, float *__restrict__ z) {
  const int k0 = 1 + 4 * (blockIdx.x * 64 + (threadIdx.x & 63));
  const int j = 1 + blockIdx.y * 4 + (threadIdx.x >> 6);
  const int i = 1 + blockIdx.z;
  if (k0 > fd.m - 2 || j > fd.l - 2) return;
  const int64_t c = ((int64_t)i * fd.l + j) * fd.m + k0;
  const int64_t pc = ((int64_t)((i + 1) >> 1) * cd.l + ((j + 1) >> 1)) * cd.m;
  const int ke = min(4, fd.m - 1 - k0);
#pragma unroll
  for (int e = 0; e < 4; e++) {
    if (e >= ke) break;
    if (fflags[c + e] & FL_UNKNOWN) z[c + e] += zc[pc + ((k0 + e + 1) >> 1)];
  }
}

// per-block partials of (a, b) and (b, c) over the unknowns (c may be null: only (a, b))
__global__ __launch_bounds__(256) void k_mg_dots(int64_t nc, const uint8_t *__restrict__ flags,
                                                  const float *__restrict__ a, const float *__restrict__ b,
                                                  const float *__restrict__ cc, double *__restrict__ part) {
  __shared__ double lds[16];
  double s0 = 0, s1 = 0;
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nc; q += (int64_t)gridDim.x * blockDim.x) {
    if (!(flags[q] & FL_UNKNOWN)) continue;
    s0 += (double)a[q] * (double)b[q];
    if (cc) s1 += (double)b[q] * (double)cc[q];
  }
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = s0;
    part[2 * blockIdx.x + 1] = s1;
  }
}

// MODE 0: scal[0] = sum of the first partials; MODE 1: scal[1], scal[2] and
// alpha = scal[1] / scal[2] (0 if not finite and positive)
template <int MODE>
__global__ void k_mg_finish(const double *__restrict__ part, int np, double *__restrict__ scal) {
  __shared__ double lds[16];
  double s0 = 0, s1 = 0;
  for (int q = threadIdx.x; q < np; q += blockDim.x) {
    s0 += part[2 * q];
    s1 += part[2 * q + 1];
  }
  s0 = block_sum(s0, lds);
  s1 = block_sum(s1, lds);
  if (threadIdx.x == 0) {
    if (MODE == 0) {
      scal[0] = s0;
    } else {
      scal[1] = s0;
      scal[2] = s1;
      const double a = s0 / s1;
      scal[3] = (s1 > 0 && a == a && fabs(a) < 1e30) ? a : 0.0;
    }
  }
}

// z *= scal[3] on the whole array
__global__ void k_mg_scale(int64_t nc, float *__restrict__ z, const double *__restrict__ scal) {
  const float a = (float)scal[3];
  for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < nc; q += (int64_t)gridDim.x * blockDim.x)
    z[q] *= a;
}

// z = A^+ r on the deepest grid, one thread per cell (<= kCoarseCells): r at
// the unknowns staged in LDS, each unknown's row of the pseudo-inverse
// (stored symmetric, read column-wise across threads) summed in FP64.
__global__ __launch_bounds__(kCoarseCells) void k_mg_coarse_dense(int ncell, int nu, const int32_t *__restrict__ uidx,
                                                                   const float *__restrict__ minv,
                                                                   const float *__restrict__ rin,
                                                                   float *__restrict__ zout) {
  __shared__ float rs[kCoarseCells];
  const int q = threadIdx.x;
  const int u = q < ncell ? uidx[q] : -1;
  if (u >= 0) rs[u] = rin[q];
  __syncthreads();
  if (q >= ncell) return;
  double acc = 0;
  if (u >= 0)
    for (int v = 0; v < nu; v++) acc += (double)minv[(size_t)v * nu + u] * (double)rs[v];
  zout[q] = (float)acc;
}

// The bottom of the cycle without a network in one workgroup: levels
// first..last (each at most kBottomCells cells) down to the dense solve and
// back up, every phase a loop over the level's interior cells separated by
// barriers (the level fields stay in global memory, warm in the caches).
// Each phase does what the multi-block path does at that level, in the same
// floating-point order: the sweeps from z = 0, the residual, the
// restriction, the dense solve, the prolongation and the sweeps after.
struct BotLevel {
  Dims d;
  const uint8_t *flags;
  const float *cx, *cy, *cz;
  float *r, *z, *z2, *t;
};
struct BotArgs {
  BotLevel lv[kMaxLevels];
  int first, last, nu, ndense;
  const int32_t *uidx;
  const float *minv;
  const float *r_in;  // r of level first
  float *z_out;       // z of level first
};

__device__ inline float bot_az(const BotLevel &L, uint8_t f, int64_t c, int i, int j, int k, const float *z) {
  const int64_t sI = (int64_t)L.d.l * L.d.m, sJ = L.d.m;
  const float cxm = L.cx[2 * i], cxp = L.cx[2 * i + 1], cym = L.cy[2 * j], cyp = L.cy[2 * j + 1];
  const float czm = L.cz[2 * k], czp = L.cz[2 * k + 1];
  const float zc = z[c];
  float tx = ((f & FL_IP) ? cxp * (z[c + sI] - zc) : 0.f) - ((f & FL_IM) ? cxm * (zc - z[c - sI]) : 0.f);
  float ty = ((f & FL_JP) ? cyp * (z[c + sJ] - zc) : 0.f) - ((f & FL_JM) ? cym * (zc - z[c - sJ]) : 0.f);
  float tz = ((f & FL_KP) ? czp * (z[c + 1] - zc) : 0.f) - ((f & FL_KM) ? czm * (zc - z[c - 1]) : 0.f);
  return -(tx + ty + tz);
}

__global__ __launch_bounds__(1024) void k_mg_bottom(BotArgs B) {
  __shared__ float lut[kMaxLevels][64];
  __shared__ float rs[kCoarseCells];
  const int tid = threadIdx.x, nt = blockDim.x;
  for (int q = tid; q < kMaxLevels * 64; q += nt) {
    const int k = q >> 6, fb = q & 63;
    if (k < B.first || k > B.last) continue;
    const float a = B.lv[k].cx[2], b = B.lv[k].cy[2], e = B.lv[k].cz[2];
    const float dg = diag_of((uint8_t)fb, a, a, b, b, e, e);
    lut[k][fb] = dg > 0 ? kOmega / dg : 0.f;
  }
  __syncthreads();
  // one pass over the interior cells of level k
  auto interior = [&](int k, auto &&body) {
    const Dims d = B.lv[k].d;
    const int N = d.n - 2, Lx = d.l - 2, M = d.m - 2;
    const int tot = N * Lx * M;
    for (int q = tid; q < tot; q += nt) {
      const int kk = 1 + q % M, jj = 1 + (q / M) % Lx, ii = 1 + q / (M * Lx);
      body(((int64_t)ii * d.l + jj) * d.m + kk, ii, jj, kk);
    }
    __syncthreads();
  };
  for (int k = B.first; k < B.last; k++) {
    const BotLevel &L = B.lv[k], &C = B.lv[k + 1];
    const float *r = k == B.first ? B.r_in : L.r;
    float *z = k == B.first ? B.z_out : L.z;
    // the first sweep from z = 0: z2 = omega D^-1 r
    interior(k, [&](int64_t c, int, int, int) {
      const uint8_t f = L.flags[c];
      L.z2[c] = (f & FL_ACTUAL) ? lut[k][f & 63] * r[c] : 0.f;
    });
    // the second: z = z2 + omega D^-1 (r - A z2); further sweeps alternate
    for (int sw = 1; sw < B.nu; sw++) {
      const float *zi = (sw & 1) ? L.z2 : z;
      float *zo = (sw & 1) ? z : L.z2;
      interior(k, [&](int64_t c, int i, int j, int kk) {
        const uint8_t f = L.flags[c];
        zo[c] = (f & FL_ACTUAL) ? zi[c] + lut[k][f & 63] * (r[c] - bot_az(L, f, c, i, j, kk, zi)) : 0.f;
      });
    }
    // t = r - A z
    interior(k, [&](int64_t c, int i, int j, int kk) {
      const uint8_t f = L.flags[c];
      L.t[c] = (f & FL_ACTUAL) ? r[c] - bot_az(L, f, c, i, j, kk, z) : 0.f;
    });
    // coarse r = mean of t over the unknown children
    interior(k + 1, [&](int64_t cc, int I, int J, int K) {
      float sum = 0;
      int cnt = 0;
      for (int a = 0; a < 2; a++)
        for (int b = 0; b < 2; b++)
          for (int e = 0; e < 2; e++) {
            const int64_t q = ((int64_t)(2 * I - 1 + a) * L.d.l + (2 * J - 1 + b)) * L.d.m + (2 * K - 1 + e);
            if (L.flags[q] & FL_UNKNOWN) {
              sum += L.t[q];
              cnt++;
            }
          }
      C.r[cc] = cnt ? sum / (float)cnt : 0.f;
    });
  }
  {
    // the deepest level: z = A^+ r
    const BotLevel &D = B.lv[B.last];
    const float *r = B.last == B.first ? B.r_in : D.r;
    float *z = B.last == B.first ? B.z_out : D.z;
    const int ncell = D.d.n * D.d.l * D.d.m;
    for (int q = tid; q < ncell; q += nt) {
      const int u = B.uidx[q];
      if (u >= 0) rs[u] = r[q];
    }
    __syncthreads();
    const int nu = B.ndense;
    for (int q = tid; q < ncell; q += nt) {
      const int u = B.uidx[q];
      double acc = 0;
      if (u >= 0)
        for (int v = 0; v < nu; v++) acc += (double)B.minv[(size_t)v * nu + u] * (double)rs[v];
      z[q] = (float)acc;
    }
    __syncthreads();
  }
  for (int k = B.last - 1; k >= B.first; k--) {
    const BotLevel &L = B.lv[k], &C = B.lv[k + 1];
    const float *r = k == B.first ? B.r_in : L.r;
    float *z = k == B.first ? B.z_out : L.z;
    const float *zc = C.z;
    // z += the coarse correction of the parent at the unknowns
    interior(k, [&](int64_t c, int i, int j, int kk) {
      if (L.flags[c] & FL_UNKNOWN) z[c] += zc[((int64_t)((i + 1) >> 1) * C.d.l + ((j + 1) >> 1)) * C.d.m + ((kk + 1) >> 1)];
    });
    for (int sw = 0; sw < B.nu; sw++) {
      const float *zi = (sw & 1) ? L.z2 : z;
      float *zo = (sw & 1) ? z : L.z2;
      interior(k, [&](int64_t c, int i, int j, int kk) {
        const uint8_t f = L.flags[c];
        zo[c] = (f & FL_ACTUAL) ? zi[c] + lut[k][f & 63] * (r[c] - bot_az(L, f, c, i, j, kk, zi)) : 0.f;
      });
    }
    if (B.nu & 1) {  // an odd sweep count leaves the result in z2
      interior(k, [&](int64_t c, int, int, int) { z[c] = L.z2[c]; });
    }
  }
}

// ------------------------------------------------------------------ host
static dim3 cell_grid(const xf_grid g) { return dim3((g->m - 2 + 63) / 64, (g->l - 2 + 3) / 4, g->n - 2); }
static Dims dims_of(const xf_grid g) { return Dims{g->n, g->l, g->m}; }

// out = sweep MODE on grid g (the vector kernel where rows allow it)
template <int MODE>
static void sweep(xf_ctx c, const xf_grid g, const float *r, const float *z, float *out,
                  const float *zcoarse = nullptr, Dims cd = Dims{0, 0, 0}) {
  const Dims d = dims_of(g);
  const int rows = g->m - 2;
  if (MODE == 4 && (rows < 128 || g->m % 2)) {
    // short rows: the prolongation as its own pass, then MODE 1
    const dim3 pg((g->m - 2 + 255) / 256, (g->l - 2 + 3) / 4, g->n - 2);
    hipLaunchKernelGGL(k_mg_prolong4, pg, dim3(256), 0, c->stream, d, cd, g->d_flags, zcoarse, const_cast<float *>(z));
    hipLaunchKernelGGL(k_mg_sweep<1>, cell_grid(g), dim3(256), 0, c->stream, d, g->d_flags, g->d_cxf, g->d_cyf,
                       g->d_czf, r, z, out);
    return;
  }
  if (rows >= 128) {
    const int cpl = rows >= 256 ? 4 : 2;
    SweepTiles tl;
    tl.kt = (rows + 64 * cpl - 1) / (64 * cpl);
    tl.jt = (g->l - 2 + 3) / 4;
    // 16 planes per tile, fewer on small grids so that >= 1024 tiles keep the chip busy
    tl.ic = 16;
    while (tl.ic > 2 && (int64_t)tl.kt * tl.jt * ((g->n - 2 + tl.ic - 1) / tl.ic) < 1024) tl.ic /= 2;
    tl.it = (g->n - 2 + tl.ic - 1) / tl.ic;
    tl.total = tl.kt * tl.jt * tl.it;
    const int nb = std::min(2048, 8 * ((tl.total + 7) / 8));
    if (cpl == 4)
      hipLaunchKernelGGL((k_mg_sweep_vec<MODE, 4>), dim3(nb), dim3(256), 0, c->stream, d, tl, g->d_flags, g->d_cxf,
                         g->d_cyf, g->d_czf, r, z, out, zcoarse, cd);
    else
      hipLaunchKernelGGL((k_mg_sweep_vec<MODE, 2>), dim3(nb), dim3(256), 0, c->stream, d, tl, g->d_flags, g->d_cxf,
                         g->d_cyf, g->d_czf, r, z, out, zcoarse, cd);
  } else if constexpr (MODE != 6) {
    hipLaunchKernelGGL(k_mg_sweep<MODE == 4 ? 1 : MODE>, cell_grid(g), dim3(256), 0, c->stream, d, g->d_flags,
                       g->d_cxf, g->d_cyf, g->d_czf, r, z, out);
  } else {
    XF_REQUIRE(false, XF_EINVAL, "multigrid: the fused restriction needs rows of >= 128 cells");
  }
}

static float *zalloc(xf_ctx c, int64_t n) {
  float *p = nullptr;
  hipError_t e = hipMalloc(&p, sizeof(float) * n);
  XF_REQUIRE(e == hipSuccess, XF_ENOMEM, "multigrid: out of device memory");
  XF_HIP(hipMemsetAsync(p, 0, sizeof(float) * n, c->stream));
  return p;
}

static bool uniform(const std::vector<double> &h) {
  for (double v : h)
    if (std::fabs(v - h[0]) > 1e-12 * std::fabs(h[0])) return false;
  return !h.empty();
}

void free_mg(MgHier *h) {
  if (!h) return;
  for (int k = 0; k < h->nlev; k++) {
    MgLevel &L = h->lev[k];
    if (k > 0) {
      hipFree(L.r);
      free_grid(L.g);
    }
    hipFree(L.z); hipFree(L.z2); hipFree(L.t);
  }
  hipFree(h->part);
  hipFree(h->scal);
  hipFree(h->minv);
  hipFree(h->uidx);
  delete h;
}

// The pseudo-inverse of the deepest grid's operator (the sweeps' A on its
// Actual interior cells, coefficients rounded to FP32 as the kernels use
// them).  Unknowns joined by open faces form components; a component with no
// open face to a boundary-value cell is singular with the constant as its
// null vector.  With P = sum over those components of 1 1^T / |component|,
// A + P is SPD and A^+ = (A + P)^-1 - P.
static void build_dense(xf_ctx c, MgHier *h, const std::vector<uint8_t> &codes) {
  const xf_grid g = h->lev[h->nlev - 1].g;
  const int n = g->n, l = g->l, m = g->m;
  const int64_t S[3] = {(int64_t)l * m, (int64_t)m, 1};
  const std::vector<double> *hs[3] = {&g->hx, &g->hy, &g->hz};
  std::vector<int32_t> uidx((size_t)g->ncell, -1);
  std::vector<int64_t> cells;
  for (int i = 1; i < n - 1; i++)
    for (int j = 1; j < l - 1; j++)
      for (int k = 1; k < m - 1; k++) {
        const int64_t q = ((int64_t)i * l + j) * m + k;
        if (codes[q] == PT_ACTUAL) {
          uidx[q] = (int32_t)cells.size();
          cells.push_back(q);
        }
      }
  const int nu = (int)cells.size();
  h->ndense = nu;
  if (nu == 0) return;
  // per-node (c-, c+) of each axis, as the grid's FP32 coefficients
  auto coef = [&](int axis, int i, int side) {
    const std::vector<double> &hh = *hs[axis];
    const double mid = (hh[i - 1] + hh[i]) / 2;
    return (double)(float)(1.0 / ((side ? hh[i] : hh[i - 1]) * mid));
  };
  std::vector<double> A((size_t)nu * nu, 0.0);
  std::vector<int> comp(nu, -1);
  std::vector<char> regular;
  for (int u = 0; u < nu; u++) {
    const int64_t q = cells[u];
    const int idx[3] = {(int)(q / S[0]), (int)((q / m) % l), (int)(q % m)};
    for (int a = 0; a < 3; a++)
      for (int side = 0; side < 2; side++) {
        const int64_t q2 = side ? q + S[a] : q - S[a];
        if (codes[q2] == PT_FICTIVE) continue;
        const double cf = coef(a, idx[a], side);
        A[(size_t)u * nu + u] += cf;
        if (uidx[q2] >= 0) A[(size_t)u * nu + uidx[q2]] -= cf;
      }
  }
  // components (breadth-first over open faces between unknowns)
  for (int u0 = 0; u0 < nu; u0++) {
    if (comp[u0] >= 0) continue;
    const int id = (int)regular.size();
    regular.push_back(0);
    std::vector<int> stack{u0};
    comp[u0] = id;
    while (!stack.empty()) {
      const int u = stack.back();
      stack.pop_back();
      const int64_t q = cells[u];
      for (int a = 0; a < 3; a++)
        for (int side = 0; side < 2; side++) {
          const int64_t q2 = side ? q + S[a] : q - S[a];
          if (codes[q2] == PT_FICTIVE) continue;
          const int v = uidx[q2];
          if (v < 0) {
            regular[id] = 1;  // an open face to a boundary-value cell
          } else if (comp[v] < 0) {
            comp[v] = id;
            stack.push_back(v);
          }
        }
    }
  }
  std::vector<int> csize(regular.size(), 0);
  for (int u = 0; u < nu; u++) csize[comp[u]]++;
  std::vector<double> P((size_t)nu * nu, 0.0);
  for (int u = 0; u < nu; u++)
    for (int v = 0; v < nu; v++)
      if (comp[u] == comp[v] && !regular[comp[u]]) P[(size_t)u * nu + v] = 1.0 / csize[comp[u]];
  // Cholesky of A + P, then its inverse column by column
  std::vector<double> Lc((size_t)nu * nu, 0.0);
  for (int i = 0; i < nu; i++)
    for (int j = 0; j <= i; j++) {
      double sum = A[(size_t)i * nu + j] + P[(size_t)i * nu + j];
      for (int k = 0; k < j; k++) sum -= Lc[(size_t)i * nu + k] * Lc[(size_t)j * nu + k];
      if (i == j) {
        XF_REQUIRE(sum > 0, XF_EINVAL, "multigrid: the deepest grid's operator is not positive definite");
        Lc[(size_t)i * nu + i] = std::sqrt(sum);
      } else {
        Lc[(size_t)i * nu + j] = sum / Lc[(size_t)j * nu + j];
      }
    }
  std::vector<float> minv((size_t)nu * nu);
  std::vector<double> col(nu);
  for (int e = 0; e < nu; e++) {
    for (int i = 0; i < nu; i++) {  // L y = e_e
      double sum = i == e ? 1.0 : 0.0;
      for (int k = 0; k < i; k++) sum -= Lc[(size_t)i * nu + k] * col[k];
      col[i] = sum / Lc[(size_t)i * nu + i];
    }
    for (int i = nu - 1; i >= 0; i--) {  // L^T x = y
      double sum = col[i];
      for (int k = i + 1; k < nu; k++) sum -= Lc[(size_t)k * nu + i] * col[k];
      col[i] = sum / Lc[(size_t)i * nu + i];
    }
    for (int i = 0; i < nu; i++) minv[(size_t)e * nu + i] = (float)(col[i] - P[(size_t)e * nu + i]);
  }
  XF_HIP(hipMalloc(&h->minv, sizeof(float) * minv.size()));
  XF_HIP(hipMemcpy(h->minv, minv.data(), sizeof(float) * minv.size(), hipMemcpyHostToDevice));
  XF_HIP(hipMalloc(&h->uidx, sizeof(int32_t) * uidx.size()));
  XF_HIP(hipMemcpy(h->uidx, uidx.data(), sizeof(int32_t) * uidx.size(), hipMemcpyHostToDevice));
}

// Builds the hierarchy of g on first use (level 0 is g itself; its r is the
// caller's residual).
static MgHier *hierarchy(xf_ctx c, xf_grid g) {
  if (g->mg) return g->mg;
  XF_REQUIRE(uniform(g->hx) && uniform(g->hy) && uniform(g->hz), XF_EINVAL,
             "XF_FCG_MGNN: the multigrid cycle needs uniform steps along each axis");
  MgHier *h = new MgHier();
  struct Guard {
    MgHier *&h;
    ~Guard() { if (h) free_mg(h); }
  } guard{h};
  h->lev[0].g = g;
  std::vector<uint8_t> fcodes((size_t)g->ncell);
  XF_HIP(hipMemcpy(fcodes.data(), g->d_codes, g->ncell, hipMemcpyDeviceToHost));
  int nlev = 1;
  h->net_lev = 0;
  while (nlev < kMaxLevels) {
    const xf_grid f = h->lev[nlev - 1].g;
    const int N = f->n - 2, L = f->l - 2, M = f->m - 2;
    if (N % 2 || L % 2 || M % 2 || N / 2 < 2 || L / 2 < 2 || M / 2 < 2) break;
    if (f->ncell <= kCoarseCells) break;  // already small enough for the dense solve
    if (!h->net_lev && M / 2 < 64) h->net_lev = nlev;  // the network's cycle ends here
    const int cn = N / 2 + 2, cl = L / 2 + 2, cm = M / 2 + 2;
    std::vector<int32_t> cc((size_t)cn * cl * cm, 0);
    auto kids = [](int I, int Nc, int *out) {  // fine indices of coarse index I (with ring)
      if (I == 0) { out[0] = 0; return 1; }
      if (I == Nc + 1) { out[0] = 2 * Nc + 1; return 1; }
      out[0] = 2 * I - 1; out[1] = 2 * I; return 2;
    };
    for (int I = 0; I < cn; I++) {
      int ki[2];
      const int ni = kids(I, cn - 2, ki);
      for (int J = 0; J < cl; J++) {
        int kj[2];
        const int nj = kids(J, cl - 2, kj);
        for (int K = 0; K < cm; K++) {
          int kk[2];
          const int nk = kids(K, cm - 2, kk);
          bool unk = false, bnd = false;
          for (int a = 0; a < ni; a++)
            for (int b = 0; b < nj; b++)
              for (int e = 0; e < nk; e++) {
                const uint8_t v = fcodes[((size_t)ki[a] * f->l + kj[b]) * f->m + kk[e]];
                if (v == PT_ACTUAL) unk = true;
                else if (v != PT_FICTIVE) bnd = true;
              }
          cc[((size_t)I * cl + J) * cm + K] = unk ? PT_ACTUAL : (bnd ? PT_DEFINED : PT_FICTIVE);
        }
      }
    }
    std::vector<double> hx(cn - 1, 2 * f->hx[0]), hy(cl - 1, 2 * f->hy[0]), hz(cm - 1, 2 * f->hz[0]);
    xf_grid cg = nullptr;
    const int st = xf_grid_create(c, cn, cl, cm, hx.data(), hy.data(), hz.data(), cc.data(), nullptr, nullptr, &cg);
    XF_REQUIRE(st == XF_OK, st, "multigrid: coarse grid: " + c->err);
    h->lev[nlev].g = cg;
    h->nlev = ++nlev;  // owned from here on
    fcodes.assign((size_t)cg->ncell, 0);
    for (size_t q = 0; q < fcodes.size(); q++) fcodes[q] = (uint8_t)cc[q];
  }
  h->nlev = nlev;
  if (!h->net_lev) h->net_lev = nlev;
  if (h->lev[nlev - 1].g->ncell <= kCoarseCells) build_dense(c, h, fcodes);
  if (h->ndense > 0) {
    h->bottom = nlev - 1;
    while (h->bottom > 0 && h->lev[h->bottom - 1].g->ncell <= kBottomCells) h->bottom--;
  }
  for (int k = 0; k < nlev; k++) {
    MgLevel &L = h->lev[k];
    const int64_t nc = L.g->ncell;
    if (k > 0) L.r = zalloc(c, nc);
    L.z = zalloc(c, nc);
    L.z2 = zalloc(c, nc);
    L.t = zalloc(c, nc);
  }
  XF_HIP(hipMalloc(&h->part, sizeof(double) * 2 * kMgBlocks));
  XF_HIP(hipMalloc(&h->scal, sizeof(double) * 4));
  XF_HIP(hipMemsetAsync(h->scal, 0, sizeof(double) * 4, c->stream));
  g->mg = h;
  h = nullptr;  // the grid owns it now
  return g->mg;
}

int mg_levels(xf_ctx c, xf_grid g) { return hierarchy(c, g)->nlev; }

static int red_blocks(int64_t nc) { return (int)std::max<int64_t>(1, std::min<int64_t>(kMgBlocks, (nc + 255) / 256)); }

// nu sweeps of damped Jacobi on A z = r starting from z = 0; result in L.z
static void smooth_from_zero(xf_ctx c, MgLevel &L, const float *r, int nu) {
  const xf_grid g = L.g;
  if (nu < 2) {
    sweep<0>(c, g, r, nullptr, L.z);
    return;
  }
  sweep<5>(c, g, r, nullptr, L.z);  // the first two sweeps in one pass
  for (int s = 2; s < nu; s++) {
    sweep<1>(c, g, r, L.z, L.z2);
    std::swap(L.z, L.z2);
  }
}

// nu sweeps of damped Jacobi continuing from L.z (plus the coarse level C's
// correction at the unknowns when C is given, folded into the first sweep);
// the last into `last` when given (else L.z holds the result)
static void smooth(xf_ctx c, MgLevel &L, const float *r, int nu, float *last = nullptr, const MgLevel *C = nullptr) {
  const xf_grid g = L.g;
  for (int s = 0; s < nu; s++) {
    float *dst = last && s == nu - 1 ? last : L.z2;
    if (s == 0 && C) sweep<4>(c, g, r, L.z, dst, C->z, dims_of(C->g));
    else sweep<1>(c, g, r, L.z, dst);
    if (dst == last) return;
    std::swap(L.z, L.z2);
  }
  if (last) XF_HIP(hipMemcpyAsync(last, L.z, sizeof(float) * g->ncell, hipMemcpyDeviceToDevice, c->stream));
}

// the deepest level of the cycle (with or without the network)
static int last_level(const MgHier *h, const struct xf_unet_s *net) { return (net ? h->net_lev : h->nlev) - 1; }

// z of the cycle's coarsest level from its r (into `last` when given)
static void coarse_solve(xf_ctx c, MgHier *h, struct xf_unet_s *net, float *last) {
  MgLevel &L = h->lev[last_level(h, net)];
  const xf_grid g = L.g;
  const float *rin = L.r;
  if (!net) {
    if (h->ndense > 0) {
      hipLaunchKernelGGL(k_mg_coarse_dense, dim3(1), dim3(kCoarseCells), 0, c->stream, (int)g->ncell, h->ndense,
                         h->uidx, h->minv, rin, last ? last : L.z);
      return;
    }
    smooth_from_zero(c, L, rin, kCoarseSweeps);
    if (last) XF_HIP(hipMemcpyAsync(last, L.z, sizeof(float) * g->ncell, hipMemcpyDeviceToDevice, c->stream));
    return;
  }
  XF_REQUIRE((g->n - 2) % 16 == 0 && (g->l - 2) % 16 == 0 && (g->m - 2) % 64 == 0, XF_EINVAL,
             "XF_FCG_MGNN: the coarsest grid does not fit the network (interior extents must be multiples of "
             "16, 16 and 64)");
  const int nb = red_blocks(g->ncell);
  // rms of r over the unknowns -> the network's input scaling
  hipLaunchKernelGGL(k_mg_dots, dim3(nb), dim3(256), 0, c->stream, g->ncell, g->d_flags, rin, rin,
                     (const float *)nullptr, h->part);
  hipLaunchKernelGGL(k_mg_finish<0>, dim3(1), dim3(256), 0, c->stream, h->part, nb, h->scal);
  unet_apply_grid(c, net, g, rin, L.z, h->scal);
  // energy-optimal scale: alpha = (r, y) / (y, A y)
  sweep<3>(c, g, nullptr, L.z, L.t);
  hipLaunchKernelGGL(k_mg_dots, dim3(nb), dim3(256), 0, c->stream, g->ncell, g->d_flags, rin, L.z, L.t, h->part);
  hipLaunchKernelGGL(k_mg_finish<1>, dim3(1), dim3(256), 0, c->stream, h->part, nb, h->scal);
  hipLaunchKernelGGL(k_mg_scale, dim3(red_blocks(g->ncell)), dim3(256), 0, c->stream, g->ncell, L.z, h->scal);
  // a few sweeps smooth what the network leaves at the grid scale
  smooth(c, L, rin, kSweeps, last);
}

// one cycle from level k; the level-0 result goes to `out` (levels below
// keep theirs in lev[k].z)
static void vcycle(xf_ctx c, MgHier *h, int k, const float *r, struct xf_unet_s *net, float *out) {
  float *last = k == 0 ? out : nullptr;
  if (!net && k == h->bottom) {
    BotArgs B{};
    for (int q = k; q < h->nlev; q++) {
      const MgLevel &L = h->lev[q];
      B.lv[q] = BotLevel{dims_of(L.g), L.g->d_flags, L.g->d_cxf, L.g->d_cyf, L.g->d_czf, L.r, L.z, L.z2, L.t};
    }
    B.first = k;
    B.last = h->nlev - 1;
    B.nu = kSweeps;
    B.ndense = h->ndense;
    B.uidx = h->uidx;
    B.minv = h->minv;
    B.r_in = r;
    B.z_out = k == 0 ? out : h->lev[k].z;
    hipLaunchKernelGGL(k_mg_bottom, dim3(1), dim3(1024), 0, c->stream, B);
    return;
  }
  if (k == last_level(h, net)) {
    if (k == 0) {
      // a single level: the coarse solve reads lev[0].r, which is the caller's r
      MgLevel &L = h->lev[0];
      float *keep = L.r;
      L.r = const_cast<float *>(r);
      coarse_solve(c, h, net, last);
      L.r = keep;
    } else {
      coarse_solve(c, h, net, last);
    }
    return;
  }
  MgLevel &L = h->lev[k], &C = h->lev[k + 1];
  const xf_grid g = L.g;
  const Dims d = dims_of(g), cd = dims_of(C.g);
  smooth_from_zero(c, L, r, kSweeps);
  if (g->m - 2 >= 128) {
    // the residual restricted in the same pass (the fine residual is not stored)
    sweep<6>(c, g, r, L.z, C.r, nullptr, cd);
  } else {
    sweep<2>(c, g, r, L.z, L.t);
    hipLaunchKernelGGL(k_mg_restrict, cell_grid(C.g), dim3(256), 0, c->stream, d, cd, g->d_flags, L.t, C.r);
  }
  vcycle(c, h, k + 1, C.r, net, nullptr);
  smooth(c, L, r, kSweeps, last, &C);
}

// z = M(r): one V-cycle on the grid's hierarchy (r, z FP32 fields of the fine grid)
void mg_apply(xf_ctx c, xf_grid g, struct xf_unet_s *net, const float *r, float *z) {
  MgHier *h = hierarchy(c, g);
  vcycle(c, h, 0, r, net, z);
  check_kernel("multigrid cycle");
}

extern "C" int xf_mg_levels(xf_ctx c, xf_grid g, int *levels) {
  return guarded(c, [&] {
    XF_REQUIRE(g && levels, XF_EINVAL, "xf_mg_levels: null argument");
    *levels = mg_levels(c, g);
  });
}

extern "C" int xf_mg_level(xf_ctx c, xf_grid g, int level, xf_grid *out) {
  return guarded(c, [&] {
    XF_REQUIRE(g && out, XF_EINVAL, "xf_mg_level: null argument");
    MgHier *h = hierarchy(c, g);
    XF_REQUIRE(level >= 0 && level < h->nlev, XF_EINVAL, "xf_mg_level: no such level");
    *out = h->lev[level].g;
  });
}

extern "C" int xf_mg_apply(xf_ctx c, xf_grid g, xf_unet net, const float *r, float *z) {
  return guarded(c, [&] {
    XF_REQUIRE(g && r && z && r != z, XF_EINVAL, "xf_mg_apply: bad argument");
    mg_apply(c, g, net, r, z);
  });
}
