// Projection-step glue around the pressure solve (TVPProblem::SolveBySplitting):
// the pressure right-hand side -div(u*)/tau (XProblem.cpp:5474-5513), the
// velocity correction u = u* - tau grad p / rho (XProblem.cpp:5598-5778) and
// the divergence monitor max |div u| (XProblem.cpp:5782-5830).
//
// Staggered fields are stored on the pressure array shape: u[i,j,k] is the
// x-face between pressure cells (i,j,k) and (i+1,j,k), v and w likewise; the
// face steps hxU[i-1], hyV[j-1], hzW[k-1] of cell (i,j,k) are the velocity
// grids' separator steps (arrays of n-1, l-1, m-1 entries).  The arithmetic
// keeps the reference's operand order, including its rho^2 precedence on the
// x term of the divergence (XProblem.cpp:5495).
#include <cmath>

#include "xf_internal.h"

__global__ void k_pressure_rhs(int n, int l, int m, const uint8_t *__restrict__ codes, const double *__restrict__ u,
                               const double *__restrict__ v, const double *__restrict__ w,
                               const double *__restrict__ rho, const double *__restrict__ hxU,
                               const double *__restrict__ hyV, const double *__restrict__ hzW,
                               const double *__restrict__ pbc, double tau, double *__restrict__ rhs) {
  // point types other than Fictive, Actual and PreDefined, and Actual points
  // on the low faces, were ruled out on the host (pressure_rhs_check)
  const int64_t nc = (int64_t)n * l * m, sI = (int64_t)l * m, sJ = m;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t code = codes[c];
    if (code == PT_FICTIVE) {
      rhs[c] = 0;
    } else if (code == PT_ACTUAL) {
      const int k = (int)(c % m), j = (int)((c / m) % l), i = (int)(c / sI);
      const double d = rho ? rho[c] : 1.0;
      const double div_ = d * d * ((u[c] - u[c - sI]) / hxU[i - 1]) + ((v[c] - v[c - sJ]) / hyV[j - 1]) +
                          ((w[c] - w[c - 1]) / hzW[k - 1]);
      rhs[c] = -(div_ / tau);
    } else if (code == PT_PREDEFINED) {
      rhs[c] = pbc ? pbc[c] : 0.0;
    }
  }
}

// pass 1: every point but PreDefined (XProblem.cpp:5636-5698)
__global__ void k_vel_correct(int n, int l, int m, int axis, const uint8_t *__restrict__ codes,
                              const double *__restrict__ us, const double *__restrict__ p,
                              const double *__restrict__ rho, const double *__restrict__ hP,
                              const double *__restrict__ bc, double tau, double *__restrict__ out, int *err) {
  // point types and stencils that leave the array were ruled out on the host
  // (vel_correct_check); err collects the data check of Defined points
  const int64_t nc = (int64_t)n * l * m;
  const int64_t S[3] = {(int64_t)l * m, (int64_t)m, 1};
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const uint8_t code = codes[c];
    if (code == PT_PREDEFINED) continue;
    if (code == PT_FICTIVE) {
      out[c] = 0;
    } else if (code == PT_ACTUAL || code == PT_NORMAL || code == PT_EQUATION) {
      const int idx[3] = {(int)(c / S[0]), (int)((c / m) % l), (int)(c % m)};
      const double gradP = (p[c + S[axis]] - p[c]) / hP[idx[axis]];
      out[c] = us[c] - tau * gradP / (rho ? rho[c] : 1.0);
    } else if (code == PT_DEFINED) {
      const double dv = bc ? bc[c] : 0.0;
      if (us[c] != dv) atomicOr(err, EB_VC_DEFINED);  // XProblem.cpp:5684-5688
      out[c] = dv;
    }
  }
}

// pass 2: PreDefined points mirror their single Actual neighbour
// (XProblem.cpp:5698-5758); one thread per cell of the grid's list of
// PreDefined cells
__global__ void k_vel_correct_predef(int n, int l, int m, const uint8_t *__restrict__ codes,
                                     const int32_t *__restrict__ list, int64_t cnt,
                                     const double *__restrict__ us, const double *__restrict__ bc,
                                     double *__restrict__ out, int *err) {
  for (int64_t t = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; t < cnt; t += (int64_t)gridDim.x * blockDim.x) {
    const int64_t c = list[t];
    const int k = (int)(c % m), j = (int)((c / m) % l), i = (int)(c / ((int64_t)l * m));
    int64_t nb = -1;
    int count = 0;
    for (int dir = 0; dir < 3; dir++)
      for (int sign = 0; sign < 2; sign++) {
        const int sh = -1 + sign * 2;
        const int a = i + (dir == 0 ? sh : 0), b = j + (dir == 1 ? sh : 0), e = k + (dir == 2 ? sh : 0);
        if (a >= 0 && a < n && b >= 0 && b < l && e >= 0 && e < m) {
          const int64_t cc = ((int64_t)a * l + b) * m + e;
          if (codes[cc] == PT_ACTUAL) {
            nb = cc;
            count++;
          }
        }
      }
    if (count != 1) continue;  // ruled out on the host
    const double dv = bc ? bc[c] : 0.0;
    if (fabs(us[c] - (2 * dv - us[nb])) > 1e-10) { atomicOr(err, EB_VC_PREDEF); continue; }
    out[c] = 2 * dv - out[nb];
  }
}

// (value, cell) pairs: the larger value wins, ties go to the lower cell
// index, and -1 (no cell) loses every tie
__device__ inline void max_pair(double &b, int64_t &bi, double v2, int64_t i2) {
  if (v2 > b || (v2 == b && i2 >= 0 && (bi < 0 || i2 < bi))) {
    b = v2;
    bi = i2;
  }
}

__device__ inline void block_max_pair(double &b, int64_t &bi, double *lv, int64_t *li) {
  lv[threadIdx.x] = b;
  li[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      double v = lv[threadIdx.x];
      int64_t i = li[threadIdx.x];
      max_pair(v, i, lv[threadIdx.x + s], li[threadIdx.x + s]);
      lv[threadIdx.x] = v;
      li[threadIdx.x] = i;
    }
    __syncthreads();
  }
  b = lv[0];
  bi = li[0];
}

// max |div u| over Actual pressure cells with the first (i->j->k) maximising
// cell: per-block partials here, folded by k_divmax_finish (no device-scope
// fences or atomics between the blocks)
__global__ void k_divmax(int n, int l, int m, const uint8_t *__restrict__ codes, const double *__restrict__ u,
                         const double *__restrict__ v, const double *__restrict__ w, const double *__restrict__ hxU,
                         const double *__restrict__ hyV, const double *__restrict__ hzW, double *__restrict__ pv,
                         int64_t *__restrict__ pi) {
  __shared__ double lv[kBlock];
  __shared__ int64_t li[kBlock];
  const int64_t nc = (int64_t)n * l * m, sI = (int64_t)l * m, sJ = m;
  double best = 0;
  int64_t bi = -1;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (codes[c] != PT_ACTUAL) continue;
    const int k = (int)(c % m), j = (int)((c / m) % l), i = (int)(c / sI);
    if (i < 1 || j < 1 || k < 1) continue;
    const double div = ((u[c] - u[c - sI]) / hxU[i - 1]) + ((v[c] - v[c - sJ]) / hyV[j - 1]) +
                       ((w[c] - w[c - 1]) / hzW[k - 1]);
    const double a = fabs(div);
    if (a > best) { best = a; bi = c; }
  }
  block_max_pair(best, bi, lv, li);
  if (threadIdx.x == 0) {
    pv[blockIdx.x] = best;
    pi[blockIdx.x] = bi;
  }
}

__global__ void k_divmax_finish(int np, const double *__restrict__ pv, const int64_t *__restrict__ pi, double *out,
                                int64_t *out_idx) {
  __shared__ double lv[kBlock];
  __shared__ int64_t li[kBlock];
  double b = 0;
  int64_t idx = -1;
  for (int q = threadIdx.x; q < np; q += blockDim.x) max_pair(b, idx, pv[q], pi[q]);
  block_max_pair(b, idx, lv, li);
  if (threadIdx.x != 0) return;
  *out = b;
  *out_idx = b > 0 ? idx : -1;
}

namespace {
int nblocks(int64_t nc) {
  int64_t b = (nc + kBlock - 1) / kBlock;
  return (int)(b > 8192 ? 8192 : (b < 1 ? 1 : b));
}

// Point types the pressure right-hand side rejects (XProblem.cpp:5502-5505):
// 1 = a type other than Fictive, Actual or PreDefined; 2 = an Actual point on
// a low face (its backward differences would leave the array).
int pressure_rhs_check(xf_grid g) {
  if (g->glue_rhs_err >= 0) return g->glue_rhs_err;
  const std::vector<uint8_t> &cd = host_codes(g);
  int e = 0;
  for (int i = 0; i < g->n; i++)
    for (int j = 0; j < g->l; j++)
      for (int k = 0; k < g->m; k++) {
        const uint8_t v = cd[((size_t)i * g->l + j) * g->m + k];
        if (v == PT_ACTUAL && (i < 1 || j < 1 || k < 1)) e |= 2;
        else if (v != PT_FICTIVE && v != PT_ACTUAL && v != PT_PREDEFINED) e |= 1;
      }
  return g->glue_rhs_err = e;
}

// Point types the velocity correction along `axis` (0..2) rejects
// (XProblem.cpp:5636-5758): 1 = an unknown type, 2 = an Actual, Normal or
// Equation point on the high face of the axis, 8 = a PreDefined point
// without exactly one Actual neighbour.
int vel_correct_check(xf_grid g, int axis) {
  if (g->glue_vc_err[axis] >= 0) return g->glue_vc_err[axis];
  const std::vector<uint8_t> &cd = host_codes(g);
  const int n = g->n, l = g->l, m = g->m;
  const int E[3] = {n, l, m};
  int e = 0;
  for (int i = 0; i < n; i++)
    for (int j = 0; j < l; j++)
      for (int k = 0; k < m; k++) {
        const uint8_t v = cd[((size_t)i * l + j) * m + k];
        const int idx[3] = {i, j, k};
        if (v == PT_ACTUAL || v == PT_NORMAL || v == PT_EQUATION) {
          if (idx[axis] + 1 >= E[axis]) e |= 2;
        } else if (v == PT_PREDEFINED) {
          int count = 0;
          for (int d = 0; d < 3; d++)
            for (int sh = -1; sh <= 1; sh += 2) {
              int q[3] = {i, j, k};
              q[d] += sh;
              if (q[d] < 0 || q[d] >= E[d]) continue;
              if (cd[((size_t)q[0] * l + q[1]) * m + q[2]] == PT_ACTUAL) count++;
            }
          if (count != 1) e |= 8;
        } else if (v != PT_FICTIVE && v != PT_DEFINED) {
          e |= 1;
        }
      }
  return g->glue_vc_err[axis] = e;
}
}  // namespace

extern "C" int xf_pressure_rhs(xf_ctx c, xf_grid gp, const double *u, const double *v, const double *w,
                               const double *rho, const double *hxU, const double *hyV, const double *hzW,
                               const double *pbc, double tau, double *rhs) {
  return guarded(c, [&] {
    XF_REQUIRE(gp && u && v && w && hxU && hyV && hzW && rhs, XF_EINVAL, "xf_pressure_rhs: null argument");
    XF_REQUIRE(tau > 0, XF_EINVAL, "xf_pressure_rhs: tau must be positive");
    XF_REQUIRE(pressure_rhs_check(gp) == 0, XF_EINVAL,
               "TVPProblem::SolveBySplitting: Method is not implemented for this border point type of pressure.");
    const double *sx = device_steps(c, gp, 0, hxU, gp->n - 1), *sy = device_steps(c, gp, 1, hyV, gp->l - 1),
                 *sz = device_steps(c, gp, 2, hzW, gp->m - 1);
    hipLaunchKernelGGL(k_pressure_rhs, dim3(nblocks(gp->ncell)), dim3(kBlock), 0, c->stream, gp->n, gp->l, gp->m,
                       gp->d_codes, u, v, w, rho, sx, sy, sz, pbc, tau, rhs);
    check_kernel("k_pressure_rhs");
  });
}

extern "C" int xf_velocity_correct(xf_ctx c, xf_grid gv, int axis, const double *u_star, const double *p,
                                   const double *rho, const double *hP, const double *bc, double tau, double *u_out) {
  return guarded(c, [&] {
    XF_REQUIRE(gv && u_star && p && hP && u_out, XF_EINVAL, "xf_velocity_correct: null argument");
    XF_REQUIRE(axis >= 1 && axis <= 3, XF_EINVAL, "Unknown velocity number.");
    XF_REQUIRE(u_out != u_star, XF_EINVAL, "xf_velocity_correct: u_out must not alias u_star");
    const int e = vel_correct_check(gv, axis - 1);
    XF_REQUIRE(!(e & 8), XF_EINVAL, "TVPProblem::SolveBySplitting: PreDefined border condition checking error - "
                                    "actual values count must be 1.");
    XF_REQUIRE(e == 0, XF_EINVAL,
               "TVPProblem::SolveBySplitting: Method is not implemented for this point border type of velocity.");
    const int E[3] = {gv->n, gv->l, gv->m};
    const double *sp = device_steps(c, gv, 3, hP, E[axis - 1] - 1);
    hipLaunchKernelGGL(k_vel_correct, dim3(nblocks(gv->ncell)), dim3(kBlock), 0, c->stream, gv->n, gv->l, gv->m,
                       axis - 1, gv->d_codes, u_star, p, rho, sp, bc, tau, u_out, c->d_err);
    if (gv->n_predef < 0) {  // the PreDefined cells, listed once per grid
      const std::vector<uint8_t> &cd = host_codes(gv);
      std::vector<int32_t> pl;
      for (int64_t q = 0; q < gv->ncell; q++)
        if (cd[(size_t)q] == PT_PREDEFINED) pl.push_back((int32_t)q);
      XF_HIP(hipMalloc(&gv->d_predef, sizeof(int32_t) * std::max<size_t>(pl.size(), 1)));
      if (!pl.empty())
        XF_HIP(hipMemcpy(gv->d_predef, pl.data(), sizeof(int32_t) * pl.size(), hipMemcpyHostToDevice));
      gv->n_predef = (int64_t)pl.size();
    }
    if (gv->n_predef > 0)
      hipLaunchKernelGGL(k_vel_correct_predef, dim3(nblocks(gv->n_predef)), dim3(kBlock), 0, c->stream, gv->n, gv->l,
                         gv->m, gv->d_codes, gv->d_predef, gv->n_predef, u_star, bc, u_out, c->d_err);
    check_kernel("k_vel_correct");
    raise_device_errors(c);
  });
}

extern "C" int xf_divergence_max(xf_ctx c, xf_grid gp, const double *u, const double *v, const double *w,
                                 const double *hxU, const double *hyV, const double *hzW, double *max_abs,
                                 int64_t *cell) {
  return guarded(c, [&] {
    XF_REQUIRE(gp && u && v && w && hxU && hyV && hzW && max_abs, XF_EINVAL, "xf_divergence_max: null argument");
    const double *sx = device_steps(c, gp, 0, hxU, gp->n - 1), *sy = device_steps(c, gp, 1, hyV, gp->l - 1),
                 *sz = device_steps(c, gp, 2, hzW, gp->m - 1);
    int nb = nblocks(gp->ncell);
    if (nb > 2048) nb = 2048;
    double *pv = c->partials;
    int64_t *pi = (int64_t *)(c->partials + kMaxPartials);
    int64_t *d_idx = (int64_t *)(c->partials + 2 * kMaxPartials);
    hipLaunchKernelGGL(k_divmax, dim3(nb), dim3(kBlock), 0, c->stream, gp->n, gp->l, gp->m, gp->d_codes, u, v, w,
                       sx, sy, sz, pv, pi);
    hipLaunchKernelGGL(k_divmax_finish, dim3(1), dim3(kBlock), 0, c->stream, nb, pv, pi, c->d_scalar, d_idx);
    check_kernel("k_divmax");
    int64_t idx;
    XF_HIP(hipMemcpyAsync(max_abs, c->d_scalar, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipMemcpyAsync(&idx, d_idx, sizeof(int64_t), hipMemcpyDeviceToHost, c->stream));
    XF_HIP(hipStreamSynchronize(c->stream));
    if (cell) *cell = idx;
  });
}

// Transport velocities, viscosity and boundary right-hand side of one
// transport equation, as TVPProblem::SolveBySplitting builds them before the
// LOD step: velocity component axis 1..3 on its own grid (point-type loop
// XProblem.cpp:5090-5303), the concentration (axis 0) on its grid
// (XProblem.cpp:5889-5944).  Operands in the reference's order; point types
// the reference rejects were ruled out on the host (xf_grid_s::tv_err).
__global__ void k_transport_vel(int n, int l, int m, int axis, const uint8_t *__restrict__ codes,
                                const double *__restrict__ u, const double *__restrict__ v,
                                const double *__restrict__ w, const double *__restrict__ vis,
                                const double *__restrict__ bc, double *__restrict__ tu, double *__restrict__ tv,
                                double *__restrict__ tw, double *__restrict__ tvis, double *__restrict__ rp) {
  const int64_t nc = (int64_t)n * l * m;
  const int64_t S[3] = {(int64_t)l * m, (int64_t)m, 1};
  const double *f[3] = {u, v, w};
  double *o[3] = {tu, tv, tw};
  const int a = axis - 1;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    const int code = codes[c];
    const double bcv = bc ? bc[c] : 0.0;
    if (code == PT_ACTUAL) {
      if (a < 0) {  // cell centres: each component over its two faces
        tu[c] = (u[c - S[0]] + u[c]) / 2;
        tv[c] = (v[c - S[1]] + v[c]) / 2;
        tw[c] = (w[c - S[2]] + w[c]) / 2;
      } else {
        for (int e = 0; e < 3; e++) {
          const double *fe = f[e];
          o[e][c] = e == a ? fe[c] : (fe[c] + fe[c + S[a]] + fe[c - S[e]] + fe[c + S[a] - S[e]]) / 4;
        }
        if (tvis) tvis[c] = vis ? (vis[c] + vis[c + S[a]]) / 2 : 0.0;
      }
      if (rp) rp[c] = 0;
    } else {
      // Fictive: 0; Defined, PreDefined, Normal: 0 and the boundary value in
      // rp (the concentration's PreDefined points: the inlet value on i = 0
      // only; its PreNormal points: 0)
      tu[c] = 0;
      tv[c] = 0;
      tw[c] = 0;
      if (tvis && a >= 0) tvis[c] = 0;
      if (rp) {
        double r = 0;
        if (a >= 0) r = code == PT_FICTIVE ? 0.0 : bcv;
        else if (code == PT_PREDEFINED && c < S[0]) r = bcv;
        rp[c] = r;
      }
    }
  }
}

// rp += the viscosity-gradient term at Actual points (XProblem.cpp:5357-5417)
// from the transport velocities and viscosity of k_transport_vel.
__global__ void k_transport_visterm(int n, int l, int m, int axis, const uint8_t *__restrict__ codes,
                                    const double *__restrict__ hx, const double *__restrict__ hy,
                                    const double *__restrict__ hz, const double *__restrict__ tu,
                                    const double *__restrict__ tv, const double *__restrict__ tw,
                                    const double *__restrict__ tvis, double *__restrict__ rp) {
  const int64_t nc = (int64_t)n * l * m;
  const int64_t S[3] = {(int64_t)l * m, (int64_t)m, 1};
  const double *tr[3] = {tu, tv, tw};
  const double *hs[3] = {hx, hy, hz};
  const int a = axis - 1;
  for (int64_t c = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; c < nc; c += (int64_t)gridDim.x * blockDim.x) {
    if (codes[c] != PT_ACTUAL) continue;
    const int idx[3] = {(int)(c / S[0]), (int)((c / m) % l), (int)(c % m)};
    const int i = idx[0], j = idx[1], k = idx[2];
    double dVisX = (tvis[c + S[0]] - tvis[c - S[0]]) / (hx[i] + hx[i - 1]);
    double dVisY = (tvis[c + S[1]] - tvis[c - S[1]]) / (hy[j] + hy[j - 1]);
    double dVisZ = (tvis[c + S[2]] - tvis[c - S[2]]) / (hz[k] + hz[k - 1]);
    const double *ha = hs[a];
    const int ia = idx[a];
    double dU = (tu[c + S[a]] - tu[c - S[a]]) / (ha[ia] + ha[ia - 1]);
    double dV = (tv[c + S[a]] - tv[c - S[a]]) / (ha[ia] + ha[ia - 1]);
    double dW = (tw[c + S[a]] - tw[c - S[a]]) / (ha[ia] + ha[ia - 1]);
    const double *velTr = tr[a];
    double dVelX = (velTr[c + S[0]] - velTr[c - S[0]]) / (hx[i] + hx[i - 1]);
    double dVelY = (velTr[c + S[1]] - velTr[c - S[1]]) / (hy[j] + hy[j - 1]);
    double dVelZ = (velTr[c + S[2]] - velTr[c - S[2]]) / (hz[k] + hz[k - 1]);
    rp[c] = rp[c] + (dVisX * (dVelX + dU) + dVisY * (dVelY + dV) + dVisZ * (dVelZ + dW));
  }
}

extern "C" int xf_transport_velocity(xf_ctx c, xf_grid g, int axis, const double *u, const double *v,
                                     const double *w, const double *vis, const double *bc, double *tu, double *tv,
                                     double *tw, double *tvis, double *rp) {
  return guarded(c, [&] {
    XF_REQUIRE(g && u && v && w && tu && tv && tw, XF_EINVAL, "xf_transport_velocity: null argument");
    XF_REQUIRE(axis >= 0 && axis <= 3, XF_EINVAL, "Error in TVPProblem::SolveBySplitting: Unknown equation number.");
    XF_REQUIRE(axis > 0 || !tvis, XF_EINVAL, "xf_transport_velocity: the concentration equation has no viscosity");
    XF_REQUIRE(g->tv_err[axis].empty(), XF_EINVAL, g->tv_err[axis]);
    hipLaunchKernelGGL(k_transport_vel, dim3(nblocks(g->ncell)), dim3(kBlock), 0, c->stream, g->n, g->l, g->m, axis,
                       g->d_codes, u, v, w, vis, bc, tu, tv, tw, tvis, rp);
    if (axis > 0 && vis && tvis && rp) {
      XF_REQUIRE(g->tv_visterm_ok[axis], XF_EINVAL,
                 "xf_transport_velocity: an Actual point on the array border (the viscosity term reads both "
                 "neighbours)");
      hipLaunchKernelGGL(k_transport_visterm, dim3(nblocks(g->ncell)), dim3(kBlock), 0, c->stream, g->n, g->l, g->m,
                         axis, g->d_codes, g->d_hx, g->d_hy, g->d_hz, tu, tv, tw, tvis, rp);
    }
    check_kernel("k_transport_vel");
  });
}
