// The reference's operator rows on the device, shared by the parity kernels
// (xf_vec.hip) and the packed rows (xf_pack.hip).
//
// The rows repeat T3dLaplaceOperator::operator() (XOperator.cpp:938-1129) and
// T3dVariableDivergentedOperator::operator() (XOperator.cpp:1901-2233) with the
// reference's operand order; with -ffp-contract=off and correctly rounded f64
// division each row is bit-identical to the reference's.
#pragma once

#include "xf_internal.h"

struct GridView {
  int n, l, m;
  const double *hx, *hy, *hz;
  const uint8_t *codes, *space;
  const double *nrm;
  const uint8_t *hasn;
  const double *rho;  // null: Laplace operator
  const int32_t *rn;
};

static inline GridView view(xf_grid g) {
  return GridView{g->n, g->l, g->m, g->d_hx, g->d_hy, g->d_hz, g->d_codes, g->d_space,
                  g->d_normals, g->d_has_normal, g->use_rho ? g->d_rho : nullptr, g->d_rn};
}

// One operator row at an unknown cell (validity was established at grid set-up),
// with the field read through an accessor f(index): a device array, or a unit
// vector when the packed rows probe coefficients.
template <class F>
__device__ inline double row_value_t(const GridView &g, const F &f, int64_t c, int i, int j, int k) {
  const uint8_t *mk = g.codes;
  const double *hx = g.hx, *hy = g.hy, *hz = g.hz;
  const int64_t sI = (int64_t)g.l * g.m, sJ = g.m;
  const uint8_t code = mk[c];
  if (code == PT_ACTUAL) {
    const double hx_2X = (hx[i - 1] + hx[i]) / 2;
    const double hx_2Y = (hy[j - 1] + hy[j]) / 2;
    const double hx_2Z = (hz[k - 1] + hz[k]) / 2;
    const double fc = f(c);
    if (!g.rho) {
      double f_i_1 = (fc - f(c - sI)) / (hx[i - 1]);
      double f_i = (f(c + sI) - fc) / (hx[i]);
      double f_j_1 = (fc - f(c - sJ)) / (hy[j - 1]);
      double f_j = (f(c + sJ) - fc) / (hy[j]);
      double f_k_1 = (fc - f(c - 1)) / (hz[k - 1]);
      double f_k = (f(c + 1) - fc) / (hz[k]);
      if (mk[c - sI] == PT_FICTIVE) f_i_1 = 0;
      if (mk[c + sI] == PT_FICTIVE) f_i = 0;
      if (mk[c - sJ] == PT_FICTIVE) f_j_1 = 0;
      if (mk[c + sJ] == PT_FICTIVE) f_j = 0;
      if (mk[c - 1] == PT_FICTIVE) f_k_1 = 0;
      if (mk[c + 1] == PT_FICTIVE) f_k = 0;
      return -((f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z);
    }
    const double *rho = g.rho;
    double f_i_1 = rho[c - sI] * (fc - f(c - sI)) / (hx[i - 1]);
    double f_i = rho[c + sI] * (f(c + sI) - fc) / (hx[i]);
    double f_j_1 = rho[c - sJ] * (fc - f(c - sJ)) / (hy[j - 1]);
    double f_j = rho[c + sJ] * (f(c + sJ) - fc) / (hy[j]);
    double f_k_1 = rho[c - 1] * (fc - f(c - 1)) / (hz[k - 1]);
    double f_k = rho[c + 1] * (f(c + 1) - fc) / (hz[k]);
    double d_i = (rho[c + sI] - rho[c]) / (hx[i - 1]);
    double d_j = (rho[c + sJ] - rho[c]) / (hy[j - 1]);
    double d_k = (rho[c + 1] - rho[c]) / (hz[k - 1]);
    if (mk[c - sI] == PT_FICTIVE) f_i_1 = 0;
    if (mk[c + sI] == PT_FICTIVE) { f_i = 0; d_i = 0; }
    if (mk[c - sJ] == PT_FICTIVE) f_j_1 = 0;
    if (mk[c + sJ] == PT_FICTIVE) { f_j = 0; d_j = 0; }
    if (mk[c - 1] == PT_FICTIVE) f_k_1 = 0;
    if (mk[c + 1] == PT_FICTIVE) { f_k = 0; d_k = 0; }
    double lap = (f_i - f_i_1) / hx_2X + (f_j - f_j_1) / hx_2Y + (f_k - f_k_1) / hx_2Z;
    double divd = d_i + d_j + d_k;
    return -(rho[c] * lap - divd * fc);  // XOperator.cpp:2025
  }
  const double nx = g.nrm[3 * c], ny = g.nrm[3 * c + 1], nz = g.nrm[3 * c + 2];
  if (code == PT_PREDEFINED) {  // XOperator.cpp:1057-1068
    double s = 0;
    if (nx > 0) s = (f(c) + f(c - sI)) / 2;
    else if (nx < 0) s = (f(c + sI) + f(c)) / 2;
    if (ny > 0) s = (f(c) + f(c - sJ)) / 2;
    else if (ny < 0) s = (f(c + sJ) + f(c)) / 2;
    if (nz > 0) s = (f(c) + f(c - 1)) / 2;
    else if (nz < 0) s = (f(c + 1) + f(c)) / 2;
    return s;
  }
  // PreNormal (XOperator.cpp:1083-1118, divergented 2113-2224)
  int sI_ = 0, sI_1 = 0, sJ_ = 0, sJ_1 = 0, sK_ = 0, sK_1 = 0;
  double hxl = 1, hyl = 1, hzl = 1;
  if (nx > 0) { sI_ = 0; sI_1 = -1; hxl = hx[i - 1]; }
  else if (nx < 0) { sI_ = 1; sI_1 = 0; hxl = hx[i]; }
  if (ny > 0) { sJ_ = 0; sJ_1 = -1; hyl = hy[j - 1]; }
  else if (ny < 0) { sJ_ = 1; sJ_1 = 0; hyl = hy[j]; }
  if (nz > 0) { sK_ = 0; sK_1 = -1; hzl = hz[k - 1]; }
  else if (nz < 0) { sK_ = 1; sK_1 = 0; hzl = hz[k]; }
#define FV(a, b, e) f(c + (int64_t)(a) * sI + (int64_t)(b) * sJ + (e))
  double fl = (FV(sI_, 0, 0) + FV(sI_, sJ_, 0) + FV(sI_, 0, sK_) + FV(sI_, sJ_, sK_)) / 4;
  double fl_1 = (FV(sI_1, 0, 0) + FV(sI_1, sJ_, 0) + FV(sI_1, 0, sK_) + FV(sI_1, sJ_, sK_)) / 4;
  double dfx = (fl - fl_1) / hxl;
  fl = (FV(0, sJ_, 0) + FV(sI_, sJ_, 0) + FV(0, sJ_, sK_) + FV(sI_, sJ_, sK_)) / 4;
  fl_1 = (FV(0, sJ_1, 0) + FV(sI_, sJ_1, 0) + FV(0, sJ_1, sK_) + FV(sI_, sJ_1, sK_)) / 4;
  double dfy = (fl - fl_1) / hyl;
  fl = (FV(0, 0, sK_) + FV(sI_, 0, sK_) + FV(0, sJ_, sK_) + FV(sI_, sJ_, sK_)) / 4;
  fl_1 = (FV(0, 0, sK_1) + FV(sI_, 0, sK_1) + FV(0, sJ_, sK_1) + FV(sI_, sJ_, sK_1)) / 4;
  double dfz = (fl - fl_1) / hzl;
#undef FV
  if (!g.rho) return dfx * nx + dfy * ny + dfz * nz;
  const double *rho = g.rho;
  // the reference's precedence: rho[i+1] - rho[i]/h (XOperator.cpp:2211-2213)
  double d_x = rho[c + sI] - rho[c] / (hx[i - 1]);
  double d_y = rho[c + sJ] - rho[c] / (hy[j - 1]);
  double d_z = rho[c + 1] - rho[c] / (hz[k - 1]);
  double lap = dfx * nx + dfy * ny + dfz * nz;
  double divd = d_x + d_y + d_z;
  return rho[c] * lap - divd * f(c);
}


struct PtrF {
  const double *p;
  __device__ double operator()(int64_t q) const { return p[q]; }
};
// the operator row on a device field
__device__ inline double row_value(const GridView &g, const double *__restrict__ f, int64_t c, int i, int j,
                                   int k) {
  return row_value_t(g, PtrF{f}, c, i, j, k);
}

__device__ inline void decompose(int64_t c, int l, int m, int &i, int &j, int &k) {
  k = (int)(c % m);
  int64_t t = c / m;
  j = (int)(t % l);
  i = (int)(t / l);
}

